// P4: unsplit MUSCL-Hancock hydro with the MC limiter and HLL fluxes (Davis speeds), 14 channels, gravity's
// half-step source with the face work, the radiation force with its share cap, and FP64 mass.
//
// The scheme is the reference's (crates/sr-physics/src/reference/hydro.rs, module doc :1-96; GPU
// crates/sr-engine/shaders/hydro.wgsl:910-1153); the kernel layout is this library's own:
//   * one 320-thread block (five wave64s) per 16x16 tile: 256 threads own the tile's cells and 64 its one-cell
//     ring (the face neighbours), so every predicted face state the tile's fluxes read is made once in the block;
//   * primitives of the tile and a two-cell halo (ghosts by the edge rules) are staged in LDS as structure of
//     arrays; the predicted faces go to LDS, each cell computes the HLL fluxes of its own four faces (a shared face
//     is computed bit-identically by both cells, so the flux differences telescope);
//   * the ten species are not carried through the main pass: their mass-weighted mean 1/mu is linear in the
//     fractions and rides along as a pseudo-species, and each species' face value and flux is made in a short loop
//     after the main fluxes, the species' shared limiter factor kept in registers;
//   * Sigma is stored and updated in FP64 (flux differences of FP32 fluxes summed in FP64), so the total mass is
//     conserved to FP64 round-off; the other channels are FP32.
// Two modes share the code: MID writes Sigma_mid = (Sigma^n + Sigma^{n+1}) / 2 from the mass fluxes
// (hydro.rs:817-836) for the half-step gravity solve; UPDATE writes the next state.
#include <cfloat>
#include <cmath>

#include "common.hpp"
#include "hydro.hpp"

namespace sr {

#define T 16
#define HALO 2
#define PW (T + 2 * HALO) // 20
#define NPRIM (PW * PW)   // 400
#define NSLOT (T * T + 4 * T) // 320
#define NS 10
#define HOT_FACE 2.0f
#ifndef SR_HYDRO_SB
#define SR_HYDRO_SB 1 // species per pass of the species phase (1, 2 or 5)
#endif

// (Z + 1) / A per species, assets/elements.json order: H He C O Ne Mg Si S Fe n (registry.rs:146-154)
__constant__ float c_imu[NS] = {2.0f, 0.75f, 7.0f / 12.0f, 9.0f / 16.0f, 11.0f / 20.0f,
                                13.0f / 24.0f, 15.0f / 28.0f, 17.0f / 32.0f, 27.0f / 56.0f, 1.0f};

struct Prim {
    float s, u, v, p, im;
    float ye, xn;  // Y_e and X_n, read only with the cold pressure on
};

__device__ __forceinline__ float mc(float l, float c, float r) {
    const float a = c - l, b = r - c;
    if (a * b <= 0.0f) return 0.0f;
    const float m = fminf(fminf(fabsf(0.5f * (a + b)), 2.0f * fabsf(a)), 2.0f * fabsf(b));
    return copysignf(m, a);
}

__device__ __forceinline__ float eps_floor(const HydroParams &P, float im) { return P.t_floor_g * im; }

__device__ __forceinline__ float sound(const HydroParams &P, float s, float p) {
    return sqrtf(fmaxf(__fdividef(P.gamma * p, s), 0.0f));
}

// E of a primitive state (hydro.rs:444-459 with gamma): eps = max((Pi - Pi_cold) / ((gamma-1) Sigma), floor),
// E = Sigma (eps + eps_cold) + kinetic
template <bool COLD> __device__ __forceinline__ float cons_e(const HydroParams &P, const Prim &q) {
    const float pc = COLD ? pi_cold(P.cold, q.s, q.ye, q.xn) : 0.0f;
    const float eps = fmaxf(__fdividef(q.p - pc, P.gm1 * q.s), eps_floor(P, q.im));
    const float ec = COLD ? eps_cold(P.cold, q.s, q.ye, q.xn) : 0.0f;
    return q.s * (eps + ec) + 0.5f * q.s * (q.u * q.u + q.v * q.v);
}

// primitives of a conserved (s, mx, my, E, s*im, s*Y_e, s*X_n); false when not a state (hydro.rs:462-479)
template <bool COLD>
__device__ __forceinline__ bool prim_of(const HydroParams &P, float s, float mx, float my, float e, float sim,
                                        float sye, float sxn, Prim &o) {
    if (!(s > 0.0f) || !isfinite(s) || !isfinite(mx) || !isfinite(my) || !isfinite(e) || !isfinite(sim))
        return false;
    const float is = 1.0f / s;
    const float im = sim * is, u = mx * is, v = my * is, ye = sye * is, xn = sxn * is;
    const float raw = e * is - 0.5f * (u * u + v * v) - (COLD ? eps_cold(P.cold, s, ye, xn) : 0.0f);
    if (!(raw > 0.0f)) return false;
    const float eps = fmaxf(raw, eps_floor(P, im));
    o.s = s;
    o.u = u;
    o.v = v;
    o.p = P.gm1 * s * eps + (COLD ? pi_cold(P.cold, s, ye, xn) : 0.0f);
    o.im = im;
    o.ye = ye;
    o.xn = xn;
    return true;
}

// normal flux of (Sigma, mx, my, E) through a face normal to dir of a state with energy e
__device__ __forceinline__ void flux4(const Prim &q, float e, int dir, float f[4]) {
    const float un = dir ? q.v : q.u, ut = dir ? q.u : q.v;
    const float m = q.s * un;
    f[0] = m;
    f[1 + dir] = m * un + q.p;
    f[2 - dir] = m * ut;
    f[3] = (e + q.p) * un;
}

// HLL between face states l and r (hydro.rs:516-540); returns the 4 fluxes, and whether the left side is upwind
__device__ __forceinline__ bool hll(const HydroParams &P, const Prim &l, float el, const Prim &r, float er, int dir,
                                    float f[4]) {
    const float unl = dir ? l.v : l.u, unr = dir ? r.v : r.u;
    const float cl = sound(P, l.s, l.p), cr = sound(P, r.s, r.p);
    const float sl = fminf(unl - cl, unr - cr), sr = fmaxf(unl + cl, unr + cr);
    float fl[4], fr[4];
    flux4(l, el, dir, fl);
    flux4(r, er, dir, fr);
    if (sl >= 0.0f) {
        for (int k = 0; k < 4; k++) f[k] = fl[k];
    } else if (sr <= 0.0f) {
        for (int k = 0; k < 4; k++) f[k] = fr[k];
    } else {
        const float ul[4] = {l.s, l.s * l.u, l.s * l.v, el}, ur[4] = {r.s, r.s * r.u, r.s * r.v, er};
        const float inv = 1.0f / (sr - sl);
        for (int k = 0; k < 4; k++) f[k] = (sr * fl[k] - sl * fr[k] + sl * sr * (ur[k] - ul[k])) * inv;
    }
    return f[0] > 0.0f;
}

// gravity's source at one face (hydro.rs:1048-1067); d = (phi_low - phi_high) / dx
__device__ inline void face_work(float q, float sl, float sh, float pl, float ph, float d, float dt, float out[3]) {
    const float a = fmaxf(sl, 0.0f), b = fmaxf(sh, 0.0f);
    const float w = a + b > 0.0f ? a / (a + b) : 0.5f;
    const float extra = dt * (q - 0.5f * (pl + ph)) * d;
    float c = 0.0f, kick = 0.0f, el = 0.0f, eh = 0.0f;
    if (extra < 0.0f && a > 0.0f && b > 0.0f) {
        const float ul = pl / a, uh = ph / b, du = ul - uh, inv = 1.0f / a + 1.0f / b;
        c = fminf(-extra, 0.25f * du * du / inv);
        if (c > 0.0f) {
            const float disc = sqrtf(fmaxf(du * du - 2.0f * inv * c, 0.0f));
            const float sg = signbit(du) ? -1.0f : 1.0f;
            kick = -2.0f * c / (du + sg * disc);
            el = ul * kick + 0.5f * kick * kick / a;
            eh = -uh * kick + 0.5f * kick * kick / b;
        }
    }
    const float rest = extra + c;
    out[0] = kick;
    out[1] = 0.5f * dt * pl * d + w * rest + el;
    out[2] = 0.5f * dt * ph * d + (1.0f - w) * rest + eh;
}

// slot index of a local tile position (x, y in -1..T), -1 for the corners
__device__ __forceinline__ int slot_of(int x, int y) {
    if (x >= 0 && x < T && y >= 0 && y < T) return y * T + x;
    if (x == -1 && y >= 0 && y < T) return T * T + y;
    if (x == T && y >= 0 && y < T) return T * T + T + y;
    if (y == -1 && x >= 0 && x < T) return T * T + 2 * T + x;
    if (y == T && x >= 0 && x < T) return T * T + 3 * T + x;
    return -1;
}

__device__ __forceinline__ void slot_pos(int t, int &x, int &y) {
    if (t < T * T) {
        x = t % T;
        y = t / T;
    } else if (t < T * T + T) {
        x = -1;
        y = t - T * T;
    } else if (t < T * T + 2 * T) {
        x = T;
        y = t - T * T - T;
    } else if (t < T * T + 3 * T) {
        x = t - T * T - 2 * T;
        y = -1;
    } else {
        x = t - T * T - 3 * T;
        y = T;
    }
}

template <bool COLD> struct Lds {
    float s[NPRIM], u[NPRIM], v[NPRIM], p[NPRIM], im[NPRIM];
    float ye[COLD ? NPRIM : 1];  // Y_e, with the cold pressure on
    float x[NS][NPRIM];
    // predicted face states of every slot: [field][face][slot], fields s u v p e; reused for species faces
    float face[5][4][NSLOT];
    float mid[3][NSLOT];
    double red[8][5];
    unsigned long long redk[8];
};

template <int MODE, bool COLD, bool SINKS = false>
__global__ void __launch_bounds__(NSLOT) k_hydro(HydroParams P, HydroState in, HydroState out,
                                                 const double *__restrict__ scal, const float *__restrict__ gx,
                                                 const float *__restrict__ gy, const float *__restrict__ sigma_mid,
                                                 const float *__restrict__ phi_mid, const float *__restrict__ rfx,
                                                 const float *__restrict__ rfy, float *__restrict__ sigma_mid_out,
                                                 float *__restrict__ pushmag, double *__restrict__ escaped,
                                                 HydroPartials part, SinkPlanes SP) {
    __shared__ Lds<COLD> L;
    const int w = P.w, h = P.h;
    const int tile = xcd_tile();
    const int bx = (tile % gridDim.x) * T, by = (tile / gridDim.x) * T;
    const int tid = threadIdx.x;
    const double dt = scal[0];
    const float dtf = (float)dt, half = 0.5f * dtf, hdx = half * P.inv_dx;
    const double dtdx = dt * (double)P.inv_dx;
    const bool grav = gx != nullptr;
    const size_t n = (size_t)w * h;

    // ---- phase 0: primitives of the tile and a two-cell halo, ghosts by the edge rules (hydro.rs:673-710)
    for (int t = tid; t < NPRIM; t += NSLOT) {
        const int lx = t % PW, ly = t / PW;
        int gxi = bx + lx - HALO, gyi = by + ly - HALO;
        bool fx = false, fy = false;
        if (gxi < 0 || gxi >= w) {
            fx = true;
            gxi = P.edge == 0 ? (gxi < 0 ? -gxi - 1 : 2 * w - 1 - gxi) : (gxi < 0 ? 0 : w - 1);
            gxi = min(max(gxi, 0), w - 1);
        }
        if (gyi < 0 || gyi >= h) {
            fy = true;
            gyi = P.edge == 0 ? (gyi < 0 ? -gyi - 1 : 2 * h - 1 - gyi) : (gyi < 0 ? 0 : h - 1);
            gyi = min(max(gyi, 0), h - 1);
        }
        const size_t i = (size_t)gyi * w + gxi;
        float xs[NS];
        float im = 0.0f, ye = 0.0f;
        for (int s = 0; s < NS; s++) {
            xs[s] = in.x[s * n + i];
            im += xs[s] * c_imu[s];
            if (COLD) ye += xs[s] * c_ye[s];
        }
        if (P.fixed_imu > 0.0f) im = P.fixed_imu;
        const double sig = in.sigma[i];
        Prim q;
        if (sig < (double)P.sigma_vac) {
            q.s = P.sigma_floor;
            q.u = q.v = 0.0f;
            q.p = P.gm1 * P.sigma_floor * eps_floor(P, im) +
                  (COLD ? pi_cold(P.cold, P.sigma_floor, ye, xs[NS - 1]) : 0.0f);
        } else {
            const float s = (float)sig, is = 1.0f / s;
            const float u = in.mx[i] * is, v = in.my[i] * is;
            const float raw = in.e[i] * is - 0.5f * (u * u + v * v) -
                              (COLD ? eps_cold(P.cold, s, ye, xs[NS - 1]) : 0.0f);
            const float eps = fmaxf(raw, eps_floor(P, im));
            q.s = s;
            q.u = u;
            q.v = v;
            q.p = P.gm1 * s * eps + (COLD ? pi_cold(P.cold, s, ye, xs[NS - 1]) : 0.0f);
        }
        if (P.edge == 0) {
            if (fx) q.u = -q.u;
            if (fy) q.v = -q.v;
        } else {
            const bool hx = gxi == w - 1 && bx + lx - HALO >= w, hy = gyi == h - 1 && by + ly - HALO >= h;
            if (fx) q.u = hx ? fmaxf(q.u, 0.0f) : fminf(q.u, 0.0f);
            if (fy) q.v = hy ? fmaxf(q.v, 0.0f) : fminf(q.v, 0.0f);
        }
        L.s[t] = q.s;
        L.u[t] = q.u;
        L.v[t] = q.v;
        L.p[t] = q.p;
        L.im[t] = im;
        if (COLD) L.ye[t] = ye;
        for (int s = 0; s < NS; s++) L.x[s][t] = xs[s];
    }
    __syncthreads();

    // ---- phase 1: every slot's slopes and Hancock predictor (hydro.rs:587-614)
    int sx_, sy_;
    slot_pos(tid, sx_, sy_);
    const int lc = (sy_ + HALO) * PW + sx_ + HALO;
    const int cgx = min(max(bx + sx_, 0), w - 1), cgy = min(max(by + sy_, 0), h - 1);
    float gcx = 0.0f, gcy = 0.0f;
    if (grav) {
        gcx = gx[(size_t)cgy * w + cgx];
        gcy = gy[(size_t)cgy * w + cgx];
    }
    float theta[2];            // the species' shared limiter factor per axis
    float fs_w[4], f0_w[4];    // reconstructed faces' Sigma and mass flux
    float s_e[4];              // evolved faces' Sigma (0 when the unevolved face is kept)
    float is_e[4];             // its reciprocal (0 when the unevolved face is kept)
    {
        const Prim c = {L.s[lc], L.u[lc], L.v[lc], L.p[lc], L.im[lc], COLD ? L.ye[lc] : 0.0f, L.x[NS - 1][lc]};
        Prim wf[4];
        for (int ax = 0; ax < 2; ax++) {
            const int off = ax ? PW : 1;
            const int il = lc - off, ir = lc + off;
            const Prim l = {L.s[il], L.u[il], L.v[il], L.p[il], L.im[il], COLD ? L.ye[il] : 0.0f, L.x[NS - 1][il]};
            const Prim r = {L.s[ir], L.u[ir], L.v[ir], L.p[ir], L.im[ir], COLD ? L.ye[ir] : 0.0f, L.x[NS - 1][ir]};
            float ss = mc(l.s, c.s, r.s), su = mc(l.u, c.u, r.u), sv = mc(l.v, c.v, r.v), sp = mc(l.p, c.p, r.p);
            float th = 1.0f;
            for (int s = 0; s < NS; s++) {
                const float xl = L.x[s][il], xc = L.x[s][lc], xr = L.x[s][ir];
                const float central = 0.5f * (xr - xl);
                if (central != 0.0f) th = fminf(th, __fdividef(mc(xl, xc, xr), central));
            }
            theta[ax] = th;
            // 1/mu, Y_e and X_n are linear in the fractions: their faces take the species' shared slope
            const float sim = th * 0.5f * (r.im - l.im);
            const float sye = COLD ? th * 0.5f * (r.ye - l.ye) : 0.0f, sxn = COLD ? th * 0.5f * (r.xn - l.xn) : 0.0f;
            const float hottest = fmaxf(fmaxf(__fdividef(l.p, l.s), __fdividef(c.p, c.s)), __fdividef(r.p, r.s));
            bool hot = false;
            for (int k = 0; k < 2; k++) {
                const float f = k ? 0.5f : -0.5f;
                const float pi = c.p + f * sp, bound = HOT_FACE * hottest * (c.s + f * ss);
                if (!(pi <= bound)) hot = true;
            }
            if (hot) {
                ss = 0.0f;
                sp = 0.0f;
            }
            for (int k = 0; k < 2; k++) {
                const float f = k ? 0.5f : -0.5f;
                Prim &q = wf[2 * ax + k];
                q.s = c.s + f * ss;
                q.u = c.u + f * su;
                q.v = c.v + f * sv;
                q.p = c.p + f * sp;
                q.im = c.im + f * sim;
                q.ye = c.ye + f * sye;
                q.xn = c.xn + f * sxn;
            }
        }
        float ue[4], fl[4][4];
        for (int k = 0; k < 4; k++) {
            ue[k] = cons_e<COLD>(P, wf[k]);
            flux4(wf[k], ue[k], k >> 1, fl[k]);
            fs_w[k] = wf[k].s;
            f0_w[k] = fl[k][0];
        }
        float delta[7];
        for (int c4 = 0; c4 < 4; c4++) delta[c4] = hdx * ((fl[0][c4] - fl[1][c4]) + (fl[2][c4] - fl[3][c4]));
        delta[4] = hdx * ((fl[0][0] * wf[0].im - fl[1][0] * wf[1].im) + (fl[2][0] * wf[2].im - fl[3][0] * wf[3].im));
        delta[5] = COLD ? hdx * ((fl[0][0] * wf[0].ye - fl[1][0] * wf[1].ye) +
                                 (fl[2][0] * wf[2].ye - fl[3][0] * wf[3].ye))
                        : 0.0f;
        delta[6] = COLD ? hdx * ((fl[0][0] * wf[0].xn - fl[1][0] * wf[1].xn) +
                                 (fl[2][0] * wf[2].xn - fl[3][0] * wf[3].xn))
                        : 0.0f;
        float m0 = c.s + delta[0], m1 = c.s * c.u + delta[1], m2 = c.s * c.v + delta[2];
        if (grav) {
            m1 += half * c.s * gcx;
            m2 += half * c.s * gcy;
        }
        L.mid[0][tid] = m0;
        L.mid[1][tid] = m1;
        L.mid[2][tid] = m2;
        for (int k = 0; k < 4; k++) {
            const Prim &q = wf[k];
            float d1 = delta[1], d2 = delta[2], d3 = delta[3];
            if (grav) {
                d1 += half * q.s * gcx;
                d2 += half * q.s * gcy;
                d3 += half * q.s * ((q.u + 0.5f * half * gcx) * gcx + (q.v + 0.5f * half * gcy) * gcy);
            }
            const float s0 = q.s + delta[0], mx = q.s * q.u + d1, my = q.s * q.v + d2, e = ue[k] + d3;
            const float sim = q.s * q.im + delta[4], sye = q.s * q.ye + delta[5], sxn = q.s * q.xn + delta[6];
            Prim o;
            float oe;
            if (prim_of<COLD>(P, s0, mx, my, e, sim, sye, sxn, o)) {
                oe = e;
                s_e[k] = s0;
                is_e[k] = 1.0f / s0;
            } else {
                o = q;
                oe = ue[k];
                s_e[k] = 0.0f;
                is_e[k] = 0.0f;
            }
            L.face[0][k][tid] = o.s;
            L.face[1][k][tid] = o.u;
            L.face[2][k][tid] = o.v;
            L.face[3][k][tid] = o.p;
            L.face[4][k][tid] = oe;
        }
    }
    __syncthreads();

    // ---- phase 2: the tile cells' four face fluxes (hydro.rs:744-778, boundary :569-582)
    const bool own = tid < T * T;
    const int gxc = bx + sx_, gyc = by + sy_;
    const bool live = own && gxc < w && gyc < h;
    float F[4][4];  // fluxes of faces x-, x+, y-, y+ along +x / +y
    bool upl[4];    // whether the low side is upwind
    bool wallf[4];  // a wall face (bounce, or leave with inflow): no species flux
    if (live) {
        for (int k = 0; k < 4; k++) {
            const int dir = k >> 1;
            const bool high = k & 1;
            const int nx = sx_ + (dir == 0 ? (high ? 1 : -1) : 0), ny = sy_ + (dir == 1 ? (high ? 1 : -1) : 0);
            const int ns = slot_of(nx, ny);
            // the low side's face toward +, the high side's face toward -
            const int ls = high ? tid : ns, lk = high ? k : k + 1;
            const int hs = high ? ns : tid, hk = high ? k - 1 : k;
            const Prim lp = {L.face[0][lk][ls], L.face[1][lk][ls], L.face[2][lk][ls], L.face[3][lk][ls], 0.0f};
            const Prim hp = {L.face[0][hk][hs], L.face[1][hk][hs], L.face[2][hk][hs], L.face[3][hk][hs], 0.0f};
            const float le = L.face[4][lk][ls], he = L.face[4][hk][hs];
            const bool edge = dir == 0 ? (high ? gxc == w - 1 : gxc == 0) : (high ? gyc == h - 1 : gyc == 0);
            wallf[k] = false;
            if (edge && P.edge == 0) {
                for (int c4 = 0; c4 < 4; c4++) F[k][c4] = 0.0f;
                F[k][1 + dir] = high ? lp.p : hp.p; // the own face's pressure
                wallf[k] = true;
                upl[k] = true;
            } else {
                upl[k] = hll(P, lp, le, hp, he, dir, F[k]);
                if (edge) {
                    const bool inflow = high ? F[k][0] < 0.0f : F[k][0] > 0.0f;
                    if (inflow) {
                        for (int c4 = 0; c4 < 4; c4++) F[k][c4] = 0.0f;
                        F[k][1 + dir] = high ? lp.p : hp.p;
                        wallf[k] = true;
                    }
                }
            }
        }
    }
    const size_t ci = (size_t)min(gyc, h - 1) * w + min(gxc, w - 1);
    const double sig0 = live ? in.sigma[ci] : 0.0;
    bool empty = false;
    if (live) {
        auto vac = [&](int x, int y) { return in.sigma[(size_t)y * w + x] < (double)P.sigma_vac; };
        empty = vac(gxc, gyc) && (gxc == 0 || vac(gxc - 1, gyc)) && (gxc + 1 == w || vac(gxc + 1, gyc)) &&
                (gyc == 0 || vac(gxc, gyc - 1)) && (gyc + 1 == h || vac(gxc, gyc + 1));
    }
    const bool vacuum = sig0 < (double)P.sigma_vac;
    const double start = vacuum ? (double)P.sigma_floor : sig0;
    const double divm = ((double)F[1][0] - (double)F[0][0]) + ((double)F[3][0] - (double)F[2][0]);

    if (MODE == HYDRO_MID) {
        if (live) {
            double s = empty ? 0.0 : 0.5 * (sig0 + (start - dtdx * divm));
            sigma_mid_out[ci] = s >= (double)P.sigma_vac ? (float)s : 0.0f;
        }
        return;
    }

    // ---- phase 3: the update (hydro.rs:864-984)
    double kick_net[2] = {0.0, 0.0}, push_net[3] = {0.0, 0.0, 0.0};
    unsigned long long dkey = 0ull;
    double u0 = start;
    float u1 = 0.f, u2 = 0.f, u3 = 0.f;
    if (live && !empty) {
        if (vacuum) {
            const Prim fq = {L.s[lc], L.u[lc], L.v[lc], L.p[lc], L.im[lc], COLD ? L.ye[lc] : 0.0f, L.x[NS - 1][lc]};
            u1 = 0.0f;
            u2 = 0.0f;
            u3 = cons_e<COLD>(P, fq);
        } else {
            u1 = in.mx[ci];
            u2 = in.my[ci];
            u3 = in.e[ci];
        }
        u0 -= dtdx * divm;
        const float dtdxf = (float)dtdx;
        u1 -= dtdxf * ((F[1][1] - F[0][1]) + (F[3][1] - F[2][1]));
        u2 -= dtdxf * ((F[1][2] - F[0][2]) + (F[3][2] - F[2][2]));
        u3 -= dtdxf * ((F[1][3] - F[0][3]) + (F[3][3] - F[2][3]));
        float phi_c = 0.0f;
        if (grav) {
            const int rw = w + 2;
            const float *pr = phi_mid + (size_t)(gyc + 1) * rw + gxc + 1;
            const float src = sigma_mid[ci];
            const float gmx = -(pr[1] - pr[-1]) * 0.5f * P.inv_dx, gmy = -(pr[rw] - pr[-rw]) * 0.5f * P.inv_dx;
            float kick[2] = {dtf * src * gmx, dtf * src * gmy};
            // the net kick taken back is the gas's own field's; the sinks' pull is the sinks' to book
            kick_net[0] = (double)kick[0];
            kick_net[1] = (double)kick[1];
            // phi_mid as the faces' work reads it: the gas's with the sinks' potential
            float pw[5] = {pr[0], pr[1], pr[-1], pr[rw], pr[-rw]};
            if constexpr (SINKS) {
                const float *ps = SP.ps + (size_t)(gyc + 1) * rw + gxc + 1;
                pw[0] += ps[0];
                pw[1] += ps[1];
                pw[2] += ps[-1];
                pw[3] += ps[rw];
                pw[4] += ps[-rw];
                kick[0] += dtf * src * SP.gsx[ci];
                kick[1] += dtf * src * SP.gsy[ci];
            }
            phi_c = pw[0];
            if (src > 0.0f) dkey = ((unsigned long long)__float_as_uint(src) << 32) | (0xffffffffu - (unsigned)ci);
            const float m0 = L.mid[0][tid], m1 = L.mid[1][tid], m2 = L.mid[2][tid];
            float r[3], l[3], dn[3], up[3];
            // neighbours' phi_mid and half-step state; the world's edge does no work (the cell's own values)
            {
                const bool has = gxc + 1 < w;
                const int ns = slot_of(sx_ + 1, sy_);
                const float pn = has ? pw[1] : phi_c;
                face_work(F[1][0], m0, has ? L.mid[0][ns] : m0, m1, has ? L.mid[1][ns] : m1, (phi_c - pn) * P.inv_dx,
                          dtf, r);
            }
            {
                const bool has = gxc > 0;
                const int ns = slot_of(sx_ - 1, sy_);
                const float pn = has ? pw[2] : phi_c;
                face_work(F[0][0], has ? L.mid[0][ns] : m0, m0, has ? L.mid[1][ns] : m1, m1, (pn - phi_c) * P.inv_dx,
                          dtf, l);
            }
            {
                const bool has = gyc + 1 < h;
                const int ns = slot_of(sx_, sy_ + 1);
                const float pn = has ? pw[3] : phi_c;
                face_work(F[3][0], m0, has ? L.mid[0][ns] : m0, m2, has ? L.mid[2][ns] : m2, (phi_c - pn) * P.inv_dx,
                          dtf, dn);
            }
            {
                const bool has = gyc > 0;
                const int ns = slot_of(sx_, sy_ - 1);
                const float pn = has ? pw[4] : phi_c;
                face_work(F[2][0], has ? L.mid[0][ns] : m0, m0, has ? L.mid[2][ns] : m2, m2, (pn - phi_c) * P.inv_dx,
                          dtf, up);
            }
            kick[0] += r[0] - l[0];
            kick[1] += dn[0] - up[0];
            u1 += kick[0];
            u2 += kick[1];
            u3 += r[1] + l[2] + dn[1] + up[2];
        }
        if (rfx) {
            // the radiation force from the previous step's flux, capped by the heat above the floor
            float ax = 0.0f, ay = 0.0f;
            if (!vacuum) {
                float im = L.im[lc];
                float kap = P.kappa0 * (1.0f + L.x[0][lc]);
                if (P.opacity == 1 || P.kappa_dust != 0.0f) {
                    const float s = (float)sig0, mx = in.mx[ci], my = in.my[ci];
                    const float raw = in.e[ci] / s - 0.5f * (mx * mx + my * my) / (s * s) -
                                      (COLD ? eps_cold(P.cold, s, L.ye[lc], L.x[NS - 1][lc]) : 0.0f);
                    const float t = P.gm1 * fmaxf(raw, eps_floor(P, im)) / im;
                    if (P.opacity == 1) {
                        kap = P.kappa0 * s * powf(t, -3.5f);
                    } else if (t < P.t_dust) {
                        const float zm = 1.0f - L.x[0][lc] - L.x[1][lc] - L.x[9][lc];
                        kap += P.kappa_dust * zm;
                    }
                }
                const float ks = kap * (float)sig0 / P.c_sb;
                ax = ks * rfx[ci];
                ay = ks * rfy[ci];
            }
            const float d0 = dtf * ax, d1 = dtf * ay, d2 = d0 * d0 + d1 * d1;
            float share = 1.0f;
            if (d2 != 0.0f) {
                const float s = (float)u0;
                // the cell's composition before the update, as for 1/mu
                const float raw = u3 / s - 0.5f * (u1 * u1 + u2 * u2) / (s * s) -
                                  (COLD ? eps_cold(P.cold, s, L.ye[lc], L.x[NS - 1][lc]) : 0.0f);
                const float a = s * (raw - eps_floor(P, L.im[lc]));
                if (!(a > 0.0f)) {
                    share = 0.0f;
                } else {
                    const float q = u1 * d0 + u2 * d1;
                    const float disc = sqrtf(q * q + 2.0f * s * a * d2);
                    share = fminf(q >= 0.0f ? 2.0f * s * a / (q + disc) : (disc - q) / d2, 1.0f);
                }
            }
            const float p0 = share * d0, p1 = share * d1;
            u1 += p0;
            u2 += p1;
            const float pm = sqrtf(p0 * p0 + p1 * p1);
            pushmag[ci] = pm;
            push_net[0] = p0;
            push_net[1] = p1;
            push_net[2] = pm;
        }
        // escaped through the world's edge (hydro.rs:954-976): dt/dx f, signed, per boundary-face slot
        {
            const int nf = 2 * (w + h);
            const int slots[4] = {gyc, h + gyc, 2 * h + gxc, 2 * h + w + gxc};
            const bool edges[4] = {gxc == 0, gxc + 1 == w, gyc == 0, gyc + 1 == h};
            for (int k = 0; k < 4; k++)
                if (edges[k]) {
                    const double sg = (k & 1) ? 1.0 : -1.0;
                    for (int c4 = 0; c4 < 4; c4++) escaped[(size_t)c4 * nf + slots[k]] += sg * dtdx * F[k][c4];
                    escaped[(size_t)3 * nf + slots[k]] += sg * dtdx * (double)F[k][0] * (double)phi_c;
                }
        }
    }
    if (live) {
        if (empty) {
            out.sigma[ci] = sig0;
            out.mx[ci] = in.mx[ci];
            out.my[ci] = in.my[ci];
            out.e[ci] = in.e[ci];
        } else {
            out.sigma[ci] = u0;
            out.mx[ci] = u1;
            out.my[ci] = u2;
            out.e[ci] = u3;
        }
    }

    const double inv_u0 = u0 != 0.0 ? 1.0 / u0 : 0.0;
    // ---- phase 4: the species, SB at a time in the freed face buffer (flux = mass flux x the upwind face's
    // fraction)
    constexpr int SB = SR_HYDRO_SB;
    static_assert(NS % SB == 0 && SB <= 5, "species batches fill the face buffer's five fields");
    for (int s0 = 0; s0 < NS; s0 += SB) {
        __syncthreads(); // the face buffer is reused
#pragma unroll
        for (int j = 0; j < SB; j++) {
            const int s = s0 + j;
            const int offs[2] = {1, PW};
            float xf[4];
            const float xc = L.x[s][lc];
            for (int ax = 0; ax < 2; ax++) {
                const float central = 0.5f * (L.x[s][lc + offs[ax]] - L.x[s][lc - offs[ax]]);
                xf[2 * ax] = xc - 0.5f * theta[ax] * central;
                xf[2 * ax + 1] = xc + 0.5f * theta[ax] * central;
            }
            const float ds = hdx * ((f0_w[0] * xf[0] - f0_w[1] * xf[1]) + (f0_w[2] * xf[2] - f0_w[3] * xf[3]));
            for (int k = 0; k < 4; k++) L.face[j][k][tid] = s_e[k] > 0.0f ? (fs_w[k] * xf[k] + ds) * is_e[k] : xf[k];
        }
        __syncthreads();
        if (live) {
#pragma unroll
            for (int j = 0; j < SB; j++) {
                const int s = s0 + j;
                const size_t pi = (size_t)s * n + ci;
                const float xs0 = L.x[s][lc];  // the cell's own fraction, staged in phase 0
                if (empty) {
                    out.x[pi] = xs0;
                    continue;
                }
                float fsx[4];
                for (int k = 0; k < 4; k++) {
                    if (wallf[k]) {
                        fsx[k] = 0.0f;
                        continue;
                    }
                    const int dir = k >> 1;
                    const bool high = k & 1;
                    const int nx = sx_ + (dir == 0 ? (high ? 1 : -1) : 0),
                              ny = sy_ + (dir == 1 ? (high ? 1 : -1) : 0);
                    const int ns = slot_of(nx, ny);
                    const int ls = high ? tid : ns, lk = high ? k : k + 1;
                    const int hs = high ? ns : tid, hk = high ? k - 1 : k;
                    fsx[k] = F[k][0] * (upl[k] ? L.face[j][lk][ls] : L.face[j][hk][hs]);
                }
                const double um =
                    start * (double)xs0 - dtdx * (((double)fsx[1] - fsx[0]) + ((double)fsx[3] - fsx[2]));
                out.x[pi] = (float)(um * inv_u0);
                const int nf = 2 * (w + h);
                const int slots[4] = {gyc, h + gyc, 2 * h + gxc, 2 * h + w + gxc};
                const bool edges[4] = {gxc == 0, gxc + 1 == w, gyc == 0, gyc + 1 == h};
                for (int k = 0; k < 4; k++)
                    if (edges[k])
                        escaped[(size_t)(4 + s) * nf + slots[k]] += ((k & 1) ? 1.0 : -1.0) * dtdx * fsx[k];
            }
        }
    }

    // ---- block partials of gravity's net kick, the densest free cell and the radiation force's net push
    if (grav || rfx) {
        const int lane = tid & 63, wid = tid >> 6;
        double v[5] = {kick_net[0], kick_net[1], push_net[0], push_net[1], push_net[2]};
        for (int k = 0; k < 5; k++) v[k] = wave_sum(v[k]);
        unsigned long long kk = dkey;
        for (int o = 32; o > 0; o >>= 1) {
            const unsigned long long other = __shfl_xor(kk, o, 64);
            kk = other > kk ? other : kk;
        }
        __syncthreads();
        if (lane == 0) {
            for (int k = 0; k < 5; k++) L.red[wid][k] = v[k];
            L.redk[wid] = kk;
        }
        __syncthreads();
        if (tid == 0) {
            double acc[5] = {0, 0, 0, 0, 0};
            unsigned long long best = 0ull;
            for (int wv = 0; wv < NSLOT / 64; wv++) {
                for (int k = 0; k < 5; k++) acc[k] += L.red[wv][k];
                best = L.redk[wv] > best ? L.redk[wv] : best;
            }
            const int b = tile;  // partials in tile order
            for (int k = 0; k < 5; k++) part.sums[(size_t)b * 5 + k] = acc[k];
            part.keys[b] = best;
        }
    }
}

// The take-backs (hydro.rs:986-1001): the radiation force's net push from every pushed cell in proportion to its
// push, gravity's net kick from the densest free cell. One block reduces the partials in a fixed order first.
__global__ void __launch_bounds__(256) k_hydro_reduce(HydroPartials part, int nblocks, double *__restrict__ takeback) {
    __shared__ double sh[4][5];
    __shared__ unsigned long long shk[4];
    double acc[5] = {0, 0, 0, 0, 0};
    unsigned long long best = 0ull;
    for (int b = threadIdx.x; b < nblocks; b += blockDim.x) {
        for (int k = 0; k < 5; k++) acc[k] += part.sums[(size_t)b * 5 + k];
        best = part.keys[b] > best ? part.keys[b] : best;
    }
    for (int k = 0; k < 5; k++) acc[k] = wave_sum(acc[k]);
    for (int o = 32; o > 0; o >>= 1) {
        const unsigned long long other = __shfl_xor(best, o, 64);
        best = other > best ? other : best;
    }
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) {
        for (int k = 0; k < 5; k++) sh[wid][k] = acc[k];
        shk[wid] = best;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double t[5] = {0, 0, 0, 0, 0};
        unsigned long long bk = 0ull;
        for (int wv = 0; wv < 4; wv++) {
            for (int k = 0; k < 5; k++) t[k] += sh[wv][k];
            bk = shk[wv] > bk ? shk[wv] : bk;
        }
        for (int k = 0; k < 5; k++) takeback[k] = t[k];
        takeback[5] = bk ? (double)(0xffffffffu - (unsigned)(bk & 0xffffffffu)) : -1.0;
    }
}

__global__ void k_hydro_takeback(HydroState st, const float *__restrict__ pushmag, const double *__restrict__ tb,
                                 int n, int radiation) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float mx = st.mx[i], my = st.my[i];
    bool changed = false;
    if (radiation && tb[4] > 0.0) {
        const float m = pushmag[i];
        if (m != 0.0f) {
            mx -= (float)((double)m * tb[2] / tb[4]);
            my -= (float)((double)m * tb[3] / tb[4]);
            changed = true;
        }
    }
    if (tb[5] >= 0.0 && (int)tb[5] == i) {
        mx -= (float)tb[0];
        my -= (float)tb[1];
        changed = true;
    }
    if (changed) {
        st.mx[i] = mx;
        st.my[i] = my;
    }
}

// ------------------------------------------------------------------------------------------------------------ host

int hydro_launch(int mode, const HydroParams &P, const HydroState &in, const HydroState &out, const double *scal,
                 const float *gx, const float *gy, const float *sigma_mid, const float *phi_mid, const float *rfx,
                 const float *rfy, float *sigma_mid_out, float *pushmag, double *escaped, const HydroPartials &part,
                 hipStream_t s, const SinkPlanes *sinks) {
    dim3 grid((P.w + T - 1) / T, (P.h + T - 1) / T);
    const SinkPlanes sp = sinks ? *sinks : SinkPlanes{nullptr, nullptr, nullptr};
    const bool with_sinks = mode == HYDRO_UPDATE && sp.ps && phi_mid;
    // the cold pressure's variant only where a pair is on, the sinks' only with sinks: the plain gas keeps the kernel
    // without them
    auto go = [&](auto kernel) {
        hipLaunchKernelGGL(kernel, grid, dim3(NSLOT), 0, s, P, in, out, scal, gx, gy, sigma_mid, phi_mid, rfx, rfy,
                           sigma_mid_out, pushmag, escaped, part, sp);
    };
    if (mode == HYDRO_MID)
        P.cold.on ? go(k_hydro<HYDRO_MID, true>) : go(k_hydro<HYDRO_MID, false>);
    else if (with_sinks)
        P.cold.on ? go(k_hydro<HYDRO_UPDATE, true, true>) : go(k_hydro<HYDRO_UPDATE, false, true>);
    else
        P.cold.on ? go(k_hydro<HYDRO_UPDATE, true>) : go(k_hydro<HYDRO_UPDATE, false>);
    SR_HIP(hipGetLastError());
    return SR_OK;
}

int hydro_takeback(const HydroParams &P, const HydroState &st, const float *pushmag, const HydroPartials &part,
                   double *takeback, bool radiation, hipStream_t s) {
    const int nb = ((P.w + T - 1) / T) * ((P.h + T - 1) / T);
    hipLaunchKernelGGL(k_hydro_reduce, dim3(1), dim3(256), 0, s, part, nb, takeback);
    const int n = P.w * P.h;
    hipLaunchKernelGGL(k_hydro_takeback, dim3((n + 255) / 256), dim3(256), 0, s, st, pushmag, takeback, n,
                       radiation ? 1 : 0);
    SR_HIP(hipGetLastError());
    return SR_OK;
}

int hydro_blocks(int w, int h) { return ((w + T - 1) / T) * ((h + T - 1) / T); }

} // namespace sr
