// P5: one implicit (backward Euler) flux-limited-diffusion step in u = a_r T^4 with the Levermore-Pomraning limiter,
// Newton passes on T(u), each pass's SPD system solved by FldSolver, then the delta-closure commit (or the limited
// commit past the ceiling). The scheme is the reference's (crates/sr-physics/src/reference/heat.rs:453-592; GPU
// crates/sr-engine/shaders/heat.wgsl:190-285, :547-565, :914-1160); the kernels are this library's own: one pass over
// the cells for the sides, one over the faces for the law, one per Newton pass for the system and one for its update,
// FP64 sums by the deterministic last-block reduction.
#include <algorithm>
#include <cmath>

#include "common.hpp"
#include "heat.hpp"

namespace sr {

#define NT 256
#define CELLS 4  // cells per thread whose loads a pass issues together (the grids hold n / (NT CELLS) blocks)
#define NS 10

__constant__ float h_imu[NS] = {2.0f, 0.75f, 7.0f / 12.0f, 9.0f / 16.0f, 11.0f / 20.0f,
                                13.0f / 24.0f, 15.0f / 28.0f, 17.0f / 32.0f, 27.0f / 56.0f, 1.0f};


__device__ __forceinline__ double u_of(const HeatParams &P, double t) { return P.cd ? t : P.a_r * t * t * t * t; }
__device__ __forceinline__ double t_of(const HeatParams &P, double u) { return P.cd ? u : sqrt(sqrt(u / P.a_r)); }
__device__ __forceinline__ double dt_du(const HeatParams &P, double t) {
    return P.cd ? 1.0 : 1.0 / (4.0 * P.a_r * t * t * t);
}

// the sides of every cell (heat.rs:202-219): gas flag, T with the floor, kappa, 1/mu, heat capacity C
__global__ void k_heat_cells(HeatParams P, const double *__restrict__ sigma, const float *__restrict__ mx,
                             const float *__restrict__ my, const float *__restrict__ e, const float *__restrict__ x,
                             HeatPlanes H, double *part) {
    const int n = P.w * P.h;
    double v[1] = {0.0};
    for (int i = blockIdx.x * NT + threadIdx.x; i < n; i += gridDim.x * NT) {
        float xs0 = x[i];
        float im = 0.0f, ye = 0.0f;
        for (int s = 0; s < NS; s++) {
            im += x[(size_t)s * n + i] * h_imu[s];
            ye += x[(size_t)s * n + i] * c_ye[s];
        }
        if (P.fixed_imu > 0.0f) im = P.fixed_imu;
        const double sig = sigma[i];
        const bool gas = sig >= (double)P.sigma_vac;
        double t, s;
        if (!gas) {
            t = P.t_floor;
            s = P.sigma_floor;
        } else {
            s = sig;
            const double m0 = mx[i], m1 = my[i];
            const double cold = P.cold.on ? (double)eps_cold(P.cold, (float)s, ye, x[(size_t)(NS - 1) * n + i]) : 0.0;
            const double raw = (double)e[i] / s - 0.5 * (m0 * m0 + m1 * m1) / (s * s) - cold;
            const double fl = (double)P.t_floor * im / P.gm1;
            const double eps = raw < fl ? fl : raw;
            t = P.gm1 * eps / im;
        }
        float kappa;
        if (P.opacity == 1) {
            kappa = (float)(P.kappa0 * s * pow(t, -3.5));
        } else {
            float dust = 0.0f;
            if (t < P.t_dust) dust = P.kappa_dust * (1.0f - xs0 - x[(size_t)n + i] - x[(size_t)9 * n + i]);
            kappa = P.kappa0 * (1.0f + xs0) + dust;
        }
        H.gas[i] = gas ? 1 : 0;
        H.t0[i] = t;
        H.t[i] = t;
        H.kappa[i] = kappa;
        H.imu[i] = im;
        H.cap[i] = (float)(s * im / P.gm1);
        H.sig[i] = (float)s;
        if (gas) v[0] += (double)H.cap[i] * t;
    }
    block_partials<1>(v, part, gridDim.x);
}

struct SideV {
    bool gas;
    double t;
    float sig, kappa, imu;
};

__device__ inline SideV side_at(const HeatPlanes &H, int i) {
    return {H.gas[i] != 0, H.t0[i], H.sig[i], H.kappa[i], H.imu[i]};
}

// D_f / dx^2 of a face between sides a and b, either standing for the vacuum beyond (heat.rs:251-282)
__device__ inline float face_k(const HeatParams &P, bool has_a, SideV a, bool has_b, SideV b) {
    SideV A, B;
    auto vac = [&](const SideV &g) { return SideV{false, (double)P.t_floor, P.sigma_floor, g.kappa, g.imu}; };
    if (has_a && has_b) {
        if (!a.gas && !b.gas) return 0.0f;
        A = a.gas ? a : vac(b);
        B = b.gas ? b : vac(a);
    } else if (has_a) {
        if (!a.gas) return 0.0f;
        A = a;
        B = vac(a);
    } else {
        if (!b.gas) return 0.0f;
        A = vac(b);
        B = b;
    }
    const double idx2 = (double)P.inv_dx * P.inv_dx;
    if (P.cd) {
        const double ca = (double)A.sig * A.imu / P.gm1, cb = (double)B.sig * B.imu / P.gm1;
        return (float)(P.chi * 0.5 * (ca + cb) * idx2);
    }
    const double ta = A.t, tb = B.t;
    const double era = P.a_r * ta * ta * ta * ta, erb = P.a_r * tb * tb * tb * tb;
    const double fer = 0.5 * (era + erb);
    const double ks = 0.5 * ((double)A.kappa * A.sig + (double)B.kappa * B.sig);
    const double r = fer > 0.0 ? fabs(erb - era) / ((double)P.dx * ks * fer) : 0.0;
    const double lam = P.limiter ? (2.0 + r) / (6.0 + 3.0 * r + r * r) : 1.0 / 3.0;
    double d = P.c_sb * lam / ks;
    if (P.k_cond != 0.0) d += P.k_cond / (P.a_r * (ta + tb) * (ta * ta + tb * tb));
    return (float)(d * idx2);
}

// the face law on the start state: raw k on every face, the couplings (both sides gas) into the solver's level 0
__global__ void k_heat_law(HeatParams P, HeatPlanes H, float *__restrict__ ax, float *__restrict__ ay) {
    const int w = P.w, h = P.h;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < (w + 1) * h) {
        const int x = t % (w + 1), y = t / (w + 1);
        const bool ha = x > 0, hb = x < w;
        const SideV a = ha ? side_at(H, y * w + x - 1) : SideV{}, b = hb ? side_at(H, y * w + x) : SideV{};
        const float k = face_k(P, ha, a, hb, b);
        H.kx[t] = k;
        ax[t] = (ha && hb && a.gas && b.gas) ? k : 0.0f;
    }
    if (t < w * (h + 1)) {
        const int x = t % w, y = t / w;
        const bool ha = y > 0, hb = y < h;
        const SideV a = ha ? side_at(H, (y - 1) * w + x) : SideV{}, b = hb ? side_at(H, y * w + x) : SideV{};
        const float k = face_k(P, ha, a, hb, b);
        H.ky[t] = k;
        ay[t] = (ha && hb && a.gas && b.gas) ? k : 0.0f;
    }
}

// the sum of k over a gas cell's vacuum and edge faces
__global__ void k_heat_kvac(HeatParams P, HeatPlanes H) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int w = P.w, h = P.h;
    if (i >= w * h) return;
    const int x = i % w, y = i / w;
    float s = 0.0f;
    if (H.gas[i]) {
        if (!(x > 0 && H.gas[i - 1])) s += H.kx[y * (w + 1) + x];
        if (!(x + 1 < w && H.gas[i + 1])) s += H.kx[y * (w + 1) + x + 1];
        if (!(y > 0 && H.gas[i - w])) s += H.ky[y * w + x];
        if (!(y + 1 < h && H.gas[i + w])) s += H.ky[(y + 1) * w + x];
    }
    H.kvac[i] = s;
}

// a Newton pass's system at the iterate t (heat.rs:474-488): own = c + k_vac, rhs = c u - (C/dt)(t - t0) + k_vac u_vac,
// x = u; the two stopping bounds eps^2 sum c u^2 and eps^2 sum (C/dt) t^2
__global__ void k_heat_system(HeatParams P, HeatPlanes H, const double *__restrict__ scal, float *__restrict__ own,
                              double *__restrict__ rhs, double *__restrict__ xs, double *part) {
    const int n = P.w * P.h;
    const double dt = scal[0];
    const double uvac = u_of(P, P.t_floor);
    double v[2] = {0.0, 0.0};
    const int stride = gridDim.x * NT;
    for (int base = blockIdx.x * NT + threadIdx.x; base < n; base += stride * CELLS) {
        // the loads of CELLS cells issued together, then the cells in order
        uint8_t g[CELLS];
        double tv[CELLS], t0v[CELLS];
        float cv[CELLS], kvv[CELLS];
#pragma unroll
        for (int k = 0; k < CELLS; k++) {
            const int i = base + k * stride;
            g[k] = i < n ? H.gas[i] : 0;
            if (i < n) {
                tv[k] = H.t[i];
                cv[k] = H.cap[i];
                kvv[k] = H.kvac[i];
                t0v[k] = H.t0[i];
            }
        }
#pragma unroll
        for (int k = 0; k < CELLS; k++) {
            const int i = base + k * stride;
            if (i >= n) break;
            if (!g[k]) {
                own[i] = 1.0f;
                rhs[i] = 0.0;
                xs[i] = 0.0;
                H.heat[i] = 0.0f;
                H.slope[i] = 0.0f;
                continue;
            }
            const double t = tv[k], heat = (double)cv[k] / dt, slope = dt_du(P, t), ci = heat * slope;
            const double u = u_of(P, t), kv = kvv[k];
            own[i] = (float)(ci + kv);
            rhs[i] = ci * u - heat * (t - t0v[k]) + kv * uvac;
            xs[i] = u;
            H.heat[i] = (float)heat;
            H.slope[i] = (float)slope;
            v[0] += ci * u * u;
            v[1] += heat * t * t;
        }
    }
    block_partials<2>(v, part, gridDim.x);
}

// the pass's update (heat.rs:502-515): t = clamp(T(max(x, u_vac)), t/2, 2t); sums C |dT| and counts the cells that
// moved more than newton_cell of their T
__global__ void k_heat_update(HeatParams P, HeatPlanes H, const double *__restrict__ xs, double newton_cell,
                              double *part) {
    const int n = P.w * P.h;
    const double uvac = u_of(P, P.t_floor);
    double v[2] = {0.0, 0.0};
    const int stride = gridDim.x * NT;
    for (int base = blockIdx.x * NT + threadIdx.x; base < n; base += stride * CELLS) {
        uint8_t g[CELLS];
        double tv[CELLS], xv[CELLS];
        float cv[CELLS];
#pragma unroll
        for (int k = 0; k < CELLS; k++) {
            const int i = base + k * stride;
            g[k] = i < n ? H.gas[i] : 0;
            if (i < n) {
                tv[k] = H.t[i];
                xv[k] = xs[i];
                cv[k] = H.cap[i];
            }
        }
#pragma unroll
        for (int k = 0; k < CELLS; k++) {
            const int i = base + k * stride;
            if (i >= n) break;
            if (!g[k]) continue;
            const double t = tv[k];
            double nx = t_of(P, fmax(xv[k], uvac));
            nx = fmin(fmax(nx, 0.5 * t), 2.0 * t);
            const double d = fabs(nx - t);
            v[0] += (double)cv[k] * d;
            if (!(d <= newton_cell * t)) v[1] += 1.0;
            H.t[i] = nx;
        }
    }
    block_partials<2>(v, part, gridDim.x);
}

// the commit's sums (heat.rs:522-530): radiated = sum dt k_vac (u - u_vac), gained = sum C (t - t0), held = sum C t;
// then delta = -(gained + radiated) / held
__global__ void k_heat_commit_sums(HeatParams P, HeatPlanes H, const double *__restrict__ scal, double *part) {
    const int n = P.w * P.h;
    const double dt = scal[0], uvac = u_of(P, P.t_floor);
    double v[3] = {0.0, 0.0, 0.0};
    for (int i = blockIdx.x * NT + threadIdx.x; i < n; i += gridDim.x * NT) {
        if (!H.gas[i]) continue;
        const double t = H.t[i], c = H.cap[i];
        v[0] += dt * (double)H.kvac[i] * (u_of(P, t) - uvac);
        v[1] += c * (t - (double)H.t0[i]);
        v[2] += c * t;
    }
    block_partials<3>(v, part, gridDim.x);
}

// F = -k dx (v_b - v_a) at a face, v the iterate's u on gas cells, u_vac elsewhere (heat.rs:340-362)
__device__ inline double face_flux(const HeatParams &P, const HeatPlanes &H, float k, int a, int b) {
    if (k == 0.0f) return 0.0;
    const double uvac = u_of(P, P.t_floor);
    const double va = (a >= 0 && H.gas[a]) ? u_of(P, H.t[a]) : uvac;
    const double vb = (b >= 0 && H.gas[b]) ? u_of(P, H.t[b]) : uvac;
    return -(double)k * P.dx * (vb - va);
}

// E += C ((1 + delta) t - t0) on the gas, and the cell-centred F the next P4 reads (heat.rs:89-100, :531-534); with
// `limited`, the ceiling's commit instead: every gas cell's outflow scaled to the heat it holds above its floor
// (heat.rs:544-592), applied by k_heat_limited.
__global__ void k_heat_commit(HeatParams P, HeatPlanes H, double delta, float *__restrict__ e, float *__restrict__ fx,
                              float *__restrict__ fy, int pinned) {
    const int w = P.w, h = P.h;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= w * h) return;
    const int x = i % w, y = i / w;
    const double fl = face_flux(P, H, H.kx[y * (w + 1) + x], x > 0 ? i - 1 : -1, i);
    const double fr = face_flux(P, H, H.kx[y * (w + 1) + x + 1], i, x + 1 < w ? i + 1 : -1);
    const double fu = face_flux(P, H, H.ky[y * w + x], y > 0 ? i - w : -1, i);
    const double fd = face_flux(P, H, H.ky[(y + 1) * w + x], i, y + 1 < h ? i + w : -1);
    fx[i] = (float)(0.5 * (fl + fr));
    fy[i] = (float)(0.5 * (fu + fd));
    if (pinned || !H.gas[i]) return;
    e[i] = (float)((double)e[i] + (double)H.cap[i] * ((1.0 + delta) * (double)H.t[i] - (double)H.t0[i]));
}

// the limited commit: scale = min(1, C (t0 - T_floor) / outflow) per gas cell (1 elsewhere)
__global__ void k_heat_scale(HeatParams P, HeatPlanes H, const double *__restrict__ scal) {
    const int w = P.w, h = P.h;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= w * h) return;
    if (!H.gas[i]) {
        H.scale[i] = 1.0f;
        return;
    }
    const int x = i % w, y = i / w;
    const double dt = scal[0];
    const double o0 = -face_flux(P, H, H.kx[y * (w + 1) + x], x > 0 ? i - 1 : -1, i);
    const double o1 = face_flux(P, H, H.kx[y * (w + 1) + x + 1], i, x + 1 < w ? i + 1 : -1);
    const double o2 = -face_flux(P, H, H.ky[y * w + x], y > 0 ? i - w : -1, i);
    const double o3 = face_flux(P, H, H.ky[(y + 1) * w + x], i, y + 1 < h ? i + w : -1);
    const double out = dt * P.inv_dx * (fmax(o0, 0.0) + fmax(o1, 0.0) + fmax(o2, 0.0) + fmax(o3, 0.0));
    const double held = (double)H.cap[i] * fmax((double)H.t0[i] - P.t_floor, 0.0);
    H.scale[i] = (float)(out > held ? held / out : 1.0);
}

__device__ inline double limited_flux(const HeatParams &P, const HeatPlanes &H, float k, int a, int b) {
    double f = face_flux(P, H, k, a, b);
    const int from = f > 0.0 ? a : b;
    return f * (from >= 0 ? (double)H.scale[from] : 1.0);
}

__global__ void k_heat_limited(HeatParams P, HeatPlanes H, const double *__restrict__ scal, float *__restrict__ e,
                               float *__restrict__ fx, float *__restrict__ fy, double *part) {
    const int w = P.w, h = P.h, n = w * h;
    const double dt = scal[0];
    double v[1] = {0.0};
    for (int i = blockIdx.x * NT + threadIdx.x; i < n; i += gridDim.x * NT) {
        const int x = i % w, y = i / w;
        const double fl = limited_flux(P, H, H.kx[y * (w + 1) + x], x > 0 ? i - 1 : -1, i);
        const double fr = limited_flux(P, H, H.kx[y * (w + 1) + x + 1], i, x + 1 < w ? i + 1 : -1);
        const double fu = limited_flux(P, H, H.ky[y * w + x], y > 0 ? i - w : -1, i);
        const double fd = limited_flux(P, H, H.ky[(y + 1) * w + x], i, y + 1 < h ? i + w : -1);
        fx[i] = (float)(0.5 * (fl + fr));
        fy[i] = (float)(0.5 * (fu + fd));
        if (!H.gas[i]) continue;
        const double o[4] = {-fl, fr, -fu, fd};
        const bool vac[4] = {!(x > 0 && H.gas[i - 1]), !(x + 1 < w && H.gas[i + 1]), !(y > 0 && H.gas[i - w]),
                             !(y + 1 < h && H.gas[i + w])};
        double em = 0.0;
        for (int f = 0; f < 4; f++)
            if (vac[f]) em += dt * P.inv_dx * o[f];
        v[0] += em;
        e[i] = (float)((double)e[i] - dt * P.inv_dx * (o[0] + o[1] + o[2] + o[3]));
    }
    block_partials<1>(v, part, gridDim.x);
}

// ------------------------------------------------------------------------------------------------------------ host

static int grid_cells(size_t n) { return (int)std::min<size_t>(4096, (n + NT * 4 - 1) / (NT * 4)); }

HeatPass::~HeatPass() {
    void *ps[] = {H.gas, H.t0, H.t, H.kappa, H.imu, H.cap, H.sig, H.kx, H.ky, H.kvac, H.heat, H.slope, H.scale,
                  rhs, xs, part};
    for (void *p : ps)
        if (p) (void)hipFree(p);
    if (host) (void)hipHostFree(host);
    for (hipEvent_t e : ev_sums)
        if (e) (void)hipEventDestroy(e);
}

int HeatPass::init(const HeatParams &p, const sr_heat_settings &hs_) {
    P = p;
    hs = hs_;
    const size_t n = (size_t)P.w * P.h, nx = (size_t)(P.w + 1) * P.h, ny = (size_t)P.w * (P.h + 1);
    nb = grid_cells(n);
    SR_HIP(hipMalloc(&H.gas, n));
    float **fp[] = {&H.kappa, &H.imu, &H.cap, &H.sig, &H.kvac, &H.heat, &H.slope, &H.scale};
    for (float **q : fp) SR_HIP(hipMalloc(q, sizeof(float) * n));
    SR_HIP(hipMalloc(&H.t0, sizeof(double) * n));
    SR_HIP(hipMalloc(&H.t, sizeof(double) * n));
    SR_HIP(hipMalloc(&H.kx, sizeof(float) * nx));
    SR_HIP(hipMalloc(&H.ky, sizeof(float) * ny));
    SR_HIP(hipMalloc(&rhs, sizeof(double) * n));
    SR_HIP(hipMalloc(&xs, sizeof(double) * n));
    SR_HIP(hipMalloc(&part, sizeof(double) * 4 * nb));
    SR_HIP(hipHostMalloc(&host, sizeof(double) * 2 * 4 * nb));
    for (hipEvent_t &e : ev_sums) SR_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    return solver.init(P.w, P.h, hs.levels, hs.nu1, hs.nu2, hs.coarse_scale);
}

int HeatPass::copy_sums(int k, hipStream_t s, int slot) {
    SR_HIP(hipMemcpyAsync(host + (size_t)slot * 4 * nb, part, sizeof(double) * k * nb, hipMemcpyDeviceToHost, s));
    SR_HIP(hipEventRecord(ev_sums[slot], s));
    return SR_OK;
}

int HeatPass::wait_sums(int k, double *out, int slot) {
    SR_HIP(hipEventSynchronize(ev_sums[slot]));
    const double *h = host + (size_t)slot * 4 * nb;
    for (int j = 0; j < k; j++) {
        double t = 0.0;
        for (int i = 0; i < nb; i++) t += h[j * nb + i];
        out[j] = t;
    }
    return SR_OK;
}

// the k sums of the last kernel's partials, in block order
int HeatPass::read_sums(int k, double *out, hipStream_t s) {
    SR_HIP(hipMemcpyAsync(host, part, sizeof(double) * k * nb, hipMemcpyDeviceToHost, s));
    SR_HIP(hipStreamSynchronize(s));
    for (int j = 0; j < k; j++) {
        double t = 0.0;
        for (int i = 0; i < nb; i++) t += host[j * nb + i];
        out[j] = t;
    }
    return SR_OK;
}

// Per cell: the cells pass (the state in: Sigma f64 + 13 f32 planes; gas, T0 and T f64, kappa, 1/mu, C, Sigma out:
// 60 + 1 + 24 + 16 = 101), the face law (both cells' T, Sigma, kappa in per face pair, kx/ky and the solver's ax/ay
// out: 16 + 16 = 32) and the vacuum sums (gas, kx, ky in, kvac out: 13); each Newton pass's system (T, T0 f64, C,
// kvac, gas in; own, heat, slope f32 and rhs, x f64 out: 25 + 28 = 53) and update (x, T f64, C, gas in; T out: 29);
// the commit (its sums: T, T0 f64, C, kvac, x, gas: 33; the commit: kx, ky, T, gas, E in; E, F_x, F_y out: 29 + 12)
// or the limited commit (the scale: 29; the limited fluxes and E: 45).
double HeatPass::work_bytes(const HeatWork &w) const {
    const double n = (double)P.w * P.h;
    return n * ((101.0 + 32.0 + 13.0) * w.steps + (53.0 + 29.0) * w.passes + (33.0 + 41.0) * w.commits +
                (29.0 + 45.0) * w.limited);
}

int HeatPass::first_system(const double *scal, hipStream_t s) {
    const size_t n = (size_t)P.w * P.h;
    SR_HIP(hipMemcpyAsync(H.t, H.t0, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
    hipLaunchKernelGGL(k_heat_system, dim3(nb), dim3(NT), 0, s, P, H, scal, solver.own0(), rhs, xs, part);
    SR_HIP(hipGetLastError());
    SR_HIP(hipMemcpyAsync(solver.gas0(), H.gas, n, hipMemcpyDeviceToDevice, s));
    return solver.rebuild(true, s);
}

int HeatPass::step(const HydroLike &st, const double *scal, double dt_host, float *fx, float *fy, int precond,
                   hipStream_t s) {
    const size_t n = (size_t)P.w * P.h;
    const int g = nb;
    report = sr_heat_report{};
    report.dt = dt_host;
    report.converged = 1;
    hipLaunchKernelGGL(k_heat_cells, dim3(g), dim3(NT), 0, s, P, st.sigma, st.mx, st.my, st.e, st.x, H, part);
    work.steps++;
    double heat0 = 0.0;
    // Lookahead (relative stopping): the face law, the vacuum sums, the first pass's system and the hierarchy's
    // rebuild are queued behind the heat sum's copy before the host waits for it; none of them writes what the commit
    // of a step that does not solve reads (the system, its right-hand side and start, the solver's coefficients).
    const bool ahead = hs.stop == SR_STOP_RELATIVE && !P.pinned && dt_host > 0.0;
    SR_TRY(copy_sums(1, s));
    const int nf = std::max((P.w + 1) * P.h, P.w * (P.h + 1));
    hipLaunchKernelGGL(k_heat_law, dim3((nf + 255) / 256), dim3(256), 0, s, P, H, solver.ax0(), solver.ay0());
    hipLaunchKernelGGL(k_heat_kvac, dim3((n + 255) / 256), dim3(256), 0, s, P, H);
    SR_HIP(hipGetLastError());
    bool built = false;  // the next pass's system built and the hierarchy rebuilt already
    if (ahead) {
        SR_HIP(hipMemcpyAsync(solver.gas0(), H.gas, n, hipMemcpyDeviceToDevice, s));
        hipLaunchKernelGGL(k_heat_system, dim3(g), dim3(NT), 0, s, P, H, scal, solver.own0(), rhs, xs, part);
        SR_TRY(solver.rebuild(true, s));
        built = true;
    }
    SR_TRY(wait_sums(1, &heat0));
    if (P.pinned || !(dt_host > 0.0) || !(heat0 > 0.0)) {
        hipLaunchKernelGGL(k_heat_commit, dim3((n + 255) / 256), dim3(256), 0, s, P, H, 0.0, st.e, fx, fy, 1);
        SR_HIP(hipGetLastError());
        return SR_OK;
    }
    if (!built) SR_HIP(hipMemcpyAsync(solver.gas0(), H.gas, n, hipMemcpyDeviceToDevice, s));
    const double eps = hs.stop == SR_STOP_REFERENCE ? hs.solve_tol : 0.0;
    int passes = 0, rounds = 0;
    bool converged = false, faces = true, commit_ahead = false;
    std::vector<int> step_rounds;
    double last_rel = 0.0;
    // inexact Newton (relative stopping only): loose early passes, the last one to solve_tol
    const bool inexact = hs.forcing > 0.0 && hs.stop == SR_STOP_RELATIVE;
    double eta = inexact ? std::max(hs.solve_tol, 1e-2) : hs.solve_tol;
    while (passes < hs.max_passes && rounds < hs.max_rounds) {
        passes++;
        work.passes++;
        if (!built)
            hipLaunchKernelGGL(k_heat_system, dim3(g), dim3(NT), 0, s, P, H, scal, solver.own0(), rhs, xs, part);
        // the reference's two bounds are needed (and read back) only under its stopping rule
        RefStop ref;
        if (hs.stop == SR_STOP_REFERENCE) {
            double b2[2];
            SR_TRY(read_sums(2, b2, s));
            ref.heat = H.heat;
            ref.slope = H.slope;
            ref.u_bound = eps * eps * b2[0];
            ref.t_bound = eps * eps * b2[1];
            if (!(std::isfinite(ref.u_bound) && std::isfinite(ref.t_bound))) {
                passes--;
                break;
            }
        }
        if (!built) SR_TRY(solver.rebuild(faces, s));
        built = false;
        faces = false;
        FldStats fs;
        const int hint = passes <= (int)pass_rounds.size() ? pass_rounds[passes - 1] : 0;
        SR_TRY(solver.solve(rhs, xs, eta, hs.max_rounds - rounds, precond,
                            hs.stop == SR_STOP_REFERENCE ? &ref : nullptr, &fs, s, hint));
        if ((int)step_rounds.size() < passes) step_rounds.resize(passes);
        step_rounds[passes - 1] = fs.iterations;
        rounds += fs.iterations;
        last_rel = fs.rel;
        if (!fs.finite) break;
        hipLaunchKernelGGL(k_heat_update, dim3(g), dim3(NT), 0, s, P, H, xs, hs.newton_cell, part);
        double mv[2];
        // Lookahead (relative stopping): after a loose pass, which the tight one always follows, the next pass's
        // system and rebuild are queued behind the sums' copy before the host waits; after a tight pass, which mostly
        // ends the loop, they are not (the commit would wait behind them)
        const bool next = hs.stop == SR_STOP_RELATIVE && eta > hs.solve_tol && passes < hs.max_passes;
        SR_TRY(copy_sums(2, s));
        // after an inexact solve's tight pass, which mostly ends the loop, the commit's sums instead (they read the
        // iterate alone; unused when the loop goes on)
        commit_ahead = inexact && !next;
        if (next) {
            hipLaunchKernelGGL(k_heat_system, dim3(g), dim3(NT), 0, s, P, H, scal, solver.own0(), rhs, xs, part);
            SR_TRY(solver.rebuild(false, s));
            built = true;
        } else if (commit_ahead) {
            hipLaunchKernelGGL(k_heat_commit_sums, dim3(g), dim3(NT), 0, s, P, H, scal, part);
            SR_TRY(copy_sums(3, s, 1));
        }
        SR_TRY(wait_sums(2, mv));
        if (mv[0] <= hs.newton_heat * heat0 && mv[1] == 0.0) {
            if (eta > hs.solve_tol) {  // converged on a loose solve: one more pass, solved to solve_tol
                eta = hs.solve_tol;
                continue;
            }
            converged = true;
            break;
        }
        if (inexact) eta = std::min(1e-2, std::max(hs.solve_tol, hs.forcing * mv[0] / heat0));
    }
    report.passes = passes;
    report.rounds = rounds;
    pass_rounds = step_rounds;
    report.final_residual = last_rel;
    if (converged) {
        work.commits++;
        double c3[3];
        if (commit_ahead) {
            SR_TRY(wait_sums(3, c3, 1));
        } else {
            hipLaunchKernelGGL(k_heat_commit_sums, dim3(g), dim3(NT), 0, s, P, H, scal, part);
            SR_TRY(read_sums(3, c3, s));
        }
        const double delta = c3[2] != 0.0 ? -(c3[1] + c3[0]) / c3[2] : 0.0;
        hipLaunchKernelGGL(k_heat_commit, dim3((n + 255) / 256), dim3(256), 0, s, P, H, delta, st.e, fx, fy, 0);
        SR_HIP(hipGetLastError());
        report.delta = delta;
        report.radiated = c3[0];
        report.converged = 1;
    } else {
        hipLaunchKernelGGL(k_heat_scale, dim3((n + 255) / 256), dim3(256), 0, s, P, H, scal);
        work.limited++;
        hipLaunchKernelGGL(k_heat_limited, dim3(g), dim3(NT), 0, s, P, H, scal, st.e, fx, fy, part);
        SR_HIP(hipGetLastError());
        double r1;
        SR_TRY(read_sums(1, &r1, s));
        report.delta = 0.0;
        report.radiated = r1;
        report.converged = 0;
    }
    return SR_OK;
}

} // namespace sr
