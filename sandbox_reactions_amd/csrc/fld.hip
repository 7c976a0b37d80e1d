// P5 linear solver for the implicit flux-limited-diffusion step: FP32 multigrid-preconditioned flexible CG inside
// FP64 iterative refinement.
//
// The operator is the reference's pass system (crates/sr-physics/src/reference/heat.rs:374-392, :474-488) in face
// form; the hierarchy is the reference GPU's Galerkin 2x2 aggregation (crates/sr-engine/shaders/heat.wgsl:1316-1400)
// and its V-cycle smooths by red-black Gauss-Seidel, red first down, black first up (heat.wgsl:1401-1445). The
// machinery is this library's own:
//   * outer loop in FP64: r = b - A x with FP64 x and b, the stopping rule on that true residual;
//   * inner loop in FP32: flexible (Polak-Ribiere) PCG on the residual, FP64 dot products, x updated in FP64;
//   * the pre-smoother of a level, its residual and its restriction are one kernel: a 32x32 tile with a halo of
//     2 nu1 + 1 cells in LDS, each half sweep recomputed redundantly on the shrinking halo, so the result is exact
//     red-black Gauss-Seidel with one pass over HBM; the prolongation and the post-smoother are another such kernel;
//   * every level of at most 1024 cells is solved inside one 1024-thread block (LDS-resident V-cycle);
//   * dot products reduce by wave64 shuffles, per-block partials summed in a fixed order by the last block to finish
//     (deterministic), which also forms alpha and beta on the device.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "common.hpp"
#include "fld.hpp"

namespace sr {

#define TILE 32
#define NT 256
#define COARSE_SWEEPS 16
#define COARSE_MAX 1024

enum {
    SC_RZ = 0, SC_PQ, SC_ALPHA, SC_BETA, SC_RR, SC_RZN, SC_RZO, SC_BB, SC_R64, SC_UMEAS, SC_HMEAS, SC_RESTART,
    SC_N = 16
};
enum { CNT_TRUE = 0, CNT_PQ, CNT_RR, CNT_DOTS, CNT_N };

// ------------------------------------------------------------------------------------------------ elementwise kernels

struct L0 {
    int w, h;
    const float *own, *ax, *ay;
    const uint8_t *gas;
};

__device__ __forceinline__ bool is_gas(const L0 &a, size_t i) { return a.gas == nullptr || a.gas[i]; }

// r = b - A x in FP64 (x zeroed off the gas), r32 = (float) r; sums ||r||^2, ||b||^2 and the reference's two measures
__global__ void __launch_bounds__(NT) k_true_residual(L0 a, const double *__restrict__ b, double *__restrict__ x,
                                                      float *__restrict__ r32,
                                                      const float *__restrict__ heat, const float *__restrict__ slope,
                                                      double *sc, double *part, unsigned *counter) {
    const int n = a.w * a.h;
    double v[4] = {0.0, 0.0, 0.0, 0.0};
    for (int i = blockIdx.x * NT + threadIdx.x; i < n; i += gridDim.x * NT) {
        const int xx = i % a.w, yy = i / a.w;
        double r = 0.0;
        if (is_gas(a, i)) {
            const double xi = x[i];
            const double al = a.ax[(size_t)yy * (a.w + 1) + xx], ar = a.ax[(size_t)yy * (a.w + 1) + xx + 1];
            const double au = a.ay[(size_t)yy * a.w + xx], ad = a.ay[(size_t)(yy + 1) * a.w + xx];
            double ax = (double)a.own[i] * xi;
            if (al != 0.0) ax += al * (xi - x[i - 1]);
            if (ar != 0.0) ax += ar * (xi - x[i + 1]);
            if (au != 0.0) ax += au * (xi - x[i - a.w]);
            if (ad != 0.0) ax += ad * (xi - x[i + a.w]);
            r = b[i] - ax;
            v[1] += b[i] * b[i];
            if (heat) {
                const double d = (double)a.own[i] + al + ar + au + ad;
                v[2] += r * r / d;
                const double e = r * (double)slope[i] / d;
                v[3] += (double)heat[i] * e * e;
            }
        } else {
            x[i] = 0.0;
        }
        r32[i] = (float)r;
        v[0] += r * r;
    }
    double out[4];
    if (reduce_finish<4>(v, part, gridDim.x, counter + CNT_TRUE, out) && threadIdx.x == 0) {
        sc[SC_R64] = out[0];
        sc[SC_BB] = out[1];
        sc[SC_UMEAS] = out[2];
        sc[SC_HMEAS] = out[3];
        sc[SC_RESTART] = 1.0;
        sc[SC_BETA] = 0.0;
    }
}

// p = z + beta p_old (into p_new), q = A p_new in FP32; sums p.q and forms alpha = rz / pq
__global__ void __launch_bounds__(NT) k_matvec_p(L0 a, const float *__restrict__ z, const float *__restrict__ pold,
                                                 float *__restrict__ pnew, float *__restrict__ q, double *sc,
                                                 double *part, unsigned *counter) {
    const int n = a.w * a.h;
    const float beta = (float)sc[SC_BETA];
    const bool fresh = sc[SC_RESTART] != 0.0;
    auto P = [&](int j) { return fresh ? z[j] : z[j] + beta * pold[j]; };
    double v[1] = {0.0};
    for (int i = blockIdx.x * NT + threadIdx.x; i < n; i += gridDim.x * NT) {
        const int xx = i % a.w, yy = i / a.w;
        const float pi = P(i);
        const float al = a.ax[(size_t)yy * (a.w + 1) + xx], ar = a.ax[(size_t)yy * (a.w + 1) + xx + 1];
        const float au = a.ay[(size_t)yy * a.w + xx], ad = a.ay[(size_t)(yy + 1) * a.w + xx];
        float s = a.own[i] * pi;
        if (al != 0.0f) s += al * (pi - P(i - 1));
        if (ar != 0.0f) s += ar * (pi - P(i + 1));
        if (au != 0.0f) s += au * (pi - P(i - a.w));
        if (ad != 0.0f) s += ad * (pi - P(i + a.w));
        pnew[i] = pi;
        q[i] = s;
        v[0] += (double)pi * (double)s;
    }
    double out[1];
    if (reduce_finish<1>(v, part, gridDim.x, counter + CNT_PQ, out) && threadIdx.x == 0) {
        sc[SC_PQ] = out[0];
        sc[SC_ALPHA] = out[0] != 0.0 ? sc[SC_RZ] / out[0] : 0.0;
        sc[SC_RESTART] = 0.0;
    }
}

// x += alpha p (FP64), r -= alpha q; sums ||r||^2
__global__ void __launch_bounds__(NT) k_update(int n, const float *__restrict__ p, const float *__restrict__ q,
                                               double *__restrict__ x, float *__restrict__ r, double *sc,
                                               double *part, unsigned *counter) {
    const double alpha = sc[SC_ALPHA];
    const float af = (float)alpha;
    double v[1] = {0.0};
    for (int i = blockIdx.x * NT + threadIdx.x; i < n; i += gridDim.x * NT) {
        x[i] += alpha * (double)p[i];
        const float ri = r[i] - af * q[i];
        r[i] = ri;
        v[0] += (double)ri * ri;
    }
    double out[1];
    if (reduce_finish<1>(v, part, gridDim.x, counter + CNT_RR, out) && threadIdx.x == 0) sc[SC_RR] = out[0];
}

// the flexible beta from r.z_new and r.z_old (sums made by the preconditioner's last kernel)
__device__ inline void finish_beta(double *sc, double rzn, double rzo) {
    const bool fresh = sc[SC_RESTART] != 0.0;
    sc[SC_RZN] = rzn;
    sc[SC_RZO] = rzo;
    sc[SC_BETA] = fresh || sc[SC_RZ] == 0.0 ? 0.0 : (rzn - rzo) / sc[SC_RZ];
    sc[SC_RZ] = rzn;
}

// Jacobi: z = r / D, D = own + sum of couplings; sums r.z and r.z_old
__global__ void __launch_bounds__(NT) k_jacobi(L0 a, const float *__restrict__ r, float *__restrict__ z,
                                               const float *__restrict__ zold, double *sc, double *part,
                                               unsigned *counter, int dots) {
    const int n = a.w * a.h;
    double v[2] = {0.0, 0.0};
    for (int i = blockIdx.x * NT + threadIdx.x; i < n; i += gridDim.x * NT) {
        const int xx = i % a.w, yy = i / a.w;
        const float d = a.own[i] + a.ax[(size_t)yy * (a.w + 1) + xx] + a.ax[(size_t)yy * (a.w + 1) + xx + 1] +
                        a.ay[(size_t)yy * a.w + xx] + a.ay[(size_t)(yy + 1) * a.w + xx];
        const float zi = is_gas(a, i) ? r[i] / d : 0.0f;
        z[i] = zi;
        v[0] += (double)r[i] * zi;
        if (zold) v[1] += (double)r[i] * zold[i];
    }
    if (!dots) return;
    double out[2];
    if (reduce_finish<2>(v, part, gridDim.x, counter + CNT_DOTS, out) && threadIdx.x == 0)
        finish_beta(sc, out[0], out[1]);
}

// r.z and r.z_old after a preconditioner that made no sums itself; forms beta
__global__ void __launch_bounds__(NT) k_dots(int n, const float *__restrict__ r, const float *__restrict__ z,
                                             const float *__restrict__ zold, double *sc, double *part,
                                             unsigned *counter) {
    double v[2] = {0.0, 0.0};
    for (int i = blockIdx.x * NT + threadIdx.x; i < n; i += gridDim.x * NT) {
        v[0] += (double)r[i] * z[i];
        if (zold) v[1] += (double)r[i] * zold[i];
    }
    double out[2];
    if (reduce_finish<2>(v, part, gridDim.x, counter + CNT_DOTS, out) && threadIdx.x == 0)
        finish_beta(sc, out[0], out[1]);
}

// ------------------------------------------------------------------------------------------------ tile smoothers

struct Lv {
    int w, h;
    const float *own, *ax, *ay;
};

// Region of a tile plus halo `hw`: LDS arrays of (T + 2 hw + 1)^2 entries (one extra column/row for the right and
// bottom faces).
template <int HW> struct Region {
    static constexpr int RW = TILE + 2 * HW;
    static constexpr int RS = RW + 1;
    float z[RS * RS], b[RS * RS], own[RS * RS], al[RS * RS], au[RS * RS];
};

template <int HW>
__device__ inline void load_region(Region<HW> &R, const Lv &L, int x0, int y0, const float *__restrict__ b) {
    constexpr int RS = Region<HW>::RS;
    for (int t = threadIdx.x; t < RS * RS; t += NT) {
        const int lx = t % RS, ly = t / RS;
        const int gx = x0 + lx - HW, gy = y0 + ly - HW;
        const bool in = gx >= 0 && gx < L.w && gy >= 0 && gy < L.h;
        const size_t i = (size_t)gy * L.w + gx;
        R.own[t] = in ? L.own[i] : 1.0f;
        R.b[t] = in && b ? b[i] : 0.0f;
        // the face on the cell's left / top; for a column or row past the grid, the grid's edge face (0)
        R.al[t] = (gy >= 0 && gy < L.h && gx >= 0 && gx <= L.w) ? L.ax[(size_t)gy * (L.w + 1) + gx] : 0.0f;
        R.au[t] = (gx >= 0 && gx < L.w && gy >= 0 && gy <= L.h) ? L.ay[(size_t)gy * L.w + gx] : 0.0f;
    }
}

// one half sweep of `color` over the region cells at least `k` cells from the region's border
template <int HW>
__device__ inline void half_sweep(Region<HW> &R, const Lv &L, int x0, int y0, int color, int k) {
    constexpr int RW = Region<HW>::RW, RS = Region<HW>::RS;
    const int span = RW - 2 * k;
    for (int t = threadIdx.x; t < span * span; t += NT) {
        const int lx = k + t % span, ly = k + t / span;
        const int gx = x0 + lx - HW, gy = y0 + ly - HW;
        if (((gx + gy) & 1) != color) continue;
        if (gx < 0 || gx >= L.w || gy < 0 || gy >= L.h) continue;
        const int c = ly * RS + lx;
        const float al = R.al[c], ar = R.al[c + 1], au = R.au[c], ad = R.au[c + RS];
        float num = R.b[c], den = R.own[c];
        if (al != 0.0f) { num += al * R.z[c - 1]; den += al; }
        if (ar != 0.0f) { num += ar * R.z[c + 1]; den += ar; }
        if (au != 0.0f) { num += au * R.z[c - RS]; den += au; }
        if (ad != 0.0f) { num += ad * R.z[c + RS]; den += ad; }
        R.z[c] = num / den;
    }
    __syncthreads();
}

// Pre-smoothing of level L from z = 0 (nu1 sweeps, red first), then the residual on the tile and its 2x2
// restriction into the next level's b. Halo 2 nu1 + 1.
template <int NU>
__global__ void __launch_bounds__(NT) k_pre(Lv L, const float *__restrict__ b, float *__restrict__ z,
                                            float *__restrict__ bc, int wc) {
    constexpr int HW = 2 * NU + 1;
    constexpr int RS = Region<HW>::RS;
    __shared__ Region<HW> R;
    const int x0 = blockIdx.x * TILE, y0 = blockIdx.y * TILE;
    load_region<HW>(R, L, x0, y0, b);
    for (int t = threadIdx.x; t < RS * RS; t += NT) R.z[t] = 0.0f;
    __syncthreads();
    for (int s = 0; s < 2 * NU; s++) half_sweep<HW>(R, L, x0, y0, s & 1, s + 1);
    // the residual of the tile's cells, summed over each 2x2 aggregate (tile origin even: aggregates lie inside)
    for (int t = threadIdx.x; t < (TILE / 2) * (TILE / 2); t += NT) {
        const int cx = t % (TILE / 2), cy = t / (TILE / 2);
        const int X = x0 / 2 + cx, Y = y0 / 2 + cy;
        float sum = 0.0f;
        bool any = false;
        for (int dy = 0; dy < 2; dy++)
            for (int dx = 0; dx < 2; dx++) {
                const int lx = HW + 2 * cx + dx, ly = HW + 2 * cy + dy;
                const int gx = x0 + lx - HW, gy = y0 + ly - HW;
                if (gx >= L.w || gy >= L.h) continue;
                any = true;
                const int c = ly * RS + lx;
                const float zc = R.z[c];
                float az = R.own[c] * zc;
                const float al = R.al[c], ar = R.al[c + 1], au = R.au[c], ad = R.au[c + RS];
                if (al != 0.0f) az += al * (zc - R.z[c - 1]);
                if (ar != 0.0f) az += ar * (zc - R.z[c + 1]);
                if (au != 0.0f) az += au * (zc - R.z[c - RS]);
                if (ad != 0.0f) az += ad * (zc - R.z[c + RS]);
                sum += R.b[c] - az;
            }
        if (any) bc[(size_t)Y * wc + X] = sum;
    }
    for (int t = threadIdx.x; t < TILE * TILE; t += NT) {
        const int lx = HW + t % TILE, ly = HW + t / TILE;
        const int gx = x0 + lx - HW, gy = y0 + ly - HW;
        if (gx < L.w && gy < L.h) z[(size_t)gy * L.w + gx] = R.z[ly * RS + lx];
    }
}

// Prolongation of the coarse correction and post-smoothing of level L (nu2 sweeps, black first). Halo 2 nu2.
// With `dots` (level 0), sums r.z_new and r.z_old and forms beta.
template <int NU>
__global__ void __launch_bounds__(NT) k_post(Lv L, const float *__restrict__ b, const float *__restrict__ zin,
                                             const float *__restrict__ zc, int wc, float *__restrict__ zout,
                                             const float *__restrict__ zold, double *sc, double *part,
                                             unsigned *counter, int dots) {
    constexpr int HW = 2 * NU;
    constexpr int RS = Region<HW>::RS;
    __shared__ Region<HW> R;
    const int x0 = blockIdx.x * TILE, y0 = blockIdx.y * TILE;
    load_region<HW>(R, L, x0, y0, b);
    for (int t = threadIdx.x; t < RS * RS; t += NT) {
        const int lx = t % RS, ly = t / RS;
        const int gx = x0 + lx - HW, gy = y0 + ly - HW;
        const bool in = gx >= 0 && gx < L.w && gy >= 0 && gy < L.h;
        R.z[t] = in ? zin[(size_t)gy * L.w + gx] + zc[(size_t)(gy >> 1) * wc + (gx >> 1)] : 0.0f;
    }
    __syncthreads();
    for (int s = 0; s < 2 * NU; s++) half_sweep<HW>(R, L, x0, y0, 1 - (s & 1), s + 1);
    double v[2] = {0.0, 0.0};
    for (int t = threadIdx.x; t < TILE * TILE; t += NT) {
        const int lx = HW + t % TILE, ly = HW + t / TILE;
        const int gx = x0 + lx - HW, gy = y0 + ly - HW;
        if (gx < L.w && gy < L.h) {
            const size_t i = (size_t)gy * L.w + gx;
            const float zn = R.z[ly * RS + lx];
            zout[i] = zn;
            if (dots) {
                const float bi = R.b[ly * RS + lx];
                v[0] += (double)bi * zn;
                if (zold) v[1] += (double)bi * zold[i];
            }
        }
    }
    if (!dots) return;
    double out[2];
    if (reduce_finish<2>(v, part, gridDim.x * gridDim.y, counter + CNT_DOTS, out) && threadIdx.x == 0)
        finish_beta(sc, out[0], out[1]);
}

// The V-cycle over every level of at most COARSE_MAX cells in one block: levels [first, last], the first one's b
// from global memory, its z written back. The coarsest takes COARSE_SWEEPS forward and backward sweeps.
struct CoarseSet {
    int n;
    int w[16], h[16];
    const float *own[16], *ax[16], *ay[16];
};

__global__ void __launch_bounds__(COARSE_MAX) k_coarse(CoarseSet C, const float *__restrict__ b0, float *__restrict__ z0,
                                                       int nu1, int nu2, int level0_gasless) {
    extern __shared__ float sm[];
    // per level: own, al ((w+1) h), au (w (h+1)), z, b
    int off[16][5];
    int o = 0;
    for (int l = 0; l < C.n; l++) {
        const int w = C.w[l], h = C.h[l];
        off[l][0] = o; o += w * h;
        off[l][1] = o; o += (w + 1) * h;
        off[l][2] = o; o += w * (h + 1);
        off[l][3] = o; o += w * h;
        off[l][4] = o; o += w * h;
    }
    for (int l = 0; l < C.n; l++) {
        const int w = C.w[l], h = C.h[l];
        for (int t = threadIdx.x; t < w * h; t += blockDim.x) {
            sm[off[l][0] + t] = C.own[l][t];
            sm[off[l][3] + t] = 0.0f;
            sm[off[l][4] + t] = l == 0 ? b0[t] : 0.0f;
        }
        for (int t = threadIdx.x; t < (w + 1) * h; t += blockDim.x) sm[off[l][1] + t] = C.ax[l][t];
        for (int t = threadIdx.x; t < w * (h + 1); t += blockDim.x) sm[off[l][2] + t] = C.ay[l][t];
    }
    __syncthreads();
    auto sweep = [&](int l, int color) {
        const int w = C.w[l], h = C.h[l];
        float *own = sm + off[l][0], *al = sm + off[l][1], *au = sm + off[l][2], *z = sm + off[l][3],
              *b = sm + off[l][4];
        for (int t = threadIdx.x; t < w * h; t += blockDim.x) {
            const int x = t % w, y = t / w;
            if (((x + y) & 1) != color) continue;
            const float a0 = al[y * (w + 1) + x], a1 = al[y * (w + 1) + x + 1], a2 = au[y * w + x],
                        a3 = au[(y + 1) * w + x];
            float num = b[t], den = own[t];
            if (a0 != 0.0f) { num += a0 * z[t - 1]; den += a0; }
            if (a1 != 0.0f) { num += a1 * z[t + 1]; den += a1; }
            if (a2 != 0.0f) { num += a2 * z[t - w]; den += a2; }
            if (a3 != 0.0f) { num += a3 * z[t + w]; den += a3; }
            z[t] = num / den;
        }
        __syncthreads();
    };
    for (int l = 0; l < C.n - 1; l++) {
        for (int k = 0; k < nu1; k++) {
            sweep(l, 0);
            sweep(l, 1);
        }
        const int w = C.w[l], h = C.h[l], wc = C.w[l + 1], hc = C.h[l + 1];
        float *own = sm + off[l][0], *al = sm + off[l][1], *au = sm + off[l][2], *z = sm + off[l][3],
              *b = sm + off[l][4];
        float *bc = sm + off[l + 1][4];
        for (int t = threadIdx.x; t < wc * hc; t += blockDim.x) {
            const int X = t % wc, Y = t / wc;
            float s = 0.0f;
            for (int dy = 0; dy < 2; dy++)
                for (int dx = 0; dx < 2; dx++) {
                    const int x = 2 * X + dx, y = 2 * Y + dy;
                    if (x >= w || y >= h) continue;
                    const int i = y * w + x;
                    const float a0 = al[y * (w + 1) + x], a1 = al[y * (w + 1) + x + 1], a2 = au[y * w + x],
                                a3 = au[(y + 1) * w + x];
                    float az = own[i] * z[i];
                    if (a0 != 0.0f) az += a0 * (z[i] - z[i - 1]);
                    if (a1 != 0.0f) az += a1 * (z[i] - z[i + 1]);
                    if (a2 != 0.0f) az += a2 * (z[i] - z[i - w]);
                    if (a3 != 0.0f) az += a3 * (z[i] - z[i + w]);
                    s += b[i] - az;
                }
            bc[t] = s;
        }
        __syncthreads();
    }
    const int lc = C.n - 1;
    for (int k = 0; k < COARSE_SWEEPS; k++) {
        sweep(lc, 0);
        sweep(lc, 1);
    }
    for (int k = 0; k < COARSE_SWEEPS; k++) {
        sweep(lc, 1);
        sweep(lc, 0);
    }
    for (int l = C.n - 2; l >= 0; l--) {
        const int w = C.w[l], h = C.h[l], wc = C.w[l + 1];
        float *z = sm + off[l][3];
        const float *zc = sm + off[l + 1][3];
        for (int t = threadIdx.x; t < w * h; t += blockDim.x) {
            const int x = t % w, y = t / w;
            z[t] += zc[(y >> 1) * wc + (x >> 1)];
        }
        __syncthreads();
        for (int k = 0; k < nu2; k++) {
            sweep(l, 1);
            sweep(l, 0);
        }
    }
    for (int t = threadIdx.x; t < C.w[0] * C.h[0]; t += blockDim.x) z0[t] = sm[off[0][3] + t];
}

// ------------------------------------------------------------------------------------------------ hierarchy build

// coarse own = the sum over the aggregate (gas cells only at level 0), 1 where the aggregate holds none
__global__ void k_coarsen_own(Lv f, const uint8_t *__restrict__ gas, float *__restrict__ oc, int wc, int hc) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= wc * hc) return;
    const int X = t % wc, Y = t / wc;
    float s = 0.0f;
    for (int dy = 0; dy < 2; dy++)
        for (int dx = 0; dx < 2; dx++) {
            const int x = 2 * X + dx, y = 2 * Y + dy;
            if (x >= f.w || y >= f.h) continue;
            const size_t i = (size_t)y * f.w + x;
            if (gas && !gas[i]) continue;
            s += f.own[i];
        }
    oc[t] = s == 0.0f ? 1.0f : s;
}

// coarse couplings: the sum of the fine faces between two aggregates
__global__ void k_coarsen_faces(Lv f, float *__restrict__ axc, float *__restrict__ ayc, int wc, int hc) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < (wc + 1) * hc) {
        const int X = t % (wc + 1), Y = t / (wc + 1);
        float s = 0.0f;
        if (X > 0 && X < wc) {
            const int x = 2 * X;
            for (int dy = 0; dy < 2; dy++) {
                const int y = 2 * Y + dy;
                if (y < f.h && x < f.w) s += f.ax[(size_t)y * (f.w + 1) + x];
            }
        }
        axc[t] = s;
    }
    if (t < wc * (hc + 1)) {
        const int X = t % wc, Y = t / wc;
        float s = 0.0f;
        if (Y > 0 && Y < hc) {
            const int y = 2 * Y;
            for (int dx = 0; dx < 2; dx++) {
                const int x = 2 * X + dx;
                if (x < f.w && y < f.h) s += f.ay[(size_t)y * f.w + x];
            }
        }
        ayc[t] = s;
    }
}

// the grid's edge faces carry no coupling (a neighbour past the edge does not exist)
__global__ void k_zero_edges(int w, int h, float *__restrict__ ax, float *__restrict__ ay) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < h) {
        ax[(size_t)t * (w + 1)] = 0.0f;
        ax[(size_t)t * (w + 1) + w] = 0.0f;
    }
    if (t < w) {
        ay[t] = 0.0f;
        ay[(size_t)h * w + t] = 0.0f;
    }
}

// out = A v in FP64
__global__ void k_matvec64(L0 a, const double *__restrict__ v, double *__restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.w * a.h) return;
    if (!is_gas(a, i)) {
        out[i] = v[i];
        return;
    }
    const int xx = i % a.w, yy = i / a.w;
    const double vi = v[i];
    const double al = a.ax[(size_t)yy * (a.w + 1) + xx], ar = a.ax[(size_t)yy * (a.w + 1) + xx + 1];
    const double au = a.ay[(size_t)yy * a.w + xx], ad = a.ay[(size_t)(yy + 1) * a.w + xx];
    double s = (double)a.own[i] * vi;
    if (al != 0.0) s += al * (vi - v[i - 1]);
    if (ar != 0.0) s += ar * (vi - v[i + 1]);
    if (au != 0.0) s += au * (vi - v[i - a.w]);
    if (ad != 0.0) s += ad * (vi - v[i + a.w]);
    out[i] = s;
}

// ------------------------------------------------------------------------------------------------ host

static bool is_coarsest(int w, int h) { return w * h <= 16 || w == 1 || h == 1; }

FldSolver::~FldSolver() {
    for (auto &l : lv) {
        hipFree(l.own);
        hipFree(l.ax);
        hipFree(l.ay);
        hipFree(l.z);
        hipFree(l.b);
        hipFree(l.zp);
    }
    hipFree(gas);
    hipFree(r64);
    hipFree(r32);
    hipFree(za);
    hipFree(zb);
    hipFree(p);
    hipFree(q);
    hipFree(sc);
    hipFree(part);
    hipFree(counter);
    if (host_sc) hipHostFree(host_sc);
    if (ev0) hipEventDestroy(ev0);
    if (ev1) hipEventDestroy(ev1);
}

int FldSolver::init(int w, int h, int levels, int nu1_, int nu2_) {
    if (w <= 0 || h <= 0) return fail(SR_ERR_ARGUMENT, "an empty FLD grid");
    if (nu1_ < 1 || nu1_ > 3 || nu2_ < 1 || nu2_ > 3) return fail(SR_ERR_ARGUMENT, "nu1, nu2 must be in 1..3");
    w0 = w;
    h0 = h;
    nu1 = nu1_;
    nu2 = nu2_;
    int cw = w, ch = h;
    const int maxl = levels > 0 ? levels : 64;
    while (true) {
        FldLevel l;
        l.w = cw;
        l.h = ch;
        SR_HIP(hipMalloc(&l.own, sizeof(float) * cw * ch));
        SR_HIP(hipMalloc(&l.ax, sizeof(float) * (cw + 1) * ch));
        SR_HIP(hipMalloc(&l.ay, sizeof(float) * cw * (ch + 1)));
        SR_HIP(hipMalloc(&l.z, sizeof(float) * cw * ch));
        SR_HIP(hipMalloc(&l.b, sizeof(float) * cw * ch));
        SR_HIP(hipMalloc(&l.zp, sizeof(float) * cw * ch));
        lv.push_back(l);
        if ((int)lv.size() >= maxl || is_coarsest(cw, ch)) break;
        cw = (cw + 1) / 2;
        ch = (ch + 1) / 2;
    }
    // the first level handled by the single-block V-cycle: every level from there down fits COARSE_MAX cells and
    // the LDS budget
    coarse_from = (int)lv.size() - 1;
    while (coarse_from > 0) {
        const int c = coarse_from - 1;
        if (lv[c].w * lv[c].h > COARSE_MAX) break;
        size_t bytes = 0;
        for (int l = c; l < (int)lv.size(); l++)
            bytes += sizeof(float) * (3 * lv[l].w * lv[l].h + (lv[l].w + 1) * lv[l].h + lv[l].w * (lv[l].h + 1));
        if (bytes > 60000 || (int)lv.size() - c > 16) break;
        coarse_from = c;
    }
    const size_t n = (size_t)w * h;
    SR_HIP(hipMalloc(&gas, n));
    SR_HIP(hipMemset(gas, 1, n));
    SR_HIP(hipMalloc(&r64, sizeof(double) * n));
    SR_HIP(hipMalloc(&r32, sizeof(float) * n));
    SR_HIP(hipMalloc(&za, sizeof(float) * n));
    SR_HIP(hipMalloc(&zb, sizeof(float) * n));
    SR_HIP(hipMalloc(&p, sizeof(float) * 2 * n));
    SR_HIP(hipMalloc(&q, sizeof(float) * n));
    SR_HIP(hipMemset(za, 0, sizeof(float) * n));
    SR_HIP(hipMemset(p, 0, sizeof(float) * 2 * n));
    SR_HIP(hipMalloc(&sc, sizeof(double) * SC_N));
    SR_HIP(hipMemset(sc, 0, sizeof(double) * SC_N));
    const size_t nb = std::max<size_t>(4096, ((w + TILE - 1) / TILE) * ((h + TILE - 1) / TILE));
    SR_HIP(hipMalloc(&part, sizeof(double) * 4 * nb));
    SR_HIP(hipMalloc(&counter, sizeof(unsigned) * CNT_N));
    SR_HIP(hipMemset(counter, 0, sizeof(unsigned) * CNT_N));
    SR_HIP(hipHostMalloc(&host_sc, sizeof(double) * SC_N));
    SR_HIP(hipEventCreate(&ev0));
    SR_HIP(hipEventCreate(&ev1));
    return SR_OK;
}

int FldSolver::set_system(const float *own, const float *ax, const float *ay, const uint8_t *g, hipStream_t s) {
    const size_t n = (size_t)w0 * h0;
    SR_HIP(hipMemcpyAsync(lv[0].own, own, sizeof(float) * n, hipMemcpyDeviceToDevice, s));
    SR_HIP(hipMemcpyAsync(lv[0].ax, ax, sizeof(float) * (w0 + 1) * h0, hipMemcpyDeviceToDevice, s));
    SR_HIP(hipMemcpyAsync(lv[0].ay, ay, sizeof(float) * w0 * (h0 + 1), hipMemcpyDeviceToDevice, s));
    const int m = std::max(w0, h0);
    hipLaunchKernelGGL(k_zero_edges, dim3((m + 255) / 256), dim3(256), 0, s, w0, h0, lv[0].ax, lv[0].ay);
    if (g)
        SR_HIP(hipMemcpyAsync(gas, g, n, hipMemcpyDeviceToDevice, s));
    else
        SR_HIP(hipMemsetAsync(gas, 1, n, s));
    return rebuild(true, s);
}

int FldSolver::rebuild(bool faces, hipStream_t s) {
    for (size_t l = 0; l + 1 < lv.size(); l++) {
        const FldLevel &f = lv[l], &c = lv[l + 1];
        Lv fl = {f.w, f.h, f.own, f.ax, f.ay};
        const int nc = c.w * c.h;
        hipLaunchKernelGGL(k_coarsen_own, dim3((nc + 255) / 256), dim3(256), 0, s, fl, l == 0 ? gas : nullptr,
                           c.own, c.w, c.h);
        if (faces) {
            const int nf = std::max((c.w + 1) * c.h, c.w * (c.h + 1));
            hipLaunchKernelGGL(k_coarsen_faces, dim3((nf + 255) / 256), dim3(256), 0, s, fl, c.ax, c.ay, c.w, c.h);
        }
    }
    SR_HIP(hipGetLastError());
    return SR_OK;
}

static int grid1(size_t n) { return (int)std::min<size_t>(4096, (n + NT * 4 - 1) / (NT * 4)); }

// z = M^-1 r on level 0 (r, z level-0 vectors); with `dots`, the last kernel sums r.z and r.zold, forms beta
int FldSolver::vcycle(const float *r, float *z, const float *zold, bool dots, hipStream_t s) {
    const int nl = (int)lv.size();
    // device levels above the single-block part
    const float *bl = r;
    for (int l = 0; l < coarse_from; l++) {
        const FldLevel &f = lv[l], &c = lv[l + 1];
        Lv L = {f.w, f.h, f.own, f.ax, f.ay};
        dim3 grid((f.w + TILE - 1) / TILE, (f.h + TILE - 1) / TILE);
        const bool t0 = timing && l == 0;
        if (t0) SR_HIP(hipEventRecord(ev0, s));
        if (nu1 == 1)
            hipLaunchKernelGGL(k_pre<1>, grid, dim3(NT), 0, s, L, bl, f.z, c.b, c.w);
        else if (nu1 == 2)
            hipLaunchKernelGGL(k_pre<2>, grid, dim3(NT), 0, s, L, bl, f.z, c.b, c.w);
        else
            hipLaunchKernelGGL(k_pre<3>, grid, dim3(NT), 0, s, L, bl, f.z, c.b, c.w);
        if (t0) {
            SR_HIP(hipEventRecord(ev1, s));
            SR_HIP(hipEventSynchronize(ev1));
            float ms = 0.f;
            SR_HIP(hipEventElapsedTime(&ms, ev0, ev1));
            smooth_ms += ms;
            smooth_launches++;
        }
        bl = c.b;
    }
    {
        CoarseSet C;
        C.n = nl - coarse_from;
        size_t bytes = 0;
        for (int l = coarse_from; l < nl; l++) {
            const int k = l - coarse_from;
            C.w[k] = lv[l].w;
            C.h[k] = lv[l].h;
            C.own[k] = lv[l].own;
            C.ax[k] = lv[l].ax;
            C.ay[k] = lv[l].ay;
            bytes += sizeof(float) * (3 * lv[l].w * lv[l].h + (lv[l].w + 1) * lv[l].h + lv[l].w * (lv[l].h + 1));
        }
        float *zc = coarse_from == 0 ? z : lv[coarse_from].zp;
        hipLaunchKernelGGL(k_coarse, dim3(1), dim3(COARSE_MAX), bytes, s, C, bl, zc, nu1, nu2, 0);
        if (coarse_from == 0 && dots) {
            const int n = w0 * h0;
            hipLaunchKernelGGL(k_dots, dim3(grid1(n)), dim3(NT), 0, s, n, r, z, zold, sc, part, counter);
        }
    }
    for (int l = coarse_from - 1; l >= 0; l--) {
        const FldLevel &f = lv[l], &c = lv[l + 1];
        Lv L = {f.w, f.h, f.own, f.ax, f.ay};
        dim3 grid((f.w + TILE - 1) / TILE, (f.h + TILE - 1) / TILE);
        const float *b = l == 0 ? r : f.b;
        float *zo = l == 0 ? z : f.zp;
        const int d = (l == 0 && dots) ? 1 : 0;
        if (nu2 == 1)
            hipLaunchKernelGGL(k_post<1>, grid, dim3(NT), 0, s, L, b, f.z, c.zp, c.w, zo, zold, sc, part, counter, d);
        else if (nu2 == 2)
            hipLaunchKernelGGL(k_post<2>, grid, dim3(NT), 0, s, L, b, f.z, c.zp, c.w, zo, zold, sc, part, counter, d);
        else
            hipLaunchKernelGGL(k_post<3>, grid, dim3(NT), 0, s, L, b, f.z, c.zp, c.w, zo, zold, sc, part, counter, d);
    }
    SR_HIP(hipGetLastError());
    return SR_OK;
}

int FldSolver::apply_precond(const float *r, float *z, int precond, hipStream_t s) {
    L0 a = {w0, h0, lv[0].own, lv[0].ax, lv[0].ay, gas};
    const size_t n = (size_t)w0 * h0;
    if (precond == 0) {
        hipLaunchKernelGGL(k_jacobi, dim3(grid1(n)), dim3(NT), 0, s, a, r, z, (const float *)nullptr, sc, part,
                           counter, 0);
        SR_HIP(hipGetLastError());
        return SR_OK;
    }
    return vcycle(r, z, nullptr, false, s);
}

int FldSolver::matvec64(const double *v, double *out, hipStream_t s) {
    L0 a = {w0, h0, lv[0].own, lv[0].ax, lv[0].ay, gas};
    const int n = w0 * h0;
    hipLaunchKernelGGL(k_matvec64, dim3((n + 255) / 256), dim3(256), 0, s, a, v, out);
    SR_HIP(hipGetLastError());
    return SR_OK;
}

int FldSolver::solve(const double *b, double *x, double tol, int max_iter, int precond, const RefStop *ref,
                     FldStats *st, hipStream_t s) {
    L0 a = {w0, h0, lv[0].own, lv[0].ax, lv[0].ay, gas};
    const size_t n = (size_t)w0 * h0;
    const int g = grid1(n);
    const bool mg = precond != 0 && lv.size() > 1;
    *st = FldStats();
    const double inner_rtol = 1e-4;
    int it = 0;
    for (int outer = 0; outer < 50; outer++) {
        hipLaunchKernelGGL(k_true_residual, dim3(g), dim3(NT), 0, s, a, b, x, r32, ref ? ref->heat : nullptr,
                           ref ? ref->slope : nullptr, sc, part, counter);
        SR_HIP(hipMemcpyAsync(host_sc, sc, sizeof(double) * SC_N, hipMemcpyDeviceToHost, s));
        SR_HIP(hipStreamSynchronize(s));
        const double rr = host_sc[SC_R64], bb = host_sc[SC_BB] > 0 ? host_sc[SC_BB] : 1.0;
        st->rel = std::sqrt(rr / bb);
        st->refinements = outer;
        if (!std::isfinite(rr) || (ref && !(std::isfinite(host_sc[SC_UMEAS]) && std::isfinite(host_sc[SC_HMEAS])))) {
            st->finite = false;
            break;
        }
        const bool done = ref ? (host_sc[SC_UMEAS] <= ref->u_bound && host_sc[SC_HMEAS] <= ref->t_bound)
                              : (std::sqrt(rr) <= tol * std::sqrt(bb));
        if (done) {
            st->converged = true;
            break;
        }
        if (it >= max_iter) break;
        // inner FP32 flexible PCG on the residual, until ||r|| falls by inner_rtol or below the outer goal
        const double goal = ref ? 0.0 : tol * std::sqrt(bb);
        const double stop2 = std::max(inner_rtol * inner_rtol * rr, 0.25 * goal * goal);
        float *zc = za, *zo = zb;
        float *pold = p, *pnew = p + n;
        SR_HIP(hipMemsetAsync(zo, 0, sizeof(float) * n, s));
        // z = M r, rz = r.z (restart: beta = 0)
        if (mg)
            SR_TRY(vcycle(r32, zc, zo, true, s));
        else
            hipLaunchKernelGGL(k_jacobi, dim3(g), dim3(NT), 0, s, a, r32, zc, zo, sc, part, counter, 1);
        while (it < max_iter) {
            hipLaunchKernelGGL(k_matvec_p, dim3(g), dim3(NT), 0, s, a, zc, pold, pnew, q, sc, part, counter);
            hipLaunchKernelGGL(k_update, dim3(g), dim3(NT), 0, s, (int)n, pnew, q, x, r32, sc, part, counter);
            it++;
            SR_HIP(hipMemcpyAsync(host_sc, sc, sizeof(double) * SC_N, hipMemcpyDeviceToHost, s));
            SR_HIP(hipStreamSynchronize(s));
            const double rri = host_sc[SC_RR];
            if (!std::isfinite(rri) || !std::isfinite(host_sc[SC_PQ])) {
                st->finite = false;
                break;
            }
            if (rri <= stop2) break;
            std::swap(zc, zo);
            std::swap(pold, pnew);
            if (mg)
                SR_TRY(vcycle(r32, zc, zo, true, s));
            else
                hipLaunchKernelGGL(k_jacobi, dim3(g), dim3(NT), 0, s, a, r32, zc, zo, sc, part, counter, 1);
        }
        if (!st->finite) break;
    }
    st->iterations = it;
    SR_HIP(hipGetLastError());
    return SR_OK;
}

// HBM bytes one inner iteration moves at level 0 (f32 unless noted), the roofline's numerator: matvec (z, p, q,
// own, ax, ay: 24), update (x f64 rw 16, p, r rw 8, q: 32), the V-cycle's level-0 kernels (pre: b, own, ax, ay, z
// out, coarse b: 21; post: b, own, ax, ay, z in, z out, coarse z, z_old: 33), times 4/3 for the coarser levels.
double FldSolver::bytes_per_iteration() const {
    const double n = (double)w0 * h0;
    return n * (24.0 + 32.0 + (21.0 + 33.0) * 4.0 / 3.0);
}

} // namespace sr

// ------------------------------------------------------------------------------------------------------------ C ABI
using namespace sr;

struct sr_fld {
    FldSolver s;
};

extern "C" int sr_fld_create(int w, int h, const sr_heat_settings *hs, sr_fld **out) {
    if (!out) return fail(SR_ERR_ARGUMENT, "sr_fld_create: null out");
    sr_heat_settings d;
    sr_heat_settings_defaults(&d);
    if (!hs) hs = &d;
    auto *f = new sr_fld();
    int c = f->s.init(w, h, hs->levels, hs->nu1, hs->nu2);
    if (c != SR_OK) {
        delete f;
        return c;
    }
    *out = f;
    return SR_OK;
}

extern "C" int sr_fld_set_system(sr_fld *f, const float *own_dev, const float *kx_dev, const float *ky_dev,
                                 const uint8_t *gas_dev, void *stream) {
    if (!f || !own_dev || !kx_dev || !ky_dev) return fail(SR_ERR_ARGUMENT, "sr_fld_set_system: null pointer");
    return f->s.set_system(own_dev, kx_dev, ky_dev, gas_dev, (hipStream_t)stream);
}

extern "C" int sr_fld_solve(sr_fld *f, const double *b_dev, double *x_dev, double tol, int max_iter, int precond,
                            int *iterations, double *final_rel, void *stream) {
    if (!f || !b_dev || !x_dev) return fail(SR_ERR_ARGUMENT, "sr_fld_solve: null pointer");
    if (precond == SR_PRECOND_LEARNED) return fail(SR_ERR_ARGUMENT, "sr_fld_solve: the learned preconditioner is not built");
    FldStats st;
    SR_TRY(f->s.solve(b_dev, x_dev, tol, max_iter, precond, nullptr, &st, (hipStream_t)stream));
    if (iterations) *iterations = st.iterations;
    if (final_rel) *final_rel = st.rel;
    if (!st.finite) return fail(SR_ERR_NON_FINITE, "sr_fld_solve: a sum left the float's range");
    return SR_OK;
}

extern "C" int sr_fld_apply_precond(sr_fld *f, const float *r_dev, float *z_dev, int precond, void *stream) {
    if (!f || !r_dev || !z_dev) return fail(SR_ERR_ARGUMENT, "sr_fld_apply_precond: null pointer");
    return f->s.apply_precond(r_dev, z_dev, precond, (hipStream_t)stream);
}

extern "C" int sr_fld_matvec(sr_fld *f, const double *v_dev, double *out_dev, void *stream) {
    if (!f || !v_dev || !out_dev) return fail(SR_ERR_ARGUMENT, "sr_fld_matvec: null pointer");
    return f->s.matvec64(v_dev, out_dev, (hipStream_t)stream);
}

extern "C" int sr_fld_levels(sr_fld *f, int *levels) {
    if (!f || !levels) return fail(SR_ERR_ARGUMENT, "sr_fld_levels: null pointer");
    *levels = f->s.nlevels();
    return SR_OK;
}

extern "C" void sr_fld_destroy(sr_fld *f) { delete f; }
