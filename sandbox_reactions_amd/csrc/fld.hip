// P5 linear solver for the implicit flux-limited-diffusion step: FP32 multigrid-preconditioned flexible CG inside
// FP64 iterative refinement.
//
// The operator is the reference's pass system (crates/sr-physics/src/reference/heat.rs:374-392, :474-488) in face
// form; the hierarchy is the reference GPU's Galerkin 2x2 aggregation (crates/sr-engine/shaders/heat.wgsl:1316-1400)
// and its V-cycle smooths by red-black Gauss-Seidel, red first down, black first up (heat.wgsl:1401-1445). The
// machinery is this library's own:
//   * outer loop in FP64: r = b - A x with FP64 x and b, the stopping rule on that true residual;
//   * inner loop in FP32: flexible (Polak-Ribiere) PCG on the residual, FP64 dot products, x updated in FP64;
//   * the pre-smoother of a level, its residual and its restriction are one kernel: a 32x32 tile (16x16 on levels
//     too small to fill the chip) with a halo of 2 nu1 + 1 cells in LDS, each half sweep recomputed redundantly on
//     the shrinking halo, so the result is exact red-black Gauss-Seidel with one pass over HBM; the prolongation and
//     the post-smoother are another such kernel;
//   * every level of at most 4096 cells is handled inside one 1024-thread block (an LDS-resident V-cycle), the
//     coarsest (at most 16 cells) by an exact dense solve;
//   * dot products reduce by wave64 shuffles into per-block partials that the next kernel sums itself in a fixed
//     order (deterministic, and no device-wide fence: a fence at agent scope writes back the whole L2 of an XCD);
//     alpha and beta are formed on the device by every block alike.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>

#include "common.hpp"
#include "fld.hpp"
#include "net.hpp"

namespace sr {

#define NT 256
#define NB_MAX 1024
#define COARSE_SWEEPS 16
#define COARSE_MAX 4096
#define COARSE_NT 1024
#define COARSE_LDS (150 * 1024)
#define DENSE_MAX 16
#define NET_RATE_AT 4 // learned iterations before its residual rate is compared with plain multigrid's
// Restarts of the inner FP32 loop from the FP64 true residual before a solve gives up: where the FP32 inner solves
// cannot carry the true residual below the goal (a goal at the edge of what they resolve), a Newton pass takes the
// iterate as it stands rather than spending the pass's whole round budget
#define MAX_REFINEMENTS 50

// ------------------------------------------------------------------------------------------------ elementwise kernels

struct L0 {
    int w, h;
    const float *own, *ax, *ay;
    const uint8_t *gas;
};

__device__ __forceinline__ bool is_gas(const L0 &a, size_t i) { return a.gas == nullptr || a.gas[i]; }

// The state of a device-decided solve (FldSolver::solve with the slots on): the slot head (k_slot_head) reads it,
// decides, and writes the next copy (double-buffered by slot parity, so no block reads a copy another block of the
// same kernel writes); the slot's other kernels read the phase (and the inner index) of the copy the head wrote.
#define PH_TRUE 0   // the next slot opens with the FP64 true residual
#define PH_INNER 1  // the inner iteration runs
#define PH_DONE 2   // converged, out of iterations, or not finite: every later slot is a no-op
struct SolveCtl {
    int phase, it, k, refinements, converged, finite, max_iter, ref;
    double tol, u_bound, t_bound, rel, rel0, stop2, rr, bb;
};

// a kernel of a slot runs only while the inner iteration does (gate: the phase the slot's head wrote; null: always)
__device__ __forceinline__ bool gated_off(const int *gate) { return gate && *gate != PH_INNER; }

#define UNROLL 4

// r = b - A x in FP64 (x zeroed off the gas), r32 = (float) r; partials of ||r||^2, ||b||^2 and the reference's two
// measures
__global__ void __launch_bounds__(NT) k_true_residual(L0 a, const double *__restrict__ b, double *__restrict__ x,
                                                      float *__restrict__ r32, const float *__restrict__ heat,
                                                      const float *__restrict__ slope, double *part) {
    const int n = a.w * a.h;
    double v[4] = {0.0, 0.0, 0.0, 0.0};
    const int stride = gridDim.x * NT;
    for (int base = blockIdx.x * NT + threadIdx.x; base < n; base += stride * UNROLL) {
        // the loads of UNROLL cells issued together (each cell's neighbours whatever its couplings, an index past the
        // grid reading the cell itself; only a nonzero coupling uses the value), then the cells in order
        double xv[UNROLL][5], bv[UNROLL];
        float cv[UNROLL][5];
        bool gv[UNROLL];
#pragma unroll
        for (int u = 0; u < UNROLL; u++) {
            const int i = base + u * stride;
            if (i < n) {
                const int xx = i % a.w, yy = i / a.w;
                gv[u] = is_gas(a, i);
                xv[u][0] = x[i];
                xv[u][1] = x[xx > 0 ? i - 1 : i];
                xv[u][2] = x[xx + 1 < a.w ? i + 1 : i];
                xv[u][3] = x[yy > 0 ? i - a.w : i];
                xv[u][4] = x[yy + 1 < a.h ? i + a.w : i];
                bv[u] = b[i];
                cv[u][0] = a.own[i];
                cv[u][1] = a.ax[(size_t)yy * (a.w + 1) + xx];
                cv[u][2] = a.ax[(size_t)yy * (a.w + 1) + xx + 1];
                cv[u][3] = a.ay[(size_t)yy * a.w + xx];
                cv[u][4] = a.ay[(size_t)(yy + 1) * a.w + xx];
            }
        }
#pragma unroll
        for (int u = 0; u < UNROLL; u++) {
            const int i = base + u * stride;
            if (i >= n) break;
            double r = 0.0;
            if (gv[u]) {
                const double xi = xv[u][0];
                const double al = cv[u][1], ar = cv[u][2], au = cv[u][3], ad = cv[u][4];
                double ax = (double)cv[u][0] * xi;
                if (al != 0.0) ax += al * (xi - xv[u][1]);
                if (ar != 0.0) ax += ar * (xi - xv[u][2]);
                if (au != 0.0) ax += au * (xi - xv[u][3]);
                if (ad != 0.0) ax += ad * (xi - xv[u][4]);
                r = bv[u] - ax;
                v[1] += bv[u] * bv[u];
                if (heat) {
                    const double d = (double)cv[u][0] + al + ar + au + ad;
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
    }
    block_partials<4>(v, part, gridDim.x);
}

// A device-decided solve's state at its start (one thread)
__global__ void k_ctl_init(SolveCtl *c, double tol, int max_iter, int ref, double u_bound, double t_bound) {
    SolveCtl v{};
    v.phase = PH_TRUE;
    v.finite = 1;
    v.max_iter = max_iter;
    v.ref = ref;
    v.tol = tol;
    v.u_bound = u_bound;
    v.t_bound = t_bound;
    v.rel0 = -1.0;
    *c = v;
}

// The head of a slot (one PCG iteration of a device-decided solve): from the state the last slot left (cin) it decides
// what this slot does and writes that into cout for the slot's other kernels.
//   * after an inner iteration: ||r||^2 from the update's partials (every block sums them alike); not finite: done;
//     at or below the inner goal, or out of iterations: the true residual follows; else the next inner iteration;
//   * at the start or after the inner loop ends: the FP64 true residual r = b - A x (as k_true_residual), and the
//     last block to finish (a ticket counter, the only device-wide fence of the solve, taken only here) sums its
//     partials and decides: converged (the relative or the reference's rule), out of iterations, not finite, or a
//     restart of the inner FP32 loop from that residual (beta = 0, the inner goal 1e-4 of it or a quarter of the
//     outer goal, as the host-decided loop).
__global__ void __launch_bounds__(NT) k_slot_head(L0 a, const double *__restrict__ b, double *__restrict__ x,
                                                  float *__restrict__ r32, const float *__restrict__ heat,
                                                  const float *__restrict__ slope, double *part,
                                                  const double *part_rr, int nrr, const SolveCtl *cin,
                                                  SolveCtl *cout, unsigned *ticket) {
    SolveCtl c = *cin;
    if (c.phase == PH_DONE) {
        if (blockIdx.x == 0 && threadIdx.x == 0) *cout = c;
        return;
    }
    if (c.phase == PH_INNER) {
        const double rr = sum_partials(part_rr, nrr);
        c.it++;
        c.rr = rr;
        if (!isfinite(rr)) {
            c.finite = 0;
            c.phase = PH_DONE;
        } else if (rr > c.stop2 && c.it < c.max_iter) {
            c.k++;
        } else {
            c.phase = PH_TRUE;
        }
        if (c.phase != PH_TRUE) {
            if (blockIdx.x == 0 && threadIdx.x == 0) *cout = c;
            return;
        }
    }
    const int n = a.w * a.h;
    double v[4] = {0.0, 0.0, 0.0, 0.0};
    const int stride = gridDim.x * NT;
    for (int base = blockIdx.x * NT + threadIdx.x; base < n; base += stride * UNROLL) {
        // the loads of UNROLL cells issued together (each cell's neighbours whatever its couplings, an index past the
        // grid reading the cell itself; only a nonzero coupling uses the value), then the cells in order
        double xv[UNROLL][5], bv[UNROLL];
        float cv[UNROLL][5];
        bool gv[UNROLL];
#pragma unroll
        for (int u = 0; u < UNROLL; u++) {
            const int i = base + u * stride;
            if (i < n) {
                const int xx = i % a.w, yy = i / a.w;
                gv[u] = is_gas(a, i);
                xv[u][0] = x[i];
                xv[u][1] = x[xx > 0 ? i - 1 : i];
                xv[u][2] = x[xx + 1 < a.w ? i + 1 : i];
                xv[u][3] = x[yy > 0 ? i - a.w : i];
                xv[u][4] = x[yy + 1 < a.h ? i + a.w : i];
                bv[u] = b[i];
                cv[u][0] = a.own[i];
                cv[u][1] = a.ax[(size_t)yy * (a.w + 1) + xx];
                cv[u][2] = a.ax[(size_t)yy * (a.w + 1) + xx + 1];
                cv[u][3] = a.ay[(size_t)yy * a.w + xx];
                cv[u][4] = a.ay[(size_t)(yy + 1) * a.w + xx];
            }
        }
#pragma unroll
        for (int u = 0; u < UNROLL; u++) {
            const int i = base + u * stride;
            if (i >= n) break;
            double r = 0.0;
            if (gv[u]) {
                const double xi = xv[u][0];
                const double al = cv[u][1], ar = cv[u][2], au = cv[u][3], ad = cv[u][4];
                double ax = (double)cv[u][0] * xi;
                if (al != 0.0) ax += al * (xi - xv[u][1]);
                if (ar != 0.0) ax += ar * (xi - xv[u][2]);
                if (au != 0.0) ax += au * (xi - xv[u][3]);
                if (ad != 0.0) ax += ad * (xi - xv[u][4]);
                r = bv[u] - ax;
                v[1] += bv[u] * bv[u];
                if (heat) {
                    const double d = (double)cv[u][0] + al + ar + au + ad;
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
    }
    block_partials<4>(v, part, gridDim.x);
    __shared__ int last;
    __syncthreads();
    if (threadIdx.x == 0) {
        __threadfence();  // this block's partials, before its ticket
        last = atomicAdd(ticket, 1u) == gridDim.x - 1;
    }
    __syncthreads();
    if (!last) return;
    __threadfence();  // every block's partials, after the last ticket
    const volatile double *vp = part;
    __shared__ double scratch[NT / 64];
    double t[4];
    for (int k = 0; k < 4; k++) {
        double sum = 0.0;
        for (int i = threadIdx.x; i < (int)gridDim.x; i += NT) sum += vp[(size_t)k * gridDim.x + i];
        t[k] = block_sum(sum, scratch);
    }
    if (threadIdx.x != 0) return;
    *ticket = 0u;
    const double rr = t[0], bb = t[1] > 0.0 ? t[1] : 1.0;
    c.rr = rr;
    c.bb = bb;
    c.rel = sqrt(rr / bb);
    if (c.rel0 < 0.0) c.rel0 = c.rel;
    const bool fin = isfinite(rr) && (!c.ref || (isfinite(t[2]) && isfinite(t[3])));
    const bool conv = c.ref ? (t[2] <= c.u_bound && t[3] <= c.t_bound) : (sqrt(rr) <= c.tol * sqrt(bb));
    if (!fin) {
        c.finite = 0;
        c.phase = PH_DONE;
    } else if (conv) {
        c.converged = 1;
        c.phase = PH_DONE;
    } else if (c.it >= c.max_iter || c.refinements >= MAX_REFINEMENTS) {
        c.phase = PH_DONE;
    } else {
        const double goal2 = c.ref ? 0.0 : c.tol * c.tol * bb;
        c.stop2 = fmax(1e-8 * rr, 0.25 * goal2);
        c.k = 0;
        c.refinements++;
        c.phase = PH_INNER;
    }
    *cout = c;
}

// beta from the preconditioner's partials: (r.z_new - r.z_old) / (r.z of the previous iteration), 0 on a restart
__device__ inline float beta_of(const double *dots_cur, const double *dots_prev, int nbd, int nbp, int restart) {
    if (restart) return 0.0f;
    const double rzn = sum_partials(dots_cur, nbd), rzo = sum_partials(dots_cur + nbd, nbd);
    const double rzp = sum_partials(dots_prev, nbp);
    return rzp != 0.0 ? (float)((rzn - rzo) / rzp) : 0.0f;
}

// p = z + beta p_old (into p_new), q = A p_new in FP32; partials of p.q. Blocks of 256 columns x `rows` rows; each
// thread walks its column down the rows with p of the rows above, at and below in registers (and the face above),
// and takes its left and right neighbours' p from the adjacent lanes, loading them only at a wave's two ends. Lanes
// past the grid's right edge shadow its last column and store nothing. Every product is the one the face form
// computes cell by cell, in the same order.
#define MV_ROWS 16
#ifndef MV_BLOCKS
#define MV_BLOCKS 2048  // the matvec's target block count: rows per block = the grid's column strips over it
#endif
__global__ void __launch_bounds__(NT) k_matvec_p(L0 a, const float *__restrict__ z, const float *__restrict__ pold,
                                                 float *__restrict__ pnew, float *__restrict__ q,
                                                 const double *dots_cur, const double *dots_prev, int nbd, int nbp,
                                                 int restart, int rows, double *part, const SolveCtl *ctl) {
    if (ctl) {  // a slot: gated by its head, a restart where the head set the inner index to 0
        if (ctl->phase != PH_INNER) return;
        restart = ctl->k == 0;
    }
    const float beta = beta_of(dots_cur, dots_prev, nbd, nbp, restart);
    const int w = a.w, h = a.h;
    const int tile = xcd_tile();
    const int x = (tile % gridDim.x) * NT + threadIdx.x, y0 = (tile / gridDim.x) * rows, lane = threadIdx.x & 63;
    const bool in = x < w;
    const int xc = in ? x : w - 1;
    // the coefficients never alias the stores, so the next row's loads may be issued ahead of this row's
    const float *__restrict__ own = a.own, *__restrict__ axf = a.ax, *__restrict__ ayf = a.ay;
    auto P = [&](size_t j) { return z[j] + beta * pold[j]; };
    double v[1] = {0.0};
    const int y1 = min(y0 + rows, h);
    if (y0 < h) {
        size_t i = (size_t)y0 * w + xc;
        float pu = y0 > 0 ? P(i - w) : 0.0f;
        float pc = P(i);
        float au = ayf[i];
        for (int y = y0; y < y1; y++, i += w) {
            const float pd = y + 1 < h ? P(i + w) : 0.0f;
            const float ad = ayf[i + w];
            const float al = axf[(size_t)y * (w + 1) + xc], ar = axf[(size_t)y * (w + 1) + xc + 1];
            float pl = __shfl_up(pc, 1, 64), pr = __shfl_down(pc, 1, 64);
            if (lane == 0 && xc > 0) pl = P(i - 1);
            if (lane == 63 && xc + 1 < w) pr = P(i + 1);
            float s = own[i] * pc;
            if (al != 0.0f) s += al * (pc - pl);
            if (ar != 0.0f) s += ar * (pc - pr);
            if (au != 0.0f) s += au * (pc - pu);
            if (ad != 0.0f) s += ad * (pc - pd);
            if (in) {
                pnew[i] = pc;
                q[i] = s;
                v[0] += (double)pc * (double)s;
            }
            pu = pc;
            pc = pd;
            au = ad;
        }
    }
    block_partials<1>(v, part, gridDim.x * gridDim.y, tile);  // partials in strip order
}

// x += alpha p (FP64), r -= alpha q, alpha = r.z / p.q from the partials; partials of ||r||^2
__global__ void __launch_bounds__(NT) k_update(int n, const float *__restrict__ p, const float *__restrict__ q,
                                               double *__restrict__ x, float *__restrict__ r, const double *dots_cur,
                                               int nbd, const double *pq_part, int nbq, double *part,
                                               const int *gate) {
    if (gated_off(gate)) return;
    const double rz = sum_partials(dots_cur, nbd), pq = sum_partials(pq_part, nbq);
    const double alpha = pq != 0.0 ? rz / pq : 0.0;
    const float af = (float)alpha;
    double v[1] = {0.0};
    for (int base = blockIdx.x * NT * UNROLL + threadIdx.x; base < n; base += gridDim.x * NT * UNROLL) {
        float pv[UNROLL], qv[UNROLL], rv[UNROLL];
        double xv[UNROLL];
#pragma unroll
        for (int u = 0; u < UNROLL; u++) {
            const int i = base + u * NT;
            if (i < n) {
                pv[u] = p[i];
                qv[u] = q[i];
                rv[u] = r[i];
                xv[u] = x[i];
            }
        }
#pragma unroll
        for (int u = 0; u < UNROLL; u++) {
            const int i = base + u * NT;
            if (i < n) {
                x[i] = xv[u] + alpha * (double)pv[u];
                const float ri = rv[u] - af * qv[u];
                r[i] = ri;
                v[0] += (double)ri * ri;
            }
        }
    }
    block_partials<1>(v, part, gridDim.x);
}

// Jacobi: z = r / D, D = own + sum of couplings; partials of r.z and r.z_old
__global__ void __launch_bounds__(NT) k_jacobi(L0 a, const float *__restrict__ r, float *__restrict__ z,
                                               const float *__restrict__ zold, double *part, const int *gate) {
    if (gated_off(gate)) return;
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
    if (part) block_partials<2>(v, part, gridDim.x);
}

// partials of r.z and r.z_old after a preconditioner that made none
__global__ void __launch_bounds__(NT) k_dots(int n, const float *__restrict__ r, const float *__restrict__ z,
                                             const float *__restrict__ zold, double *part, const int *gate) {
    if (gated_off(gate)) return;
    double v[2] = {0.0, 0.0};
    for (int i = blockIdx.x * NT + threadIdx.x; i < n; i += gridDim.x * NT) {
        v[0] += (double)r[i] * z[i];
        if (zold) v[1] += (double)r[i] * zold[i];
    }
    block_partials<2>(v, part, gridDim.x);
}

// r1 = r - A z in FP32 on the gas cells (0 elsewhere): what the V-cycle left, the learned correction's input
__global__ void __launch_bounds__(NT) k_resid32(L0 a, const float *__restrict__ r, const float *__restrict__ z,
                                                float *__restrict__ r1) {
    const int n = a.w * a.h;
    for (int i = blockIdx.x * NT + threadIdx.x; i < n; i += gridDim.x * NT) {
        const int xx = i % a.w, yy = i / a.w;
        float v = 0.0f;
        if (is_gas(a, i)) {
            const float zi = z[i];
            const float al = a.ax[(size_t)yy * (a.w + 1) + xx], ar = a.ax[(size_t)yy * (a.w + 1) + xx + 1];
            const float au = a.ay[(size_t)yy * a.w + xx], ad = a.ay[(size_t)(yy + 1) * a.w + xx];
            float s = a.own[i] * zi;
            if (al != 0.0f) s += al * (zi - z[i - 1]);
            if (ar != 0.0f) s += ar * (zi - z[i + 1]);
            if (au != 0.0f) s += au * (zi - z[i - a.w]);
            if (ad != 0.0f) s += ad * (zi - z[i + a.w]);
            v = r[i] - s;
        }
        r1[i] = v;
    }
}

// z = 0 off the gas; partials of r.z and r.z_old
__global__ void __launch_bounds__(NT) k_dots_gas(L0 a, const float *__restrict__ r, float *__restrict__ z,
                                                 const float *__restrict__ zold, double *part) {
    const int n = a.w * a.h;
    double v[2] = {0.0, 0.0};
    for (int i = blockIdx.x * NT + threadIdx.x; i < n; i += gridDim.x * NT) {
        float zi = z[i];
        if (!is_gas(a, i)) {
            zi = 0.0f;
            z[i] = 0.0f;
        }
        v[0] += (double)r[i] * zi;
        if (zold) v[1] += (double)r[i] * zold[i];
    }
    if (part) block_partials<2>(v, part, gridDim.x);
}

// ------------------------------------------------------------------------------------------------ tile smoothers

struct Lv {
    int w, h;
    const float *own, *ax, *ay;
    const float4 *cf;  // the per-cell record (tile smoothers), or null
};

// The region of a TX x TY tile plus a halo of HW cells: (TX + 2 HW) x (TY + 2 HW) entries per array. A cell at
// least one cell inside the region reads its right and bottom faces from its neighbours' left and top entries. The
// sweeps need no `own` term (the inverse diagonal carries it): only the pre-smoother's residual reads it, on the tile's
// own cells, so it is kept for those alone, and four 32x32 pre-smoothing blocks (halo 5) share a CU's 160 KB.
template <int TX, int TY, int HW> struct Region {
    static constexpr int RX = TX + 2 * HW, RY = TY + 2 * HW, N = RX * RY;
    float z[N], b[N], al[N], au[N], rd[N];
};

template <int TX, int TY, int HW>
__device__ inline void load_region(Region<TX, TY, HW> &R, const Lv &L, int x0, int y0, const float *__restrict__ b,
                                   float *own_in = nullptr) {
    using RG = Region<TX, TY, HW>;
    constexpr int IT = (RG::N + NT - 1) / NT;
    // every load of the region issued before any is used: one memory round trip per block; the coefficients are one
    // aligned 16-byte record per cell
    float4 vc[IT];
    float vb[IT];
#pragma unroll
    for (int k = 0; k < IT; k++) {
        const int t = threadIdx.x + k * NT;
        const int lx = t % RG::RX, ly = t / RG::RX;
        const int gx = x0 + lx - HW, gy = y0 + ly - HW;
        const bool in = t < RG::N && gx >= 0 && gx < L.w && gy >= 0 && gy < L.h;
        const size_t i = (size_t)gy * L.w + gx;
        vc[k] = in ? L.cf[i] : make_float4(1.0f, 0.0f, 0.0f, 1.0f);
        vb[k] = in && b ? b[i] : 0.0f;
    }
#pragma unroll
    for (int k = 0; k < IT; k++) {
        const int t = threadIdx.x + k * NT;
        if (t < RG::N) {
            if (own_in) {
                const int lx = t % RG::RX - HW, ly = t / RG::RX - HW;
                if (lx >= 0 && lx < TX && ly >= 0 && ly < TY) own_in[ly * TX + lx] = vc[k].x;
            }
            R.al[t] = vc[k].y;
            R.au[t] = vc[k].z;
            R.rd[t] = vc[k].w;
            R.b[t] = vb[k];
        }
    }
}

// One half sweep of colour `color` over the region cells at least K cells from its border (colour-compact: every
// thread takes a cell of the colour). Parity of a region cell: (lx + ly) & 1 (tile origins and 2 HW are even).
template <int TX, int TY, int HW, int K>
__device__ inline void half_sweep(Region<TX, TY, HW> &R, const Lv &L, int x0, int y0, int color) {
    using RG = Region<TX, TY, HW>;
    constexpr int SX = RG::RX - 2 * K, SY = RG::RY - 2 * K, HALF = SX / 2, TOTAL = HALF * SY;
    constexpr int IT = (TOTAL + NT - 1) / NT;
    static_assert(SX % 2 == 0, "even spans");
#pragma unroll
    for (int k = 0; k < IT; k++) {
        const int t = threadIdx.x + k * NT;
        if (t >= TOTAL) break;
        const int row = t / HALF, j = t - row * HALF;
        const int ly = K + row;
        const int lx = K + ((color - (K + ly)) & 1) + 2 * j;
        const int gx = x0 + lx - HW, gy = y0 + ly - HW;
        if (gx < 0 || gx >= L.w || gy < 0 || gy >= L.h) continue;
        const int c = ly * RG::RX + lx;
        const float al = R.al[c], ar = R.al[c + 1], au = R.au[c], ad = R.au[c + RG::RX];
        const float num = R.b[c] + al * R.z[c - 1] + ar * R.z[c + 1] + au * R.z[c - RG::RX] + ad * R.z[c + RG::RX];
        R.z[c] = num * R.rd[c];
    }
    __syncthreads();
}

template <int TX, int TY, int HW, int K0, int NH, bool DOWN>
__device__ inline void sweeps(Region<TX, TY, HW> &R, const Lv &L, int x0, int y0) {
    if constexpr (NH > 0) {
        // down: red (0) first; up: black (1) first
        constexpr int color = DOWN ? (K0 - 1) & 1 : 1 - ((K0 - 1) & 1);
        half_sweep<TX, TY, HW, K0>(R, L, x0, y0, color);
        sweeps<TX, TY, HW, K0 + 1, NH - 1, DOWN>(R, L, x0, y0);
    }
}

// The residual b - A z summed over the 2x2 cells whose top-left is the tile's cell (tx, ty), at (gx, gy) on the
// level; cells past the level's right or bottom edge add nothing. `own` holds the diagonal's own term of the tile's
// cells in rows of OS. The fused and the separate pre-smoothers both restrict through this one expression.
template <int TX, int TY, int HW, int OS>
__device__ inline float restrict_residual(const Region<TX, TY, HW> &R, const Lv &L, const float *own, int tx, int ty,
                                          int gx, int gy) {
    using RG = Region<TX, TY, HW>;
    float sum = 0.0f;
#pragma unroll
    for (int dy = 0; dy < 2; dy++)
#pragma unroll
        for (int dx = 0; dx < 2; dx++) {
            if (gx + dx >= L.w || gy + dy >= L.h) continue;
            const int c = (HW + ty + dy) * RG::RX + HW + tx + dx;
            const float zc = R.z[c];
            const float al = R.al[c], ar = R.al[c + 1], au = R.au[c], ad = R.au[c + RG::RX];
            const float o = own[(ty + dy) * OS + tx + dx];
            const float az = o * zc + al * (zc - R.z[c - 1]) + ar * (zc - R.z[c + 1]) + au * (zc - R.z[c - RG::RX]) +
                             ad * (zc - R.z[c + RG::RX]);
            sum += R.b[c] - az;
        }
    return sum;
}

// Pre-smoothing of level L from z = 0 (nu sweeps, red first), then the residual on the tile and its 2x2
// restriction into the next level's b. Halo 2 nu + 1.
template <int TX, int TY, int NU>
__global__ void __launch_bounds__(NT) k_pre(Lv L, const float *__restrict__ b, float *__restrict__ z,
                                            float *__restrict__ bc, int wc, const int *gate) {
    if (gated_off(gate)) return;
    constexpr int HW = 2 * NU + 1;
    using RG = Region<TX, TY, HW>;
    __shared__ RG R;
    __shared__ float own_in[TX * TY];
    const int tile = xcd_tile();
    const int x0 = (tile % gridDim.x) * TX, y0 = (tile / gridDim.x) * TY;
    load_region<TX, TY, HW>(R, L, x0, y0, b, own_in);
#pragma unroll 4
    for (int t = threadIdx.x; t < RG::N; t += NT) R.z[t] = 0.0f;
    __syncthreads();
    sweeps<TX, TY, HW, 1, 2 * NU, true>(R, L, x0, y0);
    constexpr int CX = TX / 2, CY = TY / 2;
    for (int t = threadIdx.x; t < CX * CY; t += NT) {
        const int cx = t % CX, cy = t / CX;
        const int X = x0 / 2 + cx, Y = y0 / 2 + cy;
        if (2 * X >= L.w || 2 * Y >= L.h) continue;
        bc[(size_t)Y * wc + X] = restrict_residual<TX, TY, HW, TX>(R, L, own_in, 2 * cx, 2 * cy, 2 * X, 2 * Y);
    }
    for (int t = threadIdx.x; t < TX * TY; t += NT) {
        const int lx = HW + t % TX, ly = HW + t / TX;
        const int gx = x0 + lx - HW, gy = y0 + ly - HW;
        if (gx < L.w && gy < L.h) z[(size_t)gy * L.w + gx] = R.z[ly * RG::RX + lx];
    }
}

// The PCG update of the iteration before, folded into the next V-cycle's level-0 pre-smoother: alpha = r.z / p.q from
// the partials, x += alpha p (FP64) and r -= alpha q in the update's own expressions, ||r||^2 partials per tile. Each
// tile computes the new residual over its whole region (its cells and the halo) from the old r and q, stores x and r
// for its own cells only, r into the other residual buffer (the tiles beside it read the old one as halo), and smooths
// from it as k_pre does. x, r and z are the separate kernels' values; the ||r||^2 partials are summed per tile.
struct PendingUpdate {
    const float *p, *q, *rold;
    float *rnew;
    double *x;
    const double *dots_cur;
    int nbd;
    const double *pq;
    int nbq;
    double *part;  // ||r||^2 per level-0 tile
};

template <int TX, int TY, int NU>
__global__ void __launch_bounds__(NT) k_pre_upd(Lv L, PendingUpdate U, float *__restrict__ z, float *__restrict__ bc,
                                                int wc) {
    constexpr int HW = 2 * NU + 1;
    using RG = Region<TX, TY, HW>;
    __shared__ RG R;
    __shared__ float own_in[TX * TY];
    const double rz = sum_partials(U.dots_cur, U.nbd), pq = sum_partials(U.pq, U.nbq);
    const double alpha = pq != 0.0 ? rz / pq : 0.0;
    const float af = (float)alpha;
    const int tile = xcd_tile();
    const int x0 = (tile % gridDim.x) * TX, y0 = (tile / gridDim.x) * TY;
    constexpr int IT = (RG::N + NT - 1) / NT;
    float4 vc[IT];
    float vr[IT], vq[IT];
#pragma unroll
    for (int k = 0; k < IT; k++) {
        const int t = threadIdx.x + k * NT;
        const int lx = t % RG::RX, ly = t / RG::RX;
        const int gx = x0 + lx - HW, gy = y0 + ly - HW;
        const bool in = t < RG::N && gx >= 0 && gx < L.w && gy >= 0 && gy < L.h;
        const size_t i = (size_t)gy * L.w + gx;
        vc[k] = in ? L.cf[i] : make_float4(1.0f, 0.0f, 0.0f, 1.0f);
        vr[k] = in ? U.rold[i] : 0.0f;
        vq[k] = in ? U.q[i] : 0.0f;
    }
    double v[1] = {0.0};
#pragma unroll
    for (int k = 0; k < IT; k++) {
        const int t = threadIdx.x + k * NT;
        if (t < RG::N) {
            const int lx = t % RG::RX - HW, ly = t / RG::RX - HW;
            const float ri = vr[k] - af * vq[k];
            const bool mine = lx >= 0 && lx < TX && ly >= 0 && ly < TY;
            if (mine) own_in[ly * TX + lx] = vc[k].x;
            R.al[t] = vc[k].y;
            R.au[t] = vc[k].z;
            R.rd[t] = vc[k].w;
            R.b[t] = ri;
            const int gx = x0 + lx, gy = y0 + ly;
            if (mine && gx < L.w && gy < L.h) {
                const size_t i = (size_t)gy * L.w + gx;
                U.rnew[i] = ri;
                U.x[i] = U.x[i] + alpha * (double)U.p[i];
                v[0] += (double)ri * ri;
            }
        }
    }
#pragma unroll 4
    for (int t = threadIdx.x; t < RG::N; t += NT) R.z[t] = 0.0f;
    __syncthreads();
    sweeps<TX, TY, HW, 1, 2 * NU, true>(R, L, x0, y0);
    constexpr int CX = TX / 2, CY = TY / 2;
    for (int t = threadIdx.x; t < CX * CY; t += NT) {
        const int cx = t % CX, cy = t / CX;
        const int X = x0 / 2 + cx, Y = y0 / 2 + cy;
        if (2 * X >= L.w || 2 * Y >= L.h) continue;
        bc[(size_t)Y * wc + X] = restrict_residual<TX, TY, HW, TX>(R, L, own_in, 2 * cx, 2 * cy, 2 * X, 2 * Y);
    }
    for (int t = threadIdx.x; t < TX * TY; t += NT) {
        const int lx = HW + t % TX, ly = HW + t / TX;
        const int gx = x0 + lx - HW, gy = y0 + ly - HW;
        if (gx < L.w && gy < L.h) z[(size_t)gy * L.w + gx] = R.z[ly * RG::RX + lx];
    }
    block_partials<1>(v, U.part, gridDim.x * gridDim.y, tile);
}

// Prolongation of the coarse correction (scaled by omega) and post-smoothing of level L (nu sweeps, black first).
// Halo 2 nu. With `part` (level 0), partials of r.z_new and r.z_old.
template <int TX, int TY, int NU>
__global__ void __launch_bounds__(NT) k_post(Lv L, const float *__restrict__ b, const float *__restrict__ zin,
                                             const float *__restrict__ zc, int wc, float omega,
                                             float *__restrict__ zout, const float *__restrict__ zold, double *part,
                                             const int *gate) {
    if (gated_off(gate)) return;
    constexpr int HW = 2 * NU;
    using RG = Region<TX, TY, HW>;
    __shared__ RG R;
    const int tile = xcd_tile();
    const int x0 = (tile % gridDim.x) * TX, y0 = (tile / gridDim.x) * TY;
    load_region<TX, TY, HW>(R, L, x0, y0, b);
    {
        constexpr int IT = (RG::N + NT - 1) / NT;
        float zv[IT];
#pragma unroll
        for (int k = 0; k < IT; k++) {
            const int t = threadIdx.x + k * NT;
            const int lx = t % RG::RX, ly = t / RG::RX;
            const int gx = x0 + lx - HW, gy = y0 + ly - HW;
            const bool in = t < RG::N && gx >= 0 && gx < L.w && gy >= 0 && gy < L.h;
            zv[k] = in ? zin[(size_t)gy * L.w + gx] + omega * zc[(size_t)(gy >> 1) * wc + (gx >> 1)] : 0.0f;
        }
#pragma unroll
        for (int k = 0; k < IT; k++) {
            const int t = threadIdx.x + k * NT;
            if (t < RG::N) R.z[t] = zv[k];
        }
    }
    __syncthreads();
    sweeps<TX, TY, HW, 1, 2 * NU, false>(R, L, x0, y0);
    double v[2] = {0.0, 0.0};
    for (int t = threadIdx.x; t < TX * TY; t += NT) {
        const int lx = HW + t % TX, ly = HW + t / TX;
        const int gx = x0 + lx - HW, gy = y0 + ly - HW;
        if (gx < L.w && gy < L.h) {
            const size_t i = (size_t)gy * L.w + gx;
            const float zn = R.z[ly * RG::RX + lx];
            zout[i] = zn;
            if (part) {
                const float bi = R.b[ly * RG::RX + lx];
                v[0] += (double)bi * zn;
                if (zold) v[1] += (double)bi * zold[i];
            }
        }
    }
    if (part) block_partials<2>(v, part, gridDim.x * gridDim.y, tile);  // partials in tile order
}

// Two adjacent tile-smoothed levels a and b = a + 1 in one launch per direction (the last two above the
// single-block set, where every launch is latency-bound). Each block redoes on its own the finer level's work over
// the cells its coarser tile's halo reaches; neighbouring blocks recompute the overlap alike. Every value a block
// uses is one the separate kernels compute with the same expressions in the same order (sweeps, residual,
// restriction, prolongation), so the fused pair gives the same bits.

// Pre: a block owns an FB x FB tile of level b. Level b's pre-smoother reads b on its region's cells at least one
// cell in (halo HW = 2 nu + 1), so level a's residual is restricted over those: level a's cells under them form the
// tile of TA = 2 FB + 4 HW - 4 cells starting at 2 (xb0 - HW + 1), smoothed from z = 0 over that tile plus its own
// halo HW. Writes level a's z and level b's b and z on the cells under this block's tile, and level c's b.
template <int FB, int NU>
__global__ void __launch_bounds__(NT) k_pre2(Lv La, const float *__restrict__ ba, float *__restrict__ za, Lv Lb,
                                             float *__restrict__ bb, float *__restrict__ zb, float *__restrict__ bc,
                                             int wc, const int *gate) {
    if (gated_off(gate)) return;
    constexpr int HW = 2 * NU + 1;
    constexpr int TA = 2 * FB + 4 * HW - 4;
    using RGA = Region<TA, TA, HW>;
    using RGB = Region<FB, FB, HW>;
    __shared__ RGA A;
    __shared__ RGB B;
    __shared__ float own_a[TA * TA];
    __shared__ float own_b[FB * FB];
    const int tile = xcd_tile();
    const int xb0 = (tile % gridDim.x) * FB, yb0 = (tile / gridDim.x) * FB;
    const int xa0 = 2 * (xb0 - HW + 1), ya0 = 2 * (yb0 - HW + 1);
    load_region<TA, TA, HW>(A, La, xa0, ya0, ba, own_a);
    load_region<FB, FB, HW>(B, Lb, xb0, yb0, nullptr, own_b);
#pragma unroll 4
    for (int t = threadIdx.x; t < RGA::N; t += NT) A.z[t] = 0.0f;
    __syncthreads();
    sweeps<TA, TA, HW, 1, 2 * NU, true>(A, La, xa0, ya0);
    // level b's b on its region, from level a's residual; the region's outer ring is never read
    for (int t = threadIdx.x; t < RGB::N; t += NT) {
        const int cx = t % RGB::RX, cy = t / RGB::RX;
        const int X = xb0 - HW + cx, Y = yb0 - HW + cy;
        float v = 0.0f;
        if (cx >= 1 && cy >= 1 && cx < RGB::RX - 1 && cy < RGB::RY - 1 && X >= 0 && Y >= 0 && 2 * X < La.w &&
            2 * Y < La.h)
            v = restrict_residual<TA, TA, HW, TA>(A, La, own_a, 2 * (cx - 1), 2 * (cy - 1), 2 * X, 2 * Y);
        B.b[t] = v;
        B.z[t] = 0.0f;
    }
    // level a's z on the cells under this block's b tile
    constexpr int OA = 2 * HW - 2;  // their offset in level a's tile
    for (int t = threadIdx.x; t < 4 * FB * FB; t += NT) {
        const int lx = HW + OA + t % (2 * FB), ly = HW + OA + t / (2 * FB);
        const int gx = xa0 + lx - HW, gy = ya0 + ly - HW;
        if (gx < La.w && gy < La.h) za[(size_t)gy * La.w + gx] = A.z[ly * RGA::RX + lx];
    }
    __syncthreads();
    for (int t = threadIdx.x; t < FB * FB; t += NT) {
        const int gx = xb0 + t % FB, gy = yb0 + t / FB;
        if (gx < Lb.w && gy < Lb.h) bb[(size_t)gy * Lb.w + gx] = B.b[(HW + t / FB) * RGB::RX + HW + t % FB];
    }
    sweeps<FB, FB, HW, 1, 2 * NU, true>(B, Lb, xb0, yb0);
    constexpr int CX = FB / 2;
    for (int t = threadIdx.x; t < CX * CX; t += NT) {
        const int cx = t % CX, cy = t / CX;
        const int X = xb0 / 2 + cx, Y = yb0 / 2 + cy;
        if (2 * X >= Lb.w || 2 * Y >= Lb.h) continue;
        bc[(size_t)Y * wc + X] = restrict_residual<FB, FB, HW, FB>(B, Lb, own_b, 2 * cx, 2 * cy, 2 * X, 2 * Y);
    }
    for (int t = threadIdx.x; t < FB * FB; t += NT) {
        const int lx = HW + t % FB, ly = HW + t / FB;
        const int gx = xb0 + lx - HW, gy = yb0 + ly - HW;
        if (gx < Lb.w && gy < Lb.h) zb[(size_t)gy * Lb.w + gx] = B.z[ly * RGB::RX + lx];
    }
}

// Post: a block owns a TA x TA tile of level a. Its region (halo HW = 2 nu) starts from level b's final z, which
// level b's post-smoother makes exact on its region's cells at least HW in: level b's tile here is the TA / 2 + HW
// cells under level a's region, starting at xa0 / 2 - HW / 2. Writes level a's result on its tile.
template <int TA, int NU>
__global__ void __launch_bounds__(NT) k_post2(Lv La, const float *__restrict__ ba, const float *__restrict__ za, Lv Lb,
                                              const float *__restrict__ bb, const float *__restrict__ zb,
                                              const float *__restrict__ zc, int wc, float omega,
                                              float *__restrict__ zout, const int *gate) {
    if (gated_off(gate)) return;
    constexpr int HW = 2 * NU;
    constexpr int TB = TA / 2 + HW;
    using RGA = Region<TA, TA, HW>;
    using RGB = Region<TB, TB, HW>;
    __shared__ RGA A;
    __shared__ RGB B;
    const int tile = xcd_tile();
    const int xa0 = (tile % gridDim.x) * TA, ya0 = (tile / gridDim.x) * TA;
    const int xb0 = xa0 / 2 - HW / 2, yb0 = ya0 / 2 - HW / 2;
    load_region<TB, TB, HW>(B, Lb, xb0, yb0, bb);
    load_region<TA, TA, HW>(A, La, xa0, ya0, ba);
    for (int t = threadIdx.x; t < RGB::N; t += NT) {
        const int gx = xb0 + t % RGB::RX - HW, gy = yb0 + t / RGB::RX - HW;
        float v = 0.0f;
        if (gx >= 0 && gx < Lb.w && gy >= 0 && gy < Lb.h)
            v = zb[(size_t)gy * Lb.w + gx] + omega * zc[(size_t)(gy >> 1) * wc + (gx >> 1)];
        B.z[t] = v;
    }
    __syncthreads();
    sweeps<TB, TB, HW, 1, 2 * NU, false>(B, Lb, xb0, yb0);
    for (int t = threadIdx.x; t < RGA::N; t += NT) {
        const int gx = xa0 + t % RGA::RX - HW, gy = ya0 + t / RGA::RX - HW;
        float v = 0.0f;
        if (gx >= 0 && gx < La.w && gy >= 0 && gy < La.h)
            v = za[(size_t)gy * La.w + gx] + omega * B.z[((gy >> 1) - yb0 + HW) * RGB::RX + (gx >> 1) - xb0 + HW];
        A.z[t] = v;
    }
    __syncthreads();
    sweeps<TA, TA, HW, 1, 2 * NU, false>(A, La, xa0, ya0);
    for (int t = threadIdx.x; t < TA * TA; t += NT) {
        const int lx = HW + t % TA, ly = HW + t / TA;
        const int gx = xa0 + lx - HW, gy = ya0 + ly - HW;
        if (gx < La.w && gy < La.h) zout[(size_t)gy * La.w + gx] = A.z[ly * RGA::RX + lx];
    }
}

// The V-cycle over every level of at most COARSE_MAX cells in one block (the "set"). The levels' own, ax and ay
// arrays, the coarsest's inverse and every cell's inverse diagonal (both formed at the rebuild, k_coarse_prep) lie in
// one packed buffer. Every phase is swept by the whole block between block barriers (measured: one wave alone, with no
// block barrier, is slower even on a 16x16 level, its LDS latency no longer hidden by the other waves). The coarsest is
// solved exactly by the inverse formed at the rebuild when it holds at most DENSE_MAX cells, else by COARSE_SWEEPS
// forward and backward sweeps.
//
// With a head (HEAD = true; sets of three levels or more whose first level fits the block's 2 TX x 2 TY cover) the
// set's two largest levels run register-resident: thread (tx, ty) holds the 2x2 block of the first level at (2 tx,
// 2 ty) (its b, z, own, four couplings and inverse diagonal per cell) and the second level's cell (tx, ty), the
// aggregate of that block, so the restriction from the first level and the prolongation into it need no exchange;
// a half sweep reads only the neighbours outside the thread's cells from LDS and publishes the cells it updated.
// The remaining ("tail") levels run from LDS as without a head. Every cell's arithmetic, and the order of every sum,
// is the one the tail levels use, so the head changes where values live, not what is computed.
struct CoarseSet {
    int n;
    int w[16], h[16];
    int pack_floats;   // the tail's own/ax/ay per level, the coarsest's inverse, the tail's inverse diagonals: what
                       // the block loads into LDS (the whole pack without a head)
    int offo[16], offx[16], offy[16], offd[16], offinv;  // offsets in the pack (= in LDS for the tail's levels)
    int offz[16], offb[16], total;  // LDS offsets of the tail levels' z and b; LDS floats in all
    int lw[16];        // log2 of the level's width when a power of two, else -1
    int head, ltx;     // register-resident levels (0 or 2); log2 TX of the block's TX x (COARSE_NT / TX) layout
    int offza, offzb, offrb;  // LDS: the head levels' published z, the second level's residuals
    int wave;          // the first level the wave tail runs, or -1
};

// (y, x) of entry t in rows of width w: a shift when w is a power of two (lw >= 0), a division otherwise
__device__ __forceinline__ int row_of(int t, int w, int lw) { return lw >= 0 ? t >> lw : t / w; }

template <bool HEAD>
__global__ void __launch_bounds__(COARSE_NT) k_coarse(CoarseSet C, const float *__restrict__ pack,
                                                      const float *__restrict__ b0, float *__restrict__ z0, int nu1,
                                                      int nu2, float omega, const int *gate) {
    if (gated_off(gate)) return;
    extern __shared__ float4 sm4[];
    float *sm = (float *)sm4;
    const int *offo = C.offo, *offx = C.offx, *offy = C.offy, *offz = C.offz, *offb = C.offb, *offd = C.offd;
    const int offinv = C.offinv, total = C.total, n0 = C.w[0] * C.h[0];
    const int T = HEAD ? 2 : 0;  // the first tail level
    const int ob0 = HEAD ? total : offb[0];
    // the head's registers: the first level's 2x2 block (cell k = 2 j + i at (x0 + i, y0 + j)), the second level's
    // cell (tx, ty)
    const int tx = HEAD ? (int)(threadIdx.x & ((1u << C.ltx) - 1)) : 0, ty = HEAD ? (int)(threadIdx.x >> C.ltx) : 0;
    const int wA = C.w[0], hA = C.h[0], wB = C.w[1], hB = C.h[1], x0 = 2 * tx, y0 = 2 * ty;
    float ab[4], az[4], ao[4], ard[4], al[4], ar[4], au[4], ad[4];
    float bb = 0.0f, zb = 0.0f, bo = 0.0f, brd = 0.0f, bl = 0.0f, br = 0.0f, bu = 0.0f, bd = 0.0f;
    const bool vb = HEAD && tx < wB && ty < hB;
    // one phase: every load issued before any is used (the pack's LDS part as float4, level 0's b or the head's
    // registers), then zeros elsewhere (every entry a sweep may read at a grid's rim is finite)
    {
        constexpr int V4 = (COARSE_LDS / 16 + COARSE_NT - 1) / COARSE_NT, VB = COARSE_MAX / COARSE_NT;
        const int np4 = C.pack_floats >> 2;
        const float4 *p4 = (const float4 *)pack;
        float4 v[V4];
        float vb0[VB];
#pragma unroll
        for (int k = 0; k < V4; k++) {
            const int t = threadIdx.x + k * COARSE_NT;
            v[k] = t < np4 ? p4[t] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        }
        if constexpr (HEAD) {
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const int x = x0 + (k & 1), y = y0 + (k >> 1);
                const bool in = x < wA && y < hA;
                const int i = y * wA + x;
                ab[k] = in ? b0[i] : 0.0f;
                ao[k] = in ? pack[offo[0] + i] : 0.0f;
                ard[k] = in ? pack[offd[0] + i] : 0.0f;
                al[k] = in ? pack[offx[0] + y * (wA + 1) + x] : 0.0f;
                ar[k] = in ? pack[offx[0] + y * (wA + 1) + x + 1] : 0.0f;
                au[k] = in ? pack[offy[0] + i] : 0.0f;
                ad[k] = in ? pack[offy[0] + i + wA] : 0.0f;
                az[k] = 0.0f;
            }
            if (vb) {
                const int i = ty * wB + tx;
                bo = pack[offo[1] + i];
                brd = pack[offd[1] + i];
                bl = pack[offx[1] + ty * (wB + 1) + tx];
                br = pack[offx[1] + ty * (wB + 1) + tx + 1];
                bu = pack[offy[1] + i];
                bd = pack[offy[1] + i + wB];
            }
        } else {
#pragma unroll
            for (int k = 0; k < VB; k++) {
                const int t = threadIdx.x + k * COARSE_NT;
                vb0[k] = t < n0 ? b0[t] : 0.0f;
            }
        }
        const int tail = 4 * np4 + (int)threadIdx.x;
        const float vt = tail < C.pack_floats ? pack[tail] : 0.0f;
#pragma unroll
        for (int k = 0; k < V4; k++) {
            const int t = threadIdx.x + k * COARSE_NT;
            if (t < np4) sm4[t] = v[k];
        }
        if constexpr (!HEAD) {
#pragma unroll
            for (int k = 0; k < VB; k++) {
                const int t = threadIdx.x + k * COARSE_NT;
                if (t < n0) sm[ob0 + t] = vb0[k];
            }
        }
        if (tail < C.pack_floats) sm[tail] = vt;
    }
    for (int t = C.pack_floats + threadIdx.x; t < total; t += COARSE_NT)
        if (t < ob0 || t >= ob0 + n0) sm[t] = 0.0f;
    __syncthreads();
    // one half sweep of `color` on tail level l, colour-compact. Couplings on the grid's rim are zero, so the
    // neighbour entries read there (the adjacent LDS arrays, all finite) contribute nothing.
    auto sweep = [&](int l, int color) {
        const int w = C.w[l], h = C.h[l], hw = (w + 1) >> 1, lh = C.lw[l] > 0 ? C.lw[l] - 1 : -1;
        const float *alx = sm + offx[l], *auy = sm + offy[l], *b = sm + offb[l], *rd = sm + offd[l];
        float *z = sm + offz[l];
        for (int t = threadIdx.x; t < hw * h; t += COARSE_NT) {
            const int y = row_of(t, hw, lh);
            const int x = 2 * (t - y * hw) + ((color + y) & 1);
            if (x >= w) continue;
            const int i = y * w + x;
            const float a0 = alx[y * (w + 1) + x], a1 = alx[y * (w + 1) + x + 1], a2 = auy[i], a3 = auy[i + w];
            z[i] = (b[i] + a0 * z[i - 1] + a1 * z[i + 1] + a2 * z[i - w] + a3 * z[i + w]) * rd[i];
        }
        __syncthreads();
    };
    // pre-smoothing of tail level l from z = 0 and the restriction of its residual into level l + 1's b
    auto down = [&](int l) {
        for (int k = 0; k < nu1; k++) {
            sweep(l, 0);
            sweep(l, 1);
        }
        const int w = C.w[l], h = C.h[l], wc = C.w[l + 1], hc = C.h[l + 1];
        const float *own = sm + offo[l], *alx = sm + offx[l], *auy = sm + offy[l], *z = sm + offz[l],
                    *b = sm + offb[l];
        float *bc = sm + offb[l + 1];
        for (int t = threadIdx.x; t < wc * hc; t += COARSE_NT) {
            const int Y = row_of(t, wc, C.lw[l + 1]), X = t - Y * wc;
            float s = 0.0f;
#pragma unroll
            for (int dy = 0; dy < 2; dy++)
#pragma unroll
                for (int dx = 0; dx < 2; dx++) {
                    const int x = 2 * X + dx, y = 2 * Y + dy;
                    if (x >= w || y >= h) continue;
                    const int i = y * w + x;
                    const float a0 = alx[y * (w + 1) + x], a1 = alx[y * (w + 1) + x + 1], a2 = auy[i],
                                a3 = auy[i + w];
                    const float zi = z[i];
                    const float azi = own[i] * zi + a0 * (zi - z[i - 1]) + a1 * (zi - z[i + 1]) +
                                      a2 * (zi - z[i - w]) + a3 * (zi - z[i + w]);
                    s += b[i] - azi;
                }
            bc[t] = s;
        }
        __syncthreads();
    };
    // the prolongation of level l + 1's correction (scaled by omega) and the post-smoothing of tail level l
    auto up = [&](int l) {
        const int w = C.w[l], h = C.h[l], wc = C.w[l + 1];
        float *z = sm + offz[l];
        const float *zc = sm + offz[l + 1];
        for (int t = threadIdx.x; t < w * h; t += COARSE_NT) {
            const int y = row_of(t, w, C.lw[l]), x = t - y * w;
            z[t] += omega * zc[(y >> 1) * wc + (x >> 1)];
        }
        __syncthreads();
        for (int k = 0; k < nu2; k++) {
            sweep(l, 1);
            sweep(l, 0);
        }
    };
    const int lc = C.n - 1, m = C.w[lc] * C.h[lc];
    // the coarsest level's solve
    auto coarsest = [&]() {
        if (m <= DENSE_MAX) {
            // the exact solve by the inverse formed at the system's rebuild
            float *z = sm + offz[lc];
            const float *b = sm + offb[lc], *ainv = sm + offinv;
            if (threadIdx.x < m) {
                float t = 0.0f;
                for (int j = 0; j < m; j++) t += ainv[threadIdx.x * m + j] * b[j];
                z[threadIdx.x] = t;
            }
            __syncthreads();
        } else {
            for (int k = 0; k < COARSE_SWEEPS; k++) {
                sweep(lc, 0);
                sweep(lc, 1);
            }
            for (int k = 0; k < COARSE_SWEEPS; k++) {
                sweep(lc, 1);
                sweep(lc, 0);
            }
        }
    };
    // the head's LDS: the first and second level's published z, the second level's residuals
    float *za = sm + C.offza, *zbl = sm + C.offzb, *rb = sm + C.offrb;
    // a first-level z outside the thread's block (clamped into the grid: a rim coupling is zero)
    auto ZA = [&](int x, int y) {
        return za[min(max(y, 0), hA - 1) * wA + min(max(x, 0), wA - 1)];
    };
    auto ZB = [&](int x, int y) {
        return zbl[min(max(y, 0), hB - 1) * wB + min(max(x, 0), wB - 1)];
    };
    auto publish_a = [&](int k) {
        const int x = x0 + (k & 1), y = y0 + (k >> 1);
        if (x < wA && y < hA) za[y * wA + x] = az[k];
    };
    // a half sweep of the first level: colour 0 holds cells 0 and 3 of the block, colour 1 cells 1 and 2
    auto sweep_a = [&](int color) {
        if (color == 0) {
            az[0] = (ab[0] + al[0] * ZA(x0 - 1, y0) + ar[0] * az[1] + au[0] * ZA(x0, y0 - 1) + ad[0] * az[2]) *
                    ard[0];
            az[3] = (ab[3] + al[3] * az[2] + ar[3] * ZA(x0 + 2, y0 + 1) + au[3] * az[1] +
                     ad[3] * ZA(x0 + 1, y0 + 2)) *
                    ard[3];
            publish_a(0);
            publish_a(3);
        } else {
            az[1] = (ab[1] + al[1] * az[0] + ar[1] * ZA(x0 + 2, y0) + au[1] * ZA(x0 + 1, y0 - 1) + ad[1] * az[3]) *
                    ard[1];
            az[2] = (ab[2] + al[2] * ZA(x0 - 1, y0 + 1) + ar[2] * az[3] + au[2] * az[0] +
                     ad[2] * ZA(x0, y0 + 2)) *
                    ard[2];
            publish_a(1);
            publish_a(2);
        }
        __syncthreads();
    };
    auto sweep_b = [&](int color) {
        if (vb && ((tx + ty) & 1) == color) {
            zb = (bb + bl * ZB(tx - 1, ty) + br * ZB(tx + 1, ty) + bu * ZB(tx, ty - 1) + bd * ZB(tx, ty + 1)) *
                 brd;
            zbl[ty * wB + tx] = zb;
        }
        __syncthreads();
    };
    if constexpr (HEAD) {
        // the first level: pre-smoothing from z = 0, then the residual of the block restricted into the thread's
        // second-level cell (the block's cells in row order, as the tail's restriction sums them)
        for (int k = 0; k < nu1; k++) {
            sweep_a(0);
            sweep_a(1);
        }
        {
            float s = 0.0f;
            const float zn[4][4] = {{ZA(x0 - 1, y0), az[1], ZA(x0, y0 - 1), az[2]},
                                    {az[0], ZA(x0 + 2, y0), ZA(x0 + 1, y0 - 1), az[3]},
                                    {ZA(x0 - 1, y0 + 1), az[3], az[0], ZA(x0, y0 + 2)},
                                    {az[2], ZA(x0 + 2, y0 + 1), az[1], ZA(x0 + 1, y0 + 2)}};
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const int x = x0 + (k & 1), y = y0 + (k >> 1);
                if (x >= wA || y >= hA) continue;
                const float zi = az[k];
                const float azi = ao[k] * zi + al[k] * (zi - zn[k][0]) + ar[k] * (zi - zn[k][1]) +
                                  au[k] * (zi - zn[k][2]) + ad[k] * (zi - zn[k][3]);
                s += ab[k] - azi;
            }
            bb = s;
        }
        // the second level: pre-smoothing, its residual published for the tail's restriction
        for (int k = 0; k < nu1; k++) {
            sweep_b(0);
            sweep_b(1);
        }
        if (vb) {
            const float zi = zb;
            const float azi = bo * zi + bl * (zi - ZB(tx - 1, ty)) + br * (zi - ZB(tx + 1, ty)) +
                              bu * (zi - ZB(tx, ty - 1)) + bd * (zi - ZB(tx, ty + 1));
            rb[ty * wB + tx] = bb - azi;
        }
        __syncthreads();
        {  // the first tail level's b: the second level's residuals summed over each aggregate
            const int wc = C.w[T], hc = C.h[T];
            float *bc = sm + offb[T];
            for (int t = threadIdx.x; t < wc * hc; t += COARSE_NT) {
                const int Y = row_of(t, wc, C.lw[T]), X = t - Y * wc;
                float s = 0.0f;
#pragma unroll
                for (int dy = 0; dy < 2; dy++)
#pragma unroll
                    for (int dx = 0; dx < 2; dx++) {
                        const int x = 2 * X + dx, y = 2 * Y + dy;
                        if (x >= wB || y >= hB) continue;
                        s += rb[y * wB + x];
                    }
                bc[t] = s;
            }
            __syncthreads();
        }
    }
    // The wave tail: the levels from C.wave on (one of at most 16 x 16 cells, then one or two more, the last solved
    // by the dense inverse) on the block's first wave alone, with no block barrier. Lane (tx, ty) holds the first
    // wave level's 2x2 block at (2 tx, 2 ty) in registers and the next level's cell (tx, ty), the block's aggregate;
    // a neighbour in another lane comes by a cross-lane shuffle. Every cell's expression and every sum's order are
    // the LDS levels', so the tail computes the same values (a zero coupling's product may differ in the sign of 0).
    auto wave_tail = [&](int s0) {
        const int lane = (int)threadIdx.x, qx = lane & 7, qy = lane >> 3;
        const int w = C.w[s0], h = C.h[s0], s1 = s0 + 1, w1 = C.w[s1], h1 = C.h[s1];
        auto nb = [&](float v, int src) { return __shfl(v, src, 64); };
        float b[4], z[4], o[4], rd[4], cl[4], cr[4], cu[4], cd[4];
        bool in[4];
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const int x = 2 * qx + (k & 1), y = 2 * qy + (k >> 1), i = y * w + x;
            in[k] = x < w && y < h;
            b[k] = in[k] ? sm[offb[s0] + i] : 0.0f;
            o[k] = in[k] ? sm[offo[s0] + i] : 0.0f;
            rd[k] = in[k] ? sm[offd[s0] + i] : 0.0f;
            cl[k] = in[k] ? sm[offx[s0] + y * (w + 1) + x] : 0.0f;
            cr[k] = in[k] ? sm[offx[s0] + y * (w + 1) + x + 1] : 0.0f;
            cu[k] = in[k] ? sm[offy[s0] + i] : 0.0f;
            cd[k] = in[k] ? sm[offy[s0] + i + w] : 0.0f;
            z[k] = 0.0f;
        }
        // colour 0 holds cells 0 and 3 of the block, colour 1 cells 1 and 2
        auto sweep0 = [&](int color) {
            if (color == 0) {
                const float L1 = nb(z[1], lane - 1), U2 = nb(z[2], lane - 8), R2 = nb(z[2], lane + 1),
                            D1 = nb(z[1], lane + 8);
                if (in[0]) z[0] = (b[0] + cl[0] * L1 + cr[0] * z[1] + cu[0] * U2 + cd[0] * z[2]) * rd[0];
                if (in[3]) z[3] = (b[3] + cl[3] * z[2] + cr[3] * R2 + cu[3] * z[1] + cd[3] * D1) * rd[3];
            } else {
                const float R0 = nb(z[0], lane + 1), U3 = nb(z[3], lane - 8), L3 = nb(z[3], lane - 1),
                            D0 = nb(z[0], lane + 8);
                if (in[1]) z[1] = (b[1] + cl[1] * z[0] + cr[1] * R0 + cu[1] * U3 + cd[1] * z[3]) * rd[1];
                if (in[2]) z[2] = (b[2] + cl[2] * L3 + cr[2] * z[3] + cu[2] * z[0] + cd[2] * D0) * rd[2];
            }
        };
        for (int k = 0; k < nu1; k++) {
            sweep0(0);
            sweep0(1);
        }
        // the residual of the block restricted into the lane's cell of the next level (cells in row order)
        float b1 = 0.0f;
        {
            const float L1 = nb(z[1], lane - 1), U2 = nb(z[2], lane - 8), R0 = nb(z[0], lane + 1),
                        U3 = nb(z[3], lane - 8), L3 = nb(z[3], lane - 1), D0 = nb(z[0], lane + 8),
                        R2 = nb(z[2], lane + 1), D1 = nb(z[1], lane + 8);
            const float zn[4][4] = {{L1, z[1], U2, z[2]}, {z[0], R0, U3, z[3]}, {L3, z[3], z[0], D0},
                                    {z[2], R2, z[1], D1}};
#pragma unroll
            for (int k = 0; k < 4; k++) {
                if (!in[k]) continue;
                const float zi = z[k];
                const float azi = o[k] * zi + cl[k] * (zi - zn[k][0]) + cr[k] * (zi - zn[k][1]) +
                                  cu[k] * (zi - zn[k][2]) + cd[k] * (zi - zn[k][3]);
                b1 += b[k] - azi;
            }
        }
        const bool v1 = qx < w1 && qy < h1;
        const int i1 = qy * w1 + qx;
        // the coarsest's z = A^-1 b in lane t for its cell t; b of cell j from lane src(j), every exchange issued
        // before the sum (in j order) starts
        auto dense = [&](float bv, auto src) {
            float bj[DENSE_MAX];
#pragma unroll
            for (int j = 0; j < DENSE_MAX; j++) bj[j] = nb(bv, src(j));
            float zt = 0.0f;
            if (lane < m) {
#pragma unroll
                for (int j = 0; j < DENSE_MAX; j++)
                    if (j < m) zt += sm[offinv + lane * m + j] * bj[j];
            }
            return zt;
        };
        if (s1 == lc) {
            // the next level is the coarsest
            const float zt = dense(b1, [&](int j) { return (j % w1) + 8 * (j / w1); });
            const float z1 = nb(zt, i1);
#pragma unroll
            for (int k = 0; k < 4; k++)
                if (in[k]) z[k] += omega * z1;
        } else {
            const float o1 = v1 ? sm[offo[s1] + i1] : 0.0f, rd1 = v1 ? sm[offd[s1] + i1] : 0.0f;
            const float l1 = v1 ? sm[offx[s1] + qy * (w1 + 1) + qx] : 0.0f;
            const float r1 = v1 ? sm[offx[s1] + qy * (w1 + 1) + qx + 1] : 0.0f;
            const float u1 = v1 ? sm[offy[s1] + i1] : 0.0f, d1 = v1 ? sm[offy[s1] + i1 + w1] : 0.0f;
            float z1 = 0.0f;
            auto sweep1 = [&](int color) {
                const float L = nb(z1, lane - 1), R = nb(z1, lane + 1), U = nb(z1, lane - 8), D = nb(z1, lane + 8);
                if (v1 && ((qx + qy) & 1) == color) z1 = (b1 + l1 * L + r1 * R + u1 * U + d1 * D) * rd1;
            };
            for (int k = 0; k < nu1; k++) {
                sweep1(0);
                sweep1(1);
            }
            // the residual, restricted into the coarsest (lane t forming its cell t), the dense solve
            float r = 0.0f;
            {
                const float L = nb(z1, lane - 1), R = nb(z1, lane + 1), U = nb(z1, lane - 8), D = nb(z1, lane + 8);
                const float azi = o1 * z1 + l1 * (z1 - L) + r1 * (z1 - R) + u1 * (z1 - U) + d1 * (z1 - D);
                r = b1 - azi;
            }
            const int w2 = C.w[lc], X = lane % w2, Y = lane / w2;
            float b2 = 0.0f;
#pragma unroll
            for (int dy = 0; dy < 2; dy++)
#pragma unroll
                for (int dx = 0; dx < 2; dx++) {
                    const int x = 2 * X + dx, y = 2 * Y + dy;
                    const float rv = nb(r, x + 8 * y);
                    if (lane < m && x < w1 && y < h1) b2 += rv;
                }
            const float zt = dense(b2, [](int j) { return j; });
            // the coarsest's correction prolonged, the post-smoothing
            const float zc = nb(zt, (qy >> 1) * w2 + (qx >> 1));
            if (v1) z1 += omega * zc;
            for (int k = 0; k < nu2; k++) {
                sweep1(1);
                sweep1(0);
            }
#pragma unroll
            for (int k = 0; k < 4; k++)
                if (in[k]) z[k] += omega * z1;
        }
        for (int k = 0; k < nu2; k++) {
            sweep0(1);
            sweep0(0);
        }
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const int x = 2 * qx + (k & 1), y = 2 * qy + (k >> 1);
            if (in[k]) sm[offz[s0] + y * w + x] = z[k];
        }
    };
    const int S = C.wave >= 0 ? C.wave : lc;
    for (int l = T; l < S; l++) down(l);
    if (C.wave >= 0) {
        if (threadIdx.x < 64) wave_tail(S);
        __syncthreads();
    } else {
        coarsest();
    }
    for (int l = S - 1; l >= T; l--) up(l);
    if constexpr (HEAD) {
        // the second level: the tail's correction prolonged, published, post-smoothed
        if (vb) {
            zb += omega * sm[offz[T] + (ty >> 1) * C.w[T] + (tx >> 1)];
            zbl[ty * wB + tx] = zb;
        }
        __syncthreads();
        for (int k = 0; k < nu2; k++) {
            sweep_b(1);
            sweep_b(0);
        }
        // the first level: the second level's correction (the thread's own cell) prolonged, published,
        // post-smoothed, written out
#pragma unroll
        for (int k = 0; k < 4; k++) {
            az[k] += omega * zb;
            publish_a(k);
        }
        __syncthreads();
        for (int k = 0; k < nu2; k++) {
            sweep_a(1);
            sweep_a(0);
        }
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const int x = x0 + (k & 1), y = y0 + (k >> 1);
            if (x < wA && y < hA) z0[y * wA + x] = az[k];
        }
    } else {
        for (int t = threadIdx.x; t < n0; t += COARSE_NT) z0[t] = sm[offz[0] + t];
    }
}

// ------------------------------------------------------------------------------------------------ hierarchy build

// the tile smoothers' per-cell record {own, left face, up face, 1 / (own + all four faces)}
__global__ void k_pack_cf(Lv L, float4 *__restrict__ cf) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= L.w * L.h) return;
    const int x = i % L.w, y = i / L.w;
    const float al = L.ax[(size_t)y * (L.w + 1) + x], ar = L.ax[(size_t)y * (L.w + 1) + x + 1];
    const float au = L.ay[i], ad = L.ay[(size_t)i + L.w];
    const float own = L.own[i];
    cf[i] = make_float4(own, al, au, 1.0f / (own + al + ar + au + ad));
}

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

// Once per rebuild, for the single-block V-cycle: every cell's inverse diagonal 1 / (own + its four couplings) on
// every level of the set, and, when the coarsest holds at most DENSE_MAX cells, the inverse of its dense matrix by
// Gauss-Jordan elimination (the matrix is SPD, so no pivoting); both into the pack. The elimination runs on the first
// wave alone, row i of [A | I] in lane i's registers and the pivot row read across lanes, so it needs no block
// barrier; the other waves form the inverse diagonals meanwhile.
__global__ void __launch_bounds__(COARSE_NT) k_coarse_prep(CoarseSet C, float *__restrict__ pack) {
    const int lc = C.n - 1, wl = C.w[lc], m = wl * C.h[lc];
    const bool dense = m <= DENSE_MAX;
    if (dense && threadIdx.x < 64) {
        const int i = threadIdx.x;
        const float *own = pack + C.offo[lc], *ax = pack + C.offx[lc], *ay = pack + C.offy[lc];
        float A[2 * DENSE_MAX];
#pragma unroll
        for (int j = 0; j < 2 * DENSE_MAX; j++) A[j] = 0.0f;
        if (i < m) {
            const int x = i % wl, y = i / wl;
            const float a[4] = {ax[y * (wl + 1) + x], ax[y * (wl + 1) + x + 1], ay[y * wl + x], ay[(y + 1) * wl + x]};
            const int nb[4] = {i - 1, i + 1, i - wl, i + wl};
            float d = own[i];
#pragma unroll
            for (int f = 0; f < 4; f++)
                if (a[f] != 0.0f) {
                    d += a[f];
#pragma unroll
                    for (int j = 0; j < DENSE_MAX; j++)
                        if (j == nb[f]) A[j] -= a[f];
                }
#pragma unroll
            for (int j = 0; j < DENSE_MAX; j++)
                if (j == i) A[j] += d;
#pragma unroll
            for (int j = 0; j < 2 * DENSE_MAX; j++)
                if (j == m + i) A[j] = 1.0f;
        }
#pragma unroll
        for (int k = 0; k < DENSE_MAX; k++) {
            if (k >= m) break;
            const float piv = __shfl(A[k], k, 64);
            const float f = (i < m && i != k) ? A[k] / piv : 0.0f;
#pragma unroll
            for (int j = 0; j < 2 * DENSE_MAX; j++) {
                const float pk = __shfl(A[j], k, 64);
                if (i < m && i != k) A[j] -= f * pk;
            }
        }
        if (i < m) {
            float d = 0.0f;
#pragma unroll
            for (int j = 0; j < DENSE_MAX; j++)
                if (j == i) d = A[j];
            float *ainv = pack + C.offinv;
#pragma unroll
            for (int j = 0; j < 2 * DENSE_MAX; j++)
                if (j >= m && j < 2 * m) ainv[i * m + j - m] = A[j] / d;
        }
    }
    const int t0 = dense ? (int)threadIdx.x - 64 : (int)threadIdx.x, nt = dense ? COARSE_NT - 64 : COARSE_NT;
    if (t0 < 0) return;
    for (int l = 0; l < C.n; l++) {
        const int w = C.w[l], h = C.h[l];
        const float *own = pack + C.offo[l], *al = pack + C.offx[l], *au = pack + C.offy[l];
        float *rd = pack + C.offd[l];
        for (int t = t0; t < w * h; t += nt) {
            const int y = t / w, x = t - y * w;
            rd[t] = 1.0f / (own[t] + al[y * (w + 1) + x] + al[y * (w + 1) + x + 1] + au[t] + au[t + w]);
        }
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

#define PTX 64
#define PTY 32

static bool is_coarsest(int w, int h) { return w * h <= DENSE_MAX; }

FldSolver::~FldSolver() {
    for (size_t l = 0; l < lv.size(); l++) {
        if ((int)l < coarse_from) {
            (void)hipFree(lv[l].own);
            (void)hipFree(lv[l].ax);
            (void)hipFree(lv[l].ay);
        }
        (void)hipFree(lv[l].z);
        (void)hipFree(lv[l].b);
        (void)hipFree(lv[l].zp);
        if (lv[l].cf) (void)hipFree(lv[l].cf);
    }
    void *ps[] = {gas, r32, r32b, za, zb, p, q, pack, part_true, part_pq, part_rr, part_rr_t, part_dots};
    for (void *x : ps)
        if (x) (void)hipFree(x);
    if (host_part) (void)hipHostFree(host_part);
    if (host_ctl) (void)hipHostFree(host_ctl);
    if (ev_part) (void)hipEventDestroy(ev_part);
    for (hipGraphExec_t g : rebuild_exec)
        if (g) (void)hipGraphExecDestroy(g);
    if (ctl) (void)hipFree(ctl);
    if (ticket) (void)hipFree(ticket);
    if (slot_exec) (void)hipGraphExecDestroy(slot_exec);
    for (hipEvent_t e : evs) (void)hipEventDestroy(e);
}

static size_t level_floats(int w, int h) { return (size_t)w * h + (size_t)(w + 1) * h + (size_t)w * (h + 1); }

int FldSolver::init(int w, int h, int levels, int nu1_, int nu2_, double omega_) {
    if (w <= 0 || h <= 0) return fail(SR_ERR_ARGUMENT, "an empty FLD grid");
    if (nu1_ < 1 || nu1_ > 3 || nu2_ < 1 || nu2_ > 3) return fail(SR_ERR_ARGUMENT, "nu1, nu2 must be in 1..3");
    if (omega_ == 0.0) omega_ = 1.0;
    if (!(omega_ > 0.0 && omega_ < 2.0))
        return fail(SR_ERR_ARGUMENT, "the coarse correction scale must lie in (0, 2)");
    omega = (float)omega_;
    w0 = w;
    h0 = h;
    nu1 = nu1_;
    nu2 = nu2_;
    // the hierarchy's shape: halve until a level holds at most DENSE_MAX cells (or the level cap)
    int cw = w, ch = h;
    const int maxl = levels > 0 ? levels : 64;
    while (true) {
        FldLevel l;
        l.w = cw;
        l.h = ch;
        lv.push_back(l);
        if ((int)lv.size() >= maxl || is_coarsest(cw, ch)) break;
        cw = (cw + 1) / 2;
        ch = (ch + 1) / 2;
    }
    // the first level handled by the single-block V-cycle: from there down every level fits COARSE_MAX cells and
    // the packed arrays plus z and b fit the block's LDS
    coarse_from = (int)lv.size() - 1;
    while (coarse_from > 0) {
        const int c = coarse_from - 1;
        if (lv[c].w * lv[c].h > COARSE_MAX || (int)lv.size() - c > 16) break;
        size_t floats = DENSE_MAX * DENSE_MAX;
        for (int l = c; l < (int)lv.size(); l++) floats += level_floats(lv[l].w, lv[l].h) + 3 * lv[l].w * lv[l].h;
        if (floats * sizeof(float) > COARSE_LDS) break;
        coarse_from = c;
    }
    {
        const FldLevel &c = lv.back();
        size_t floats = DENSE_MAX * DENSE_MAX + level_floats(c.w, c.h) + 3 * c.w * c.h;
        if (c.w * c.h > COARSE_MAX || floats * sizeof(float) > COARSE_LDS)
            return fail(SR_ERR_ARGUMENT, "too few multigrid levels: the coarsest level must fit one block (" +
                                             std::to_string(COARSE_MAX) + " cells)");
    }
    // the register-resident head: the set's two largest levels when the set has three or more and its first level
    // fits the block's cover of 2 TX x 2 TY cells, TX a power of two and TX TY = COARSE_NT
    head = 0;
    ltx = 0;
    if ((int)lv.size() - coarse_from >= 3) {
        const int wa = lv[coarse_from].w, ha = lv[coarse_from].h;
        int tx = 1;
        while (2 * tx < wa) tx *= 2;
        if (tx <= COARSE_NT && 2 * (COARSE_NT / tx) >= ha) {
            head = 2;
            ltx = ilog2(tx);
        }
    }
    pack_floats = DENSE_MAX * DENSE_MAX;
    for (int l = coarse_from; l < (int)lv.size(); l++)
        pack_floats += (int)level_floats(lv[l].w, lv[l].h) + lv[l].w * lv[l].h;  // + the inverse diagonals
    SR_HIP(hipMalloc(&pack, sizeof(float) * pack_floats));
    SR_HIP(hipMemset(pack, 0, sizeof(float) * pack_floats));
    {
        const CoarseSet C = coarse_set();
        for (int k = 0; k < C.n; k++) {
            FldLevel &L = lv[coarse_from + k];
            L.own = pack + C.offo[k];
            L.ax = pack + C.offx[k];
            L.ay = pack + C.offy[k];
        }
    }
    for (int l = 0; l < (int)lv.size(); l++) {
        FldLevel &L = lv[l];
        if (l < coarse_from) {
            SR_HIP(hipMalloc(&L.own, sizeof(float) * L.w * L.h));
            SR_HIP(hipMalloc(&L.ax, sizeof(float) * (L.w + 1) * L.h));
            SR_HIP(hipMalloc(&L.ay, sizeof(float) * L.w * (L.h + 1)));
        }
        SR_HIP(hipMalloc(&L.z, sizeof(float) * L.w * L.h));
        SR_HIP(hipMalloc(&L.b, sizeof(float) * L.w * L.h));
        SR_HIP(hipMalloc(&L.zp, sizeof(float) * L.w * L.h));
        if (l < coarse_from) SR_HIP(hipMalloc(&L.cf, sizeof(float4) * L.w * L.h));
    }
    const size_t n = (size_t)w * h;
    nbe = (int)std::max<size_t>(1, std::min<size_t>(NB_MAX, (n + NT * UNROLL - 1) / (NT * UNROLL)));
    // until the autotuner runs: 32x32 tiles, 16x16 on levels below level 0 that would launch fewer than 512 blocks of
    // 32x32 (four times the blocks and a quarter of the serial work per thread where the launch is latency-bound)
    pre_tile.assign(coarse_from, TILE_32);
    post_tile.assign(coarse_from, TILE_32);
    for (int l = 1; l < coarse_from; l++)
        if (((lv[l].w + 31) / 32) * ((lv[l].h + 31) / 32) < 512) pre_tile[l] = post_tile[l] = TILE_16;
    ntiles0 = ((w + 31) / 32) * ((h + 31) / 32);
    nbd_max = std::max(nbe, ntiles0);
    SR_HIP(hipMalloc(&gas, n));
    SR_HIP(hipMemset(gas, 1, n));
    SR_HIP(hipMalloc(&r32, sizeof(float) * n));
    SR_HIP(hipMalloc(&r32b, sizeof(float) * n));
    SR_HIP(hipMalloc(&za, sizeof(float) * n));
    SR_HIP(hipMalloc(&zb, sizeof(float) * n));
    SR_HIP(hipMalloc(&p, sizeof(float) * 2 * n));
    SR_HIP(hipMalloc(&q, sizeof(float) * n));
    SR_HIP(hipMemset(za, 0, sizeof(float) * n));
    SR_HIP(hipMemset(zb, 0, sizeof(float) * n));
    SR_HIP(hipMemset(p, 0, sizeof(float) * 2 * n));
    SR_HIP(hipMalloc(&part_true, sizeof(double) * 4 * nbe));
    SR_HIP(hipMalloc(&part_pq, sizeof(double) * ((w + NT - 1) / NT) * h));
    SR_HIP(hipMalloc(&part_rr, sizeof(double) * nbe));
    SR_HIP(hipMalloc(&part_rr_t, sizeof(double) * ntiles0));
    SR_HIP(hipMalloc(&part_dots, sizeof(double) * 2 * 2 * nbd_max));
    SR_HIP(hipMemset(part_dots, 0, sizeof(double) * 2 * 2 * nbd_max));
    SR_HIP(hipHostMalloc(&host_part, sizeof(double) * std::max(4 * nbe, ntiles0)));
    SR_HIP(hipEventCreateWithFlags(&ev_part, hipEventDisableTiming));
    SR_HIP(hipMalloc(&ctl, sizeof(SolveCtl) * 2));
    SR_HIP(hipMemset(ctl, 0, sizeof(SolveCtl) * 2));
    SR_HIP(hipMalloc(&ticket, sizeof(unsigned)));
    SR_HIP(hipMemset(ticket, 0, sizeof(unsigned)));
    SR_HIP(hipHostMalloc(&host_ctl, sizeof(SolveCtl)));
    SR_HIP(hipFuncSetAttribute((const void *)k_coarse<false>, hipFuncAttributeMaxDynamicSharedMemorySize,
                               COARSE_LDS));
    SR_HIP(hipFuncSetAttribute((const void *)k_coarse<true>, hipFuncAttributeMaxDynamicSharedMemorySize,
                               COARSE_LDS));
    return SR_OK;
}

int FldSolver::set_system(const float *own, const float *ax, const float *ay, const uint8_t *g, hipStream_t s) {
    plain_rate = 0.0;
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

// The single-block levels' layout. The pack holds first the tail levels' own, ax, ay, the coarsest's inverse and the
// tail levels' inverse diagonals (the part the block copies into LDS, at the same offsets there), then the head
// levels' own, ax, ay and inverse diagonals (read into registers). In LDS after the tail part: z and b per tail level,
// then the head's published z of its two levels and the second level's residuals.
CoarseSet FldSolver::coarse_set() const {
    const int nl = (int)lv.size();
    CoarseSet C;
    C.n = nl - coarse_from;
    C.head = head;
    C.ltx = ltx;
    int o = 0;
    auto level = [&](int k) {
        const int w = lv[coarse_from + k].w, h = lv[coarse_from + k].h;
        C.offo[k] = o;
        o += w * h;
        C.offx[k] = o;
        o += (w + 1) * h;
        C.offy[k] = o;
        o += w * (h + 1);
    };
    for (int k = 0; k < C.n; k++) {
        C.w[k] = lv[coarse_from + k].w;
        C.h[k] = lv[coarse_from + k].h;
        C.lw[k] = is_pow2(C.w[k]) ? ilog2(C.w[k]) : -1;
    }
    for (int k = head; k < C.n; k++) level(k);
    C.offinv = o;
    o += DENSE_MAX * DENSE_MAX;
    for (int k = head; k < C.n; k++) {
        C.offd[k] = o;
        o += C.w[k] * C.h[k];
    }
    C.pack_floats = o;
    for (int k = 0; k < head; k++) level(k);
    for (int k = 0; k < head; k++) {
        C.offd[k] = o;
        o += C.w[k] * C.h[k];
    }
    // o is the whole pack's size here
    o = C.pack_floats;
    for (int k = head; k < C.n; k++) {
        C.offz[k] = o;
        o += C.w[k] * C.h[k];
        C.offb[k] = o;
        o += C.w[k] * C.h[k];
    }
    C.offza = C.offzb = C.offrb = o;
    if (head) {
        C.offza = o;
        o += C.w[0] * C.h[0];
        C.offzb = o;
        o += C.w[1] * C.h[1];
        C.offrb = o;
        o += C.w[1] * C.h[1];
    }
    C.total = o;
    // the wave tail: from the first level past the head of at most 16 x 16 cells, when one or two levels follow it
    // and the last is solved by the dense inverse
    C.wave = -1;
    if (coarse_wave) {
        int s = head;
        while (s < C.n && (C.w[s] > 16 || C.h[s] > 16)) s++;
        const int lc = C.n - 1;
        if (s < C.n && (lc - s == 1 || lc - s == 2) && C.w[lc] * C.h[lc] <= DENSE_MAX) C.wave = s;
    }
    return C;
}

// The hierarchy from level 0's arrays: a chain of small launches (the coarsenings level by level, the tile levels'
// records, the single block's preparation) whose arguments are fixed for the solver's life, so on a stream it is
// captured once per variant (with or without the faces) and layout of the single-block set, and replayed: the
// device then runs the chain back to back instead of waiting on each launch from the host.
int FldSolver::rebuild(bool faces, hipStream_t s) {
    if (faces)
        work.builds_faces++;
    else
        work.builds_own++;
    auto launch = [&]() -> int {
        for (size_t l = 0; l + 1 < lv.size(); l++) {
            const FldLevel &f = lv[l], &c = lv[l + 1];
            Lv fl = {f.w, f.h, f.own, f.ax, f.ay, nullptr};
            const int nc = c.w * c.h;
            hipLaunchKernelGGL(k_coarsen_own, dim3((nc + 255) / 256), dim3(256), 0, s, fl, l == 0 ? gas : nullptr,
                               c.own, c.w, c.h);
            if (faces) {
                const int nf = std::max((c.w + 1) * c.h, c.w * (c.h + 1));
                hipLaunchKernelGGL(k_coarsen_faces, dim3((nf + 255) / 256), dim3(256), 0, s, fl, c.ax, c.ay, c.w,
                                   c.h);
            }
        }
        for (int l = 0; l < coarse_from; l++) {
            const FldLevel &f = lv[l];
            Lv fl = {f.w, f.h, f.own, f.ax, f.ay, nullptr};
            hipLaunchKernelGGL(k_pack_cf, dim3((f.w * f.h + 255) / 256), dim3(256), 0, s, fl, f.cf);
        }
        hipLaunchKernelGGL(k_coarse_prep, dim3(1), dim3(COARSE_NT), 0, s, coarse_set(), pack);
        SR_HIP(hipGetLastError());
        return SR_OK;
    };
    hipGraphExec_t &g = rebuild_exec[faces ? 1 : 0];
    if (g && rebuild_key[faces ? 1 : 0] != tiles_version) {
        (void)hipGraphExecDestroy(g);
        g = nullptr;
    }
    bool replayed = false;
    bool capturing = false;
    if (s != nullptr && !g) {
        capturing = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) == hipSuccess;
        if (!capturing) (void)hipGetLastError();  // a stream that cannot be captured: direct launches
    }
    if (capturing) {
        hipGraph_t graph = nullptr;
        const int rc = launch();
        const hipError_t e = hipStreamEndCapture(s, &graph);
        SR_TRY(rc);
        if (e != hipSuccess) return fail(SR_ERR_DEVICE, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
        const hipError_t ei = hipGraphInstantiate(&g, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (ei != hipSuccess) {
            g = nullptr;
            return fail(SR_ERR_DEVICE, std::string("hipGraphInstantiate: ") + hipGetErrorString(ei));
        }
        rebuild_key[faces ? 1 : 0] = tiles_version;
    }
    if (s != nullptr && g) {
        SR_HIP(hipGraphLaunch(g, s));
        replayed = true;
    }
    if (!replayed) SR_TRY(launch());
    if (!tuned) {
        tuned = true;
        if (autotune) SR_TRY(tune(s));
        tiles_version++;
    }
    return SR_OK;
}

// The autotuner: on the system just built, each tile-smoothed level's pre- and post-smoother below level 0 is timed
// in every tile shape (one launch to warm, then the best of three means of eight launches, each between two events)
// and keeps the fastest; a shape must beat the one kept by 3 % to replace it, so near-ties keep the default order's
// choice. Level 0 keeps 32x32 tiles: it is the HBM-bound level the shape was chosen for by measurement, its
// post-smoother's dot partials are counted per tile, and its pre-smoother is the kernel the bench's roofline names.
// The smoothers compute exact red-black Gauss-Seidel sweeps whatever the tiling, so the choice changes no result bit.
// The launches write only the hierarchy's own work arrays (z, the coarse b, zp), which every V-cycle rewrites before
// it reads them.
int FldSolver::tune(hipStream_t s) {
    if (coarse_from == 0) return SR_OK;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    SR_HIP(hipEventCreate(&e0));
    SR_HIP(hipEventCreate(&e1));
    std::vector<int> shapes = {TILE_32, TILE_16};
    if (nu1 <= 2 && nu2 <= 2) shapes.push_back(TILE_64);
    auto time_of = [&](auto launch, float *ms) -> int {
        SR_TRY(launch());
        *ms = 0.0f;
        for (int trial = 0; trial < 3; trial++) {
            SR_HIP(hipEventRecord(e0, s));
            for (int k = 0; k < 8; k++) SR_TRY(launch());
            SR_HIP(hipEventRecord(e1, s));
            SR_HIP(hipEventSynchronize(e1));
            float t = 0.0f;
            SR_HIP(hipEventElapsedTime(&t, e0, e1));
            if (trial == 0 || t < *ms) *ms = t;
        }
        return SR_OK;
    };
    int rc = SR_OK;
    for (int l = 1; l < coarse_from && rc == SR_OK; l++) {
        const FldLevel &f = lv[l];
        const float *b = f.b;
        for (int kind = 0; kind < 2 && rc == SR_OK; kind++) {
            std::vector<int> &tile = kind == 0 ? pre_tile : post_tile;
            int best = tile[l];
            float best_ms = 0.0f;
            for (int shape : shapes) {
                tile[l] = shape;
                float ms = 0.0f;
                rc = kind == 0 ? time_of([&] { return launch_pre_level(l, b, s); }, &ms)
                               : time_of([&] { return launch_post_level(l, b, f.zp, nullptr, nullptr, nullptr, s); },
                                         &ms);
                if (rc != SR_OK) break;
                if (shape == shapes[0] || ms < 0.97f * best_ms) {
                    best = shape;
                    best_ms = ms;
                }
            }
            tile[l] = best;
        }
    }
    // the last two tile-smoothed levels: their two launches with the shapes just chosen against the fused pair in
    // each of its tilings (the large one for nu <= 2), per direction, the fused form kept when it is 3 % faster
    const int pa = pair_level();
    if (pa >= 1 && !fuse_fixed && rc == SR_OK) {
        std::vector<int> forms = {FUSE_OFF, FUSE_SMALL};
        if (nu1 <= 2 && nu2 <= 2) forms.push_back(FUSE_LARGE);
        const FldLevel &fa = lv[pa], &fb = lv[pa + 1];
        for (int kind = 0; kind < 2 && rc == SR_OK; kind++) {
            int &form = kind == 0 ? fuse_pre : fuse_post;
            int best = FUSE_OFF;
            float best_ms = 0.0f;
            for (int f : forms) {
                form = f;
                float ms = 0.0f;
                if (kind == 0)
                    rc = time_of(
                        [&] {
                            if (f != FUSE_OFF) return launch_pre_pair(pa, fa.b, s);
                            SR_TRY(launch_pre_level(pa, fa.b, s));
                            return launch_pre_level(pa + 1, fb.b, s);
                        },
                        &ms);
                else
                    rc = time_of(
                        [&] {
                            if (f != FUSE_OFF) return launch_post_pair(pa, s);
                            SR_TRY(launch_post_level(pa + 1, fb.b, fb.zp, nullptr, nullptr, nullptr, s));
                            return launch_post_level(pa, fa.b, fa.zp, nullptr, nullptr, nullptr, s);
                        },
                        &ms);
                if (rc != SR_OK) break;
                if (f == forms[0] || ms < 0.97f * best_ms) {
                    best = f;
                    best_ms = ms;
                }
            }
            form = best;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return rc;
}

int FldSolver::tiles(int *pre, int *post, int max_levels, int *levels) const {
    if (levels) *levels = coarse_from;
    for (int l = 0; l < std::min(max_levels, coarse_from); l++) {
        if (pre) pre[l] = pre_tile[l];
        if (post) post[l] = post_tile[l];
    }
    return SR_OK;
}

int FldSolver::set_tiles(const int *pre, const int *post, int levels) {
    if (!pre || !post || levels != coarse_from)
        return fail(SR_ERR_ARGUMENT, "set_tiles: one pre and one post shape per tile-smoothed level (" +
                                         std::to_string(coarse_from) + ")");
    for (int l = 0; l < levels; l++) {
        for (int v : {pre[l], post[l]})
            if (v != TILE_16 && v != TILE_32 && !(v == TILE_64 && nu1 <= 2 && nu2 <= 2))
                return fail(SR_ERR_ARGUMENT, "set_tiles: an unknown tile shape (or 64x32 with more than 2 sweeps)");
    }
    if (levels > 0 && (pre[0] != TILE_32 || post[0] != TILE_32))
        return fail(SR_ERR_ARGUMENT, "set_tiles: level 0 keeps 32x32 tiles");
    pre_tile.assign(pre, pre + levels);
    post_tile.assign(post, post + levels);
    tuned = true;
    autotune = false;
    tiles_version++;
    return SR_OK;
}

int FldSolver::fusion(int *pre, int *post) const {
    if (pre) *pre = pair_level() >= 1 ? fuse_pre : FUSE_OFF;
    if (post) *post = pair_level() >= 1 ? fuse_post : FUSE_OFF;
    return SR_OK;
}

int FldSolver::set_fusion(int pre, int post) {
    for (int v : {pre, post})
        if (v != FUSE_OFF && v != FUSE_SMALL && v != FUSE_LARGE)
            return fail(SR_ERR_ARGUMENT, "set_fusion: SR_FUSE_OFF, SR_FUSE_SMALL or SR_FUSE_LARGE");
    if ((pre == FUSE_LARGE && nu1 > 2) || (post == FUSE_LARGE && nu2 > 2))
        return fail(SR_ERR_ARGUMENT, "set_fusion: the large fused tiles serve nu1, nu2 <= 2 only");
    fuse_pre = pre;
    fuse_post = post;
    fuse_fixed = true;
    tiles_version++;
    return SR_OK;
}

// the 64x32 tile serves nu <= 2 only (its nu = 3 region would not leave room for two blocks per CU)
template <int TX, int TY>
static void launch_pre(int nu, hipStream_t s, Lv L, const float *b, float *z, float *bc, int wc, const int *gate) {
    const dim3 grid((L.w + TX - 1) / TX, (L.h + TY - 1) / TY);
    if (nu == 1)
        hipLaunchKernelGGL((k_pre<TX, TY, 1>), grid, dim3(NT), 0, s, L, b, z, bc, wc, gate);
    else if (nu == 2)
        hipLaunchKernelGGL((k_pre<TX, TY, 2>), grid, dim3(NT), 0, s, L, b, z, bc, wc, gate);
    else if constexpr (TX <= 32)
        hipLaunchKernelGGL((k_pre<TX, TY, 3>), grid, dim3(NT), 0, s, L, b, z, bc, wc, gate);
}

template <int TX, int TY>
static void launch_pre_upd(int nu, hipStream_t s, Lv L, const PendingUpdate &U, float *z, float *bc, int wc) {
    const dim3 grid((L.w + TX - 1) / TX, (L.h + TY - 1) / TY);
    if (nu == 1)
        hipLaunchKernelGGL((k_pre_upd<TX, TY, 1>), grid, dim3(NT), 0, s, L, U, z, bc, wc);
    else if (nu == 2)
        hipLaunchKernelGGL((k_pre_upd<TX, TY, 2>), grid, dim3(NT), 0, s, L, U, z, bc, wc);
    else
        hipLaunchKernelGGL((k_pre_upd<TX, TY, 3>), grid, dim3(NT), 0, s, L, U, z, bc, wc);
}

template <int TX, int TY>
static int launch_post(int nu, hipStream_t s, Lv L, const float *b, const float *zin, const float *zc, int wc,
                       float omega, float *zout, const float *zold, double *part, const int *gate) {
    const dim3 grid((L.w + TX - 1) / TX, (L.h + TY - 1) / TY);
    if (nu == 1)
        hipLaunchKernelGGL((k_post<TX, TY, 1>), grid, dim3(NT), 0, s, L, b, zin, zc, wc, omega, zout, zold, part,
                           gate);
    else if (nu == 2)
        hipLaunchKernelGGL((k_post<TX, TY, 2>), grid, dim3(NT), 0, s, L, b, zin, zc, wc, omega, zout, zold, part,
                           gate);
    else if constexpr (TX <= 32)
        hipLaunchKernelGGL((k_post<TX, TY, 3>), grid, dim3(NT), 0, s, L, b, zin, zc, wc, omega, zout, zold, part,
                           gate);
    return (int)(grid.x * grid.y);
}

// the fused pair's launches: level b's FB x FB tiles before, level a's TA x TA tiles after
template <int FB>
static void launch_pre2(int nu, hipStream_t s, Lv La, const float *ba, float *za, Lv Lb, float *bb, float *zb,
                        float *bc, int wc, const int *gate) {
    const dim3 grid((Lb.w + FB - 1) / FB, (Lb.h + FB - 1) / FB);
    if (nu == 1)
        hipLaunchKernelGGL((k_pre2<FB, 1>), grid, dim3(NT), 0, s, La, ba, za, Lb, bb, zb, bc, wc, gate);
    else if (nu == 2)
        hipLaunchKernelGGL((k_pre2<FB, 2>), grid, dim3(NT), 0, s, La, ba, za, Lb, bb, zb, bc, wc, gate);
    else if constexpr (FB <= 8)
        hipLaunchKernelGGL((k_pre2<FB, 3>), grid, dim3(NT), 0, s, La, ba, za, Lb, bb, zb, bc, wc, gate);
}

template <int TA>
static void launch_post2(int nu, hipStream_t s, Lv La, const float *ba, const float *za, Lv Lb, const float *bb,
                         const float *zb, const float *zc, int wc, float omega, float *zout, const int *gate) {
    const dim3 grid((La.w + TA - 1) / TA, (La.h + TA - 1) / TA);
    if (nu == 1)
        hipLaunchKernelGGL((k_post2<TA, 1>), grid, dim3(NT), 0, s, La, ba, za, Lb, bb, zb, zc, wc, omega, zout, gate);
    else if (nu == 2)
        hipLaunchKernelGGL((k_post2<TA, 2>), grid, dim3(NT), 0, s, La, ba, za, Lb, bb, zb, zc, wc, omega, zout, gate);
    else
        hipLaunchKernelGGL((k_post2<TA, 3>), grid, dim3(NT), 0, s, La, ba, za, Lb, bb, zb, zc, wc, omega, zout, gate);
}

// levels a and a + 1 pre-smoothed in one launch: level a's b in, level a + 2's b out
int FldSolver::launch_pre_pair(int a, const float *ba, hipStream_t s, const int *gate) {
    const FldLevel &fa = lv[a], &fb = lv[a + 1], &fc = lv[a + 2];
    const Lv La = {fa.w, fa.h, fa.own, fa.ax, fa.ay, fa.cf}, Lb = {fb.w, fb.h, fb.own, fb.ax, fb.ay, fb.cf};
    if (fuse_pre == FUSE_LARGE)
        launch_pre2<16>(nu1, s, La, ba, fa.z, Lb, fb.b, fb.z, fc.b, fc.w, gate);
    else
        launch_pre2<8>(nu1, s, La, ba, fa.z, Lb, fb.b, fb.z, fc.b, fc.w, gate);
    return SR_OK;
}

// levels a + 1 and a post-smoothed in one launch: level a + 2's correction in, level a's zp out
int FldSolver::launch_post_pair(int a, hipStream_t s, const int *gate) {
    const FldLevel &fa = lv[a], &fb = lv[a + 1], &fc = lv[a + 2];
    const Lv La = {fa.w, fa.h, fa.own, fa.ax, fa.ay, fa.cf}, Lb = {fb.w, fb.h, fb.own, fb.ax, fb.ay, fb.cf};
    if (fuse_post == FUSE_LARGE)
        launch_post2<32>(nu2, s, La, fa.b, fa.z, Lb, fb.b, fb.z, fc.zp, fc.w, omega, fa.zp, gate);
    else
        launch_post2<16>(nu2, s, La, fa.b, fa.z, Lb, fb.b, fb.z, fc.zp, fc.w, omega, fa.zp, gate);
    return SR_OK;
}

// level l's pre-smoother, residual and restriction into level l + 1's b, in the level's tile shape
int FldSolver::launch_pre_level(int l, const float *b, hipStream_t s, const int *gate) {
    const FldLevel &f = lv[l], &c = lv[l + 1];
    const Lv L = {f.w, f.h, f.own, f.ax, f.ay, f.cf};
    switch (pre_tile[l]) {
    case TILE_16: launch_pre<16, 16>(nu1, s, L, b, f.z, c.b, c.w, gate); break;
    case TILE_64: launch_pre<PTX, PTY>(nu1, s, L, b, f.z, c.b, c.w, gate); break;
    default: launch_pre<32, 32>(nu1, s, L, b, f.z, c.b, c.w, gate); break;
    }
    return SR_OK;
}

// level l's prolongation of level l + 1's correction and post-smoother into zout; the launch's block count (the
// number of dot partials when part is given) into *nblocks
int FldSolver::launch_post_level(int l, const float *b, float *zout, const float *zold, double *part, int *nblocks,
                                 hipStream_t s, const int *gate) {
    const FldLevel &f = lv[l], &c = lv[l + 1];
    const Lv L = {f.w, f.h, f.own, f.ax, f.ay, f.cf};
    int nb = 0;
    switch (post_tile[l]) {
    case TILE_16: nb = launch_post<16, 16>(nu2, s, L, b, f.z, c.zp, c.w, omega, zout, zold, part, gate); break;
    case TILE_64: nb = launch_post<PTX, PTY>(nu2, s, L, b, f.z, c.zp, c.w, omega, zout, zold, part, gate); break;
    default: nb = launch_post<32, 32>(nu2, s, L, b, f.z, c.zp, c.w, omega, zout, zold, part, gate); break;
    }
    if (nblocks) *nblocks = nb;
    return SR_OK;
}

// z = M^-1 r on level 0; with `dots`, partials of r.z and r.zold into `dots` (their count returned in *nbd)
// The fold pays where the update alone is latency-bound and loses where it streams: at 1024^2 the folded launch takes
// 21.7 us against 25.2 for the update and the pre-smoother apart (the stiff solve 91.2 against 91.8 us per iteration);
// at 1280^2 the solve takes 129.3 against 126.0 us per iteration, at 1448^2 159.3 against 146.7, at 2048^2 the folded
// launch 88.8 us against 61.7, the update then moving its bytes faster on its own (5.2 TB/s) than inside the tile
// kernel. So it is taken up to FOLD_MAX_CELLS (about 1145^2).
#define FOLD_MAX_CELLS ((1u << 20) + (1u << 18))
bool FldSolver::update_fusable() const {
    return fuse_update && coarse_from > 0 && pair_level() != 0 && !pre_tile.empty() && pre_tile[0] == TILE_32 &&
           nu1 >= 1 && nu1 <= 3 && (size_t)w0 * h0 <= FOLD_MAX_CELLS;
}

int FldSolver::vcycle(const float *r, float *z, const float *zold, double *dots, int *nbd, hipStream_t s,
                      const int *gate, const PendingUpdate *upd, int stage) {
    const int pa = pair_level();
    const float *bl = r;
    for (int l = 0; l < coarse_from;) {
        if (l == pa && fuse_pre != FUSE_OFF) {
            SR_TRY(launch_pre_pair(l, bl, s, gate));
            bl = lv[l + 2].b;
            l += 2;
            continue;
        }
        const FldLevel &c = lv[l + 1];
        if (l > 0 || stage != 2) {
            if (l == 0 && upd) {
                const int slot = tick(CLK_UPDATE_PRE0, s);
                const FldLevel &f = lv[0];
                launch_pre_upd<32, 32>(nu1, s, Lv{f.w, f.h, f.own, f.ax, f.ay, f.cf}, *upd, f.z, c.b, c.w);
                SR_TRY(tock(slot, s));
            } else {
                const int slot = l == 0 ? tick(CLK_PRE0, s) : -1;
                SR_TRY(launch_pre_level(l, bl, s, gate));
                SR_TRY(tock(slot, s));
            }
            if (l == 0 && stage == 1) {
                SR_HIP(hipGetLastError());
                return SR_OK;
            }
        }
        bl = c.b;
        l++;
    }
    {
        const CoarseSet C = coarse_set();
        const size_t floats = C.total;
        float *zc = coarse_from == 0 ? z : lv[coarse_from].zp;
        const int slot = tick(CLK_COARSE, s);
        if (C.head)
            hipLaunchKernelGGL(k_coarse<true>, dim3(1), dim3(COARSE_NT), sizeof(float) * floats, s, C,
                               (const float *)pack, bl, zc, nu1, nu2, omega, gate);
        else
            hipLaunchKernelGGL(k_coarse<false>, dim3(1), dim3(COARSE_NT), sizeof(float) * floats, s, C,
                               (const float *)pack, bl, zc, nu1, nu2, omega, gate);
        SR_TRY(tock(slot, s));
        if (coarse_from == 0 && dots) {
            const int n = w0 * h0;
            hipLaunchKernelGGL(k_dots, dim3(nbe), dim3(NT), 0, s, n, r, z, zold, dots, gate);
            *nbd = nbe;
        }
    }
    for (int l = coarse_from - 1; l >= 0; l--) {
        if (l == pa + 1 && pa >= 1 && fuse_post != FUSE_OFF) {
            SR_TRY(launch_post_pair(pa, s, gate));
            l--;
            continue;
        }
        const FldLevel &f = lv[l];
        const float *b = l == 0 ? r : f.b;
        float *zo = l == 0 ? z : f.zp;
        double *pp = (l == 0 && dots) ? dots : nullptr;
        int nb = 0;
        const int slot = l == 0 ? tick(CLK_POST0, s) : -1;
        SR_TRY(launch_post_level(l, b, zo, zold, pp, &nb, s, gate));
        SR_TRY(tock(slot, s));
        if (l == 0 && dots) *nbd = nb;
    }
    SR_HIP(hipGetLastError());
    return SR_OK;
}

int FldSolver::tick(int kernel, hipStream_t s) {
    if (!timing) return -1;
    if (ev_used + 2 > 8192) SR_TRY(fold_clocks());  // a bounded pool: fold what is recorded into the sums first
    if (ev_used + 2 > evs.size()) {                  // grow the pool; recorded pairs are summed when read
        for (int k = 0; k < 64; k++) {
            hipEvent_t e;
            // timing only: no system-scope fence (an L2 write-back) around the timed kernel
            SR_HIP(hipEventCreateWithFlags(&e, hipEventDisableSystemFence));
            evs.push_back(e);
        }
        ev_kernel.resize(evs.size() / 2);
    }
    const int slot = (int)ev_used;
    ev_kernel[slot / 2] = kernel;
    ev_used += 2;
    SR_HIP(hipEventRecord(evs[slot], s));
    return slot;
}

int FldSolver::tock(int slot, hipStream_t s) {
    if (slot >= 0) SR_HIP(hipEventRecord(evs[slot + 1], s));
    return SR_OK;
}

int FldSolver::fold_clocks() {
    for (size_t k = 0; k + 1 < ev_used; k += 2) {
        SR_HIP(hipEventSynchronize(evs[k + 1]));
        float t = 0.f;
        SR_HIP(hipEventElapsedTime(&t, evs[k], evs[k + 1]));
        clock_ms[ev_kernel[k / 2]] += t;
        clock_n[ev_kernel[k / 2]]++;
    }
    ev_used = 0;
    return SR_OK;
}

int FldSolver::clock(int kernel, double *ms, unsigned long long *launches) {
    if (kernel < 0 || kernel >= CLK_N) return fail(SR_ERR_ARGUMENT, "no such solver clock");
    SR_TRY(fold_clocks());
    *ms = clock_ms[kernel];
    *launches = clock_n[kernel];
    return SR_OK;
}

void FldSolver::reset_clocks() {
    ev_used = 0;
    for (int k = 0; k < CLK_N; k++) {
        clock_ms[k] = 0.0;
        clock_n[k] = 0;
    }
}

// Per launch at level 0 (f32 unless noted): the pre-smoother reads b, own, ax, ay and writes z and the coarse b (21
// B per fine cell); the post-smoother reads b, own, ax, ay, z, the coarse z and z_old and writes z (33); the matvec
// reads z and p and own, ax, ay and writes p and q (28); the update reads p, q, r and writes r, and reads and writes x
// in f64 (32); the update folded into the next pre-smoother reads r, q, p and the coefficients, reads and writes x in
// f64 and writes r, z and the coarse b (49); the single-block V-cycle reads the pack and its first level's b and writes
// that level's z.
double FldSolver::clock_bytes(int kernel) const {
    const double n = (double)w0 * h0;
    switch (kernel) {
    case CLK_PRE0: return 21.0 * n;
    case CLK_POST0: return 33.0 * n;
    case CLK_MATVEC: return 28.0 * n;
    case CLK_UPDATE: return 32.0 * n;
    case CLK_UPDATE_PRE0: return 49.0 * n;  // the update's 32 and the pre-smoother's 21 less its read of b
    case CLK_COARSE: return 4.0 * (pack_floats + 2.0 * lv[coarse_from].w * lv[coarse_from].h);
    default: return 0.0;
    }
}

int FldSolver::set_network(Net *n, int fallback) {
    if (n && (w0 % 4 || h0 % 4))
        return fail(SR_ERR_ARGUMENT, "the learned preconditioner needs grid sides that are multiples of 4");
    net = n;
    fallback_it = std::max(0, fallback);
    return SR_OK;
}

// the learned preconditioner: z = V r + N(r - A V r), z = 0 off the gas; with `dots`, partials of r.z and r.zold
int FldSolver::learned(const float *r, float *z, const float *zold, double *dots, hipStream_t s) {
    if (!net) return fail(SR_ERR_ARGUMENT, "the learned preconditioner has no network");
    L0 a = {w0, h0, lv[0].own, lv[0].ax, lv[0].ay, gas};
    int nbd = 0;
    SR_TRY(vcycle(r, z, nullptr, nullptr, &nbd, s));
    hipLaunchKernelGGL(k_resid32, dim3(nbe), dim3(NT), 0, s, a, r, (const float *)z, q);
    SR_TRY(net->apply(w0, h0, lv[0].own, lv[0].ax, lv[0].ay, q, z, z, s));
    hipLaunchKernelGGL(k_dots_gas, dim3(nbe), dim3(NT), 0, s, a, r, z, zold, dots);
    SR_HIP(hipGetLastError());
    return SR_OK;
}

int FldSolver::apply_precond(const float *r, float *z, int precond, hipStream_t s) {
    L0 a = {w0, h0, lv[0].own, lv[0].ax, lv[0].ay, gas};
    if (precond == 0) {
        hipLaunchKernelGGL(k_jacobi, dim3(nbe), dim3(NT), 0, s, a, r, z, (const float *)nullptr, (double *)nullptr,
                           (const int *)nullptr);
        SR_HIP(hipGetLastError());
        return SR_OK;
    }
    if (precond == SR_PRECOND_LEARNED) return learned(r, z, nullptr, nullptr, s);
    int nbd = 0;
    return vcycle(r, z, nullptr, nullptr, &nbd, s);
}

int FldSolver::matvec64(const double *v, double *out, hipStream_t s) {
    L0 a = {w0, h0, lv[0].own, lv[0].ax, lv[0].ay, gas};
    const int n = w0 * h0;
    hipLaunchKernelGGL(k_matvec64, dim3((n + 255) / 256), dim3(256), 0, s, a, v, out);
    SR_HIP(hipGetLastError());
    return SR_OK;
}

int FldSolver::read_partials(const double *part, int k, int nb, double *out, hipStream_t s) {
    SR_TRY(copy_partials(part, k, nb, s));
    return wait_partials(k, nb, out);
}

// the partials' copy to the host queued, with the event that marks it done
int FldSolver::copy_partials(const double *part, int k, int nb, hipStream_t s) {
    SR_HIP(hipMemcpyAsync(host_part, part, sizeof(double) * k * nb, hipMemcpyDeviceToHost, s));
    SR_HIP(hipEventRecord(ev_part, s));
    return SR_OK;
}

// the host's wait for that copy alone (work queued after it keeps the device busy), then the sums in block order
int FldSolver::wait_partials(int k, int nb, double *out) {
    SR_HIP(hipEventSynchronize(ev_part));
    for (int j = 0; j < k; j++) {
        double t = 0.0;
        for (int i = 0; i < nb; i++) t += host_part[j * nb + i];
        out[j] = t;
    }
    return SR_OK;
}

int FldSolver::solve(const double *b, double *x, double tol, int max_iter, int precond, const RefStop *ref,
                     FldStats *st, hipStream_t s, int expect) {
    L0 a = {w0, h0, lv[0].own, lv[0].ax, lv[0].ay, gas};
    const size_t n = (size_t)w0 * h0;
    if (precond == SR_PRECOND_LEARNED && !net)
        return fail(SR_ERR_ARGUMENT, "the learned preconditioner has no network (sr_fld_set_network)");
    if ((solve_mode == SOLVE_DEVICE || solve_mode == SOLVE_GRAPH) && precond != SR_PRECOND_LEARNED)
        return solve_slots(b, x, tol, max_iter, precond, ref, st, s, expect);
    bool use_net = precond == SR_PRECOND_LEARNED;
    const bool mg = precond != 0 && lv.size() > 1;
    *st = FldStats();
    // the current residual's buffer and the other one: an update folded into the next pre-smoother writes the new
    // residual into the other buffer (the tiles read the old one as halo), and the two swap
    float *rc = r32, *ro = r32b;
    // z = M r with partials of r.z and r.z_old into `dots` (their count into nbd)
    auto precondition = [&](float *zc, const float *zo, double *dots, int *nbd) -> int {
        if (use_net) {
            SR_TRY(learned(rc, zc, zo, dots, s));
            *nbd = nbe;
            return SR_OK;
        }
        if (mg) return vcycle(rc, zc, zo, dots, nbd, s);
        L0 aa = {w0, h0, lv[0].own, lv[0].ax, lv[0].ay, gas};
        hipLaunchKernelGGL(k_jacobi, dim3(nbe), dim3(NT), 0, s, aa, rc, zc, zo, dots, (const int *)nullptr);
        *nbd = nbe;
        return SR_OK;
    };
    const double inner_rtol = 1e-4;
    int it = 0, net_it = 0;
    double rel0 = -1.0, rr_net0 = 0.0;
    double *dots[2] = {part_dots, part_dots + 2 * nbd_max};
    // rows per matvec block: about 2048 blocks, 1 to 16 rows each
    const int gxm = (w0 + NT - 1) / NT;
    const int mv_rows = std::max(1, std::min(MV_ROWS, (h0 * gxm + MV_BLOCKS - 1) / MV_BLOCKS));
    const dim3 mv_grid(gxm, (h0 + mv_rows - 1) / mv_rows);
    const int nmv = (int)(mv_grid.x * mv_grid.y);
    bool again = false;  // the last inner loop stopped at inner_rtol, short of the outer goal
    for (int outer = 0; outer < MAX_REFINEMENTS; outer++) {
        hipLaunchKernelGGL(k_true_residual, dim3(nbe), dim3(NT), 0, s, a, b, x, rc, ref ? ref->heat : nullptr,
                           ref ? ref->slope : nullptr, part_true);
        work.true_residuals++;
        double tr[4];
        // Lookahead: on the first refinement, and after an inner loop that stopped at inner_rtol rather than at the
        // goal, the next inner loop's first z = M r is queued behind the true residual's copy before the host waits;
        // it writes z and its dot partials only, which go unused when the check ends the solve.
        const bool ahead = solve_mode == SOLVE_HOST && !use_net && !timing && it < max_iter && (outer == 0 || again);
        int nbd_ahead = nbe;
        SR_TRY(copy_partials(part_true, 4, nbe, s));
        if (ahead) SR_TRY(precondition(za, nullptr, dots[0], &nbd_ahead));
        again = false;
        SR_TRY(wait_partials(4, nbe, tr));
        const double rr = tr[0], bb = tr[1] > 0 ? tr[1] : 1.0;
        st->rel = std::sqrt(rr / bb);
        st->refinements = outer;
        if (rel0 < 0.0) rel0 = st->rel;
        if (!std::isfinite(rr) || (ref && !(std::isfinite(tr[2]) && std::isfinite(tr[3])))) {
            st->finite = false;
            break;
        }
        const bool done =
            ref ? (tr[2] <= ref->u_bound && tr[3] <= ref->t_bound) : (std::sqrt(rr) <= tol * std::sqrt(bb));
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
        int k = 0, nbd = nbe, nbp = nbe;
        // z = M r, partials of r.z (restart: beta = 0, so no z_old and its partial, which stays 0, is not read)
        if (ahead)
            nbd = nbd_ahead;
        else
            SR_TRY(precondition(zc, nullptr, dots[k & 1], &nbd));
        work.precond++;
        int last_check = it, interval = 1;
        double rr_last = rr, left = HUGE_VAL;  // iterations left, as the last check's rate predicted them
        if (use_net && net_it == 0) rr_net0 = rr;
        // how much of the next iteration's z = M r is queued: 0 none, 1 its first launch (this iteration's update
        // folded in), 2 all of it (the lookahead below)
        int queued = 0;
        while (it < max_iter) {
            int slot = tick(CLK_MATVEC, s);
            hipLaunchKernelGGL(k_matvec_p, mv_grid, dim3(NT), 0, s, a, zc, pold, pnew, q, dots[k & 1],
                               dots[(k + 1) & 1], nbd, nbp, k == 0 ? 1 : 0, mv_rows, part_pq,
                               (const SolveCtl *)nullptr);
            SR_TRY(tock(slot, s));
            // the update: folded into the next z = M r's level-0 pre-smoother where the V-cycle allows it (the new
            // residual into the other buffer), else on its own, in place
            const bool fold = mg && !use_net && update_fusable();
            int nbd_next = nbe;
            if (fold) {
                const PendingUpdate U = {pnew, q, rc, ro, x, dots[k & 1], nbd, part_pq, nmv, part_rr_t};
                SR_TRY(vcycle(ro, zo, zc, dots[(k + 1) & 1], &nbd_next, s, nullptr, &U, 1));
                std::swap(rc, ro);
                queued = 1;
            } else {
                slot = tick(CLK_UPDATE, s);
                hipLaunchKernelGGL(k_update, dim3(nbe), dim3(NT), 0, s, (int)n, pnew, q, x, rc, dots[k & 1], nbd,
                                   part_pq, nmv, part_rr, (const int *)nullptr);
                SR_TRY(tock(slot, s));
            }
            it++;
            work.iterations++;
            if (use_net) net_it++;
            if (it - last_check >= interval || it >= max_iter) {
                // Lookahead: where the last rate predicts more than one iteration beyond this check, the rest of the
                // next iteration's z = M r is queued behind the partials' copy before the host waits for them, so the
                // device works while the host decides. It writes the next z and its dot partials only (x and r are
                // as this check reads them): when the check ends the loop after all, that z goes unused.
                const bool ahead = solve_mode == SOLVE_HOST && !use_net && !timing && it < max_iter &&
                                   left - (double)(it - last_check) > 1.5;
                const int nrr = fold ? ntiles0 : nbe;
                SR_TRY(copy_partials(fold ? part_rr_t : part_rr, 1, nrr, s));
                if (ahead) {
                    if (fold)
                        SR_TRY(vcycle(rc, zo, zc, dots[(k + 1) & 1], &nbd_next, s, nullptr, nullptr, 2));
                    else
                        SR_TRY(precondition(zo, zc, dots[(k + 1) & 1], &nbd_next));
                }
                double rri;
                SR_TRY(wait_partials(1, nrr, &rri));
                if (!std::isfinite(rri)) {
                    st->finite = false;
                    break;
                }
                if (rri <= stop2) {
                    again = inner_rtol * inner_rtol * rr > 0.25 * goal * goal;
                    break;
                }
                // with a fallback budget, the learned preconditioner gives way to plain multigrid once it has used
                // the budget, or when from NET_RATE_AT iterations on it reduces the residual more slowly per
                // iteration than plain multigrid did on its last solve of this system (checked at every check)
                const bool slower = fallback_it > 0 && plain_rate < 0.0 && net_it >= NET_RATE_AT &&
                                    0.5 * std::log(std::max(rri, 1e-300) / std::max(rr_net0, 1e-300)) / net_it >
                                        plain_rate;
                if (use_net && ((fallback_it > 0 && it >= fallback_it) || slower)) {
                    use_net = false;
                    st->fell_back = true;
                    break;
                }
                // iterations left at the rate seen since the last check: check again a little before then
                const double rate =
                    std::log(std::max(rri, 1e-300) / std::max(rr_last, 1e-300)) / (double)(it - last_check);
                left = rate < 0.0 ? std::log(stop2 / rri) / rate : 1.0;
                interval = std::max(1, std::min(8, (int)(0.7 * left)));
                if (use_net && fallback_it > 0 && plain_rate < 0.0 && net_it < NET_RATE_AT)
                    interval = std::min(interval, NET_RATE_AT - net_it);  // the first rate check comes on time
                last_check = it;
                rr_last = rri;
                if (ahead) queued = 2;
            }
            std::swap(zc, zo);
            std::swap(pold, pnew);
            k++;
            nbp = nbd;
            if (queued == 2)
                nbd = nbd_next;
            else if (queued == 1)
                SR_TRY(vcycle(rc, zc, zo, dots[k & 1], &nbd, s, nullptr, nullptr, 2));
            else
                SR_TRY(precondition(zc, zo, dots[k & 1], &nbd));
            queued = 0;
            work.precond++;
        }
        if (!st->finite) break;
    }
    st->iterations = it;
    st->learned_iterations = net_it;
    if (precond == SR_PRECOND_MULTIGRID && st->converged && it > 0 && rel0 > 0.0 && st->rel > 0.0)
        plain_rate = std::log(st->rel / rel0) / it;
    SR_HIP(hipGetLastError());
    return SR_OK;
}

// One slot of a device-decided solve, parity `par` (its state read from ctl[par], written to ctl[1 - par]; the z, p and
// dot-partial buffers alternate with the parity as the host-decided loop alternates them with the inner index): the
// head, then z = M r, p = z + beta p with q = A p, and x, r updated, each skipping itself unless the head left the
// inner iteration running.
int FldSolver::launch_slot(int par, const double *b, double *x, const RefStop *ref, int precond, hipStream_t s) {
    L0 a = {w0, h0, lv[0].own, lv[0].ax, lv[0].ay, gas};
    const size_t n = (size_t)w0 * h0;
    SolveCtl *cin = ctl + par, *cout = ctl + (1 - par);
    const int *gate = &cout->phase;
    hipLaunchKernelGGL(k_slot_head, dim3(nbe), dim3(NT), 0, s, a, b, x, r32, ref ? ref->heat : nullptr,
                       ref ? ref->slope : nullptr, part_true, (const double *)part_rr, nbe, (const SolveCtl *)cin,
                       cout, ticket);
    float *zc = par ? zb : za, *zo = par ? za : zb;
    float *pold = par ? p + n : p, *pnew = par ? p : p + n;
    double *dcur = part_dots + (size_t)par * 2 * nbd_max, *dprev = part_dots + (size_t)(1 - par) * 2 * nbd_max;
    int nbd = nbe;
    if (precond != 0 && lv.size() > 1) {
        SR_TRY(vcycle(r32, zc, zo, dcur, &nbd, s, gate));
    } else {
        hipLaunchKernelGGL(k_jacobi, dim3(nbe), dim3(NT), 0, s, a, (const float *)r32, zc, (const float *)zo, dcur,
                           gate);
    }
    const int gxm = (w0 + NT - 1) / NT;
    const int mv_rows = std::max(1, std::min(MV_ROWS, (h0 * gxm + MV_BLOCKS - 1) / MV_BLOCKS));
    const dim3 mv_grid(gxm, (h0 + mv_rows - 1) / mv_rows);
    const int nmv = (int)(mv_grid.x * mv_grid.y);
    int slot = tick(CLK_MATVEC, s);
    hipLaunchKernelGGL(k_matvec_p, mv_grid, dim3(NT), 0, s, a, (const float *)zc, (const float *)pold, pnew, q,
                       (const double *)dcur, (const double *)dprev, nbd, nbd, 0, mv_rows, part_pq,
                       (const SolveCtl *)cout);
    SR_TRY(tock(slot, s));
    slot = tick(CLK_UPDATE, s);
    hipLaunchKernelGGL(k_update, dim3(nbe), dim3(NT), 0, s, (int)n, (const float *)pnew, (const float *)q, x, r32,
                       (const double *)dcur, nbd, (const double *)part_pq, nmv, part_rr, gate);
    SR_TRY(tock(slot, s));
    work.precond++;  // an upper count: a slot that skips itself is counted too
    work.iterations++;
    SR_HIP(hipGetLastError());
    return SR_OK;
}

// The device-decided solve: the state set, then pairs of slots (replayed as one captured graph, or launched one by
// one with the clocks on or in SOLVE_DEVICE mode) until the state says done. The host reads the state between bursts
// of pairs, each burst sized from the residual's rate at the last two reads (about 0.7 of the iterations it predicts
// are left, as the host-decided loop spaces its checks); the pairs past the end cost their launches only.
int FldSolver::solve_slots(const double *b, double *x, double tol, int max_iter, int precond, const RefStop *ref,
                           FldStats *st, hipStream_t s, int expect) {
    *st = FldStats();
    const FldWork w_before = work;
    hipLaunchKernelGGL(k_ctl_init, dim3(1), dim3(1), 0, s, ctl, tol, max_iter, ref ? 1 : 0,
                       ref ? ref->u_bound : 0.0, ref ? ref->t_bound : 0.0);
    SR_HIP(hipGetLastError());
    // the graph is captured on the caller's stream; the legacy default stream cannot be captured, so there (and with
    // the clocks on) the same slots are launched one by one
    bool graph = solve_mode == SOLVE_GRAPH && !timing && s != nullptr;
    if (graph) {
        GraphKey k;
        k.b = b;
        k.x = x;
        k.heat = ref ? ref->heat : nullptr;
        k.slope = ref ? ref->slope : nullptr;
        k.precond = precond;
        k.tiles = tiles_version;
        k.s = s;
        if (!slot_exec || memcmp(&k, &slot_key, sizeof k) != 0) {
            if (slot_exec) SR_HIP(hipGraphExecDestroy(slot_exec));
            slot_exec = nullptr;
            if (hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) != hipSuccess) {
                (void)hipGetLastError();
                graph = false;
            }
        }
        if (graph && !slot_exec) {
            hipGraph_t g = nullptr;
            const int c0 = launch_slot(0, b, x, ref, precond, s);
            const int c1 = c0 == SR_OK ? launch_slot(1, b, x, ref, precond, s) : c0;
            const hipError_t e = hipStreamEndCapture(s, &g);
            SR_TRY(c1);
            if (e != hipSuccess) return fail(SR_ERR_DEVICE, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
            const hipError_t ei = hipGraphInstantiate(&slot_exec, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (ei != hipSuccess) {
                slot_exec = nullptr;
                return fail(SR_ERR_DEVICE, std::string("hipGraphInstantiate: ") + hipGetErrorString(ei));
            }
            slot_key = k;
        }
    }
    // the first burst: most of what a similar solve took (the caller's hint, else this solver's last solve); then
    // bursts sized from the residual's rate, one pair once the goal is near
    const int guess = expect > 0 ? expect : last_iterations;
    int pairs = std::max(1, (int)(0.45 * guess));
    int prev_it = -1;
    double prev_rr = 0.0;
    SolveCtl c{};
    for (;;) {
        for (int i = 0; i < pairs; i++) {
            if (graph) {
                SR_HIP(hipGraphLaunch(slot_exec, s));
                work.precond += 2;
                work.iterations += 2;
            } else {
                SR_TRY(launch_slot(0, b, x, ref, precond, s));
                SR_TRY(launch_slot(1, b, x, ref, precond, s));
            }
        }
        SR_HIP(hipMemcpyAsync(host_ctl, ctl, sizeof(SolveCtl), hipMemcpyDeviceToHost, s));
        SR_HIP(hipStreamSynchronize(s));
        c = *host_ctl;
        if (c.phase == PH_DONE) break;
        // iterations left at the rate since the last read, toward the outer goal (the inner one under the
        // reference's rule)
        const double target = c.ref ? c.stop2 : c.tol * c.tol * c.bb;
        double left = 2.0;
        if (prev_it >= 0 && c.it > prev_it && c.rr > 0.0 && prev_rr > 0.0 && c.rr < prev_rr && target > 0.0) {
            const double rate = std::log(c.rr / prev_rr) / (double)(c.it - prev_it);
            left = std::log(target / c.rr) / rate;
        }
        pairs = std::max(1, std::min(16, (int)(0.4 * left)));
        prev_it = c.it;
        prev_rr = c.rr;
    }
    // the counters: what ran, not what was launched
    work = w_before;
    work.iterations += c.it;
    work.precond += c.it;  // one application opens each inner iteration; none is left over at the end
    work.true_residuals += c.refinements + 1;
    st->iterations = c.it;
    st->refinements = c.refinements;
    st->rel = c.rel;
    st->converged = c.converged != 0;
    st->finite = c.finite != 0;
    last_iterations = c.it;
    if (precond == SR_PRECOND_MULTIGRID && st->converged && c.it > 0 && c.rel0 > 0.0 && c.rel > 0.0)
        plain_rate = std::log(c.rel / c.rel0) / c.it;
    return SR_OK;
}

// HBM bytes one inner iteration moves at level 0 by its algorithm (f32 unless noted), the roofline's numerator:
// matvec (z, p in, p out, q out, own, ax, ay: 28), update (x f64 rw 16, p, q, r rw 8: 32), the V-cycle's level-0
// kernels (pre: b, own, ax, ay in, z out, coarse b out 1: 21; post: b, own, ax, ay, z in, coarse z 1, z out, z_old:
// 33), the coarser levels adding a third of level 0's smoothing.
double FldSolver::bytes_per_iteration() const {
    const double n = (double)w0 * h0;
    return n * (28.0 + 32.0 + (21.0 + 33.0) * 4.0 / 3.0);
}

// Per fine cell: an iteration's matvec and update (28 + 32); a preconditioner application's V-cycle, level 0's pre- and
// post-smoother with a third more for the levels below ((21 + 33) 4/3); the FP64 true residual (b, x f64 in, r f64
// and r32 out, own/ax/ay in: 44); a hierarchy build (each coarsening reads level l's own, ax, ay and writes level
// l + 1's: 12 (1 + 1/4) per fine cell with the faces, own alone 5; the smoothers' packed records read own, ax, ay and
// write the 16-byte record: 28; a third more for the levels below).
double FldSolver::work_bytes(const FldWork &w) const {
    const double n = (double)w0 * h0;
    const double vcycle = (21.0 + 33.0) * 4.0 / 3.0;
    return n * ((28.0 + 32.0) * w.iterations + vcycle * w.precond + 44.0 * w.true_residuals +
                (12.0 * 1.25 + 28.0) * 4.0 / 3.0 * w.builds_faces + (5.0 * 1.25 + 28.0) * 4.0 / 3.0 * w.builds_own);
}

} // namespace sr

// ------------------------------------------------------------------------------------------------------------ C ABI
using namespace sr;

struct sr_fld {
    FldSolver s;
};

namespace sr {
Net *net_of(sr_net *p);
}

extern "C" int sr_fld_create(int w, int h, const sr_heat_settings *hs, sr_fld **out) {
    if (!out) return fail(SR_ERR_ARGUMENT, "sr_fld_create: null out");
    sr_heat_settings d;
    sr_heat_settings_defaults(&d);
    if (!hs) hs = &d;
    auto *f = new sr_fld();
    int c = f->s.init(w, h, hs->levels, hs->nu1, hs->nu2, hs->coarse_scale);
    if (c != SR_OK) {
        delete f;
        return c;
    }
    *out = f;
    return SR_OK;
}

extern "C" int sr_fld_tiles(sr_fld *f, int *pre, int *post, int max_levels, int *levels) {
    if (!f) return fail(SR_ERR_ARGUMENT, "sr_fld_tiles: null solver");
    return f->s.tiles(pre, post, max_levels, levels);
}

extern "C" int sr_fld_set_tiles(sr_fld *f, const int *pre, const int *post, int levels) {
    if (!f) return fail(SR_ERR_ARGUMENT, "sr_fld_set_tiles: null solver");
    return f->s.set_tiles(pre, post, levels);
}

extern "C" int sr_fld_set_coarse_wave(sr_fld *f, int on) {
    if (!f) return fail(SR_ERR_ARGUMENT, "sr_fld_set_coarse_wave: null solver");
    f->s.set_coarse_wave(on != 0);
    return SR_OK;
}

extern "C" int sr_fld_set_update_fusion(sr_fld *f, int on) {
    if (!f) return fail(SR_ERR_ARGUMENT, "sr_fld_set_update_fusion: null solver");
    f->s.set_update_fusion(on != 0);
    return SR_OK;
}

extern "C" int sr_fld_fusion(sr_fld *f, int *pre, int *post) {
    if (!f) return fail(SR_ERR_ARGUMENT, "sr_fld_fusion: null solver");
    return f->s.fusion(pre, post);
}

extern "C" int sr_fld_set_fusion(sr_fld *f, int pre, int post) {
    if (!f) return fail(SR_ERR_ARGUMENT, "sr_fld_set_fusion: null solver");
    return f->s.set_fusion(pre, post);
}

extern "C" int sr_fld_set_system(sr_fld *f, const float *own_dev, const float *kx_dev, const float *ky_dev,
                                 const uint8_t *gas_dev, void *stream) {
    if (!f || !own_dev || !kx_dev || !ky_dev) return fail(SR_ERR_ARGUMENT, "sr_fld_set_system: null pointer");
    return f->s.set_system(own_dev, kx_dev, ky_dev, gas_dev, (hipStream_t)stream);
}

extern "C" int sr_fld_solve(sr_fld *f, const double *b_dev, double *x_dev, double tol, int max_iter, int precond,
                            int *iterations, double *final_rel, void *stream) {
    if (!f || !b_dev || !x_dev) return fail(SR_ERR_ARGUMENT, "sr_fld_solve: null pointer");
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

extern "C" int sr_fld_set_network(sr_fld *f, sr_net *n, int fallback_iterations) {
    if (!f) return fail(SR_ERR_ARGUMENT, "sr_fld_set_network: null solver");
    return f->s.set_network(net_of(n), fallback_iterations);
}

extern "C" int sr_fld_solve_stats(sr_fld *f, const double *b_dev, double *x_dev, double tol, int max_iter,
                                  int precond, int *iterations, double *final_rel, int *learned_iterations,
                                  int *fell_back, void *stream) {
    if (!f || !b_dev || !x_dev) return fail(SR_ERR_ARGUMENT, "sr_fld_solve_stats: null pointer");
    FldStats st;
    SR_TRY(f->s.solve(b_dev, x_dev, tol, max_iter, precond, nullptr, &st, (hipStream_t)stream));
    if (iterations) *iterations = st.iterations;
    if (final_rel) *final_rel = st.rel;
    if (learned_iterations) *learned_iterations = st.learned_iterations;
    if (fell_back) *fell_back = st.fell_back ? 1 : 0;
    if (!st.finite) return fail(SR_ERR_NON_FINITE, "sr_fld_solve: a sum left the float's range");
    return SR_OK;
}

extern "C" int sr_fld_set_solve_mode(sr_fld *f, int mode) {
    if (!f) return fail(SR_ERR_ARGUMENT, "sr_fld_set_solve_mode: null solver");
    if (mode != SR_SOLVE_HOST && mode != SR_SOLVE_DEVICE && mode != SR_SOLVE_GRAPH && mode != SR_SOLVE_HOST_WAIT)
        return fail(SR_ERR_ARGUMENT, "sr_fld_set_solve_mode: mode must be SR_SOLVE_HOST, _DEVICE, _GRAPH or _HOST_WAIT");
    f->s.set_solve_mode(mode);
    return SR_OK;
}

extern "C" int sr_fld_levels(sr_fld *f, int *levels) {
    if (!f || !levels) return fail(SR_ERR_ARGUMENT, "sr_fld_levels: null pointer");
    *levels = f->s.nlevels();
    return SR_OK;
}

extern "C" void sr_fld_destroy(sr_fld *f) { delete f; }
