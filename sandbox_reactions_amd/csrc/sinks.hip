// P3 sinks on the GPU (crates/sr-physics/src/reference/sinks.rs; the reference's shader pass, step/sinks.rs) in FP64,
// the sinks on the device only. Six launches a pass, each skipping itself with no sink to move and nothing to form:
// the candidates (one 16 x 16 tile per block, each tile's best by a fixed tree: all four criteria, Sigma >= Sigma_BH
// or the mass test at a local Sigma maximum), the update (one block: the tiles' best in a fixed order, the formation
// taking its cell or its disk, then accretion sink by sink in index order, each sink's M and v frozen for its own
// cells, the cells of one sink taken by the block's threads at once and their mass, momentum and energy summed by a
// fixed tree), the gas's pull on every sink at the pass's start positions (per block, by a fixed tree), the half kick
// and the drift (the blocks summed in a fixed order, the pairs' pull softened by the larger r_acc), the pull at the
// drifted positions, and the closing half kick with merging, the lowest pair first. No atomic: the same state gives
// the same bits. The state's Sigma is its FP64 plane, the momentum and energy its f32 planes.
#include <cfloat>
#include <climits>
#include <cmath>

#include "common.hpp"
#include "sinks.hpp"

namespace sr {

#define NSP 10
constexpr int STILE = 16;  // the candidates' tile side
constexpr int SNT = 256;
constexpr int PULL_BLOCKS = 512;
constexpr double R_MIN = 0.5, R_CORE = 4.0, R_Y_E = 3.0, CORE_Y_E = 0.40;  // sinks.rs:48, :55-57

// (Z + 1) / A and Z / A per species, assets/elements.json order
__constant__ double k_zp1_a[NSP] = {2.0, 3.0 / 4.0, 7.0 / 12.0, 9.0 / 16.0, 11.0 / 20.0, 13.0 / 24.0, 15.0 / 28.0,
                                    17.0 / 32.0, 27.0 / 56.0, 1.0};
__constant__ double k_z_a[NSP] = {1.0, 0.5, 0.5, 0.5, 0.5, 0.5, 0.5, 0.5, 26.0 / 56.0, 0.0};

__device__ inline double r_acc_d(const SinkParams &K, double m) {
    return fmax(2.0 * K.g * m / (K.c_sb * K.c_sb), 2.0);
}

// 1/mu and Y_e of cell i's fractions renormalised as the floors do (floors.rs:45-60)
__device__ inline void composition(const HydroState &st, size_t n, size_t i, double &inv_mu, double &y_e, double &x_n) {
    double x[NSP], sum = 0.0;
#pragma unroll
    for (int s = 0; s < NSP; s++) {
        x[s] = fmax((double)st.x[(size_t)s * n + i], 0.0);
        sum += x[s];
    }
    double im = 0.0, ye = 0.0;
    if (sum > 0.0) {
#pragma unroll
        for (int s = 0; s < NSP; s++) {
            x[s] /= sum;
            im += x[s] * k_zp1_a[s];
            ye += x[s] * k_z_a[s];
        }
    } else {
        x[9] = 0.0;
        im = k_zp1_a[0];
        ye = k_z_a[0];
    }
    inv_mu = im;
    y_e = ye;
    x_n = x[9];
}

// the gas's eps_th (eos.rs:249-260): E / Sigma - |u|^2 / 2 - eps_cold, floored at T_floor / mu / (gamma - 1)
__device__ inline double eps_th(const SinkParams &K, const HydroState &st, size_t n, size_t i) {
    const double sigma = st.sigma[i], mx = st.mx[i], my = st.my[i], e = st.e[i];
    double im, ye, xn;
    composition(st, n, i, im, ye, xn);
    if (K.fixed_imu > 0.0) im = K.fixed_imu;
    const double kinetic = 0.5 * (mx * mx + my * my) / (sigma * sigma);
    const double cold = K.cold.on ? (double)eps_cold(K.cold, (float)sigma, (float)ye, (float)xn) : 0.0;
    const double raw = e / sigma - kinetic - cold;
    const double floor = K.t_floor * im / K.gm1;
    return raw < floor ? floor : raw;
}

__device__ inline int peak(const SinkParams &K, const HydroState &st, int x, int y) {
    const double s = st.sigma[(size_t)y * K.w + x];
    for (int ny = y - 1; ny <= y + 1; ny++)
        for (int nx = x - 1; nx <= x + 1; nx++)
            if (!(s >= st.sigma[(size_t)ny * K.w + nx])) return 0;
    return 1;
}

// the mass test (sinks.rs:221-236): the disk r <= 4 holds m_tov_sb, the disk r <= 3 is neutronised
__device__ inline int core(const SinkParams &K, const HydroState &st, int x, int y) {
    const size_t n = (size_t)K.w * K.h;
    double mass = 0.0;
    for (int dy = -4; dy <= 4; dy++)
        for (int dx = -4; dx <= 4; dx++) {
            const int cx = x + dx, cy = y + dy;
            if (cx >= 0 && cy >= 0 && cx < K.w && cy < K.h && (double)(dx * dx + dy * dy) <= R_CORE * R_CORE)
                mass += st.sigma[(size_t)cy * K.w + cx];
        }
    if (mass < K.m_tov) return 0;
    double m = 0.0, ye = 0.0;
    for (int dy = -3; dy <= 3; dy++)
        for (int dx = -3; dx <= 3; dx++) {
            const int cx = x + dx, cy = y + dy;
            if (cx >= 0 && cy >= 0 && cx < K.w && cy < K.h && (double)(dx * dx + dy * dy) <= R_Y_E * R_Y_E) {
                const size_t i = (size_t)cy * K.w + cx;
                double im, y_e, xn;
                composition(st, n, i, im, y_e, xn);
                m += st.sigma[i];
                ye += st.sigma[i] * y_e;
            }
        }
    return ye <= CORE_Y_E * m;
}

__device__ inline void velocity(const SinkParams &K, const HydroState &st, size_t i, double &ux, double &uy) {
    const double s = st.sigma[i];
    if (s < K.sigma_vac) {
        ux = uy = 0.0;
        return;
    }
    ux = st.mx[i] / s;
    uy = st.my[i] / s;
}

__device__ inline float phi_at(const SinkParams &K, const float *phi, int x, int y) {
    return phi[(size_t)(y + 1) * (K.w + 2) + x + 1];
}

// all four criteria at cell (x, y) (sinks.rs:238-283), an inner cell
__device__ int candidate(const SinkParams &K, const HydroState &st, const float *phi, int x, int y) {
    const size_t n = (size_t)K.w * K.h, i = (size_t)y * K.w + x;
    const double sigma = st.sigma[i];
    if (sigma < K.sigma_bh && !peak(K, st, x, y)) return 0;
    double ur[2], ul[2], uu[2], ud[2];
    velocity(K, st, i + 1, ur[0], ur[1]);
    velocity(K, st, i - 1, ul[0], ul[1]);
    velocity(K, st, i + K.w, uu[0], uu[1]);
    velocity(K, st, i - K.w, ud[0], ud[1]);
    const double div = (ur[0] - ul[0]) / 2.0 + (uu[1] - ud[1]) / 2.0;
    if (div >= 0.0) return 0;
    const float pc = phi_at(K, phi, x, y);
    for (int ny = y - 1; ny <= y + 1; ny++)
        for (int nx = x - 1; nx <= x + 1; nx++)
            if (!(pc <= phi_at(K, phi, nx, ny))) return 0;
    const double ux = st.mx[i] / sigma, uy = st.my[i] / sigma;
    if (0.5 * (ux * ux + uy * uy) + eps_th(K, st, n, i) + (double)pc >= 0.0) return 0;
    if (sigma < K.sigma_bh && !core(K, st, x, y)) return 0;
    return 1;
}

// the share 1 - Sigma_floor / Sigma of cell i (sinks.rs:205-219); returns mass, momentum and the thermal and cold energy
__device__ inline void take(const SinkParams &K, const HydroState &st, size_t i, double &dm, double &px, double &py,
                            double &sw) {
    const double sigma = st.sigma[i], mx = st.mx[i], my = st.my[i], e = st.e[i];
    const double f = K.sigma_floor / sigma, taken = 1.0 - K.sigma_floor / sigma;
    sw = taken * (e - 0.5 * (mx * mx + my * my) / sigma);
    st.sigma[i] = sigma * f;
    st.mx[i] = (float)(mx * f);
    st.my[i] = (float)(my * f);
    st.e[i] = (float)(e * f);
    dm = taken * sigma;
    px = taken * mx;
    py = taken * my;
}

// a block's sum of four values per thread by a fixed tree; thread 0 gets the sums
__device__ inline void block_sum4(double (&v)[4]) {
    __shared__ double sh[4][SNT / 64];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < 4; k++) {
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) v[k] += __shfl_xor(v[k], o, 64);
    }
    __syncthreads();
    if (lane == 0)
#pragma unroll
        for (int k = 0; k < 4; k++) sh[k][wid] = v[k];
    __syncthreads();
    if (threadIdx.x == 0)
#pragma unroll
        for (int k = 0; k < 4; k++) v[k] = sh[k][0] + sh[k][1] + sh[k][2] + sh[k][3];
}

// ----------------------------------------------------------------------------------------------------- kernels

__global__ void __launch_bounds__(SNT) k_sink_field(SinkParams K, const SinkDev *__restrict__ S, float *phi,
                                                    float *gx, float *gy, float *ps, float *gsx, float *gsy) {
    const int rw = K.w + 2, rh = K.h + 2;
    const int ri = blockIdx.x * SNT + threadIdx.x;
    if (ri >= rw * rh) return;
    const int x = ri % rw - 1, y = ri / rw - 1;
    const int cnt = S->count;
    double pot = 0.0, ax = 0.0, ay = 0.0;
    for (int q = 0; q < cnt; q++) {
        const double dx = x + 0.5 - S->x[q], dy = y + 0.5 - S->y[q];
        const double ra = r_acc_d(K, S->m[q]);
        const double soft = dx * dx + dy * dy + ra * ra, rs = sqrt(soft);
        pot -= K.g * S->m[q] / rs;
        const double pull = K.g * S->m[q] / (soft * rs);
        ax -= pull * dx;
        ay -= pull * dy;
    }
    ps[ri] = (float)pot;
    if (cnt && phi) phi[ri] += (float)pot;
    if (x >= 0 && y >= 0 && x < K.w && y < K.h) {
        const size_t i = (size_t)y * K.w + x;
        gsx[i] = (float)ax;
        gsy[i] = (float)ay;
        if (cnt && gx) {
            gx[i] += (float)ax;
            gy[i] += (float)ay;
        }
    }
}

__global__ void __launch_bounds__(SNT) k_sink_candidates(SinkParams K, const SinkDev *__restrict__ S, HydroState st,
                                                         const float *__restrict__ phi, double *cand_s, int *cand_i) {
    const int ntx = (K.w + STILE - 1) / STILE;
    const int x = (blockIdx.x % ntx) * STILE + threadIdx.x % STILE, y = (blockIdx.x / ntx) * STILE + threadIdx.x / STILE;
    double best = -1.0;
    int bi = INT_MAX;
    if (S->count < SR_MAX_SINKS && x >= 1 && y >= 1 && x < K.w - 1 && y < K.h - 1 && candidate(K, st, phi, x, y)) {
        best = st.sigma[(size_t)y * K.w + x];
        bi = y * K.w + x;
    }
    // the denser, then the lower cell index (the lowest y, then the lowest x)
    __shared__ double ss[SNT];
    __shared__ int si[SNT];
    ss[threadIdx.x] = best;
    si[threadIdx.x] = bi;
    __syncthreads();
    for (int o = SNT / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) {
            const double a = ss[threadIdx.x], b = ss[threadIdx.x + o];
            const int ia = si[threadIdx.x], ib = si[threadIdx.x + o];
            if (b > a || (b == a && ib < ia)) {
                ss[threadIdx.x] = b;
                si[threadIdx.x] = ib;
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        cand_s[blockIdx.x] = ss[0];
        cand_i[blockIdx.x] = si[0];
    }
}

// the formation and the accretion, one block
__global__ void __launch_bounds__(SNT) k_sink_update(SinkParams K, SinkDev *S, HydroState st,
                                                     const double *__restrict__ cand_s, const int *__restrict__ cand_i,
                                                     int tiles) {
    __shared__ double bs_sh[SNT];
    __shared__ int bi_sh[SNT];
    __shared__ int formed_cell, disk_take;
    double best = -1.0;
    int bi = INT_MAX;
    for (int t = threadIdx.x; t < tiles; t += SNT) {
        const double s = cand_s[t];
        const int i = cand_i[t];
        if (s > best || (s == best && i < bi)) {
            best = s;
            bi = i;
        }
    }
    bs_sh[threadIdx.x] = best;
    bi_sh[threadIdx.x] = bi;
    __syncthreads();
    for (int o = SNT / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) {
            const double a = bs_sh[threadIdx.x], b = bs_sh[threadIdx.x + o];
            const int ia = bi_sh[threadIdx.x], ib = bi_sh[threadIdx.x + o];
            if (b > a || (b == a && ib < ia)) {
                bs_sh[threadIdx.x] = b;
                bi_sh[threadIdx.x] = ib;
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        S->formed_x = S->formed_y = -1;
        S->merged = 0;
        S->accreted = 0.0;
        formed_cell = bs_sh[0] >= 0.0 ? bi_sh[0] : -1;
        disk_take = 0;
        if (formed_cell >= 0) {
            const int fx = formed_cell % K.w, fy = formed_cell / K.w;
            disk_take = peak(K, st, fx, fy) && core(K, st, fx, fy);
        }
    }
    __syncthreads();
    // formation: the cell, or its disk r <= 4 (the cells with Sigma >= Sigma_vac above the floor), its 81-cell square
    // one cell a thread
    if (formed_cell >= 0) {
        const int fx = formed_cell % K.w, fy = formed_cell / K.w;
        double v[4] = {0.0, 0.0, 0.0, 0.0};
        if ((int)threadIdx.x < 81) {
            const int dx = (int)threadIdx.x % 9 - 4, dy = (int)threadIdx.x / 9 - 4;
            const int cx = fx + dx, cy = fy + dy;
            bool mine;
            if (disk_take) {
                mine = cx >= 0 && cy >= 0 && cx < K.w && cy < K.h && (double)(dx * dx + dy * dy) <= R_CORE * R_CORE;
                if (mine) {
                    const double s = st.sigma[(size_t)cy * K.w + cx];
                    mine = s >= K.sigma_vac && s > K.sigma_floor;
                }
            } else {
                mine = dx == 0 && dy == 0;
            }
            if (mine) take(K, st, (size_t)cy * K.w + cx, v[0], v[1], v[2], v[3]);
        }
        block_sum4(v);
        if (threadIdx.x == 0) {
            const int q = S->count++;
            S->x[q] = fx + 0.5;
            S->y[q] = fy + 0.5;
            S->vx[q] = v[1] / v[0];
            S->vy[q] = v[2] / v[0];
            S->m[q] = v[0];
            S->swallowed += v[3];
            S->formed_x = fx;
            S->formed_y = fy;
        }
        __syncthreads();
    }
    // accretion, sink by sink in index order (sinks.rs:289-321)
    const size_t n = (size_t)K.w * K.h;
    const int cnt = S->count;
    double total = 0.0;
    for (int q = 0; q < cnt; q++) {
        const double sx = S->x[q], sy = S->y[q], svx = S->vx[q], svy = S->vy[q], sm = S->m[q];
        const double ra = r_acc_d(K, sm);
        const int x0 = (int)fmax(floor(sx - ra), 0.0), y0 = (int)fmax(floor(sy - ra), 0.0);
        const int x1 = (int)fmin(fmax(ceil(sx + ra), 0.0), (double)(K.w - 1));
        const int y1 = (int)fmin(fmax(ceil(sy + ra), 0.0), (double)(K.h - 1));
        const int ww = x1 - x0 + 1, cells = (x1 >= x0 && y1 >= y0) ? ww * (y1 - y0 + 1) : 0;
        double v[4] = {0.0, 0.0, 0.0, 0.0};
        for (int c = threadIdx.x; c < cells; c += SNT) {
            const int x = x0 + c % ww, y = y0 + c / ww;
            const size_t i = (size_t)y * K.w + x;
            const double cx = x + 0.5 - sx, cy = y + 0.5 - sy;
            const double r = hypot(cx, cy);
            const double sigma = st.sigma[i];
            if (r > ra || sigma < K.sigma_vac || sigma <= K.sigma_floor) continue;
            const double rx = st.mx[i] / sigma - svx, ry = st.my[i] / sigma - svy;
            if (0.5 * (rx * rx + ry * ry) + eps_th(K, st, n, i) - K.g * sm / fmax(r, R_MIN) >= 0.0) continue;
            double dm, px, py, sw;
            take(K, st, i, dm, px, py, sw);
            v[0] += dm;
            v[1] += px;
            v[2] += py;
            v[3] += sw;
        }
        block_sum4(v);
        if (threadIdx.x == 0) {
            const double m = sm + v[0];
            S->m[q] = m;
            S->vx[q] = (sm * svx + v[1]) / m;
            S->vy[q] = (sm * svy + v[2]) / m;
            S->swallowed += v[3];
            total += v[0];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) S->accreted = total;
}

// the gas's pull on every sink at its current position (sinks.rs:322-340): per block and sink, x and y
__global__ void __launch_bounds__(SNT) k_sink_pull(SinkParams K, const SinkDev *__restrict__ S, HydroState st,
                                                   double *part) {
    const int cnt = S->count;
    if (cnt == 0) return;
    const int n = K.w * K.h;
    __shared__ double at[SR_MAX_SINKS][3];  // position and r_acc^2
    if ((int)threadIdx.x < cnt) {
        const double ra = r_acc_d(K, S->m[threadIdx.x]);
        at[threadIdx.x][0] = S->x[threadIdx.x];
        at[threadIdx.x][1] = S->y[threadIdx.x];
        at[threadIdx.x][2] = ra * ra;
    }
    __syncthreads();
    double ax[SR_MAX_SINKS], ay[SR_MAX_SINKS];
#pragma unroll
    for (int q = 0; q < SR_MAX_SINKS; q++) ax[q] = ay[q] = 0.0;
    for (int i = blockIdx.x * SNT + threadIdx.x; i < n; i += gridDim.x * SNT) {
        const double sigma = st.sigma[i];
        if (sigma < K.sigma_vac) continue;
        const double cx = (i % K.w) + 0.5, cy = (i / K.w) + 0.5;
#pragma unroll
        for (int q = 0; q < SR_MAX_SINKS; q++) {
            if (q >= cnt) break;
            const double dx = cx - at[q][0], dy = cy - at[q][1];
            const double soft = dx * dx + dy * dy + at[q][2];
            const double pull = K.g * sigma / (soft * sqrt(soft));
            ax[q] += pull * dx;
            ay[q] += pull * dy;
        }
    }
    __shared__ double sh[SR_MAX_SINKS][2][SNT / 64];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < SR_MAX_SINKS; q++) {
        if (q < cnt) {
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) {
                ax[q] += __shfl_xor(ax[q], o, 64);
                ay[q] += __shfl_xor(ay[q], o, 64);
            }
            if (lane == 0) {
                sh[q][0][wid] = ax[q];
                sh[q][1][wid] = ay[q];
            }
        }
    }
    __syncthreads();
    if ((int)threadIdx.x < 2 * cnt) {
        const int q = threadIdx.x >> 1, c = threadIdx.x & 1;
        part[((size_t)blockIdx.x * SR_MAX_SINKS + q) * 2 + c] = sh[q][c][0] + sh[q][c][1] + sh[q][c][2] + sh[q][c][3];
    }
}

// the half kick and the drift (phase 0), or the closing half kick and the merging (phase 1), one block
__global__ void __launch_bounds__(SNT) k_sink_kick(SinkParams K, SinkDev *S, const double *__restrict__ part, int nb,
                                                   double dt_arg, const double *__restrict__ scal, int phase) {
    const int cnt = S->count;
    if (cnt == 0) return;
    const double dt = dt_arg >= 0.0 ? dt_arg : scal[0];
    // the blocks' sums in block order: 8 threads per (sink, axis), each a strided run, joined in order
    __shared__ double runs[SR_MAX_SINKS * 2][8];
    __shared__ double acc[SR_MAX_SINKS][2];
    {
        const int pair = threadIdx.x >> 3, r = threadIdx.x & 7;
        if (pair < 2 * cnt) {
            double s = 0.0;
            for (int b = r; b < nb; b += 8) s += part[((size_t)b * SR_MAX_SINKS + (pair >> 1)) * 2 + (pair & 1)];
            runs[pair][r] = s;
        }
    }
    __syncthreads();
    if ((int)threadIdx.x < 2 * cnt) {
        double s = 0.0;
        for (int r = 0; r < 8; r++) s += runs[threadIdx.x][r];
        acc[threadIdx.x >> 1][threadIdx.x & 1] = s;
    }
    __syncthreads();
    if (threadIdx.x != 0) return;
    double a[SR_MAX_SINKS][2];
    for (int i = 0; i < cnt; i++) {  // plus the other sinks' pull, a pair softened by its larger r_acc
        double bx = 0.0, by = 0.0;
        for (int j = 0; j < cnt; j++) {
            if (j == i) continue;
            const double dx = S->x[j] - S->x[i], dy = S->y[j] - S->y[i];
            const double rp = fmax(r_acc_d(K, S->m[i]), r_acc_d(K, S->m[j]));
            const double soft = dx * dx + dy * dy + rp * rp;
            const double pull = K.g * S->m[j] / (soft * sqrt(soft));
            bx += pull * dx;
            by += pull * dy;
        }
        a[i][0] = acc[i][0] + bx;
        a[i][1] = acc[i][1] + by;
    }
    for (int i = 0; i < cnt; i++) {
        S->vx[i] += 0.5 * dt * a[i][0];
        S->vy[i] += 0.5 * dt * a[i][1];
        if (phase == 0) {
            S->x[i] += dt * S->vx[i];
            S->y[i] += dt * S->vy[i];
        }
    }
    if (phase == 0) return;
    int lost = 0;  // merging (sinks.rs:376-399)
    bool again = true;
    while (again) {
        again = false;
        for (int i = 0; i < S->count && !again; i++)
            for (int j = i + 1; j < S->count; j++) {
                const double ax_ = S->x[i], ay_ = S->y[i], bx_ = S->x[j], by_ = S->y[j];
                if (hypot(ax_ - bx_, ay_ - by_) < fmax(r_acc_d(K, S->m[i]), r_acc_d(K, S->m[j]))) {
                    const double ma = S->m[i], mb = S->m[j], m = ma + mb;
                    S->x[i] = (ma * ax_ + mb * bx_) / m;
                    S->y[i] = (ma * ay_ + mb * by_) / m;
                    S->vx[i] = (ma * S->vx[i] + mb * S->vx[j]) / m;
                    S->vy[i] = (ma * S->vy[i] + mb * S->vy[j]) / m;
                    S->m[i] = m;
                    for (int q = j; q + 1 < S->count; q++) {
                        S->x[q] = S->x[q + 1];
                        S->y[q] = S->y[q + 1];
                        S->vx[q] = S->vx[q + 1];
                        S->vy[q] = S->vy[q + 1];
                        S->m[q] = S->m[q + 1];
                    }
                    S->count--;
                    lost++;
                    again = true;
                    break;
                }
            }
    }
    S->merged = lost;
}

// ------------------------------------------------------------------------------------------------------------ host

int sinks_alloc_sizes(int w, int h, int *tiles, int *pull_blocks) {
    *tiles = ((w + STILE - 1) / STILE) * ((h + STILE - 1) / STILE);
    *pull_blocks = std::max(1, std::min(PULL_BLOCKS, (w * h + SNT - 1) / SNT));
    return SR_OK;
}

int sinks_field(const SinkParams &K, const SinkBufs &B, float *phi, float *gx, float *gy, hipStream_t s) {
    const int nr = (K.w + 2) * (K.h + 2);
    hipLaunchKernelGGL(k_sink_field, dim3((nr + SNT - 1) / SNT), dim3(SNT), 0, s, K, B.dev, phi, gx, gy, B.ps, B.gsx,
                       B.gsy);
    SR_HIP(hipGetLastError());
    return SR_OK;
}

int sinks_pass(const SinkParams &K, const SinkBufs &B, const HydroState &st, const float *phi, double dt,
               const double *scal, hipStream_t s) {
    hipLaunchKernelGGL(k_sink_candidates, dim3(B.tiles), dim3(SNT), 0, s, K, B.dev, st, phi, B.cand_s, B.cand_i);
    hipLaunchKernelGGL(k_sink_update, dim3(1), dim3(SNT), 0, s, K, B.dev, st, B.cand_s, B.cand_i, B.tiles);
    hipLaunchKernelGGL(k_sink_pull, dim3(B.pull_blocks), dim3(SNT), 0, s, K, B.dev, st, B.pull);
    hipLaunchKernelGGL(k_sink_kick, dim3(1), dim3(SNT), 0, s, K, B.dev, B.pull, B.pull_blocks, dt, scal, 0);
    hipLaunchKernelGGL(k_sink_pull, dim3(B.pull_blocks), dim3(SNT), 0, s, K, B.dev, st, B.pull);
    hipLaunchKernelGGL(k_sink_kick, dim3(1), dim3(SNT), 0, s, K, B.dev, B.pull, B.pull_blocks, dt, scal, 1);
    SR_HIP(hipGetLastError());
    return SR_OK;
}

} // namespace sr
