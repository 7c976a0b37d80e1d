// The world: device state, the step driver (P1 -> P2 -> P4 -> P5 -> P8 -> P9, reference/mod.rs:264-297) and the
// C ABI of include/sr_hip.h for it. The reference's Step records every pass into one submission and keeps dt, the
// sim time and the solver's round counts on the GPU (crates/sr-engine/src/step/mod.rs:610-718); here dt and the sim
// time live in a small device array that every kernel reads, so a step with no heat pass issues its kernels
// back to back on the world's stream without a host round trip.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <mutex>
#include <numeric>
#include <string>
#include <thread>
#include <vector>

#include "common.hpp"
#include "grav.hpp"
#include "heat.hpp"
#include "hydro.hpp"
#include "react.hpp"
#include "sinks.hpp"

namespace sr {

#define NT 256
#define NS 10

enum { SCAL_DT = 0, SCAL_DT_NEXT, SCAL_SIM_TIME, SCAL_CAP, SCAL_RATE, SCAL_GMAX, SCAL_NF, SCAL_NF_STEP, SCAL_STEPS,
       SCAL_N = 16 };

__constant__ float w_imu[NS] = {2.0f, 0.75f, 7.0f / 12.0f, 9.0f / 16.0f, 11.0f / 20.0f,
                                13.0f / 24.0f, 15.0f / 28.0f, 17.0f / 32.0f, 27.0f / 56.0f, 1.0f};

struct FloorParams {
    int w, h;
    float gamma, gm1, t_floor_g, fixed_imu;
    ColdEos cold;
    double sigma_vac;          // the hydro's test (f64)
    float svac32, sfloor32;    // the floors' f32 constants (floors.rs:110)
    double cfl, eta_g, dx;
};

__global__ void k_upload(const float *__restrict__ planes, HydroState st, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    st.sigma[i] = (double)planes[i];
    st.mx[i] = planes[(size_t)n + i];
    st.my[i] = planes[(size_t)2 * n + i];
    st.e[i] = planes[(size_t)3 * n + i];
    for (int s = 0; s < NS; s++) st.x[(size_t)s * n + i] = planes[(size_t)(4 + s) * n + i];
}

__global__ void k_readback(HydroState st, float *__restrict__ planes, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    planes[i] = (float)st.sigma[i];
    planes[(size_t)n + i] = st.mx[i];
    planes[(size_t)2 * n + i] = st.my[i];
    planes[(size_t)3 * n + i] = st.e[i];
    for (int s = 0; s < NS; s++) planes[(size_t)(4 + s) * n + i] = st.x[(size_t)s * n + i];
}

// P1: this step's dt is the last P9's (capped), and the sim time advances by it (reference/mod.rs:266-270)
__global__ void k_p1(double *scal) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        double dt = scal[SCAL_DT_NEXT];
        if (scal[SCAL_CAP] > 0.0 && scal[SCAL_CAP] < dt) dt = scal[SCAL_CAP];
        scal[SCAL_DT] = dt;
        scal[SCAL_SIM_TIME] += dt;
    }
}

// P8 floors (floors.rs:124-175, no hold) then P9's dt (hydro.rs:618-634) and the non-finite guard: the largest
// wave-speed rate and |g| by a max reduction, the first non-finite (cell, channel) as the max of -key.
__global__ void __launch_bounds__(NT) k_floors_dt(FloorParams P, HydroState st, const float *__restrict__ gx,
                                                  const float *__restrict__ gy, int apply, double *part,
                                                  const int *__restrict__ box) {
    const int n = P.w * P.h;
    // the active box: P8 and P9 read and write its cells only (the whole world when the box is held at it)
    const int bx0 = box[0], by0 = box[1], bw = box[2], bh = box[3];
    double v[3] = {0.0, 0.0, -HUGE_VAL};
    for (int i = blockIdx.x * NT + threadIdx.x; i < n; i += gridDim.x * NT) {
        {
            const int x = i % P.w - bx0, y = i / P.w - by0;
            if (x < 0 || y < 0 || x >= bw || y >= bh) continue;
        }
        float xs[NS];
        bool finite = true;
        for (int s = 0; s < NS; s++) {
            xs[s] = st.x[(size_t)s * n + i];
            finite &= isfinite(xs[s]);
        }
        double sig = st.sigma[i];
        float mx = st.mx[i], my = st.my[i], e = st.e[i];
        if (apply && finite) {
            float sum = 0.0f;
            for (int s = 0; s < NS; s++) {
                xs[s] = fmaxf(xs[s], 0.0f);
                sum += xs[s];
            }
            if (sum > 0.0f) {
                for (int s = 0; s < NS; s++) xs[s] /= sum;
            } else {
                for (int s = 0; s < NS; s++) xs[s] = s == 0 ? 1.0f : 0.0f;
            }
            float im = 0.0f, ye = 0.0f;
            for (int s = 0; s < NS; s++) {
                im += xs[s] * w_imu[s];
                ye += xs[s] * c_ye[s];
            }
            if (P.fixed_imu > 0.0f) im = P.fixed_imu;
            const float efl = P.t_floor_g * im;
            if (sig < (double)P.svac32) {
                // floors.rs:128-139: the floor state at rest, its thermal energy at T_floor and its cold energy
                sig = (double)P.sfloor32;
                mx = my = 0.0f;
                e = P.sfloor32 * (efl + (P.cold.on ? eps_cold(P.cold, P.sfloor32, ye, xs[NS - 1]) : 0.0f));
            } else {
                const float s = (float)sig;
                const float raw = e / s - 0.5f * (mx * mx + my * my) / (s * s) -
                                  (P.cold.on ? eps_cold(P.cold, s, ye, xs[NS - 1]) : 0.0f);
                if (raw < efl) e = e + s * (efl - raw);
            }
            st.sigma[i] = sig;
            st.mx[i] = mx;
            st.my[i] = my;
            st.e[i] = e;
            for (int s = 0; s < NS; s++) st.x[(size_t)s * n + i] = xs[s];
        }
        // the rate as P9 reads the cell (hydro.rs:424-441): the floor state below sigma_vac
        float im = 0.0f, ye = 0.0f;
        for (int s = 0; s < NS; s++) {
            im += xs[s] * w_imu[s];
            ye += xs[s] * c_ye[s];
        }
        if (P.fixed_imu > 0.0f) im = P.fixed_imu;
        float u = 0.0f, vv = 0.0f, c2;
        // c^2 = gamma Pi / Sigma with Pi = Pi_th + Pi_e + Pi_n (hydro.rs:481-484, eos.rs:233-241)
        if (sig < P.sigma_vac) {
            c2 = P.gamma * P.gm1 * P.t_floor_g * im;
            if (P.cold.on) c2 += P.gamma * pi_cold(P.cold, P.sfloor32, ye, xs[NS - 1]) / P.sfloor32;
        } else {
            const float s = (float)sig;
            u = mx / s;
            vv = my / s;
            const float cold = P.cold.on ? eps_cold(P.cold, s, ye, xs[NS - 1]) : 0.0f;
            const float raw = e / s - 0.5f * (u * u + vv * vv) - cold;
            const float eps = fmaxf(raw, P.t_floor_g * im);
            c2 = P.gamma * P.gm1 * eps;
            if (P.cold.on) c2 += P.gamma * pi_cold(P.cold, s, ye, xs[NS - 1]) / s;
        }
        const float c = sqrtf(fmaxf(c2, 0.0f));
        const double rate = (double)((fabsf(u) + c) + (fabsf(vv) + c));
        v[0] = fmax(v[0], rate);
        if (gx) v[1] = fmax(v[1], (double)hypotf(gx[i], gy[i]));
        if (!(finite && isfinite(sig) && isfinite(mx) && isfinite(my) && isfinite(e))) {
            int ch = 0;
            if (!isfinite(sig)) ch = 0;
            else if (!isfinite(mx)) ch = 1;
            else if (!isfinite(my)) ch = 2;
            else if (!isfinite(e)) ch = 3;
            else
                for (int s = NS - 1; s >= 0; s--)
                    if (!isfinite(xs[s])) ch = 4 + s;
            v[2] = fmax(v[2], -((double)i * 16.0 + ch));
        }
        if (!isfinite(rate)) v[0] = HUGE_VAL;
    }
    block_partials<3, 1>(v, part, gridDim.x);
}

// P9's dt_{n+1} from the maxima: min(C dx / rate, eta_g sqrt(dx / |g|max)) (hydro.rs:618-634)
// The device scalars at an upload: zero, and no guard finding yet
__global__ void k_scal_reset(double *scal) {
    if (threadIdx.x < SCAL_N) scal[threadIdx.x] = threadIdx.x == SCAL_NF ? -HUGE_VAL : 0.0;
}

// (the step index a guard finding names is the device's count of the steps whose P9 has run; `count`: this P9 closes a
// step, so the count moves on)
__global__ void k_finalize_dt(FloorParams P, const double *__restrict__ part, int nb, double *scal, int count) {
    __shared__ double sh[3][4];
    double v[3] = {0.0, 0.0, -HUGE_VAL};
    for (int i = threadIdx.x; i < nb; i += blockDim.x)
        for (int k = 0; k < 3; k++) v[k] = fmax(v[k], part[(size_t)k * nb + i]);
    for (int k = 0; k < 3; k++)
        for (int o = 32; o > 0; o >>= 1) v[k] = fmax(v[k], __shfl_xor(v[k], o, 64));
    if ((threadIdx.x & 63) == 0)
        for (int k = 0; k < 3; k++) sh[k][threadIdx.x >> 6] = v[k];
    __syncthreads();
    if (threadIdx.x == 0) {
        double o[3];
        for (int k = 0; k < 3; k++) o[k] = fmax(fmax(sh[k][0], sh[k][1]), fmax(sh[k][2], sh[k][3]));
        scal[SCAL_RATE] = o[0];
        scal[SCAL_GMAX] = o[1];
        // the guard keeps its first finding and the index of the step whose P9 made it (step/dt.rs:40-47, NonFinite)
        if (o[2] > -HUGE_VAL && !(scal[SCAL_NF] > -HUGE_VAL)) {
            scal[SCAL_NF] = o[2];
            scal[SCAL_NF_STEP] = scal[SCAL_STEPS];
        }
        if (count) scal[SCAL_STEPS] += 1.0;
        double dt = P.cfl * P.dx / o[0];
        if (o[1] > 0.0) dt = fmin(dt, P.eta_g * sqrt(P.dx / o[1]));
        scal[SCAL_DT_NEXT] = dt;
    }
}

// The active box (step/boxfit.rs:1-14, the test oracle of tests/gpu/box.rs:21-90): the bounding box of the cells with
// Sigma >= Sigma_vac, grown by BOX_MARGIN cells, rounded out to multiples of BOX_ROUND, clamped to the world, in that
// order; empty (all 0) with no such cell; the FFT size per axis the smallest power of two >= 2 x the box's, at least
// BOX_FFT_MIN. Fit before step 0 (at the upload) and after P9 at step indices = 0 (mod BOX_REFIT_EVERY); the record
// holds x0, y0, w, h, fft_w, fft_h, the first step that runs inside it, and the fits written since the upload.
constexpr int BOX_MARGIN = 8, BOX_ROUND = 8, BOX_FFT_MIN = 32, BOX_REFIT_EVERY = 16;
enum { BOX_AT_UPLOAD = 0, BOX_AFTER_P9 = 1 };

// after a P9: whether the step whose P9 just ran (index = the count of P9s closing a step, less one) refits the box
__device__ inline bool box_due(const double *scal) {
    const long long k = (long long)scal[SCAL_STEPS] - 1;
    return k >= 0 && k % BOX_REFIT_EVERY == 0;
}

// per block: the min x, min y, max x + 1, max y + 1 of its cells holding matter (min > max when none)
__global__ void __launch_bounds__(NT) k_box_cells(const double *__restrict__ sigma, int w, int h, double svac,
                                                  const double *__restrict__ scal, int when, int *__restrict__ part) {
    if (when == BOX_AFTER_P9 && !box_due(scal)) return;
    __shared__ int sh[4][NT / 64];
    int v[4] = {w, h, 0, 0};
    const int n = w * h;
    for (int i = blockIdx.x * NT + threadIdx.x; i < n; i += gridDim.x * NT)
        if (sigma[i] >= svac) {
            const int x = i % w, y = i / w;
            v[0] = min(v[0], x);
            v[1] = min(v[1], y);
            v[2] = max(v[2], x + 1);
            v[3] = max(v[3], y + 1);
        }
    for (int o = 32; o > 0; o >>= 1) {
        v[0] = min(v[0], __shfl_xor(v[0], o, 64));
        v[1] = min(v[1], __shfl_xor(v[1], o, 64));
        v[2] = max(v[2], __shfl_xor(v[2], o, 64));
        v[3] = max(v[3], __shfl_xor(v[3], o, 64));
    }
    if ((threadIdx.x & 63) == 0)
        for (int k = 0; k < 4; k++) sh[k][threadIdx.x >> 6] = v[k];
    __syncthreads();
    if (threadIdx.x < 4) {
        const int k = threadIdx.x;
        int r = sh[k][0];
        for (int j = 1; j < NT / 64; j++) r = k < 2 ? min(r, sh[k][j]) : max(r, sh[k][j]);
        part[4 * blockIdx.x + k] = r;
    }
}

__device__ inline int box_fft(int side) {
    int f = 1;
    while (f < 2 * side) f <<= 1;
    return max(f, BOX_FFT_MIN);
}

// one block: the partials joined (integer min and max: no order to fix) and the record written
__global__ void __launch_bounds__(NT) k_box_fit(const int *__restrict__ part, int nb, int w, int h, int whole,
                                                const double *__restrict__ scal, int when, int *__restrict__ box) {
    if (when == BOX_AFTER_P9 && !box_due(scal)) return;
    __shared__ int sh[4][NT / 64];
    int v[4] = {w, h, 0, 0};
    for (int b = threadIdx.x; b < nb; b += NT) {
        v[0] = min(v[0], part[4 * b]);
        v[1] = min(v[1], part[4 * b + 1]);
        v[2] = max(v[2], part[4 * b + 2]);
        v[3] = max(v[3], part[4 * b + 3]);
    }
    for (int o = 32; o > 0; o >>= 1) {
        v[0] = min(v[0], __shfl_xor(v[0], o, 64));
        v[1] = min(v[1], __shfl_xor(v[1], o, 64));
        v[2] = max(v[2], __shfl_xor(v[2], o, 64));
        v[3] = max(v[3], __shfl_xor(v[3], o, 64));
    }
    if ((threadIdx.x & 63) == 0)
        for (int k = 0; k < 4; k++) sh[k][threadIdx.x >> 6] = v[k];
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int j = 1; j < NT / 64; j++) {
            sh[0][0] = min(sh[0][0], sh[0][j]);
            sh[1][0] = min(sh[1][0], sh[1][j]);
            sh[2][0] = max(sh[2][0], sh[2][j]);
            sh[3][0] = max(sh[3][0], sh[3][j]);
        }
        int x0 = 0, y0 = 0, bw = 0, bh = 0;
        if (whole) {
            bw = w;
            bh = h;
        } else if (sh[0][0] < sh[2][0]) {
            // v - MARGIN rounded down and v + MARGIN rounded up to the 8-grid, then clamped (a negative quotient
            // truncates toward 0 where the floor would go lower, and both clamp to 0)
            x0 = max(0, (sh[0][0] - BOX_MARGIN) / BOX_ROUND * BOX_ROUND);
            y0 = max(0, (sh[1][0] - BOX_MARGIN) / BOX_ROUND * BOX_ROUND);
            bw = min(w, (sh[2][0] + BOX_MARGIN + BOX_ROUND - 1) / BOX_ROUND * BOX_ROUND) - x0;
            bh = min(h, (sh[3][0] + BOX_MARGIN + BOX_ROUND - 1) / BOX_ROUND * BOX_ROUND) - y0;
        }
        box[0] = x0;
        box[1] = y0;
        box[2] = bw;
        box[3] = bh;
        box[4] = box_fft(bw);
        box[5] = box_fft(bh);
        box[6] = when == BOX_AT_UPLOAD ? 0 : (int)scal[SCAL_STEPS];
        box[7] = when == BOX_AT_UPLOAD ? 1 : box[7] + 1;
    }
}

// FP64 totals: Sigma, Sigma u, Sigma v, E, Sigma X_s (hydro.rs:210-224), and the escaped sums per component
__global__ void __launch_bounds__(NT) k_totals(HydroState st, int n, double *part) {
    double v[14];
    for (int k = 0; k < 14; k++) v[k] = 0.0;
    for (int i = blockIdx.x * NT + threadIdx.x; i < n; i += gridDim.x * NT) {
        const double s = st.sigma[i];
        v[0] += s;
        v[1] += st.mx[i];
        v[2] += st.my[i];
        v[3] += st.e[i];
        for (int k = 0; k < NS; k++) v[4 + k] += s * (double)st.x[(size_t)k * n + i];
    }
    block_partials<14>(v, part, gridDim.x);
}

__global__ void k_escaped_sum(const double *__restrict__ esc, int nf, double *out) {
    __shared__ double sh[4];
    const int k = blockIdx.x;
    double s = 0.0;
    for (int i = threadIdx.x; i < nf; i += blockDim.x) s += esc[(size_t)k * nf + i];
    s = wave_sum(s);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) out[k] = sh[0] + sh[1] + sh[2] + sh[3];
}

__global__ void k_extract_box(const float *__restrict__ ringed, int w, int h, float *__restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= w * h) return;
    const int x = i % w, y = i / w;
    out[i] = ringed[(size_t)(y + 1) * (w + 2) + x + 1];
}

} // namespace sr

using namespace sr;

namespace sr {
Net *net_of(sr_net *p);
}

struct KernelClock {
    unsigned long long launches = 0;
    double ms = 0.0;
    double bytes = 0.0;
};

struct sr_world {
    int w = 0, h = 0;
    size_t n = 0;
    sr_physics ph{};
    sr_switches sw{};
    sr_heat_settings hs{};
    int precond = SR_PRECOND_MULTIGRID;
    sr_net *net = nullptr;
    hipStream_t s = nullptr;
    HydroState st[2]{};
    int cur = 0;
    HydroParams hp{};
    FloorParams fp{};
    double *scal = nullptr, *host_scal = nullptr;
    double *takeback = nullptr;
    HydroPartials part{};
    double *rpart = nullptr;
    unsigned *counters = nullptr;
    Gravity grav;
    float *phi = nullptr, *gx = nullptr, *gy = nullptr, *phi_mid = nullptr, *sigma_mid = nullptr;
    float *pushmag = nullptr, *rfx = nullptr, *rfy = nullptr;
    bool have_flux = false;
    bool heat_ran = false;      // a step with heat has run since the upload (its first system can be rebuilt)
    bool first_built = false;   // the solver holds that first system (sr_heat_first_system), not a later pass's
    double *escaped = nullptr, *totals = nullptr;
    HeatPass heat;
    uint64_t steps = 0;
    sr_heat_report last_heat{};
    bool timing = false;
    hipEvent_t ev[14]{};
    sr_step_cost cost{};
    KernelClock k_update, k_mid, k_grav, k_pre0;
    // the step's fixed parts captured as graphs (with the clocks off): P1..P4 per current buffer and radiation flag,
    // P8/P9 per current buffer
    bool graphs = true;
    hipGraphExec_t g_front[2][2] = {}, g_tail[2] = {};
    // the step's launches outside the heat pass, counted as they are made (sr_read_work)
    struct {
        unsigned long long steps = 0, hydro = 0, mid = 0, conv_first = 0, conv_mid = 0, radiation = 0;
    } wk;
    // P6: the records and constants on the device (null: off), the last pass's per-cell bookings and partials
    ReactDev *react = nullptr, *react_dev_mem = nullptr;
    double *react_nuc = nullptr, *react_lost = nullptr, *react_snu = nullptr, *react_psum = nullptr,
           *react_pmax = nullptr;
    int react_nb = 0;
    bool sinks_on = false;  // P3 in the step
    SinkParams sk_par{};
    SinkBufs sk{};
    // the active box: its record (x0, y0, w, h, fft_w, fft_h, from_step, fits) and the cell pass's partials
    bool box_fit = false;
    int *box = nullptr, *box_part = nullptr;
    int box_nb = 1;
    std::vector<void *> allocs;
};

static int world_alloc(sr_world *wd, void **p, size_t bytes) {
    SR_HIP(hipMalloc(p, bytes));
    SR_HIP(hipMemset(*p, 0, bytes));
    wd->allocs.push_back(*p);
    return SR_OK;
}

static int hydro_state_alloc(sr_world *wd, HydroState &st) {
    SR_TRY(world_alloc(wd, (void **)&st.sigma, sizeof(double) * wd->n));
    SR_TRY(world_alloc(wd, (void **)&st.mx, sizeof(float) * wd->n));
    SR_TRY(world_alloc(wd, (void **)&st.my, sizeof(float) * wd->n));
    SR_TRY(world_alloc(wd, (void **)&st.e, sizeof(float) * wd->n));
    SR_TRY(world_alloc(wd, (void **)&st.x, sizeof(float) * wd->n * NS));
    return SR_OK;
}

static int grid_n(size_t n) { return (int)std::min<size_t>(4096, (n + NT * 4 - 1) / (NT * 4)); }

extern "C" void sr_destroy(sr_world *wd) {
    if (!wd) return;
    if (wd->s) (void)hipStreamSynchronize(wd->s);
    for (void *p : wd->allocs) (void)hipFree(p);
    if (wd->host_scal) (void)hipHostFree(wd->host_scal);
    for (auto &e : wd->ev)
        if (e) (void)hipEventDestroy(e);
    for (auto &row : wd->g_front)
        for (auto &g : row)
            if (g) (void)hipGraphExecDestroy(g);
    for (auto &g : wd->g_tail)
        if (g) (void)hipGraphExecDestroy(g);
    if (wd->s) (void)hipStreamDestroy(wd->s);
    delete wd;
}

extern "C" int sr_world_create(int w, int h, const sr_physics *ph, const sr_switches *sw, const sr_heat_settings *hs,
                               sr_world **out) {
    if (!out || !ph || !sw) return fail(SR_ERR_ARGUMENT, "sr_world_create: null pointer");
    *out = nullptr;
    if (w < 8 || h < 8 || w > 4096 || h > 4096 || w % 8 || h % 8)
        return fail(SR_ERR_WORLD_SIZE, "a " + std::to_string(w) + "x" + std::to_string(h) +
                                            " world: each axis a multiple of 8 in [8, 4096]");
    if (!(ph->gamma > 1.0) || !(ph->dx > 0.0) || !(ph->t_floor > 0.0))
        return fail(SR_ERR_ARGUMENT, "gamma > 1, dx > 0 and t_floor > 0 are required");
    // a cold pressure pair is on or off as a whole (eos.rs:172-183)
    for (const double *k : {&ph->k1e, &ph->k1n}) {
        const double k1 = k[0], k2 = k[1];
        if (!std::isfinite(k1) || !std::isfinite(k2) || k1 < 0.0 || k2 < 0.0 || ((k1 > 0.0) != (k2 > 0.0)))
            return fail(SR_ERR_ARGUMENT, "a cold pressure pair (k1e, k2e or k1n, k2n) must be both > 0 or both 0");
    }
    auto *wd = new sr_world();
    wd->w = w;
    wd->h = h;
    wd->n = (size_t)w * h;
    wd->ph = *ph;
    wd->sw = *sw;
    if (hs)
        wd->hs = *hs;
    else
        sr_heat_settings_defaults(&wd->hs);
    auto bail = [&](int c) {
        sr_destroy(wd);
        return c;
    };
    int c;
    if (hipStreamCreateWithFlags(&wd->s, hipStreamNonBlocking) != hipSuccess)
        return bail(fail(SR_ERR_DEVICE, "hipStreamCreate"));
    // the pass clocks' events only time: no system-scope fence (an L2 write-back) at each mark
    for (auto &e : wd->ev)
        if (hipEventCreateWithFlags(&e, hipEventDisableSystemFence) != hipSuccess)
            return bail(fail(SR_ERR_DEVICE, "hipEventCreate"));
    const size_t n = wd->n;
    if ((c = hydro_state_alloc(wd, wd->st[0])) || (c = hydro_state_alloc(wd, wd->st[1]))) return bail(c);
    if ((c = world_alloc(wd, (void **)&wd->scal, sizeof(double) * SCAL_N))) return bail(c);
    if (hipHostMalloc(&wd->host_scal, sizeof(double) * SCAL_N) != hipSuccess)
        return bail(fail(SR_ERR_DEVICE, "hipHostMalloc"));
    const int nb = hydro_blocks(w, h);
    wd->box_nb = (int)std::min<size_t>(1024, (n + NT - 1) / NT);
    if ((c = world_alloc(wd, (void **)&wd->takeback, sizeof(double) * 8)) ||
        (c = world_alloc(wd, (void **)&wd->part.sums, sizeof(double) * 5 * nb)) ||
        (c = world_alloc(wd, (void **)&wd->part.keys, sizeof(unsigned long long) * nb)) ||
        (c = world_alloc(wd, (void **)&wd->rpart, sizeof(double) * 16 * 4096)) ||
        (c = world_alloc(wd, (void **)&wd->counters, sizeof(unsigned) * 8)) ||
        (c = world_alloc(wd, (void **)&wd->box, sizeof(int) * 8)) ||
        (c = world_alloc(wd, (void **)&wd->box_part, sizeof(int) * 4 * 1024)) ||
        (c = world_alloc(wd, (void **)&wd->phi, sizeof(float) * (w + 2) * (h + 2))) ||
        (c = world_alloc(wd, (void **)&wd->phi_mid, sizeof(float) * (w + 2) * (h + 2))) ||
        (c = world_alloc(wd, (void **)&wd->gx, sizeof(float) * n)) ||
        (c = world_alloc(wd, (void **)&wd->gy, sizeof(float) * n)) ||
        (c = world_alloc(wd, (void **)&wd->sigma_mid, sizeof(float) * n)) ||
        (c = world_alloc(wd, (void **)&wd->pushmag, sizeof(float) * n)) ||
        (c = world_alloc(wd, (void **)&wd->rfx, sizeof(float) * n)) ||
        (c = world_alloc(wd, (void **)&wd->rfy, sizeof(float) * n)) ||
        (c = world_alloc(wd, (void **)&wd->escaped, sizeof(double) * 14 * 2 * (w + h))) ||
        (c = world_alloc(wd, (void **)&wd->totals, sizeof(double) * 32)))
        return bail(c);
    if (sw->gravity) {
        if ((c = wd->grav.init(w, h, ph->G, ph->dx, ph->sigma_vac, ph->max_padded > 0 ? ph->max_padded : 4096)))
            return bail(c);
        // the row transforms' rows per block, timed on the (zeroed) mid-step arrays
        if ((c = wd->grav.tune(wd->sigma_mid, wd->phi_mid, wd->s))) return bail(c);
        if (hipStreamSynchronize(wd->s) != hipSuccess) return bail(fail(SR_ERR_DEVICE, "hipStreamSynchronize"));
    } else {
        int px;
        if ((c = padded_size(w, ph->max_padded > 0 ? ph->max_padded : 4096, &px))) return bail(c);
        if ((c = padded_size(h, ph->max_padded > 0 ? ph->max_padded : 4096, &px))) return bail(c);
    }
    ColdEos cold{};
    cold.k1e = (float)ph->k1e;
    cold.k2e = (float)ph->k2e;
    cold.k1n = (float)ph->k1n;
    cold.k2n = (float)ph->k2n;
    cold.on = ph->k1e > 0.0 || ph->k1n > 0.0;
    if (cold.on) {
        std::vector<float> t(2 * COLD_POINTS, 0.0f);
        if (ph->k1e > 0.0) cold_table(ph->k1e, ph->k2e, t.data());
        if (ph->k1n > 0.0) cold_table(ph->k1n, ph->k2n, t.data() + COLD_POINTS);
        float *d = nullptr;
        if ((c = world_alloc(wd, (void **)&d, sizeof(float) * t.size()))) return bail(c);
        if (hipMemcpy(d, t.data(), sizeof(float) * t.size(), hipMemcpyHostToDevice) != hipSuccess)
            return bail(fail(SR_ERR_DEVICE, "hipMemcpy (cold tables)"));
        cold.lnu = d;
    }
    HydroParams &hp = wd->hp;
    hp.cold = cold;
    hp.w = w;
    hp.h = h;
    hp.gamma = (float)ph->gamma;
    hp.gm1 = (float)(ph->gamma - 1.0);
    hp.t_floor_g = (float)(ph->t_floor / (ph->gamma - 1.0));
    hp.sigma_vac = (float)ph->sigma_vac;
    hp.sigma_floor = (float)ph->sigma_floor;
    hp.fixed_imu = ph->mu > 0 ? (float)(1.0 / ph->mu) : 0.0f;
    hp.edge = ph->edge;
    hp.inv_dx = (float)(1.0 / ph->dx);
    hp.kappa0 = (float)ph->kappa0;
    hp.c_sb = (float)ph->c_sb;
    hp.kappa_dust = (float)ph->kappa_dust;
    hp.t_dust = (float)ph->t_dust;
    hp.opacity = ph->opacity;
    FloorParams &fp = wd->fp;
    fp.cold = cold;
    fp.w = w;
    fp.h = h;
    fp.gamma = hp.gamma;
    fp.gm1 = hp.gm1;
    fp.t_floor_g = hp.t_floor_g;
    fp.fixed_imu = hp.fixed_imu;
    fp.sigma_vac = ph->sigma_vac;
    fp.svac32 = (float)ph->sigma_vac;
    fp.sfloor32 = (float)ph->sigma_floor;
    fp.cfl = ph->cfl;
    fp.eta_g = ph->eta_g;
    fp.dx = ph->dx;
    if (sw->heat) {
        HeatParams P{};
        P.w = w;
        P.h = h;
        P.dx = ph->dx;
        P.inv_dx = (float)(1.0 / ph->dx);
        P.gm1 = ph->gamma - 1.0;
        P.t_floor = ph->t_floor;
        P.sigma_vac = (float)ph->sigma_vac;
        P.sigma_floor = (float)ph->sigma_floor;
        P.fixed_imu = hp.fixed_imu;
        P.a_r = ph->a_r;
        P.c_sb = ph->c_sb;
        P.k_cond = ph->k_cond;
        P.kappa0 = (float)ph->kappa0;
        P.kappa_dust = (float)ph->kappa_dust;
        P.t_dust = (float)ph->t_dust;
        P.opacity = ph->opacity;
        P.cd = ph->constant_diffusivity > 0.0;
        P.chi = ph->constant_diffusivity;
        P.limiter = ph->flux_limiter;
        P.pinned = ph->pin_temperature;
        P.cold = cold;
        if ((c = wd->heat.init(P, wd->hs))) return bail(c);
    }
    *out = wd;
    return SR_OK;
}

static int state_ok(sr_world *wd, size_t n) {
    if (!wd) return fail(SR_ERR_ARGUMENT, "null world");
    if (n != SR_COMPONENTS * wd->n)
        return fail(SR_ERR_LENGTH, "planes hold " + std::to_string(n) + " values, expected " +
                                       std::to_string(SR_COMPONENTS * wd->n));
    return SR_OK;
}

static int p2(sr_world *wd) {
    if (!wd->sw.gravity) return SR_OK;
    SR_TRY(wd->grav.potential<double>(wd->st[wd->cur].sigma, wd->phi, wd->s));
    return wd->grav.gradient(wd->phi, wd->gx, wd->gy, wd->s);
}

// the box fit on the current state: at the upload, or after a P9 (each kernel skipping itself unless the step index
// is one the box is refit at, so the sequence is the same every step and can be captured)
static int box_refit(sr_world *wd, int when) {
    hipLaunchKernelGGL(k_box_cells, dim3(wd->box_nb), dim3(NT), 0, wd->s, wd->st[wd->cur].sigma, wd->w, wd->h,
                       (double)wd->fp.svac32, wd->scal, when, wd->box_part);
    hipLaunchKernelGGL(k_box_fit, dim3(1), dim3(NT), 0, wd->s, wd->box_part, wd->box_nb, wd->w, wd->h,
                       wd->box_fit ? 0 : 1, wd->scal, when, wd->box);
    SR_HIP(hipGetLastError());
    return SR_OK;
}

static int p89(sr_world *wd, int apply) {
    HydroState &st = wd->st[wd->cur];
    const int nb = grid_n(wd->n);
    hipLaunchKernelGGL(k_floors_dt, dim3(nb), dim3(NT), 0, wd->s, wd->fp, st, wd->sw.gravity ? wd->gx : nullptr,
                       wd->sw.gravity ? wd->gy : nullptr, apply, wd->rpart, (const int *)wd->box);
    hipLaunchKernelGGL(k_finalize_dt, dim3(1), dim3(256), 0, wd->s, wd->fp, wd->rpart, nb, wd->scal, apply);
    SR_HIP(hipGetLastError());
    if (apply) SR_TRY(box_refit(wd, BOX_AFTER_P9));
    return SR_OK;
}

extern "C" int sr_upload(sr_world *wd, const float *planes, size_t n) {
    SR_TRY(state_ok(wd, n));
    if (!planes) return fail(SR_ERR_ARGUMENT, "sr_upload: null planes");
    float *tmp = nullptr;
    SR_HIP(hipMallocAsync((void **)&tmp, sizeof(float) * n, wd->s));
    SR_HIP(hipMemcpyAsync(tmp, planes, sizeof(float) * n, hipMemcpyHostToDevice, wd->s));
    wd->cur = 0;
    hipLaunchKernelGGL(k_upload, dim3((wd->n + 255) / 256), dim3(256), 0, wd->s, tmp, wd->st[0], (int)wd->n);
    SR_HIP(hipFreeAsync(tmp, wd->s));
    hipLaunchKernelGGL(k_scal_reset, dim3(1), dim3(64), 0, wd->s, wd->scal);
    SR_HIP(hipMemsetAsync(wd->escaped, 0, sizeof(double) * 14 * 2 * (wd->w + wd->h), wd->s));
    wd->have_flux = false;
    wd->heat_ran = false;
    wd->first_built = false;
    wd->steps = 0;
    wd->last_heat = sr_heat_report{};
    wd->wk = {};
    if (wd->sw.heat) {
        wd->heat.work = HeatWork{};
        wd->heat.solver.work = FldWork{};
    }
    // the box fit before step 0, ahead of the initial dt (step/boxfit.rs:10); then the twin's start: P2 on the
    // state and dt_0 from it (reference/mod.rs:173-175), no floors
    SR_TRY(box_refit(wd, BOX_AT_UPLOAD));
    SR_TRY(p2(wd));
    SR_TRY(p89(wd, 0));
    SR_HIP(hipStreamSynchronize(wd->s));
    return SR_OK;
}

extern "C" int sr_readback(sr_world *wd, float *planes, size_t n) {
    SR_TRY(state_ok(wd, n));
    if (!planes) return fail(SR_ERR_ARGUMENT, "sr_readback: null planes");
    float *tmp = nullptr;
    SR_HIP(hipMallocAsync((void **)&tmp, sizeof(float) * n, wd->s));
    hipLaunchKernelGGL(k_readback, dim3((wd->n + 255) / 256), dim3(256), 0, wd->s, wd->st[wd->cur], tmp, (int)wd->n);
    SR_HIP(hipMemcpyAsync(planes, tmp, sizeof(float) * n, hipMemcpyDeviceToHost, wd->s));
    SR_HIP(hipFreeAsync(tmp, wd->s));
    SR_HIP(hipStreamSynchronize(wd->s));
    return SR_OK;
}

__global__ void k_sigma_in(const double *__restrict__ src, double *__restrict__ sigma, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) sigma[i] = src[i];
}

extern "C" int sr_resume(sr_world *wd, const sr_resume_records *rec, const float *flux, size_t flux_n,
                         const double *sigma, size_t sigma_n) {
    if (!wd || !rec) return fail(SR_ERR_ARGUMENT, "sr_resume: null pointer");
    if (!(rec->dt >= 0.0) || !(rec->dt_next > 0.0) || !std::isfinite(rec->dt) || !std::isfinite(rec->dt_next) ||
        !std::isfinite(rec->sim_time))
        return fail(SR_ERR_ARGUMENT, "sr_resume: dt >= 0, dt_next > 0 and a finite sim time are required");
    if (flux && flux_n != 2 * wd->n)
        return fail(SR_ERR_LENGTH, "sr_resume: F holds " + std::to_string(flux_n) + " values, expected " +
                                       std::to_string(2 * wd->n));
    if (sigma && sigma_n != wd->n)
        return fail(SR_ERR_LENGTH, "sr_resume: Sigma holds " + std::to_string(sigma_n) + " values, expected " +
                                       std::to_string(wd->n));
    double h[3] = {rec->dt, rec->dt_next, rec->sim_time};
    static_assert(SCAL_DT == 0 && SCAL_DT_NEXT == 1 && SCAL_SIM_TIME == 2, "the clock's slots are the first three");
    SR_HIP(hipMemcpyAsync(wd->scal, h, sizeof h, hipMemcpyHostToDevice, wd->s));
    if (flux && wd->sw.heat) {
        SR_HIP(hipMemcpyAsync(wd->rfx, flux, sizeof(float) * wd->n, hipMemcpyHostToDevice, wd->s));
        SR_HIP(hipMemcpyAsync(wd->rfy, flux + wd->n, sizeof(float) * wd->n, hipMemcpyHostToDevice, wd->s));
        wd->have_flux = true;
    }
    if (sigma) {
        double *tmp = nullptr;
        SR_HIP(hipMallocAsync((void **)&tmp, sizeof(double) * wd->n, wd->s));
        SR_HIP(hipMemcpyAsync(tmp, sigma, sizeof(double) * wd->n, hipMemcpyHostToDevice, wd->s));
        hipLaunchKernelGGL(k_sigma_in, dim3((wd->n + 255) / 256), dim3(256), 0, wd->s, (const double *)tmp,
                           wd->st[wd->cur].sigma, (int)wd->n);
        SR_HIP(hipGetLastError());
        SR_HIP(hipFreeAsync(tmp, wd->s));
    }
    SR_HIP(hipStreamSynchronize(wd->s));
    return SR_OK;
}

extern "C" int sr_readback_sigma(sr_world *wd, double *sigma, size_t n) {
    if (!wd || !sigma) return fail(SR_ERR_ARGUMENT, "sr_readback_sigma: null pointer");
    if (n != wd->n) return fail(SR_ERR_LENGTH, "sr_readback_sigma: Sigma holds w*h values");
    SR_HIP(hipMemcpyAsync(sigma, wd->st[wd->cur].sigma, sizeof(double) * n, hipMemcpyDeviceToHost, wd->s));
    SR_HIP(hipStreamSynchronize(wd->s));
    return SR_OK;
}

extern "C" int sr_heat_first_system(sr_world *wd, float *own, float *kx, float *ky, uint8_t *gas, size_t n) {
    if (!wd) return fail(SR_ERR_ARGUMENT, "sr_heat_first_system: null world");
    if (!wd->sw.heat || !wd->heat_ran)
        return fail(SR_ERR_ARGUMENT, "sr_heat_first_system: no step with heat has run since the upload");
    if (n != wd->n) return fail(SR_ERR_LENGTH, "sr_heat_first_system: n must be w*h");
    FldSolver &f = wd->heat.solver;
    SR_TRY(wd->heat.first_system(wd->scal, wd->s));
    const size_t nx = (size_t)(wd->w + 1) * wd->h, ny = (size_t)wd->w * (wd->h + 1);
    if (own) SR_HIP(hipMemcpyAsync(own, f.own0(), sizeof(float) * n, hipMemcpyDeviceToHost, wd->s));
    if (kx) SR_HIP(hipMemcpyAsync(kx, f.ax0(), sizeof(float) * nx, hipMemcpyDeviceToHost, wd->s));
    if (ky) SR_HIP(hipMemcpyAsync(ky, f.ay0(), sizeof(float) * ny, hipMemcpyDeviceToHost, wd->s));
    if (gas) SR_HIP(hipMemcpyAsync(gas, f.gas0(), n, hipMemcpyDeviceToHost, wd->s));
    SR_HIP(hipStreamSynchronize(wd->s));
    wd->first_built = true;
    return SR_OK;
}

extern "C" int sr_apply_heat_precond(sr_world *wd, const float *r, float *z, size_t n) {
    if (!wd || !r || !z) return fail(SR_ERR_ARGUMENT, "sr_apply_heat_precond: null pointer");
    if (!wd->first_built) return fail(SR_ERR_ARGUMENT, "sr_apply_heat_precond: no sr_heat_first_system yet");
    if (n != wd->n) return fail(SR_ERR_LENGTH, "sr_apply_heat_precond: n must be w*h");
    float *d = nullptr;
    SR_HIP(hipMallocAsync((void **)&d, sizeof(float) * 2 * n, wd->s));
    SR_HIP(hipMemcpyAsync(d, r, sizeof(float) * n, hipMemcpyHostToDevice, wd->s));
    SR_TRY(wd->heat.solver.apply_precond(d, d + n, wd->precond, wd->s));
    std::vector<uint8_t> gas(n);
    SR_HIP(hipMemcpyAsync(z, d + n, sizeof(float) * n, hipMemcpyDeviceToHost, wd->s));
    SR_HIP(hipMemcpyAsync(gas.data(), wd->heat.solver.gas0(), n, hipMemcpyDeviceToHost, wd->s));
    SR_HIP(hipFreeAsync(d, wd->s));
    SR_HIP(hipStreamSynchronize(wd->s));
    for (size_t i = 0; i < n; i++)
        if (!gas[i]) z[i] = 0.0f;  // rows off the gas are no part of the system
    return SR_OK;
}

static void mark(sr_world *wd, int k) {
    if (wd->timing) (void)hipEventRecord(wd->ev[k], wd->s);
}

static double between(sr_world *wd, int a, int b) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, wd->ev[a], wd->ev[b]) != hipSuccess) return 0.0;
    return ms;
}

// P1 .. P4 and the take-backs: launches only (no host decision), so the sequence can be captured
static int step_front(sr_world *wd) {
    hipStream_t s = wd->s;
    mark(wd, 0);
    hipLaunchKernelGGL(k_p1, dim3(1), dim3(64), 0, s, wd->scal);
    mark(wd, 1);
    SR_TRY(p2(wd));
    mark(wd, 13);
    if (wd->sinks_on) {  // P3 after P2's field, with the sinks' part added to it
        if (wd->sw.gravity) SR_TRY(sinks_field(wd->sk_par, wd->sk, wd->phi, wd->gx, wd->gy, s));
        SR_TRY(sinks_pass(wd->sk_par, wd->sk, wd->st[wd->cur], wd->phi, -1.0, wd->scal, s));
    }
    mark(wd, 2);
    if (wd->sw.hydro) {
        HydroState &in = wd->st[wd->cur], &out = wd->st[1 - wd->cur];
        const bool grav = wd->sw.gravity, rad = wd->sw.heat && wd->have_flux;
        if (grav) {
            mark(wd, 8);
            SR_TRY(hydro_launch(HYDRO_MID, wd->hp, in, out, wd->scal, wd->gx, wd->gy, nullptr, nullptr, nullptr,
                                nullptr, wd->sigma_mid, nullptr, wd->escaped, wd->part, s));
            mark(wd, 9);
            SR_TRY(wd->grav.potential<float>(wd->sigma_mid, wd->phi_mid, s));
        }
        mark(wd, 10);
        const SinkPlanes sp{wd->sk.ps, wd->sk.gsx, wd->sk.gsy};
        SR_TRY(hydro_launch(HYDRO_UPDATE, wd->hp, in, out, wd->scal, grav ? wd->gx : nullptr,
                            grav ? wd->gy : nullptr, grav ? wd->sigma_mid : nullptr, grav ? wd->phi_mid : nullptr,
                            rad ? wd->rfx : nullptr, rad ? wd->rfy : nullptr, nullptr, wd->pushmag, wd->escaped,
                            wd->part, s, wd->sinks_on && grav ? &sp : nullptr));
        mark(wd, 11);
        if (grav || rad) SR_TRY(hydro_takeback(wd->hp, out, wd->pushmag, wd->part, wd->takeback, rad, s));
    }
    return SR_OK;
}

// `launch` captured on the world's stream into *g (once), then the graph launched
template <typename F> static int replay(sr_world *wd, hipGraphExec_t *g, F launch) {
    if (!*g) {
        hipGraph_t graph = nullptr;
        SR_HIP(hipStreamBeginCapture(wd->s, hipStreamCaptureModeThreadLocal));
        const int c = launch();
        const hipError_t e = hipStreamEndCapture(wd->s, &graph);
        SR_TRY(c);
        if (e != hipSuccess) return fail(SR_ERR_DEVICE, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
        const hipError_t ei = hipGraphInstantiate(g, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (ei != hipSuccess) {
            *g = nullptr;
            return fail(SR_ERR_DEVICE, std::string("hipGraphInstantiate: ") + hipGetErrorString(ei));
        }
    }
    SR_HIP(hipGraphLaunch(*g, wd->s));
    return SR_OK;
}

static int one_step(sr_world *wd) {
    hipStream_t s = wd->s;
    const bool graphs = wd->graphs && !wd->timing;
    const bool grav = wd->sw.gravity, rad = wd->sw.heat && wd->have_flux;
    if (graphs)
        SR_TRY(replay(wd, &wd->g_front[wd->cur][rad ? 1 : 0], [&] { return step_front(wd); }));
    else
        SR_TRY(step_front(wd));
    wd->wk.steps++;
    if (grav) wd->wk.conv_first++;
    if (wd->sw.hydro) {
        if (grav) {
            wd->wk.mid++;
            wd->wk.conv_mid++;
        }
        wd->wk.hydro++;
        if (rad) wd->wk.radiation++;
        wd->cur = 1 - wd->cur;
    }
    mark(wd, 3);
    if (wd->sw.heat) {
        SR_HIP(hipMemcpyAsync(wd->host_scal, wd->scal, sizeof(double) * SCAL_N, hipMemcpyDeviceToHost, s));
        SR_HIP(hipStreamSynchronize(s));
        HydroState &st = wd->st[wd->cur];
        HydroLike hl = {st.sigma, st.mx, st.my, st.e, st.x};
        SR_TRY(wd->heat.step(hl, wd->scal, wd->host_scal[SCAL_DT], wd->rfx, wd->rfy, wd->precond, s));
        wd->last_heat = wd->heat.report;
        wd->have_flux = true;
        wd->heat_ran = true;
        wd->first_built = false;
    }
    mark(wd, 12);
    if (wd->react) {  // P6 at the step's dt
        HydroState &st = wd->st[wd->cur];
        SR_TRY(react_launch(wd->react, st, (int)wd->n, -1.0, wd->scal, wd->react_nuc, wd->react_lost, wd->react_snu,
                            wd->react_psum, wd->react_pmax, &wd->react_nb, s));
    }
    mark(wd, 4);
    if (graphs)
        SR_TRY(replay(wd, &wd->g_tail[wd->cur], [&] { return p89(wd, 1); }));
    else
        SR_TRY(p89(wd, 1));
    mark(wd, 5);
    wd->steps++;
    if (wd->timing) {
        SR_HIP(hipEventSynchronize(wd->ev[5]));
        wd->cost.pass_ms[1] += between(wd, 0, 1);
        wd->cost.pass_ms[2] += between(wd, 1, 13);
        wd->cost.pass_ms[3] += between(wd, 13, 2);
        wd->cost.pass_ms[4] += between(wd, 2, 3);
        wd->cost.pass_ms[5] += between(wd, 3, 12);
        wd->cost.pass_ms[6] += between(wd, 12, 4);
        wd->cost.pass_ms[8] += between(wd, 4, 5);
        wd->cost.step_ms += between(wd, 0, 5);
        wd->cost.steps++;
        if (wd->sw.hydro) {
            wd->k_update.ms += between(wd, 10, 11);
            wd->k_update.launches++;
            if (wd->sw.gravity) {
                wd->k_mid.ms += between(wd, 8, 9);
                wd->k_mid.launches++;
            }
        }
    }
    return SR_OK;
}

extern "C" int sr_run(sr_world *wd, uint32_t steps) {
    if (!wd) return fail(SR_ERR_ARGUMENT, "null world");
    for (uint32_t k = 0; k < steps; k++) SR_TRY(one_step(wd));
    SR_HIP(hipStreamSynchronize(wd->s));
    return SR_OK;
}

// Independent worlds stepped together. Each world stays on one host thread for the call (its stream and its heat
// pass's host decisions are that thread's), the worlds shared out largest first onto the least-loaded thread; a
// thread takes its worlds one step each in turn, so while one waits on a heat-pass decision the others' kernels,
// on their own streams, keep the device busy. Every world's result is the one it reaches stepped alone.
extern "C" int sr_ensemble_run(sr_world **worlds, int count, uint32_t steps, int threads) {
    if (!worlds || count <= 0) return fail(SR_ERR_ARGUMENT, "sr_ensemble_run: no worlds");
    if (threads <= 0 || threads > 256) return fail(SR_ERR_ARGUMENT, "sr_ensemble_run: threads must be in 1..256");
    std::vector<sr_world *> ws(worlds, worlds + count);
    for (sr_world *w : ws)
        if (!w) return fail(SR_ERR_ARGUMENT, "sr_ensemble_run: a null world");
    {
        std::vector<sr_world *> sorted = ws;
        std::sort(sorted.begin(), sorted.end());
        if (std::adjacent_find(sorted.begin(), sorted.end()) != sorted.end())
            return fail(SR_ERR_ARGUMENT, "sr_ensemble_run: a world appears twice");
    }
    const int nt = std::min(threads, count);
    std::vector<int> order(count);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return ws[a]->n > ws[b]->n; });
    std::vector<std::vector<sr_world *>> mine(nt);
    std::vector<size_t> load(nt, 0);
    for (int i : order) {
        const int t = (int)(std::min_element(load.begin(), load.end()) - load.begin());
        mine[t].push_back(ws[i]);
        load[t] += ws[i]->n;
    }
    int dev = 0;
    SR_HIP(hipGetDevice(&dev));
    std::atomic<bool> stop{false};
    std::mutex mu;
    int err = SR_OK;
    std::string msg;
    auto body = [&](int t) {
        auto bail = [&](int c) {
            std::lock_guard<std::mutex> g(mu);
            if (err == SR_OK) {
                err = c;
                msg = sr_last_error();
            }
            stop = true;
        };
        if (hipSetDevice(dev) != hipSuccess) return bail(fail(SR_ERR_DEVICE, "hipSetDevice"));
        for (uint32_t k = 0; k < steps && !stop.load(); k++) {
            for (sr_world *wd : mine[t]) {
                const int c = one_step(wd);
                if (c != SR_OK) return bail(c);
            }
            for (sr_world *wd : mine[t])
                if (hipStreamSynchronize(wd->s) != hipSuccess) return bail(fail(SR_ERR_DEVICE, "hipStreamSynchronize"));
        }
    };
    if (nt == 1) {
        body(0);
    } else {
        std::vector<std::thread> pool;
        for (int t = 0; t < nt; t++) pool.emplace_back(body, t);
        for (auto &th : pool) th.join();
    }
    if (err != SR_OK) return fail(err, msg);
    return SR_OK;
}

extern "C" int sr_read_dt(sr_world *wd, sr_dt_report *out) {
    if (!wd || !out) return fail(SR_ERR_ARGUMENT, "null pointer");
    SR_HIP(hipMemcpyAsync(wd->host_scal, wd->scal, sizeof(double) * SCAL_N, hipMemcpyDeviceToHost, wd->s));
    SR_HIP(hipStreamSynchronize(wd->s));
    const double *h = wd->host_scal;
    std::memset(out, 0, sizeof *out);
    out->dt = h[SCAL_DT];
    out->dt_next = h[SCAL_DT_NEXT];
    out->sim_time = h[SCAL_SIM_TIME];
    out->steps = wd->steps;
    if (h[SCAL_NF] > -HUGE_VAL) {
        const long key = (long)(-h[SCAL_NF]);
        out->non_finite = 1;
        const long cell = key / 16;
        out->nf_x = (int)(cell % wd->w);
        out->nf_y = (int)(cell / wd->w);
        out->nf_channel = (int)(key % 16);
        out->nf_step = (uint64_t)h[SCAL_NF_STEP];
    }
    return SR_OK;
}

extern "C" int sr_read_heat(sr_world *wd, sr_heat_report *out) {
    if (!wd || !out) return fail(SR_ERR_ARGUMENT, "null pointer");
    *out = wd->last_heat;
    return SR_OK;
}

extern "C" int sr_read_field(sr_world *wd, int field, float *out, size_t n) {
    if (!wd || !out) return fail(SR_ERR_ARGUMENT, "null pointer");
    if (n != wd->n) return fail(SR_ERR_LENGTH, "a side field holds w*h values");
    const float *src = nullptr;
    float *tmp = nullptr;
    switch (field) {
    case SR_FIELD_PHI:
    case SR_FIELD_PHI_MID:
        SR_HIP(hipMallocAsync((void **)&tmp, sizeof(float) * n, wd->s));
        hipLaunchKernelGGL(k_extract_box, dim3((n + 255) / 256), dim3(256), 0, wd->s,
                           field == SR_FIELD_PHI ? wd->phi : wd->phi_mid, wd->w, wd->h, tmp);
        src = tmp;
        break;
    case SR_FIELD_GX: src = wd->gx; break;
    case SR_FIELD_GY: src = wd->gy; break;
    case SR_FIELD_FLUX_X: src = wd->rfx; break;
    case SR_FIELD_FLUX_Y: src = wd->rfy; break;
    case SR_FIELD_SIGMA_MID: src = wd->sigma_mid; break;
    default: return fail(SR_ERR_ARGUMENT, "unknown field");
    }
    SR_HIP(hipMemcpyAsync(out, src, sizeof(float) * n, hipMemcpyDeviceToHost, wd->s));
    if (tmp) SR_HIP(hipFreeAsync(tmp, wd->s));
    SR_HIP(hipStreamSynchronize(wd->s));
    return SR_OK;
}

extern "C" int sr_read_totals(sr_world *wd, double *totals, double *escaped) {
    if (!wd) return fail(SR_ERR_ARGUMENT, "null world");
    const int nb = grid_n(wd->n);
    hipLaunchKernelGGL(k_totals, dim3(nb), dim3(NT), 0, wd->s, wd->st[wd->cur], (int)wd->n, wd->rpart);
    hipLaunchKernelGGL(k_escaped_sum, dim3(14), dim3(256), 0, wd->s, wd->escaped, 2 * (wd->w + wd->h),
                       wd->totals + 16);
    SR_HIP(hipGetLastError());
    std::vector<double> part((size_t)14 * nb);
    double h[32];
    SR_HIP(hipMemcpyAsync(part.data(), wd->rpart, sizeof(double) * 14 * nb, hipMemcpyDeviceToHost, wd->s));
    SR_HIP(hipMemcpyAsync(h, wd->totals, sizeof(double) * 32, hipMemcpyDeviceToHost, wd->s));
    SR_HIP(hipStreamSynchronize(wd->s));
    for (int k = 0; k < 14; k++) {
        double t = 0.0;
        for (int i = 0; i < nb; i++) t += part[(size_t)k * nb + i];
        h[k] = t;
    }
    if (totals) std::memcpy(totals, h, sizeof(double) * 14);
    if (escaped) std::memcpy(escaped, h + 16, sizeof(double) * 14);
    return SR_OK;
}

extern "C" int sr_set_precond(sr_world *wd, int precond) {
    if (!wd) return fail(SR_ERR_ARGUMENT, "null world");
    if (precond != SR_PRECOND_JACOBI && precond != SR_PRECOND_MULTIGRID && precond != SR_PRECOND_LEARNED)
        return fail(SR_ERR_ARGUMENT, "unknown preconditioner");
    if (precond == SR_PRECOND_LEARNED && !wd->net)
        return fail(SR_ERR_ARGUMENT, "the learned preconditioner needs a network (sr_set_network)");
    wd->precond = precond;
    return SR_OK;
}

extern "C" int sr_set_network(sr_world *wd, sr_net *n, int fallback_iterations) {
    if (!wd) return fail(SR_ERR_ARGUMENT, "null world");
    if (!wd->sw.heat) return fail(SR_ERR_ARGUMENT, "sr_set_network: the world has no heat pass");
    SR_TRY(wd->heat.solver.set_network(net_of(n), fallback_iterations));
    wd->net = n;
    if (!n && wd->precond == SR_PRECOND_LEARNED) wd->precond = SR_PRECOND_MULTIGRID;
    return SR_OK;
}

extern "C" int sr_set_solve_mode(sr_world *wd, int mode) {
    if (!wd) return fail(SR_ERR_ARGUMENT, "null world");
    if (mode != SR_SOLVE_HOST && mode != SR_SOLVE_DEVICE && mode != SR_SOLVE_GRAPH && mode != SR_SOLVE_HOST_WAIT)
        return fail(SR_ERR_ARGUMENT, "sr_set_solve_mode: mode must be SR_SOLVE_HOST, _DEVICE, _GRAPH or _HOST_WAIT");
    if (wd->sw.heat) wd->heat.solver.set_solve_mode(mode);
    return SR_OK;
}

extern "C" int sr_set_reactions(sr_world *wd, const sr_react_params *p) {
    if (!wd) return fail(SR_ERR_ARGUMENT, "null world");
    if (!p) {
        wd->react = nullptr;  // its device memory stays with the world until sr_destroy
        return SR_OK;
    }
    if (react_check(p) != SR_OK)
        return fail(SR_ERR_ARGUMENT, "sr_set_reactions: a record out of range (species, group, order, terms, gate) "
                                     "or N_Fe's closed form on a record not of its shape");
    const size_t n = wd->n;
    if (!wd->react_nuc) {
        void *d = nullptr;
        SR_TRY(world_alloc(wd, &d, sizeof(ReactDev)));
        wd->react = nullptr;
        wd->react_dev_mem = (ReactDev *)d;
        SR_TRY(world_alloc(wd, (void **)&wd->react_nuc, sizeof(double) * SR_REACT_GROUPS * n));
        SR_TRY(world_alloc(wd, (void **)&wd->react_lost, sizeof(double) * n));
        SR_TRY(world_alloc(wd, (void **)&wd->react_snu, sizeof(double) * n));
        SR_TRY(world_alloc(wd, (void **)&wd->react_psum, sizeof(double) * REACT_SUMS * REACT_BLOCKS));
        SR_TRY(world_alloc(wd, (void **)&wd->react_pmax, sizeof(double) * REACT_MAXES * REACT_BLOCKS));
    }
    ReactDev *h = new ReactDev;
    react_build(p, wd->ph, h);
    const hipError_t e = hipMemcpyAsync(wd->react_dev_mem, h, sizeof(ReactDev), hipMemcpyHostToDevice, wd->s);
    const hipError_t e2 = e == hipSuccess ? hipStreamSynchronize(wd->s) : e;
    delete h;
    if (e2 != hipSuccess) return fail(SR_ERR_DEVICE, std::string("sr_set_reactions: ") + hipGetErrorString(e2));
    wd->react = wd->react_dev_mem;
    return SR_OK;
}

extern "C" int sr_react_apply(sr_world *wd, double dt) {
    if (!wd) return fail(SR_ERR_ARGUMENT, "null world");
    if (!wd->react) return fail(SR_ERR_ARGUMENT, "sr_react_apply: no reactions (sr_set_reactions)");
    if (!(dt >= 0.0) || !std::isfinite(dt)) return fail(SR_ERR_ARGUMENT, "sr_react_apply: dt >= 0 is required");
    SR_TRY(react_launch(wd->react, wd->st[wd->cur], (int)wd->n, dt, wd->scal, wd->react_nuc, wd->react_lost,
                        wd->react_snu, wd->react_psum, wd->react_pmax, &wd->react_nb, wd->s));
    SR_HIP(hipStreamSynchronize(wd->s));
    return SR_OK;
}

extern "C" int sr_read_react(sr_world *wd, sr_react_report *out) {
    if (!wd || !out) return fail(SR_ERR_ARGUMENT, "sr_read_react: null pointer");
    *out = sr_react_report{};
    const int nb = wd->react_nb;
    if (!wd->react_nuc || nb == 0) return SR_OK;
    std::vector<double> ps((size_t)REACT_SUMS * nb), pm((size_t)REACT_MAXES * nb);
    SR_HIP(hipMemcpyAsync(ps.data(), wd->react_psum, sizeof(double) * ps.size(), hipMemcpyDeviceToHost, wd->s));
    SR_HIP(hipMemcpyAsync(pm.data(), wd->react_pmax, sizeof(double) * pm.size(), hipMemcpyDeviceToHost, wd->s));
    SR_HIP(hipStreamSynchronize(wd->s));
    double sum[REACT_SUMS] = {}, mx[REACT_MAXES] = {};
    for (int k = 0; k < REACT_SUMS; k++)
        for (int b = 0; b < nb; b++) sum[k] += ps[(size_t)k * nb + b];
    for (int k = 0; k < REACT_MAXES; k++)
        for (int b = 0; b < nb; b++) mx[k] = std::max(mx[k], pm[(size_t)k * nb + b]);
    out->substeps = (uint64_t)sum[0];
    out->retried = (uint64_t)sum[1];
    out->unfinished = (uint64_t)sum[2];
    for (int g = 0; g < SR_REACT_GROUPS; g++) out->nuclear[g] = sum[3 + g];
    out->neutrino_lost = sum[3 + SR_REACT_GROUPS];
    out->max_substeps = (uint64_t)mx[0];
    out->max_dx = mx[1];
    out->max_dt_rel = mx[2];
    out->latch = mx[3] > 0.0;
    return SR_OK;
}

extern "C" int sr_read_react_planes(sr_world *wd, double *nuclear, double *lost, double *s_nu, size_t n) {
    if (!wd) return fail(SR_ERR_ARGUMENT, "null world");
    if (n != wd->n) return fail(SR_ERR_LENGTH, "sr_read_react_planes: w h values per plane");
    if (!wd->react_nuc) return fail(SR_ERR_ARGUMENT, "sr_read_react_planes: no reactions (sr_set_reactions)");
    if (nuclear)
        SR_HIP(hipMemcpyAsync(nuclear, wd->react_nuc, sizeof(double) * SR_REACT_GROUPS * n, hipMemcpyDeviceToHost,
                              wd->s));
    if (lost) SR_HIP(hipMemcpyAsync(lost, wd->react_lost, sizeof(double) * n, hipMemcpyDeviceToHost, wd->s));
    if (s_nu) SR_HIP(hipMemcpyAsync(s_nu, wd->react_snu, sizeof(double) * n, hipMemcpyDeviceToHost, wd->s));
    SR_HIP(hipStreamSynchronize(wd->s));
    return SR_OK;
}

// the captured step front holds P3's launches and the hydro variant: a change of P3 captures it again
static void drop_front_graphs(sr_world *wd) {
    for (auto &row : wd->g_front)
        for (auto &g : row) {
            if (g) (void)hipGraphExecDestroy(g);
            g = nullptr;
        }
}

extern "C" int sr_set_sinks(sr_world *wd, const sr_sinks_params *p) {
    if (!wd) return fail(SR_ERR_ARGUMENT, "sr_set_sinks: null world");
    SR_HIP(hipStreamSynchronize(wd->s));
    if (!p) {
        if (wd->sinks_on) drop_front_graphs(wd);
        wd->sinks_on = false;
        return SR_OK;
    }
    if (!(p->sigma_bh > 0.0) || !(p->m_tov > 0.0))
        return fail(SR_ERR_ARGUMENT, "sr_set_sinks: Sigma_BH and m_tov_sb must be > 0 (m_tov_sb may be infinite)");
    if (wd->ph.dx != 1.0) return fail(SR_ERR_ARGUMENT, "sr_set_sinks: the sinks' law is in cells: dx must be 1");
    if (!wd->sk.dev) {
        SinkBufs &B = wd->sk;
        sinks_alloc_sizes(wd->w, wd->h, &B.tiles, &B.pull_blocks);
        SR_TRY(world_alloc(wd, (void **)&B.dev, sizeof(SinkDev)));
        SR_TRY(world_alloc(wd, (void **)&B.cand_s, sizeof(double) * B.tiles));
        SR_TRY(world_alloc(wd, (void **)&B.cand_i, sizeof(int) * B.tiles));
        SR_TRY(world_alloc(wd, (void **)&B.pull, sizeof(double) * 2 * SR_MAX_SINKS * B.pull_blocks));
        SR_TRY(world_alloc(wd, (void **)&B.ps, sizeof(float) * (wd->w + 2) * (wd->h + 2)));
        SR_TRY(world_alloc(wd, (void **)&B.gsx, sizeof(float) * wd->n));
        SR_TRY(world_alloc(wd, (void **)&B.gsy, sizeof(float) * wd->n));
    }
    SinkParams &K = wd->sk_par;
    K.w = wd->w;
    K.h = wd->h;
    K.g = wd->ph.G;
    K.sigma_bh = p->sigma_bh;
    K.sigma_vac = wd->ph.sigma_vac;
    K.sigma_floor = wd->ph.sigma_floor;
    K.c_sb = wd->ph.c_sb;
    K.m_tov = p->m_tov;
    K.t_floor = wd->ph.t_floor;
    K.gm1 = wd->ph.gamma - 1.0;
    K.fixed_imu = wd->ph.mu > 0.0 ? 1.0 / wd->ph.mu : 0.0;
    K.cold = wd->fp.cold;
    SinkDev h{};
    h.formed_x = h.formed_y = -1;
    SR_HIP(hipMemcpy(wd->sk.dev, &h, sizeof h, hipMemcpyHostToDevice));
    if (!wd->sinks_on) drop_front_graphs(wd);
    wd->sinks_on = true;
    return SR_OK;
}

extern "C" int sr_put_sinks(sr_world *wd, const sr_sink *sinks, int count) {
    if (!wd) return fail(SR_ERR_ARGUMENT, "sr_put_sinks: null world");
    if (!wd->sinks_on) return fail(SR_ERR_ARGUMENT, "sr_put_sinks: no P3 (sr_set_sinks)");
    if (count < 0 || count > SR_MAX_SINKS || (count > 0 && !sinks))
        return fail(SR_ERR_ARGUMENT, "sr_put_sinks: 0 to 16 sinks");
    SinkDev h{};
    h.count = count;
    h.formed_x = h.formed_y = -1;
    for (int q = 0; q < count; q++) {
        if (!(sinks[q].m > 0.0) || !std::isfinite(sinks[q].x) || !std::isfinite(sinks[q].y) ||
            !std::isfinite(sinks[q].vx) || !std::isfinite(sinks[q].vy) || !std::isfinite(sinks[q].m))
            return fail(SR_ERR_ARGUMENT, "sr_put_sinks: a sink's mass must be > 0 and every value finite");
        h.x[q] = sinks[q].x;
        h.y[q] = sinks[q].y;
        h.vx[q] = sinks[q].vx;
        h.vy[q] = sinks[q].vy;
        h.m[q] = sinks[q].m;
    }
    SR_HIP(hipStreamSynchronize(wd->s));
    SR_HIP(hipMemcpy(wd->sk.dev, &h, sizeof h, hipMemcpyHostToDevice));
    return SR_OK;
}

extern "C" int sr_read_sinks(sr_world *wd, sr_sink *out, int max, sr_sinks_report *rep) {
    if (!wd) return fail(SR_ERR_ARGUMENT, "sr_read_sinks: null world");
    SinkDev h{};
    h.formed_x = h.formed_y = -1;
    if (wd->sk.dev) {
        SR_HIP(hipMemcpyAsync(&h, wd->sk.dev, sizeof h, hipMemcpyDeviceToHost, wd->s));
        SR_HIP(hipStreamSynchronize(wd->s));
    }
    if (!wd->sinks_on) h.count = 0;
    for (int q = 0; q < h.count && q < max && out; q++) out[q] = sr_sink{h.x[q], h.y[q], h.vx[q], h.vy[q], h.m[q]};
    if (rep) *rep = sr_sinks_report{h.count, h.formed_x, h.formed_y, h.merged, h.accreted, h.swallowed};
    return SR_OK;
}

extern "C" int sr_sinks_apply(sr_world *wd, double dt, const float *phi, size_t n) {
    if (!wd) return fail(SR_ERR_ARGUMENT, "sr_sinks_apply: null world");
    if (!wd->sinks_on) return fail(SR_ERR_ARGUMENT, "sr_sinks_apply: no P3 (sr_set_sinks)");
    if (!(dt >= 0.0) || !std::isfinite(dt)) return fail(SR_ERR_ARGUMENT, "sr_sinks_apply: dt >= 0 is required");
    if (phi) {
        if (n != wd->n) return fail(SR_ERR_LENGTH, "sr_sinks_apply: phi holds w h values");
        SR_HIP(hipMemcpy2DAsync(wd->phi + (wd->w + 2) + 1, sizeof(float) * (wd->w + 2), phi, sizeof(float) * wd->w,
                                sizeof(float) * wd->w, wd->h, hipMemcpyHostToDevice, wd->s));
    }
    SR_TRY(sinks_pass(wd->sk_par, wd->sk, wd->st[wd->cur], wd->phi, dt, wd->scal, wd->s));
    SR_HIP(hipStreamSynchronize(wd->s));
    return SR_OK;
}

extern "C" int sr_set_box_mode(sr_world *wd, int mode) {
    if (!wd) return fail(SR_ERR_ARGUMENT, "null world");
    if (mode != SR_BOX_WHOLE_WORLD && mode != SR_BOX_FIT)
        return fail(SR_ERR_ARGUMENT, "sr_set_box_mode: SR_BOX_WHOLE_WORLD or SR_BOX_FIT");
    wd->box_fit = mode == SR_BOX_FIT;
    return SR_OK;
}

extern "C" int sr_read_box(sr_world *wd, sr_box_report *out) {
    if (!wd || !out) return fail(SR_ERR_ARGUMENT, "sr_read_box: null pointer");
    int r[8];
    SR_HIP(hipMemcpyAsync(r, wd->box, sizeof r, hipMemcpyDeviceToHost, wd->s));
    SR_HIP(hipStreamSynchronize(wd->s));
    sr_box_report b{};
    b.x0 = (uint32_t)r[0];
    b.y0 = (uint32_t)r[1];
    b.w = (uint32_t)r[2];
    b.h = (uint32_t)r[3];
    b.fft[0] = (uint32_t)r[4];
    b.fft[1] = (uint32_t)r[5];
    b.from_step = (uint32_t)r[6];
    b.fits = (uint32_t)r[7];
    b.tiles[0] = b.w / 8;
    b.tiles[1] = b.h / 8;
    b.tiles[2] = 1;
    b.cells[0] = (b.w * b.h + 255) / 256;
    b.cells[1] = b.cells[2] = 1;
    *out = b;
    return SR_OK;
}

extern "C" int sr_resume_box(sr_world *wd, const uint32_t *record, uint64_t step) {
    if (!wd || !record) return fail(SR_ERR_ARGUMENT, "sr_resume_box: null pointer");
    const uint32_t x0 = record[0], y0 = record[1], bw = record[2], bh = record[3];
    if (x0 % 8 || y0 % 8 || bw % 8 || bh % 8 || (uint64_t)x0 + bw > (uint64_t)wd->w ||
        (uint64_t)y0 + bh > (uint64_t)wd->h || (bw == 0) != (bh == 0) || step > (1ull << 52))
        return fail(SR_ERR_ARGUMENT, "sr_resume_box: a box of multiples of 8 inside the world is required");
    int r[8];
    for (int k = 0; k < 8; k++) r[k] = (int)record[k];
    SR_HIP(hipMemcpyAsync(wd->box, r, sizeof r, hipMemcpyHostToDevice, wd->s));
    const double st = (double)step;
    SR_HIP(hipMemcpyAsync(wd->scal + SCAL_STEPS, &st, sizeof st, hipMemcpyHostToDevice, wd->s));
    SR_HIP(hipStreamSynchronize(wd->s));
    wd->steps = step;
    return SR_OK;
}

extern "C" int sr_set_fusion(sr_world *wd, int pre, int post) {
    if (!wd) return fail(SR_ERR_ARGUMENT, "null world");
    return wd->sw.heat ? wd->heat.solver.set_fusion(pre, post) : SR_OK;
}

extern "C" int sr_set_update_fusion(sr_world *wd, int on) {
    if (!wd) return fail(SR_ERR_ARGUMENT, "null world");
    if (wd->sw.heat) wd->heat.solver.set_update_fusion(on != 0);
    return SR_OK;
}

extern "C" int sr_set_dt_cap(sr_world *wd, double cap) {
    if (!wd) return fail(SR_ERR_ARGUMENT, "null world");
    if (!(cap > 0.0) || !std::isfinite(cap)) return fail(SR_ERR_ARGUMENT, "a dt cap must be positive and finite");
    SR_HIP(hipMemcpyAsync(wd->scal + SCAL_CAP, &cap, sizeof(double), hipMemcpyHostToDevice, wd->s));
    SR_HIP(hipStreamSynchronize(wd->s));
    return SR_OK;
}

extern "C" int sr_set_timing(sr_world *wd, int on) {
    if (!wd) return fail(SR_ERR_ARGUMENT, "null world");
    wd->timing = on != 0;
    wd->cost = sr_step_cost{};
    wd->k_update = wd->k_mid = wd->k_grav = wd->k_pre0 = KernelClock{};
    if (wd->sw.heat) {
        wd->heat.solver.set_timing(wd->timing);
        wd->heat.solver.reset_clocks();
    }
    return SR_OK;
}

extern "C" int sr_read_cost(sr_world *wd, sr_step_cost *out) {
    if (!wd || !out) return fail(SR_ERR_ARGUMENT, "null pointer");
    *out = wd->cost;
    if (wd->cost.steps) {
        out->step_ms = wd->cost.step_ms / wd->cost.steps;
        for (int k = 0; k < 10; k++) out->pass_ms[k] = wd->cost.pass_ms[k] / wd->cost.steps;
    }
    return SR_OK;
}

// Per cell: P2's convolution (the box's Sigma in, f64 for the first and f32 for the mid-step one; the half-length row
// spectra, (P/2 + 1) h complex values, written and read; the column outputs, (P/2 + 1)(h + 2), written and read; the
// kernel's quarter spectrum; phi with its ring out: about 40 B at P = 2n) and the gradient (phi in, g_x, g_y out: 12); P4's
// mass-flux pass (the state in, g in, Sigma_mid out: 60 + 8 + 4) and update (the state in and out, g, Sigma_mid and
// phi_mid in, F and |push| in with radiation: 120 + 16 + 12); P8/P9 (the state in and out, g in: 128).
extern "C" int sr_read_work(sr_world *wd, sr_step_work *out, int reset) {
    if (!wd || !out) return fail(SR_ERR_ARGUMENT, "sr_read_work: null pointer");
    const double n = (double)wd->n;
    sr_step_work w{};
    w.steps = wd->wk.steps;
    w.hydro_steps = wd->wk.hydro;
    w.convolutions = wd->wk.conv_first + wd->wk.conv_mid;
    double conv = 0.0;
    if (wd->sw.gravity) {
        const double P = (double)wd->grav.px, Q = (double)wd->grav.py;
        const double fft = (2.0 * 8.0 * (P / 2.0 + 1.0) * wd->h + 2.0 * 8.0 * (P / 2.0 + 1.0) * (wd->h + 2.0) +
                            4.0 * (P / 2.0 + 1.0) * (Q / 2.0 + 1.0) + 4.0 * (wd->w + 2.0) * (wd->h + 2.0)) /
                           n;
        conv = (8.0 + fft + 12.0) * wd->wk.conv_first + (4.0 + fft) * wd->wk.conv_mid;
    }
    w.bytes_p2 = n * conv;
    w.bytes_p4 = n * (72.0 * wd->wk.mid + (120.0 + (wd->sw.gravity ? 16.0 : 0.0)) * wd->wk.hydro + 12.0 * wd->wk.radiation);
    w.bytes_p89 = n * 128.0 * wd->wk.steps;
    if (wd->sw.heat) {
        const FldWork &f = wd->heat.solver.work;
        const HeatWork &h = wd->heat.work;
        w.heat_steps = h.steps;
        w.newton_passes = h.passes;
        w.pcg_iterations = f.iterations;
        w.precond_applications = f.precond;
        w.true_residuals = f.true_residuals;
        w.hierarchy_builds = f.builds_faces + f.builds_own;
        w.bytes_p5 = wd->heat.work_bytes(h) + wd->heat.solver.work_bytes(f);
    }
    w.algorithmic_bytes = w.bytes_p2 + w.bytes_p4 + w.bytes_p5 + w.bytes_p89;
    *out = w;
    if (reset) {
        wd->wk = {};
        if (wd->sw.heat) {
            wd->heat.work = HeatWork{};
            wd->heat.solver.work = FldWork{};
        }
    }
    return SR_OK;
}

extern "C" int sr_kernel_time(sr_world *wd, const char *kernel, uint64_t *launches, double *mean_ms,
                              double *bytes_per_launch) {
    if (!wd || !kernel) return fail(SR_ERR_ARGUMENT, "null pointer");
    const std::string k = kernel;
    const double n = (double)wd->n;
    KernelClock c;
    double bytes = 0.0;
    if (k == "hydro_update") {
        c = wd->k_update;
        // state in and out (Sigma f64 + 13 f32 planes) + g + Sigma_mid + phi_mid (+ F and |push| with radiation)
        bytes = n * (2 * 60.0 + (wd->sw.gravity ? 16.0 : 0.0) + (wd->sw.heat ? 12.0 : 0.0));
    } else if (k == "hydro_mid") {
        c = wd->k_mid;
        bytes = n * (60.0 + 8.0 + 4.0);
    } else if (k == "mg_pre0" || k == "mg_post0" || k == "mg_coarse" || k == "fld_matvec" || k == "fld_update" ||
               k == "fld_update_pre0") {
        if (!wd->sw.heat) return fail(SR_ERR_ARGUMENT, "no heat pass");
        const int id = k == "mg_pre0"      ? FldSolver::CLK_PRE0
                       : k == "mg_post0"   ? FldSolver::CLK_POST0
                       : k == "mg_coarse"  ? FldSolver::CLK_COARSE
                       : k == "fld_matvec" ? FldSolver::CLK_MATVEC
                       : k == "fld_update" ? FldSolver::CLK_UPDATE
                                           : FldSolver::CLK_UPDATE_PRE0;
        SR_TRY(wd->heat.solver.clock(id, &c.ms, &c.launches));
        bytes = wd->heat.solver.clock_bytes(id);
    } else {
        return fail(SR_ERR_ARGUMENT, "unknown kernel " + k);
    }
    if (launches) *launches = c.launches;
    if (mean_ms) *mean_ms = c.launches ? c.ms / c.launches : 0.0;
    if (bytes_per_launch) *bytes_per_launch = bytes;
    return SR_OK;
}
