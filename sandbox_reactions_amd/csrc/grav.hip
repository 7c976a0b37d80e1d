// P2: zero-padded (Hockney-Eastwood) FFT convolution of the surface density with the thin-sheet kernel, and
// g = -grad phi by central differences. The design is this library's own: a real-to-complex row pass that packs
// each real row as a half-length complex sequence, one LDS-resident column pass per spectral column that also
// multiplies by the kernel's real-even spectrum and transforms back, and a complex-to-real row pass that writes only
// the box and its one-cell ring. The reference computes the same convolution (crates/sr-physics/src/reference/
// fft.rs:238-282, grav.rs:115-121; GPU crates/sr-engine/shaders/grav.wgsl:126-282) with complex transforms of every
// padded row.
#include <algorithm>
#include <cmath>
#include <complex>
#include <map>
#include <mutex>
#include <thread>
#include <tuple>
#include <vector>

#include "common.hpp"
#include "grav.hpp"

namespace sr {

// -------------------------------------------------------------------------------------------------------- kernels

// Rows y in [0, h) of sigma (masked at sigma_vac), zero-padded to px reals, forward real FFT; spectrum bins
// k in [0, px/2] written column-major: spec[k * ld + y].
template <typename TIn>
__global__ void __launch_bounds__(256) k_rows_r2c(const TIn *__restrict__ sigma, int w, int h, TIn sigma_vac,
                                                  int px, int logL, int rb, const float2 *__restrict__ tw,
                                                  int tw_stride_half, int tw_stride_full, float2 *__restrict__ spec,
                                                  int ld) {
    extern __shared__ float2 lds[];
    const int L = 1 << logL;
    float2 *a = lds, *b = lds + rb * L;
    // row groups remapped so that each XCD takes one contiguous run of them: the groups that share a bin's cache
    // line of the column-major spectrum then write it through the same L2
    const int y0 = xcd_tile() * rb;
    for (int t = threadIdx.x; t < rb * L; t += blockDim.x) {
        const int r = t >> logL, m = t & (L - 1), y = y0 + r;
        float2 v = make_float2(0.f, 0.f);
        if (y < h) {
            const int x0 = 2 * m, x1 = 2 * m + 1;
            if (x0 < w) {
                const TIn s = sigma[(size_t)y * w + x0];
                v.x = s >= sigma_vac ? (float)s : 0.f;
            }
            if (x1 < w) {
                const TIn s = sigma[(size_t)y * w + x1];
                v.y = s >= sigma_vac ? (float)s : 0.f;
            }
        }
        a[t] = v;
    }
    __syncthreads();
    float2 *Z = lds_fft(a, b, L, logL, rb, tw, tw_stride_half, false);
    // X[k] = (Z_k + conj Z_{L-k}) / 2 - i/2 W^k (Z_k - conj Z_{L-k}),  W = e^{-2 pi i / px}
    // Consecutive lanes take consecutive rows of one bin, so each bin's rb rows go out as one contiguous segment of
    // the column-major spectrum (rather than one 8-byte value per lane at stride ld).
    const int nk = L + 1;
    for (int t = threadIdx.x; t < rb * nk; t += blockDim.x) {
        const int k = t / rb, r = t - k * rb, y = y0 + r;
        if (y >= h) continue;
        const float2 zk = Z[r * L + (k & (L - 1))];
        const float2 zc = cconj(Z[r * L + ((L - k) & (L - 1))]);
        const float2 e = make_float2(0.5f * (zk.x + zc.x), 0.5f * (zk.y + zc.y));
        const float2 d = make_float2(0.5f * (zk.x - zc.x), 0.5f * (zk.y - zc.y));
        const float2 wk = tw[k * tw_stride_full];
        // -i * W * d
        const float2 wd = cmul(wk, d);
        spec[(size_t)k * ld + y] = make_float2(e.x + wd.y, e.y - wd.x);
    }
}

// One spectral column per block: forward FFT of length py (rows >= h are zero), multiply by the kernel spectrum
// (real, even; quarter table with (py/2 + 1) entries per column, already divided by px * py), inverse FFT, write the
// ring rows -1..h as out[k * ldo + j], j = y + 1.
__global__ void __launch_bounds__(256) k_columns(const float2 *__restrict__ spec, int ld, int h, int py, int logpy,
                                                 const float *__restrict__ khat, const float2 *__restrict__ tw,
                                                 int tw_stride, float2 *__restrict__ out, int ldo) {
    extern __shared__ float2 lds[];
    float2 *a = lds, *b = lds + py;
    const int k = blockIdx.x;
    const float2 *col = spec + (size_t)k * ld;
    for (int y = threadIdx.x; y < py; y += blockDim.x) a[y] = y < h ? col[y] : make_float2(0.f, 0.f);
    __syncthreads();
    float2 *F = lds_fft(a, b, py, logpy, 1, tw, tw_stride, false);
    float2 *G = (F == a) ? b : a;
    const float *kq = khat + (size_t)k * (py / 2 + 1);
    for (int ky = threadIdx.x; ky < py; ky += blockDim.x) {
        const float s = kq[ky <= py / 2 ? ky : py - ky];
        const float2 f = F[ky];
        F[ky] = make_float2(f.x * s, f.y * s);
    }
    __syncthreads();
    float2 *R = lds_fft(F, G, py, logpy, 1, tw, tw_stride, true);
    float2 *o = out + (size_t)k * ldo;
    for (int j = threadIdx.x; j < h + 2; j += blockDim.x) {
        const int y = j == 0 ? py - 1 : j - 1;
        o[j] = R[y];
    }
}

// Rows j in [0, h+2) of the ringed output: gather bins k in [0, px/2] from the column-major buffer, inverse real
// FFT (unnormalised), write phi_ringed[j][i] for i = 0 (x = -1), 1..w (x = 0..w-1), w+1 (x = w).
__global__ void __launch_bounds__(256) k_rows_c2r(const float2 *__restrict__ in, int ldi, int nrows, int w, int px,
                                                  int logL, int rb, const float2 *__restrict__ tw, int tw_stride_half,
                                                  int tw_stride_full, float *__restrict__ phi) {
    extern __shared__ float2 lds[];
    const int L = 1 << logL;
    float2 *a = lds, *b = lds + rb * L;
    const int j0 = xcd_tile() * rb;  // as in k_rows_r2c: a bin's cache line read through one XCD's L2
    // Z_k = (X_k + conj X_{L-k}) + i W^{-k} (X_k - conj X_{L-k}), k in [0, L). Consecutive lanes take consecutive
    // rows of one bin, so the gather reads each bin's rb rows (and its mirror's) as contiguous segments of the
    // column-major buffer.
    for (int t = threadIdx.x; t < rb * L; t += blockDim.x) {
        const int k = t / rb, r = t - k * rb, j = j0 + r;
        float2 v = make_float2(0.f, 0.f);
        if (j < nrows) {
            const float2 xk = in[(size_t)k * ldi + j];
            const float2 xc = cconj(in[(size_t)(L - k) * ldi + j]);
            const float2 s = cadd(xk, xc), d = csub(xk, xc);
            const float2 wk = cconj(tw[k * tw_stride_full]);
            const float2 wd = cmul(wk, d);
            v = make_float2(s.x - wd.y, s.y + wd.x); // s + i * wd
        }
        a[r * L + k] = v;
    }
    __syncthreads();
    float2 *z = lds_fft(a, b, L, logL, rb, tw, tw_stride_half, true);
    const int rw = w + 2;
    for (int t = threadIdx.x; t < rb * rw; t += blockDim.x) {
        const int r = t / rw, i = t - r * rw, j = j0 + r;
        if (j >= nrows) continue;
        const int n = i == 0 ? px - 1 : i - 1;
        const float2 v = z[r * L + (n >> 1)];
        phi[(size_t)j * rw + i] = (n & 1) ? v.y : v.x;
    }
}

__global__ void k_gradient(const float *__restrict__ phi, int w, int h, float inv2dx, float *__restrict__ gx,
                           float *__restrict__ gy) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (x >= w) return;
    const int rw = w + 2;
    const float *row = phi + (size_t)(y + 1) * rw;
    if (gx) gx[(size_t)y * w + x] = -(row[x + 2] - row[x]) * inv2dx;
    if (gy) gy[(size_t)y * w + x] = -(phi[(size_t)(y + 2) * rw + x + 1] - phi[(size_t)y * rw + x + 1]) * inv2dx;
}

// General complex transforms for sr_fft2: rows in place (rb rows per block), then columns one per block (strided).
__global__ void __launch_bounds__(256) k_c2c_rows(float2 *data, int nx, int ny, int logn, int rb,
                                                  const float2 *__restrict__ tw, int tw_stride, int inverse,
                                                  float scale) {
    extern __shared__ float2 lds[];
    const int n = 1 << logn;
    float2 *a = lds, *b = lds + rb * n;
    const int y0 = blockIdx.x * rb;
    for (int t = threadIdx.x; t < rb * n; t += blockDim.x) {
        const int r = t >> logn, x = t & (n - 1), y = y0 + r;
        a[t] = y < ny ? data[(size_t)y * nx + x] : make_float2(0.f, 0.f);
    }
    __syncthreads();
    float2 *o = lds_fft(a, b, n, logn, rb, tw, tw_stride, inverse != 0);
    for (int t = threadIdx.x; t < rb * n; t += blockDim.x) {
        const int r = t >> logn, x = t & (n - 1), y = y0 + r;
        if (y < ny) data[(size_t)y * nx + x] = make_float2(o[t].x * scale, o[t].y * scale);
    }
}

__global__ void __launch_bounds__(256) k_c2c_cols(float2 *data, int nx, int ny, int logn,
                                                  const float2 *__restrict__ tw, int tw_stride, int inverse,
                                                  float scale) {
    extern __shared__ float2 lds[];
    float2 *a = lds, *b = lds + ny;
    const int x = blockIdx.x;
    for (int y = threadIdx.x; y < ny; y += blockDim.x) a[y] = data[(size_t)y * nx + x];
    __syncthreads();
    float2 *o = lds_fft(a, b, ny, logn, 1, tw, tw_stride, inverse != 0);
    for (int y = threadIdx.x; y < ny; y += blockDim.x)
        data[(size_t)y * nx + x] = make_float2(o[y].x * scale, o[y].y * scale);
}

// -------------------------------------------------------------------------------------------------------- host

static const double SELF_CELL = 3.525494348078172; // 4 ln(1 + sqrt 2): the potential of a unit square at its centre

// In-place f64 radix-2 transform on the host (the kernel's spectrum is built once per padded size in f64).
static void host_fft(std::complex<double> *d, int n, const std::vector<std::complex<double>> &tw, int tw_stride) {
    for (int i = 1, j = 0; i < n; i++) {
        int bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) std::swap(d[i], d[j]);
    }
    for (int len = 2; len <= n; len <<= 1) {
        const int half = len >> 1, step = (n / len) * tw_stride;
        for (int s = 0; s < n; s += len)
            for (int k = 0; k < half; k++) {
                const std::complex<double> t = d[s + k + half] * tw[(size_t)k * step];
                d[s + k + half] = d[s + k] - t;
                d[s + k] += t;
            }
    }
}

template <class F> static void parallel_for(int n, F f) {
    const int nt = std::max(1, std::min<int>(n, (int)std::thread::hardware_concurrency()));
    std::vector<std::thread> ts;
    for (int t = 0; t < nt; t++)
        ts.emplace_back([&, t] {
            for (int i = t; i < n; i += nt) f(i);
        });
    for (auto &t : ts) t.join();
}

// The kernel's spectrum on the quarter [0, px/2] x [0, py/2], real and even, divided by px * py; K(i, j) =
// -G dx / sqrt(i^2 + j^2) at signed offsets, K(0, 0) = -G dx 4 ln(1 + sqrt 2) (reference/grav.rs:28-34).
static std::vector<float> kernel_spectrum(int px, int py, double G, double dx) {
    const int nmax = std::max(px, py);
    std::vector<std::complex<double>> tw(nmax);
    for (int t = 0; t < nmax; t++) tw[t] = std::polar(1.0, -2.0 * M_PI * t / nmax);
    const int qx = px / 2 + 1, qy = py / 2 + 1;
    std::vector<double> rowspec((size_t)qy * qx);
    parallel_for(qy, [&](int j) {
        std::vector<std::complex<double>> row(px);
        for (int i = 0; i < px; i++) {
            const long si = i < px / 2 ? i : i - px;
            const double r2 = (double)(si * si + (long)j * j);
            row[i] = (si == 0 && j == 0) ? -G * dx * SELF_CELL : -G * dx / std::sqrt(r2);
        }
        host_fft(row.data(), px, tw, nmax / px);
        for (int k = 0; k < qx; k++) rowspec[(size_t)j * qx + k] = row[k].real();
    });
    std::vector<float> q((size_t)qx * qy);
    const double scale = 1.0 / ((double)px * (double)py);
    parallel_for(qx, [&](int k) {
        std::vector<std::complex<double>> col(py);
        for (int j = 0; j < py; j++) col[j] = rowspec[(size_t)(j <= py / 2 ? j : py - j) * qx + k];
        host_fft(col.data(), py, tw, nmax / py);
        for (int ky = 0; ky < qy; ky++) q[(size_t)k * qy + ky] = (float)(col[ky].real() * scale);
    });
    return q;
}

// The spectrum for (px, py, G, dx), built once per process: an ensemble's worlds of one size share the host work
static std::vector<float> kernel_spectrum_cached(int px, int py, double G, double dx) {
    static std::mutex mu;
    static std::map<std::tuple<int, int, double, double>, std::vector<float>> cache;
    const auto key = std::make_tuple(px, py, G, dx);
    {
        std::lock_guard<std::mutex> g(mu);
        auto it = cache.find(key);
        if (it != cache.end()) return it->second;
    }
    std::vector<float> q = kernel_spectrum(px, py, G, dx);
    std::lock_guard<std::mutex> g(mu);
    if (cache.size() >= 16) cache.clear();
    cache[key] = q;
    return q;
}

int padded_size(int len, int max_padded, int *out) {
    if (len <= 0) return fail(SR_ERR_BOX_AXIS, "a box axis of 0 cells");
    int n = 1;
    while (n < 2 * len) n <<= 1;
    n = std::max(n, 32);
    if (n > max_padded)
        return fail(SR_ERR_BOX_AXIS, "box axis of " + std::to_string(len) + " cells pads past " +
                                         std::to_string(max_padded));
    *out = n;
    return SR_OK;
}

static int max_lds_bytes() {
    static int v = 0;
    if (!v) {
        int dev = 0;
        hipGetDevice(&dev);
        hipDeviceProp_t p;
        if (hipGetDeviceProperties(&p, dev) == hipSuccess)
            v = (int)p.sharedMemPerBlock;
        else
            v = 65536;
    }
    return v;
}

Gravity::~Gravity() {
    if (tw) hipFree(tw);
    if (khat) hipFree(khat);
    if (spec) hipFree(spec);
    if (cols) hipFree(cols);
    if (phi_tmp) hipFree(phi_tmp);
}

int Gravity::init(int w_, int h_, double G_, double dx_, double sigma_vac_, int max_padded) {
    w = w_;
    h = h_;
    G = G_;
    dx = dx_;
    sigma_vac = sigma_vac_;
    SR_TRY(padded_size(w, max_padded, &px));
    SR_TRY(padded_size(h, max_padded, &py));
    if ((size_t)py * 2 * sizeof(float2) > (size_t)max_lds_bytes())
        return fail(SR_ERR_BOX_AXIS, "a padded column of " + std::to_string(py) + " does not fit the LDS");
    nmax = std::max(px, py);
    std::vector<float2> t(nmax);
    for (int i = 0; i < nmax; i++) {
        const double a = -2.0 * M_PI * (double)i / (double)nmax;
        t[i] = make_float2((float)std::cos(a), (float)std::sin(a));
    }
    SR_HIP(hipMalloc(&tw, sizeof(float2) * nmax));
    SR_HIP(hipMemcpy(tw, t.data(), sizeof(float2) * nmax, hipMemcpyHostToDevice));
    std::vector<float> q = kernel_spectrum_cached(px, py, G, dx);
    SR_HIP(hipMalloc(&khat, sizeof(float) * q.size()));
    SR_HIP(hipMemcpy(khat, q.data(), sizeof(float) * q.size(), hipMemcpyHostToDevice));
    ld_spec = h;
    ld_cols = h + 2;
    SR_HIP(hipMalloc(&spec, sizeof(float2) * (size_t)(px / 2 + 1) * ld_spec));
    SR_HIP(hipMalloc(&cols, sizeof(float2) * (size_t)(px / 2 + 1) * ld_cols));
    return SR_OK;
}

static int rows_per_block(int L) { return std::max(1, std::min(8, 4096 / L)); }

template <typename TIn> int Gravity::potential(const TIn *sigma, float *phi_ringed, hipStream_t s) {
    const int L = px / 2, logL = ilog2(L), logpy = ilog2(py);
    const int rb = this->rb > 0 ? this->rb : rows_per_block(L);
    const int stride_half = nmax / L, stride_full = nmax / px;
    hipLaunchKernelGGL(k_rows_r2c<TIn>, dim3((h + rb - 1) / rb), dim3(256), sizeof(float2) * 2 * rb * L, s, sigma, w,
                       h, (TIn)sigma_vac, px, logL, rb, tw, stride_half, stride_full, spec, ld_spec);
    hipLaunchKernelGGL(k_columns, dim3(px / 2 + 1), dim3(256), sizeof(float2) * 2 * py, s, spec, ld_spec, h, py,
                       logpy, khat, tw, nmax / py, cols, ld_cols);
    hipLaunchKernelGGL(k_rows_c2r, dim3((h + 2 + rb - 1) / rb), dim3(256), sizeof(float2) * 2 * rb * L, s, cols,
                       ld_cols, h + 2, w, px, logL, rb, tw, stride_half, stride_full, phi_ringed);
    SR_HIP(hipGetLastError());
    return SR_OK;
}

// Rows per block 1, 2, 4 or 8 (as the LDS allows two rb x px/2 complex arrays), each timed as the best of three means
// of four convolutions between two events after one to warm; the heuristic's count is kept unless another is 3 %
// faster.
int Gravity::tune(const float *sigma, float *phi_ringed, hipStream_t s) {
    const int L = px / 2, heur = rows_per_block(L);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    SR_HIP(hipEventCreate(&e0));
    SR_HIP(hipEventCreate(&e1));
    int best = heur;
    float best_ms = 0.0f;
    int rc = SR_OK;
    std::vector<int> cand = {heur};
    for (int r : {1, 2, 4, 8})
        if (r != heur && (size_t)2 * r * L * sizeof(float2) <= (size_t)max_lds_bytes()) cand.push_back(r);
    for (int r : cand) {
        rb = r;
        float ms = 0.0f;
        if ((rc = potential<float>(sigma, phi_ringed, s)) != SR_OK) break;
        for (int trial = 0; trial < 3 && rc == SR_OK; trial++) {
            if (hipEventRecord(e0, s) != hipSuccess) rc = fail(SR_ERR_DEVICE, "hipEventRecord");
            for (int k = 0; k < 4 && rc == SR_OK; k++) rc = potential<float>(sigma, phi_ringed, s);
            if (rc != SR_OK) break;
            if (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess)
                rc = fail(SR_ERR_DEVICE, "hipEventSynchronize");
            float t = 0.0f;
            if (rc == SR_OK && hipEventElapsedTime(&t, e0, e1) != hipSuccess) rc = fail(SR_ERR_DEVICE, "timing");
            if (trial == 0 || t < ms) ms = t;
        }
        if (rc != SR_OK) break;
        if (r == heur || ms < 0.97f * best_ms) {
            best = r;
            best_ms = ms;
        }
    }
    rb = best;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return rc;
}

template int Gravity::potential<float>(const float *, float *, hipStream_t);
template int Gravity::potential<double>(const double *, float *, hipStream_t);

int Gravity::gradient(const float *phi_ringed, float *gx, float *gy, hipStream_t s) {
    dim3 grid((w + 255) / 256, h);
    hipLaunchKernelGGL(k_gradient, grid, dim3(256), 0, s, phi_ringed, w, h, (float)(0.5 / dx), gx, gy);
    SR_HIP(hipGetLastError());
    return SR_OK;
}

// Bytes one convolution moves by its own algorithm (the roofline's numerator): sigma in, row spectra out and in,
// column output out and in, the kernel quarter, phi out.
double Gravity::algorithmic_bytes() const {
    const double bins = px / 2 + 1;
    return 4.0 * w * h + 2 * 8.0 * bins * h + 2 * 8.0 * bins * (h + 2) + 4.0 * bins * (py / 2 + 1) +
           4.0 * (w + 2) * (h + 2);
}

} // namespace sr

// ------------------------------------------------------------------------------------------------------------ C ABI
using namespace sr;

struct sr_gravity {
    Gravity g;
};

extern "C" int sr_gravity_create(int w, int h, double G, double dx, double sigma_vac, int max_padded,
                                 sr_gravity **out) {
    if (!out || w <= 0 || h <= 0) return fail(SR_ERR_ARGUMENT, "sr_gravity_create: bad arguments");
    auto *g = new sr_gravity();
    int c = g->g.init(w, h, G, dx, sigma_vac, max_padded > 0 ? max_padded : 4096);
    if (c != SR_OK) {
        delete g;
        return c;
    }
    *out = g;
    return SR_OK;
}

extern "C" int sr_gravity_field(sr_gravity *g, const float *sigma_dev, float *phi_ringed_dev, float *gx_dev,
                                float *gy_dev, void *stream) {
    if (!g || !sigma_dev || !phi_ringed_dev) return fail(SR_ERR_ARGUMENT, "sr_gravity_field: null pointer");
    hipStream_t s = (hipStream_t)stream;
    SR_TRY(g->g.potential(sigma_dev, phi_ringed_dev, s));
    if (gx_dev || gy_dev) SR_TRY(g->g.gradient(phi_ringed_dev, gx_dev, gy_dev, s));
    return SR_OK;
}

extern "C" void sr_gravity_destroy(sr_gravity *g) { delete g; }

// The twiddles exp(-2 pi i k / n), k < n, of the complex transforms: formed in f64 on the device, stored f32.
__global__ void k_twiddles(float2 *__restrict__ tw, int n) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) {
        double sn, cs;
        sincospi(-2.0 * (double)k / (double)n, &sn, &cs);
        tw[k] = make_float2((float)cs, (float)sn);
    }
}

static int fft2_check(const float *data_dev, int nx, int ny) {
    if (!data_dev) return fail(SR_ERR_ARGUMENT, "sr_fft2: null pointer");
    if (!is_pow2(nx) || !is_pow2(ny)) return fail(SR_ERR_NOT_POW2, "sr_fft2: axes must be powers of two");
    if ((size_t)std::max(nx, ny) * 2 * sizeof(float2) > (size_t)max_lds_bytes())
        return fail(SR_ERR_LENGTH, "sr_fft2: axis too long for the LDS");
    return SR_OK;
}

// the twiddle table of the longer axis (tw, nmax entries), then the rows in place and the columns in place
static int fft2_launch(float2 *d, int nx, int ny, int inverse, float2 *tw, hipStream_t s) {
    const int nmax = std::max(nx, ny);
    hipLaunchKernelGGL(k_twiddles, dim3((nmax + 255) / 256), dim3(256), 0, s, tw, nmax);
    const int rb = std::max(1, std::min(8, 4096 / nx));
    hipLaunchKernelGGL(k_c2c_rows, dim3((ny + rb - 1) / rb), dim3(256), sizeof(float2) * 2 * rb * nx, s, d, nx, ny,
                       ilog2(nx), rb, (const float2 *)tw, nmax / nx, inverse, inverse ? 1.0f / nx : 1.0f);
    hipLaunchKernelGGL(k_c2c_cols, dim3(nx), dim3(256), sizeof(float2) * 2 * ny, s, d, nx, ny, ilog2(ny),
                       (const float2 *)tw, nmax / ny, inverse, inverse ? 1.0f / ny : 1.0f);
    SR_HIP(hipGetLastError());
    return SR_OK;
}

extern "C" int sr_fft2(float *data_dev, int nx, int ny, int inverse, void *stream) {
    SR_TRY(fft2_check(data_dev, nx, ny));
    hipStream_t s = (hipStream_t)stream;
    float2 *tw = nullptr;
    SR_HIP(hipMallocAsync((void **)&tw, sizeof(float2) * std::max(nx, ny), s));
    const int c = fft2_launch((float2 *)data_dev, nx, ny, inverse, tw, s);
    SR_HIP(hipFreeAsync(tw, s));
    SR_TRY(c);
    SR_HIP(hipStreamSynchronize(s));
    return SR_OK;
}

// The allocation-free variant's twiddle table: a file-scope device array, one copy per GPU at its own address. Host
// code never passes the array's name to a kernel (on the host it is a stand-in, not a device address): the first call
// on each device looks the device's address up with hipGetSymbolAddress and keeps it.
constexpr int FFT_SCRATCH_MAX = 8192;
static __device__ __align__(256) float2 scratch[FFT_SCRATCH_MAX];
constexpr int MAX_DEVICES = 64;
static std::mutex scratch_mu;
static float2 *scratch_at[MAX_DEVICES] = {};

static int scratch_of_current_device(float2 **out) {
    int dev = -1;
    SR_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= MAX_DEVICES) return fail(SR_ERR_DEVICE, "sr_fft2_noalloc: device index out of range");
    std::lock_guard<std::mutex> g(scratch_mu);
    if (!scratch_at[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(SR_ERR_DEVICE, std::string("sr_fft2_noalloc: hipGetSymbolAddress: ") + hipGetErrorString(e));
        scratch_at[dev] = (float2 *)p;
    }
    *out = scratch_at[dev];
    return SR_OK;
}

extern "C" int sr_fft2_noalloc(float *data_dev, int nx, int ny, int inverse, void *stream) {
    if (std::max(nx, ny) > FFT_SCRATCH_MAX)
        return fail(SR_ERR_LENGTH, "sr_fft2_noalloc: an axis longer than " + std::to_string(FFT_SCRATCH_MAX) +
                                       " (the twiddle scratch's maximum)");
    SR_TRY(fft2_check(data_dev, nx, ny));
    float2 *tw = nullptr;
    SR_TRY(scratch_of_current_device(&tw));
    return fft2_launch((float2 *)data_dev, nx, ny, inverse, tw, (hipStream_t)stream);
}
