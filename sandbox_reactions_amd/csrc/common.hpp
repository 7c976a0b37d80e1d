// Shared device and host helpers of the sr-hip library (gfx950, wave64).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>

#include "../../include/sr_hip.h"

namespace sr {

// ---------------------------------------------------------------------------------------------------------- errors
void set_error(const std::string &msg);
int fail(int code, const std::string &msg);

#define SR_HIP(call)                                                                                                   \
    do {                                                                                                               \
        hipError_t _e = (call);                                                                                        \
        if (_e != hipSuccess)                                                                                          \
            return ::sr::fail(SR_ERR_DEVICE, std::string(#call) + ": " + hipGetErrorString(_e));                     \
    } while (0)

#define SR_TRY(expr)                                                                                                   \
    do {                                                                                                               \
        int _c = (expr);                                                                                               \
        if (_c != SR_OK) return _c;                                                                                    \
    } while (0)

// ---------------------------------------------------------------------------------------------------------- complex
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cconj(float2 a) { return make_float2(a.x, -a.y); }
// multiply by -i (forward) or +i (inverse)
__device__ __forceinline__ float2 crot(float2 a, bool inverse) {
    return inverse ? make_float2(-a.y, a.x) : make_float2(a.y, -a.x);
}

// Stockham autosort FFT of `nseq` independent sequences of length L (a power of two), laid out back to back in LDS
// buffer `a` (ping-pong partner `b`), using every thread of the block. Radix-4 passes, one radix-2 pass when log2 L
// is odd; natural order in and out. `tw` holds e^{-2 pi i t / Nmax} for t < Nmax, `tw_stride` = Nmax / L. The
// inverse is the conjugate transform, unnormalised. Returns the buffer holding the result. Ends with a barrier.
__device__ inline float2 *lds_fft(float2 *a, float2 *b, int L, int log2L, int nseq, const float2 *__restrict__ tw,
                                  int tw_stride, bool inverse) {
    int Ns = 1, logNs = 0;
    while (Ns < L) {
        const int rem_log = log2L - logNs;
        const int R = rem_log >= 2 ? 4 : 2;
        const int logR = rem_log >= 2 ? 2 : 1;
        const int lognb = log2L - logR;
        const int nb = 1 << lognb;
        const int total = nseq << lognb;
        for (int t = threadIdx.x; t < total; t += blockDim.x) {
            const int s = t >> lognb, j = t & (nb - 1);
            const float2 *in = a + s * L;
            float2 *out = b + s * L;
            const int k = j & (Ns - 1);
            const int ts = (k << (log2L - logNs - logR)) * tw_stride; // k * L / (Ns R) in table units
            const int d = ((j >> logNs) << (logNs + logR)) + k;
            if (R == 4) {
                float2 v0 = in[j], v1 = in[j + nb], v2 = in[j + 2 * nb], v3 = in[j + 3 * nb];
                if (k) {
                    float2 w1 = tw[ts], w2 = tw[2 * ts], w3 = tw[3 * ts];
                    if (inverse) {
                        w1 = cconj(w1);
                        w2 = cconj(w2);
                        w3 = cconj(w3);
                    }
                    v1 = cmul(v1, w1);
                    v2 = cmul(v2, w2);
                    v3 = cmul(v3, w3);
                }
                const float2 a0 = cadd(v0, v2), a1 = csub(v0, v2), a2 = cadd(v1, v3), a3 = crot(csub(v1, v3), inverse);
                out[d] = cadd(a0, a2);
                out[d + Ns] = cadd(a1, a3);
                out[d + 2 * Ns] = csub(a0, a2);
                out[d + 3 * Ns] = csub(a1, a3);
            } else {
                float2 v0 = in[j], v1 = in[j + nb];
                if (k) {
                    float2 w1 = tw[ts];
                    if (inverse) w1 = cconj(w1);
                    v1 = cmul(v1, w1);
                }
                out[d] = cadd(v0, v1);
                out[d + Ns] = csub(v0, v1);
            }
        }
        __syncthreads();
        float2 *tmp = a;
        a = b;
        b = tmp;
        Ns <<= logR;
        logNs += logR;
    }
    return a;
}

// ---------------------------------------------------------------------------------------------------------- reductions
// wave64 reduction by cross-lane shuffles (the compiler lowers __shfl_xor to DPP / ds_swizzle on gfx950)
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
    return v;
}

// Block sum of one double per thread, result valid in thread 0. `scratch` holds blockDim.x / 64 doubles.
__device__ inline double block_sum(double v, double *scratch) {
    v = wave_sum(v);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
    __syncthreads();
    if (lane == 0) scratch[wid] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0)
        for (int i = 0; i < nw; i++) s += scratch[i];
    return s;
}

// The tile a block of a 2D tiled launch takes. The dispatcher deals consecutive block ids to the 8 XCDs in turn, so
// the ids are remapped to give each XCD one contiguous run of tiles in row order: neighbouring tiles, whose regions
// overlap by their halos, then read the shared lines through the same XCD's L2 instead of each from HBM. A bijection on
// [0, nb) for any nb (an XCD holding one more block when nb is not a multiple of 8).
__device__ __forceinline__ int xcd_tile() {
    const int nb = gridDim.x * gridDim.y, lin = blockIdx.y * gridDim.x + blockIdx.x;
    const int q = nb >> 3, r = nb & 7, xcd = lin & 7, slot = lin >> 3;
    return (xcd < r ? xcd * (q + 1) : r * (q + 1) + (xcd - r) * q) + slot;
}

// ------------------------------------------------------------------------------------------- deterministic reduction
// Block sums (OP = 0) or maxima (OP = 1) of K values into part[k * nb + block] (block: the linear block id, or `slot`
// when a kernel orders its partials itself). No fence and no counter: the partials
// are read by a later kernel or by the host, which sum them in block order (so the result is deterministic). A
// device-wide fence here would write back the whole L2 of an XCD in every block.
template <int K, int OP = 0>
__device__ inline void block_partials(double (&v)[K], double *part, int nb, int slot = -1) {
    __shared__ double sh[K][16];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
    const int bid = slot >= 0 ? slot : blockIdx.y * gridDim.x + blockIdx.x;
#pragma unroll
    for (int k = 0; k < K; k++)
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            const double t = __shfl_xor(v[k], o, 64);
            v[k] = OP ? fmax(v[k], t) : v[k] + t;
        }
    if (lane == 0)
#pragma unroll
        for (int k = 0; k < K; k++) sh[k][wid] = v[k];
    __syncthreads();
    if (threadIdx.x < K) {
        double s = sh[threadIdx.x][0];
        for (int i = 1; i < nw; i++) s = OP ? fmax(s, sh[threadIdx.x][i]) : s + sh[threadIdx.x][i];
        part[(size_t)threadIdx.x * nb + bid] = s;
    }
}

// The sum of nb partials in a fixed order; every thread of the block gets the result.
__device__ inline double sum_partials(const double *part, int nb) {
    __shared__ double shp[16];
    __shared__ double res;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
    double s = 0.0;
    for (int i = threadIdx.x; i < nb; i += blockDim.x) s += part[i];
    s = wave_sum(s);
    __syncthreads();
    if (lane == 0) shp[wid] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int i = 0; i < nw; i++) t += shp[i];
        res = t;
    }
    __syncthreads();
    return res;
}

inline int ilog2(long n) {
    int l = 0;
    while ((1L << l) < n) l++;
    return l;
}

inline bool is_pow2(long n) { return n > 0 && (n & (n - 1)) == 0; }

} // namespace sr
