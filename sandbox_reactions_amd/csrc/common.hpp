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

// ------------------------------------------------------------------------------------------- deterministic reduction
// Each block adds its K sums to part[k * nb + block]; the last block to arrive sums them in block order into out[k].
// Returns true in every thread of the last block (after out is written and visible to it).
template <int K>
__device__ inline bool reduce_finish(double (&v)[K], double *part, int nb, unsigned *counter, double *out) {
    __shared__ double sh[K][16];
    __shared__ bool last;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
    const int bid = blockIdx.y * gridDim.x + blockIdx.x;
#pragma unroll
    for (int k = 0; k < K; k++) v[k] = wave_sum(v[k]);
    if (lane == 0)
#pragma unroll
        for (int k = 0; k < K; k++) sh[k][wid] = v[k];
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 0; k < K; k++) {
            double s = 0.0;
            for (int i = 0; i < nw; i++) s += sh[k][i];
            part[(size_t)k * nb + bid] = s;
        }
        __threadfence();
        const unsigned prev = atomicAdd(counter, 1u);
        last = prev == (unsigned)(nb - 1);
    }
    __syncthreads();
    if (!last) return false;
    __threadfence();
    for (int k = 0; k < K; k++) {
        double s = 0.0;
        for (int i = threadIdx.x; i < nb; i += blockDim.x)
            s += __hip_atomic_load(&part[(size_t)k * nb + i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        s = wave_sum(s);
        __syncthreads();
        if (lane == 0) sh[k][wid] = s;
        __syncthreads();
        if (threadIdx.x == 0) {
            double t = 0.0;
            for (int i = 0; i < nw; i++) t += sh[k][i];
            out[k] = t;
        }
    }
    if (threadIdx.x == 0) *counter = 0u;
    __syncthreads();
    return true;
}

// As reduce_finish with max in place of sum (order-independent, so deterministic by construction).
template <int K>
__device__ inline bool reduce_max_finish(double (&v)[K], double *part, int nb, unsigned *counter, double *out) {
    __shared__ double shm[K][16];
    __shared__ bool lastm;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
    const int bid = blockIdx.y * gridDim.x + blockIdx.x;
#pragma unroll
    for (int k = 0; k < K; k++)
        for (int o = 32; o > 0; o >>= 1) v[k] = fmax(v[k], __shfl_xor(v[k], o, 64));
    if (lane == 0)
#pragma unroll
        for (int k = 0; k < K; k++) shm[k][wid] = v[k];
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 0; k < K; k++) {
            double s = shm[k][0];
            for (int i = 1; i < nw; i++) s = fmax(s, shm[k][i]);
            part[(size_t)k * nb + bid] = s;
        }
        __threadfence();
        const unsigned prev = atomicAdd(counter, 1u);
        lastm = prev == (unsigned)(nb - 1);
    }
    __syncthreads();
    if (!lastm) return false;
    __threadfence();
    for (int k = 0; k < K; k++) {
        double s = -HUGE_VAL;
        for (int i = threadIdx.x; i < nb; i += blockDim.x)
            s = fmax(s, __hip_atomic_load(&part[(size_t)k * nb + i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
        for (int o = 32; o > 0; o >>= 1) s = fmax(s, __shfl_xor(s, o, 64));
        __syncthreads();
        if (lane == 0) shm[k][wid] = s;
        __syncthreads();
        if (threadIdx.x == 0) {
            double t = shm[k][0];
            for (int i = 1; i < nw; i++) t = fmax(t, shm[k][i]);
            out[k] = t;
        }
    }
    if (threadIdx.x == 0) *counter = 0u;
    __syncthreads();
    return true;
}

inline int ilog2(long n) {
    int l = 0;
    while ((1L << l) < n) l++;
    return l;
}

inline bool is_pow2(long n) { return n > 0 && (n & (n - 1)) == 0; }

} // namespace sr
