// Device copies of a known size, the calibration of the HBM-traffic counters (FETCH_SIZE, WRITE_SIZE) and the
// measured copy bandwidth beside the nominal peak: one float per lane (the access pattern of the path's stencil
// kernels) and one float4 per lane. Not part of the grid-physics ABI; the bench loads this library by itself.
#include <hip/hip_runtime.h>

#include <cstddef>

__global__ void __launch_bounds__(256) k_calib_copy1(const float *__restrict__ src, float *__restrict__ dst, size_t n) {
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) dst[i] = src[i];
}

__global__ void __launch_bounds__(256) k_calib_copy4(const float4 *__restrict__ src, float4 *__restrict__ dst,
                                                     size_t n4) {
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n4; i += (size_t)gridDim.x * 256) dst[i] = src[i];
}

// kind 1: a float per lane; kind 4: a float4 per lane (n a multiple of 4 and 16-byte aligned pointers)
extern "C" int sr_calib_copy(const float *src, float *dst, size_t n, int kind, void *stream) {
    if (!src || !dst || n == 0) return -5;
    const int blocks = 256 * 32;
    if (kind == 4) {
        if (n % 4 || ((size_t)src | (size_t)dst) % 16) return -5;
        hipLaunchKernelGGL(k_calib_copy4, dim3(blocks), dim3(256), 0, (hipStream_t)stream, (const float4 *)src,
                           (float4 *)dst, n / 4);
    } else {
        hipLaunchKernelGGL(k_calib_copy1, dim3(blocks), dim3(256), 0, (hipStream_t)stream, src, dst, n);
    }
    return hipGetLastError() == hipSuccess ? 0 : -4;
}
