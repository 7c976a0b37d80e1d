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

__global__ void k_calib_empty() {}

// What a pair of timing events adds around a kernel that follows another on a stream: the mean elapsed time of
// `reps` brackets (created as the library's timing events are, without the system-scope fence) around an empty
// kernel, each after an empty predecessor, read after one synchronisation. A kernel's event-timed mean less this is
// what a profiler's kernel trace reports as its duration.
extern "C" int sr_calib_event_bracket(int reps, void *stream, double *mean_ms) {
    if (reps <= 0 || reps > 4096 || !mean_ms) return -5;
    hipStream_t s = (hipStream_t)stream;
    hipEvent_t *ev = new hipEvent_t[2 * reps]();
    int rc = 0;
    for (int k = 0; k < 2 * reps && rc == 0; k++)
        if (hipEventCreateWithFlags(&ev[k], hipEventDisableSystemFence) != hipSuccess) rc = -4;
    for (int r = 0; r < reps && rc == 0; r++) {
        hipLaunchKernelGGL(k_calib_empty, dim3(1), dim3(64), 0, s);
        if (hipEventRecord(ev[2 * r], s) != hipSuccess) rc = -4;
        hipLaunchKernelGGL(k_calib_empty, dim3(1), dim3(64), 0, s);
        if (hipEventRecord(ev[2 * r + 1], s) != hipSuccess) rc = -4;
    }
    double sum = 0.0;
    if (rc == 0 && hipStreamSynchronize(s) != hipSuccess) rc = -4;
    for (int r = 0; r < reps && rc == 0; r++) {
        float ms = 0.0f;
        if (hipEventElapsedTime(&ms, ev[2 * r], ev[2 * r + 1]) != hipSuccess) rc = -4;
        sum += ms;
    }
    for (int k = 0; k < 2 * reps; k++)
        if (ev[k]) (void)hipEventDestroy(ev[k]);
    delete[] ev;
    if (rc == 0) *mean_ms = sum / reps;
    return rc;
}
