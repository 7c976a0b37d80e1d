// P2 gravity solver: the plan of one box's zero-padded FFT convolution.
#pragma once
#include <hip/hip_runtime.h>

namespace sr {

int padded_size(int len, int max_padded, int *out);

struct Gravity {
    int w = 0, h = 0, px = 0, py = 0, nmax = 0;
    double G = 1.0, dx = 1.0, sigma_vac = 0.0;
    float2 *tw = nullptr;    // e^{-2 pi i t / nmax}
    float *khat = nullptr;   // kernel spectrum quarter, (px/2+1) x (py/2+1), / (px py)
    float2 *spec = nullptr;  // row spectra, column-major (px/2+1) x h
    float2 *cols = nullptr;  // column outputs, column-major (px/2+1) x (h+2)
    float *phi_tmp = nullptr;
    int ld_spec = 0, ld_cols = 0;
    int rb = 0;  // rows per block of the row transforms: the autotuner's choice (0 until it runs: a heuristic)

    ~Gravity();
    int init(int w, int h, double G, double dx, double sigma_vac, int max_padded);
    // phi on the box grown by one ring, (w+2) x (h+2), from sigma (w x h), cells below sigma_vac no source
    template <typename TIn> int potential(const TIn *sigma, float *phi_ringed, hipStream_t s);
    int gradient(const float *phi_ringed, float *gx, float *gy, hipStream_t s);
    double algorithmic_bytes() const;
    // the autotuner: the row transforms' rows per block, timed on the convolution of `sigma` into `phi_ringed`
    // (scratch device arrays of the box's and the ringed box's size); each row's transform is the same whichever
    // block takes it, so the choice changes no result bit
    int tune(const float *sigma, float *phi_ringed, hipStream_t s);
};

} // namespace sr
