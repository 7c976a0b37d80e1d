// The cold pressure of the equation of state (crates/sr-physics/src/eos.rs:24-140; the GPU's table,
// crates/sr-engine/src/step/mod.rs:798-844): P(x) = 1 / (1 / (K1 x^2) + 1 / (K2 x^(3/2))) with x = Y_e Sigma
// (electrons) or X_n Sigma (neutron matter), and its energy per area u(x) = x int_0^x P(s) / s^2 ds tabulated as ln u
// on 512 points log-spaced over [1e-8, 1e8] (built in f64 on the host by Simpson's rule in t = sqrt(s), stored f32),
// read by linear interpolation in (ln x, ln u), the K1 x^2 power law below the table. A pair with a K of 0 is off.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>
#include <vector>

namespace sr {

#define COLD_POINTS 512

struct ColdEos {
    int on;                  // either pair on
    float k1e, k2e, k1n, k2n;  // 0: that pair off
    const float *lnu;        // device: electrons' ln u then neutrons' (COLD_POINTS each)
};

// (Z / A) per species, assets/elements.json order: H He C O Ne Mg Si S Fe n (registry.rs:146-154)
static __constant__ float c_ye[10] = {1.0f, 0.5f, 0.5f, 0.5f, 0.5f, 0.5f, 0.5f, 0.5f, 26.0f / 56.0f, 0.0f};

__device__ __forceinline__ float cold_p(float x, float k1, float k2) {
    if (!(x > 0.0f) || !(k1 > 0.0f)) return 0.0f;
    return 1.0f / (1.0f / (k1 * x * x) + 1.0f / (k2 * x * sqrtf(x)));
}

__device__ __forceinline__ float cold_u(const float *__restrict__ lnu, float x) {
    if (!(x > 0.0f)) return 0.0f;
    // ln x_min = ln 1e-8; (COLD_POINTS - 1) / (ln 1e8 - ln 1e-8)
    const float pos = (logf(x) + 18.420680743952367f) * 13.870280015785104f;
    if (pos < 0.0f) {
        const float r = x * 1e8f;
        return expf(lnu[0]) * r * r;
    }
    const int i = max(0, min((int)pos, COLD_POINTS - 2));
    const float frac = pos - (float)i;
    return expf(lnu[i] + frac * (lnu[i + 1] - lnu[i]));
}

// Pi_e + Pi_n of a cell or face state
__device__ __forceinline__ float pi_cold(const ColdEos &C, float s, float ye, float xn) {
    return cold_p(ye * s, C.k1e, C.k2e) + cold_p(xn * s, C.k1n, C.k2n);
}

// eps_cold = (u_e + u_n) / Sigma
__device__ __forceinline__ float eps_cold(const ColdEos &C, float s, float ye, float xn) {
    float u = 0.0f;
    if (C.k1e > 0.0f) u += cold_u(C.lnu, ye * s);
    if (C.k1n > 0.0f) u += cold_u(C.lnu + COLD_POINTS, xn * s);
    return u / s;
}

// the host's table of one pair (built in f64, stored as T), ln u at the nodes
template <typename T> inline void cold_table(double k1, double k2, T *out) {
    auto p = [&](double x) { return x <= 0.0 ? 0.0 : 1.0 / (1.0 / (k1 * x * x) + 1.0 / (k2 * std::pow(x, 1.5))); };
    auto seg = [&](double t0, double t1, int n) {
        const double h = (t1 - t0) / n;
        double sum = 0.0;
        for (int i = 0; i <= n; i++) {
            const double t = t0 + h * i, s = t * t;
            const double f = t > 0.0 ? p(s) / (s * s) * 2.0 * t : 0.0;
            sum += (i == 0 || i == n ? 1.0 : (i % 2 ? 4.0 : 2.0)) * f;
        }
        return sum * h / 3.0;
    };
    double tp = 0.0, j = 0.0;
    for (int i = 0; i < COLD_POINTS; i++) {
        const double x = std::exp(std::log(1e-8) + (double)i / (COLD_POINTS - 1) * (std::log(1e8) - std::log(1e-8)));
        const double t = std::sqrt(x);
        j += seg(tp, t, i == 0 ? 512 : 64);
        out[i] = (T)std::log(x * j);
        tp = t;
    }
}

} // namespace sr
