/* Oracle (test infrastructure only): f64 FFT and zero-padded convolution.
 * Restates crates/sr-physics/src/reference/fft.rs:
 *   padded_size            fft.rs:92-101
 *   FftPlan::new/transform fft.rs:114-173 (radix-2 DIT, bit reversal, twiddles each from their own sin/cos)
 *   fft2                   fft.rs:177-196 (rows, then columns)
 *   signed_offset          fft.rs:200-206
 *   convolve_ringed        fft.rs:238-282
 *   pseudo_random (LCG)    fft.rs:292-300
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "sro.h"

static int is_pow2(size_t n) { return n && !(n & (n - 1)); }

int sro_padded_size(size_t len, size_t max_padded, size_t *out) {
    if (len == 0) return SRO_ERR_BOX;
    size_t n = 1;
    while (n < 2 * len) n <<= 1;
    if (n < SRO_MIN_PADDED) n = SRO_MIN_PADDED;
    if (n > max_padded) return SRO_ERR_BOX;
    *out = n;
    return SRO_OK;
}

int sro_fft(double *d, size_t n, int inverse) {
    if (!is_pow2(n)) return SRO_ERR_POW2;
    unsigned bits = 0;
    while (((size_t)1 << bits) < n) bits++;
    /* bit-reversal permutation */
    for (size_t i = 0; i < n; i++) {
        size_t j = 0;
        for (unsigned b = 0; b < bits; b++)
            if (i & ((size_t)1 << b)) j |= (size_t)1 << (bits - 1 - b);
        if (i < j) {
            double tr = d[2 * i], ti = d[2 * i + 1];
            d[2 * i] = d[2 * j];
            d[2 * i + 1] = d[2 * j + 1];
            d[2 * j] = tr;
            d[2 * j + 1] = ti;
        }
    }
    double *tw = (double *)malloc(sizeof(double) * (n > 1 ? n : 2));
    for (size_t k = 0; k < n / 2; k++) {
        double a = -2.0 * M_PI * (double)k / (double)n;
        tw[2 * k] = cos(a);
        tw[2 * k + 1] = sin(a);
    }
    for (size_t half = 1; half < n; half *= 2) {
        size_t step = n / (2 * half);
        for (size_t start = 0; start < n; start += 2 * half) {
            for (size_t k = 0; k < half; k++) {
                double wr = tw[2 * k * step], wi = tw[2 * k * step + 1];
                if (inverse) wi = -wi;
                double *a = d + 2 * (start + k), *b = d + 2 * (start + k + half);
                double br = b[0] * wr - b[1] * wi, bi = b[0] * wi + b[1] * wr;
                double ar = a[0], ai = a[1];
                a[0] = ar + br;
                a[1] = ai + bi;
                b[0] = ar - br;
                b[1] = ai - bi;
            }
        }
    }
    free(tw);
    if (inverse) {
        double s = 1.0 / (double)n;
        for (size_t i = 0; i < 2 * n; i++) d[i] *= s;
    }
    return SRO_OK;
}

int sro_fft2(double *d, size_t nx, size_t ny, int inverse) {
    if (!is_pow2(nx) || !is_pow2(ny)) return SRO_ERR_POW2;
#pragma omp parallel for schedule(static)
    for (size_t y = 0; y < ny; y++) sro_fft(d + 2 * y * nx, nx, inverse);
#pragma omp parallel
    {
        double *col = (double *)malloc(sizeof(double) * 2 * ny);
#pragma omp for schedule(static)
        for (size_t x = 0; x < nx; x++) {
            for (size_t y = 0; y < ny; y++) {
                col[2 * y] = d[2 * (y * nx + x)];
                col[2 * y + 1] = d[2 * (y * nx + x) + 1];
            }
            sro_fft(col, ny, inverse);
            for (size_t y = 0; y < ny; y++) {
                d[2 * (y * nx + x)] = col[2 * y];
                d[2 * (y * nx + x) + 1] = col[2 * y + 1];
            }
        }
        free(col);
    }
    return SRO_OK;
}

static long signed_offset(size_t i, size_t n) { return i < n / 2 ? (long)i : (long)i - (long)n; }

int sro_convolve_ringed(const double *field, size_t w, size_t h, size_t ring, sro_kernel_fn kern, void *ctx,
                        size_t max_padded, double *out) {
    size_t px, py;
    int e;
    if ((e = sro_padded_size(w, max_padded, &px)) != SRO_OK) return e;
    if ((e = sro_padded_size(h, max_padded, &py)) != SRO_OK) return e;
    double *k = (double *)calloc(2 * px * py, sizeof(double));
    double *p = (double *)calloc(2 * px * py, sizeof(double));
    if (!k || !p) {
        free(k);
        free(p);
        return SRO_ERR_ARG;
    }
#pragma omp parallel for schedule(static)
    for (size_t j = 0; j < py; j++)
        for (size_t i = 0; i < px; i++) k[2 * (j * px + i)] = kern(signed_offset(i, px), signed_offset(j, py), ctx);
    sro_fft2(k, px, py, 0);
    for (size_t y = 0; y < h; y++)
        for (size_t x = 0; x < w; x++) p[2 * (y * px + x)] = field[y * w + x];
    sro_fft2(p, px, py, 0);
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < px * py; i++) {
        double ar = p[2 * i], ai = p[2 * i + 1], br = k[2 * i], bi = k[2 * i + 1];
        p[2 * i] = ar * br - ai * bi;
        p[2 * i + 1] = ar * bi + ai * br;
    }
    sro_fft2(p, px, py, 1);
    size_t rw = w + 2 * ring, rh = h + 2 * ring;
    for (size_t y = 0; y < rh; y++)
        for (size_t x = 0; x < rw; x++) {
            size_t sx = (x + px - ring) % px, sy = (y + py - ring) % py;
            out[y * rw + x] = p[2 * (sy * px + sx)];
        }
    free(k);
    free(p);
    return SRO_OK;
}

void sro_lcg_fill(uint64_t seed, size_t count, double *out) {
    uint64_t s = seed;
    for (size_t i = 0; i < count; i++) {
        s = s * 6364136223846793005ULL + 1442695040888963407ULL;
        out[i] = (double)(s >> 11) / 9007199254740992.0;
    }
}
