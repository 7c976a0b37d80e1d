/* Oracle (test infrastructure only): P2 gravity in f64.
 * Restates crates/sr-physics/src/reference/grav.rs:
 *   kernel        grav.rs:28-34 (K = -G / r, K(0,0) = -G 4 ln(1 + sqrt 2)); scaled by dx for a cell of side dx
 *   source        grav.rs:37-39 (Sigma where Sigma >= sigma_vac, else 0)
 *   acceleration  grav.rs:51-67 (central differences over the box grown by one ring)
 *   field         grav.rs:115-121
 */
#include <math.h>
#include <stdlib.h>

#include "sro.h"

double sro_grav_kernel(double G, double dx, long i, long j) {
    if (i == 0 && j == 0) return -G * dx * SRO_SELF_CELL;
    double r2 = (double)(i * i + j * j);
    return -G * dx / sqrt(r2);
}

typedef struct {
    double G, dx;
} kctx;

static double kfn(long i, long j, void *c) {
    kctx *k = (kctx *)c;
    return sro_grav_kernel(k->G, k->dx, i, j);
}

int sro_gravity_field(const sro_physics *ph, const double *sigma, size_t w, size_t h, double *phi, double *gx,
                      double *gy, double *phi_ringed) {
    size_t n = w * h, rw = w + 2, rh = h + 2;
    double *src = (double *)malloc(sizeof(double) * n);
    double *ringed = (double *)malloc(sizeof(double) * rw * rh);
    for (size_t i = 0; i < n; i++) src[i] = sigma[i] >= ph->sigma_vac ? sigma[i] : 0.0;
    kctx c = {ph->G, ph->dx};
    int e = sro_convolve_ringed(src, w, h, 1, kfn, &c, (size_t)ph->max_padded, ringed);
    if (e == SRO_OK) {
        double inv2dx = 1.0 / (2.0 * ph->dx);
        for (size_t y = 0; y < h; y++)
            for (size_t x = 0; x < w; x++) {
                size_t i = y * w + x;
                if (phi) phi[i] = ringed[(y + 1) * rw + x + 1];
                if (gx) gx[i] = -(ringed[(y + 1) * rw + x + 2] - ringed[(y + 1) * rw + x]) * inv2dx;
                if (gy) gy[i] = -(ringed[(y + 2) * rw + x + 1] - ringed[y * rw + x + 1]) * inv2dx;
            }
        if (phi_ringed)
            for (size_t i = 0; i < rw * rh; i++) phi_ringed[i] = ringed[i];
    }
    free(src);
    free(ringed);
    return e;
}
