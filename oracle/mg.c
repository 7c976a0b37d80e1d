/* Oracle (test infrastructure only): an f64 model of the FLD linear solve.
 *
 * The operator is the reference's pass system (crates/sr-physics/src/reference/heat.rs:374-392, :474-488):
 *   (A v)_i = diag_i v_i - sum_{gas faces} k_f v_j,  diag_i = c_i + sum_f k_f,
 * written here in the equivalent face form own_i v_i + sum_{gas faces} k_f (v_i - v_j) with own_i = c_i + k_vac,i,
 * non-gas rows the identity (the reference excludes them from the system).
 *
 * Preconditioners: Jacobi (the reference twin's, heat.rs:419) or a multigrid V-cycle whose coarse operators are the
 * Galerkin products of a piecewise-constant 2x2 aggregation (the reference's GPU hierarchy, heat.wgsl:1160-1526,
 * builds the same aggregation), coarsened until a level holds at most 16 cells. Smoothing is exact red-black
 * Gauss-Seidel, red first on the way down and black first on the way up (the adjoint order, so the preconditioner
 * is symmetric); the coarsest level is solved exactly (dense Cholesky) when it holds at most 16 cells, else (a
 * capped level count) by 16 forward and 16 backward sweeps. The coarse correction may be scaled by omega (1 is the
 * reference's plain prolongation; the preconditioner stays SPD for 0 < omega < 2, sro_mg_set_coarse_scale). The CG is
 * the flexible (Polak-Ribiere) form. The device solver applies the same operator in f32; this model pins it.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "sro.h"

/* the coarse correction's scale (1: the reference's plain prolongation) */
static double g_omega = 1.0;

void sro_mg_set_coarse_scale(double omega) { g_omega = omega; }

#define COARSE_SWEEPS 16

typedef struct {
    size_t w, h;
    double *own; /* w*h */
    double *ax;  /* (w+1)*h couplings, 0 at the edges and to non-gas cells */
    double *ay;  /* w*(h+1) */
    double *z, *b, *r;
} level_t;

typedef struct {
    int n;
    level_t lv[24];
    const uint8_t *gas;
} hier_t;

static void level_alloc(level_t *l, size_t w, size_t h) {
    l->w = w;
    l->h = h;
    l->own = (double *)calloc(w * h, sizeof(double));
    l->ax = (double *)calloc((w + 1) * h, sizeof(double));
    l->ay = (double *)calloc(w * (h + 1), sizeof(double));
    l->z = (double *)calloc(w * h, sizeof(double));
    l->b = (double *)calloc(w * h, sizeof(double));
    l->r = (double *)calloc(w * h, sizeof(double));
}

static void level_free(level_t *l) {
    free(l->own);
    free(l->ax);
    free(l->ay);
    free(l->z);
    free(l->b);
    free(l->r);
}

static int is_coarsest(size_t w, size_t h) { return w * h <= 16; }

/* level 0 from the reference-form system, then Galerkin aggregation down */
static void hier_build(hier_t *H, size_t w, size_t h, const double *own, const double *kx, const double *ky,
                       const uint8_t *gas, int levels) {
    H->gas = gas;
    level_t *l0 = &H->lv[0];
    level_alloc(l0, w, h);
#define G(i) (gas ? gas[(i)] : 1)
    for (size_t i = 0; i < w * h; i++) l0->own[i] = G(i) ? own[i] : 1.0;
    for (size_t y = 0; y < h; y++)
        for (size_t x = 1; x < w; x++) {
            size_t a = y * w + x - 1, b = a + 1;
            l0->ax[y * (w + 1) + x] = (G(a) && G(b)) ? kx[y * (w + 1) + x] : 0.0;
        }
    for (size_t y = 1; y < h; y++)
        for (size_t x = 0; x < w; x++) {
            size_t a = (y - 1) * w + x, b = y * w + x;
            l0->ay[y * w + x] = (G(a) && G(b)) ? ky[y * w + x] : 0.0;
        }
    int n = 1;
    while (n < levels && !is_coarsest(H->lv[n - 1].w, H->lv[n - 1].h) && n < 24) {
        level_t *f = &H->lv[n - 1], *c = &H->lv[n];
        size_t cw = (f->w + 1) / 2, ch = (f->h + 1) / 2;
        level_alloc(c, cw, ch);
        for (size_t y = 0; y < f->h; y++)
            for (size_t x = 0; x < f->w; x++) {
                size_t i = y * f->w + x;
                int g = (n - 1 == 0) ? G(i) : 1;
                if (g) c->own[(y / 2) * cw + x / 2] += f->own[i];
            }
        for (size_t i = 0; i < cw * ch; i++)
            if (c->own[i] == 0.0) c->own[i] = 1.0;
        for (size_t y = 0; y < f->h; y++)
            for (size_t X = 1; X < cw; X++) {
                size_t fxi = 2 * X;
                if (fxi < f->w) c->ax[(y / 2) * (cw + 1) + X] += f->ax[y * (f->w + 1) + fxi];
            }
        for (size_t Y = 1; Y < ch; Y++)
            for (size_t x = 0; x < f->w; x++) {
                size_t fyi = 2 * Y;
                if (fyi < f->h) c->ay[Y * cw + x / 2] += f->ay[fyi * f->w + x];
            }
        n++;
    }
    H->n = n;
#undef G
}

static void hier_free(hier_t *H) {
    for (int i = 0; i < H->n; i++) level_free(&H->lv[i]);
}

static void level_matvec(const level_t *l, const double *v, double *out) {
    size_t w = l->w, h = l->h;
#pragma omp parallel for schedule(static)
    for (size_t y = 0; y < h; y++)
        for (size_t x = 0; x < w; x++) {
            size_t i = y * w + x;
            double s = l->own[i] * v[i];
            double a;
            if ((a = l->ax[y * (w + 1) + x]) != 0.0) s += a * (v[i] - v[i - 1]);
            if ((a = l->ax[y * (w + 1) + x + 1]) != 0.0) s += a * (v[i] - v[i + 1]);
            if ((a = l->ay[y * w + x]) != 0.0) s += a * (v[i] - v[i - w]);
            if ((a = l->ay[(y + 1) * w + x]) != 0.0) s += a * (v[i] - v[i + w]);
            out[i] = s;
        }
}

/* one half sweep of colour c: z_i = (b_i + sum a z_j) / (own_i + sum a) */
static void half_sweep(level_t *l, int color) {
    size_t w = l->w, h = l->h;
#pragma omp parallel for schedule(static)
    for (size_t y = 0; y < h; y++)
        for (size_t x = (y + color) & 1; x < w; x += 2) {
            size_t i = y * w + x;
            double num = l->b[i], den = l->own[i], a;
            if ((a = l->ax[y * (w + 1) + x]) != 0.0) { num += a * l->z[i - 1]; den += a; }
            if ((a = l->ax[y * (w + 1) + x + 1]) != 0.0) { num += a * l->z[i + 1]; den += a; }
            if ((a = l->ay[y * w + x]) != 0.0) { num += a * l->z[i - w]; den += a; }
            if ((a = l->ay[(y + 1) * w + x]) != 0.0) { num += a * l->z[i + w]; den += a; }
            l->z[i] = num / den;
        }
}

static void vcycle(hier_t *H, int L, int nu1, int nu2) {
    level_t *l = &H->lv[L];
    size_t n = l->w * l->h;
    memset(l->z, 0, sizeof(double) * n);
    if (L == H->n - 1 && n <= 16) {
        /* the exact solve: A = L L^T on the level's dense matrix */
        double A[16 * 16], y[16];
        size_t w = l->w, m = n;
        memset(A, 0, sizeof A);
        for (size_t i = 0; i < m; i++) {
            size_t x = i % w, yy = i / w;
            double a[4] = {l->ax[yy * (w + 1) + x], l->ax[yy * (w + 1) + x + 1], l->ay[yy * w + x],
                           l->ay[(yy + 1) * w + x]};
            long nb[4] = {(long)i - 1, (long)i + 1, (long)i - (long)w, (long)i + (long)w};
            double d = l->own[i];
            for (int f = 0; f < 4; f++)
                if (a[f] != 0.0) {
                    d += a[f];
                    A[i * m + nb[f]] -= a[f];
                }
            A[i * m + i] += d;
        }
        for (size_t j = 0; j < m; j++) {
            double s = A[j * m + j];
            for (size_t k = 0; k < j; k++) s -= A[j * m + k] * A[j * m + k];
            double d = sqrt(s > 1e-300 ? s : 1e-300);
            A[j * m + j] = d;
            for (size_t i = j + 1; i < m; i++) {
                double t = A[i * m + j];
                for (size_t k = 0; k < j; k++) t -= A[i * m + k] * A[j * m + k];
                A[i * m + j] = t / d;
            }
        }
        for (size_t i = 0; i < m; i++) {
            double t = l->b[i];
            for (size_t k = 0; k < i; k++) t -= A[i * m + k] * y[k];
            y[i] = t / A[i * m + i];
        }
        for (size_t i = m; i-- > 0;) {
            double t = y[i];
            for (size_t k = i + 1; k < m; k++) t -= A[k * m + i] * l->z[k];
            l->z[i] = t / A[i * m + i];
        }
        return;
    }
    if (L == H->n - 1) {
        for (int k = 0; k < COARSE_SWEEPS; k++) {
            half_sweep(l, 0);
            half_sweep(l, 1);
        }
        for (int k = 0; k < COARSE_SWEEPS; k++) {
            half_sweep(l, 1);
            half_sweep(l, 0);
        }
        return;
    }
    for (int k = 0; k < nu1; k++) {
        half_sweep(l, 0);
        half_sweep(l, 1);
    }
    level_matvec(l, l->z, l->r);
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < n; i++) l->r[i] = l->b[i] - l->r[i];
    level_t *c = &H->lv[L + 1];
    memset(c->b, 0, sizeof(double) * c->w * c->h);
    /* by coarse rows (each coarse cell sums its fine cells in row order, whatever the threads) */
#pragma omp parallel for schedule(static)
    for (size_t Y = 0; Y < c->h; Y++)
        for (size_t y = 2 * Y; y < 2 * Y + 2 && y < l->h; y++)
            for (size_t x = 0; x < l->w; x++) {
                size_t i = y * l->w + x;
                if (L == 0 && H->gas && !H->gas[i]) continue;
                c->b[Y * c->w + x / 2] += l->r[i];
            }
    vcycle(H, L + 1, nu1, nu2);
#pragma omp parallel for schedule(static)
    for (size_t y = 0; y < l->h; y++)
        for (size_t x = 0; x < l->w; x++) {
            size_t i = y * l->w + x;
            if (L == 0 && H->gas && !H->gas[i]) continue;
            l->z[i] += g_omega * c->z[(y / 2) * c->w + x / 2];
        }
    for (int k = 0; k < nu2; k++) {
        half_sweep(l, 1);
        half_sweep(l, 0);
    }
}

/* Jacobi on the reference diagonal own + sum of couplings (heat.rs:419) */
static void jacobi(hier_t *H, const double *r, double *z) {
    level_t *l0 = &H->lv[0];
    size_t n = l0->w * l0->h, w = l0->w;
    for (size_t i = 0; i < n; i++) {
        size_t x = i % w, y = i / w;
        double d = l0->own[i] + l0->ax[y * (w + 1) + x] + l0->ax[y * (w + 1) + x + 1] + l0->ay[y * w + x] +
                   l0->ay[(y + 1) * w + x];
        z[i] = (H->gas && !H->gas[i]) ? 0.0 : r[i] / d;
    }
}

static int g_nu1 = 2, g_nu2 = 2;

static void apply_m(hier_t *H, int kind, const double *r, double *z) {
    level_t *l0 = &H->lv[0];
    size_t n = l0->w * l0->h;
    if (kind == 0) {
        jacobi(H, r, z);
        return;
    }
    memcpy(l0->b, r, sizeof(double) * n);
    vcycle(H, 0, g_nu1, g_nu2);
    for (size_t i = 0; i < n; i++) z[i] = (H->gas && !H->gas[i]) ? 0.0 : l0->z[i];
}

void sro_fld_matvec(size_t w, size_t h, const double *own, const double *kx, const double *ky, const uint8_t *gas,
                    const double *v, double *out) {
    hier_t H;
    hier_build(&H, w, h, own, kx, ky, gas, 1);
    level_matvec(&H.lv[0], v, out);
    hier_free(&H);
}

int sro_mg_apply(size_t w, size_t h, const double *own, const double *kx, const double *ky, const uint8_t *gas,
                 const double *r, double *z, int levels, int nu1, int nu2) {
    hier_t H;
    hier_build(&H, w, h, own, kx, ky, gas, levels > 0 ? levels : 64);
    g_nu1 = nu1;
    g_nu2 = nu2;
    apply_m(&H, 1, r, z);
    hier_free(&H);
    return H.n;
}

static double dot(const double *a, const double *b, size_t n) {
    double s = 0;
#pragma omp parallel for schedule(static) reduction(+ : s)
    for (size_t i = 0; i < n; i++) s += a[i] * b[i];
    return s;
}

int sro_fld_pcg(size_t w, size_t h, const double *own, const double *kx, const double *ky, const uint8_t *gas,
                const double *b, double *x, double tol, int max_iter, int kind, int levels, int nu1, int nu2,
                double *final_rel) {
    size_t n = w * h;
    hier_t H;
    hier_build(&H, w, h, own, kx, ky, gas, kind == 1 ? (levels > 0 ? levels : 64) : 1);
    g_nu1 = nu1;
    g_nu2 = nu2;
    double *r = (double *)malloc(sizeof(double) * n), *z = (double *)malloc(sizeof(double) * n);
    double *zo = (double *)malloc(sizeof(double) * n), *p = (double *)malloc(sizeof(double) * n);
    double *q = (double *)malloc(sizeof(double) * n);
    for (size_t i = 0; i < n; i++)
        if (gas && !gas[i]) x[i] = 0.0;
    level_matvec(&H.lv[0], x, q);
    for (size_t i = 0; i < n; i++) r[i] = (gas && !gas[i]) ? 0.0 : b[i] - q[i];
    double bn = sqrt(dot(b, b, n));
    if (bn == 0.0) bn = 1.0;
    double rn = sqrt(dot(r, r, n));
    int it = 0;
    if (rn <= tol * bn) goto done;
    apply_m(&H, kind, r, z);
    memcpy(p, z, sizeof(double) * n);
    double rz = dot(r, z, n);
    for (it = 1; it <= max_iter; it++) {
        level_matvec(&H.lv[0], p, q);
        for (size_t i = 0; i < n; i++)
            if (gas && !gas[i]) q[i] = 0.0;
        double pq = dot(p, q, n);
        if (!(pq > 0.0) || !isfinite(pq)) {
            it = -it;
            break;
        }
        double alpha = rz / pq;
        for (size_t i = 0; i < n; i++) {
            x[i] += alpha * p[i];
            r[i] -= alpha * q[i];
        }
        rn = sqrt(dot(r, r, n));
        if (rn <= tol * bn) break;
        memcpy(zo, z, sizeof(double) * n);
        apply_m(&H, kind, r, z);
        double rzn = dot(r, z, n), rzo = dot(r, zo, n);
        double beta = (rzn - rzo) / rz;
        for (size_t i = 0; i < n; i++) p[i] = z[i] + beta * p[i];
        rz = rzn;
    }
    if (it > max_iter) it = max_iter;
done:
    level_matvec(&H.lv[0], x, q);
    for (size_t i = 0; i < n; i++) q[i] = (gas && !gas[i]) ? 0.0 : b[i] - q[i];
    if (final_rel) *final_rel = sqrt(dot(q, q, n)) / bn;
    free(r);
    free(z);
    free(zo);
    free(p);
    free(q);
    hier_free(&H);
    return it;
}
