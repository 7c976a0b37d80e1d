/* Oracle (test infrastructure only): P5 implicit flux-limited diffusion in f64.
 * Restates crates/sr-physics/src/reference/heat.rs:
 *   constants MAX_PASSES/MAX_ROUNDS/NEWTON_HEAT/NEWTON_CELL/SOLVE_TOL  heat.rs:55-60
 *   Flux::cells                  heat.rs:89-100
 *   lambda (Levermore-Pomraning) heat.rs:132-134
 *   side / kappa / vacuum_beyond heat.rs:202-224
 *   u_of / t_of / dt_du          heat.rs:227-248
 *   diffusivity                  heat.rs:251-265
 *   network / law                heat.rs:269-311
 *   Net::faces/flux/k_sum/k_vac/matvec  heat.rs:328-392
 *   System::solve (Jacobi PCG)   heat.rs:413-447
 *   Heat::step                   heat.rs:453-536
 *   commit_limited               heat.rs:544-592
 * With cell size dx a face's coefficient is k = D / dx^2 and its flux F = -D (u_b - u_a) / dx; R uses the gradient
 * |u_b - u_a| / dx. Heat capacity C = Sigma / (mu (gamma - 1)) (the reference's Sigma / mu at gamma = 2).
 * Pinned by the tests: the Newton passes on a single gas cell in vacuum (a scalar balance, so each pass is a step of
 * Newton's method and each solve exact) take as many passes as Newton's method takes steps under the reference's
 * stopping rule, with and without the [T/2, 2T] clamp engaging, and commit the balance's root; delta by conservation
 * (the grid loses exactly what it books as radiated); the converged result by the nonlinear balance cell by cell on
 * the reference's radiating slab, the constant-diffusivity moments, the solve against a direct solve, and
 * commit_limited by the reference's ceiling dump.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "sro.h"

#define NS SRO_NS
#define MAX_PASSES 128
#define MAX_ROUNDS 4096
#define NEWTON_HEAT 1e-12
#define NEWTON_CELL 1e-10
#define SOLVE_TOL 1e-12

/* the linear solver of Heat::step: 0 (the reference's Jacobi PCG under its energy-weighted rule, heat.rs:413-447), or
 * > 0: multigrid PCG (sro_fld_pcg, V(2,2)) to this relative residual, so that the twin finishes on grids where Jacobi
 * would take tens of thousands of rounds (a test-side switch; the reference has none) */
static double g_mg_rtol = 0.0;
void sro_heat_set_mg(double rtol) { g_mg_rtol = rtol > 0.0 ? rtol : 0.0; }

typedef struct {
    int gas;
    double t, sigma, kappa, inv_mu;
} side_t;

static double lp_lambda(double r) { return (2.0 + r) / (6.0 + 3.0 * r + r * r); }

static double limiter(const sro_physics *ph, double r) { return ph->flux_limiter ? lp_lambda(r) : 1.0 / 3.0; }

static double ks(const side_t *s) { return s->kappa * s->sigma; }
static double cap(const sro_physics *ph, const side_t *s) { return s->sigma * s->inv_mu / (ph->gamma - 1.0); }

static double kappa_of(const sro_physics *ph, const double *x, double sigma, double t) {
    if (ph->opacity == SRO_OPACITY_KRAMERS) return ph->kappa0 * sigma * pow(t, -3.5);
    double zm;
    sro_composition(x, NULL, NULL, &zm);
    double dust = t < ph->t_dust ? ph->kappa_dust * zm : 0.0;
    return ph->kappa0 * (1.0 + x[0]) + dust;
}

static double inv_mu_of(const sro_physics *ph, const double *x) {
    if (ph->mu > 0) return 1.0 / ph->mu;
    double im;
    sro_composition(x, &im, NULL, NULL);
    return im;
}

static side_t side_of(const sro_physics *ph, const double *planes, size_t n, size_t i) {
    double x[NS];
    for (int s = 0; s < NS; s++) x[s] = planes[(4 + s) * n + i];
    side_t sd;
    sd.inv_mu = inv_mu_of(ph, x);
    double sigma = planes[i];
    if (sigma < ph->sigma_vac) {
        sd.gas = 0;
        sd.t = ph->t_floor;
        sd.sigma = ph->sigma_floor;
        sd.kappa = kappa_of(ph, x, ph->sigma_floor, ph->t_floor);
        return sd;
    }
    double mx = planes[n + i], my = planes[2 * n + i], y_e;
    sro_composition(x, NULL, &y_e, NULL);
    double raw = planes[3 * n + i] / sigma - 0.5 * (mx * mx + my * my) / (sigma * sigma) -
                 sro_eps_cold(ph, sigma, y_e, x[NS - 1]);
    double fl = ph->t_floor * sd.inv_mu / (ph->gamma - 1.0);
    double eps = raw < fl ? fl : raw;
    sd.gas = 1;
    sd.t = (ph->gamma - 1.0) * eps / sd.inv_mu;
    sd.sigma = sigma;
    sd.kappa = kappa_of(ph, x, sigma, sd.t);
    return sd;
}

static side_t vacuum_beyond(const sro_physics *ph, const side_t *gas) {
    side_t s = {0, ph->t_floor, ph->sigma_floor, gas->kappa, gas->inv_mu};
    return s;
}

static int cd_on(const sro_physics *ph) { return ph->constant_diffusivity > 0.0; }
static double u_of(const sro_physics *ph, double t) { return cd_on(ph) ? t : ph->a_r * t * t * t * t; }
static double t_of(const sro_physics *ph, double u) { return cd_on(ph) ? u : sqrt(sqrt(u / ph->a_r)); }
static double dt_du(const sro_physics *ph, double t) { return cd_on(ph) ? 1.0 : 1.0 / (4.0 * ph->a_r * t * t * t); }

/* face coefficient k = D_f / dx^2 between sides a and b (heat.rs:251-265) */
static double diffusivity(const sro_physics *ph, const side_t *a, const side_t *b) {
    if (!a->gas && !b->gas) return 0.0;
    double idx2 = 1.0 / (ph->dx * ph->dx);
    if (cd_on(ph)) return ph->constant_diffusivity * (0.5 * (cap(ph, a) + cap(ph, b))) * idx2;
    double era = ph->a_r * pow(a->t, 4), erb = ph->a_r * pow(b->t, 4);
    double face_er = 0.5 * (era + erb);
    double k = 0.5 * (ks(a) + ks(b));
    double r = face_er > 0.0 ? fabs(erb - era) / (ph->dx * k * face_er) : 0.0;
    double radiative = ph->c_sb * limiter(ph, r) / k;
    double conduction = ph->k_cond != 0.0 ? ph->k_cond / (ph->a_r * (a->t + b->t) * (a->t * a->t + b->t * b->t)) : 0.0;
    return (radiative + conduction) * idx2;
}

typedef struct {
    size_t w, h;
    uint8_t *gas;
    double *c, *kx, *ky;
    double vac;
} net_t;

static double face_coef(const sro_physics *ph, const side_t *a, const side_t *b) {
    side_t A, B;
    if (a && b) {
        if (a->gas || b->gas) {
            A = a->gas ? *a : vacuum_beyond(ph, b);
            B = b->gas ? *b : vacuum_beyond(ph, a);
        } else {
            A = *a;
            B = *a;
        }
    } else if (a) {
        A = *a;
        B = vacuum_beyond(ph, a);
    } else {
        A = vacuum_beyond(ph, b);
        B = *b;
    }
    return diffusivity(ph, &A, &B);
}

static side_t *network(const sro_physics *ph, const double *planes, size_t w, size_t h, double vac, net_t *net) {
    size_t n = w * h;
    side_t *sides = (side_t *)malloc(sizeof(side_t) * n);
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < n; i++) sides[i] = side_of(ph, planes, n, i);
    net->w = w;
    net->h = h;
    net->gas = (uint8_t *)malloc(n);
    net->c = (double *)malloc(sizeof(double) * n);
    net->kx = (double *)malloc(sizeof(double) * (w + 1) * h);
    net->ky = (double *)malloc(sizeof(double) * w * (h + 1));
    net->vac = vac;
    for (size_t i = 0; i < n; i++) {
        net->gas[i] = (uint8_t)sides[i].gas;
        net->c[i] = cap(ph, &sides[i]);
    }
#pragma omp parallel for schedule(static)
    for (size_t y = 0; y < h; y++)
        for (size_t x = 0; x <= w; x++)
            net->kx[y * (w + 1) + x] =
                face_coef(ph, x > 0 ? &sides[y * w + x - 1] : NULL, x < w ? &sides[y * w + x] : NULL);
#pragma omp parallel for schedule(static)
    for (size_t y = 0; y <= h; y++)
        for (size_t x = 0; x < w; x++)
            net->ky[y * w + x] =
                face_coef(ph, y > 0 ? &sides[(y - 1) * w + x] : NULL, y < h ? &sides[y * w + x] : NULL);
    return sides;
}

static void net_free(net_t *n) {
    free(n->gas);
    free(n->c);
    free(n->kx);
    free(n->ky);
}

/* the four faces of (x, y): coefficient, neighbour index (-1 past the edge) — heat.rs:328-337 */
static void net_faces(const net_t *net, size_t x, size_t y, double k[4], long nb[4]) {
    size_t w = net->w, h = net->h, i = y * w + x;
    k[0] = net->kx[y * (w + 1) + x];
    nb[0] = x > 0 ? (long)i - 1 : -1;
    k[1] = net->kx[y * (w + 1) + x + 1];
    nb[1] = x + 1 < w ? (long)i + 1 : -1;
    k[2] = net->ky[y * w + x];
    nb[2] = y > 0 ? (long)(i - w) : -1;
    k[3] = net->ky[(y + 1) * w + x];
    nb[3] = y + 1 < h ? (long)(i + w) : -1;
}

/* F = -k dx (v_b - v_a) at every face (heat.rs:340-362 with the dx of a face flux) */
static void net_flux(const sro_physics *ph, const net_t *net, const double *v, double *fxo, double *fyo) {
    size_t w = net->w, h = net->h;
#define VAL(i) (((i) >= 0 && net->gas[(i)]) ? v[(i)] : net->vac)
    for (size_t y = 0; y < h; y++)
        for (size_t x = 0; x <= w; x++) {
            long a = x > 0 ? (long)(y * w + x - 1) : -1, b = x < w ? (long)(y * w + x) : -1;
            double k = net->kx[y * (w + 1) + x];
            fxo[y * (w + 1) + x] = k == 0.0 ? 0.0 : -k * ph->dx * (VAL(b) - VAL(a));
        }
    for (size_t y = 0; y <= h; y++)
        for (size_t x = 0; x < w; x++) {
            long a = y > 0 ? (long)((y - 1) * w + x) : -1, b = y < h ? (long)(y * w + x) : -1;
            double k = net->ky[y * w + x];
            fyo[y * w + x] = k == 0.0 ? 0.0 : -k * ph->dx * (VAL(b) - VAL(a));
        }
#undef VAL
}

static double k_sum(const net_t *net, size_t i) {
    double k[4];
    long nb[4];
    net_faces(net, i % net->w, i / net->w, k, nb);
    return k[0] + k[1] + k[2] + k[3];
}

static double k_vac(const net_t *net, size_t i) {
    double k[4], s = 0;
    long nb[4];
    net_faces(net, i % net->w, i / net->w, k, nb);
    for (int f = 0; f < 4; f++)
        if (!(nb[f] >= 0 && net->gas[nb[f]])) s += k[f];
    return s;
}

static void net_matvec(const net_t *net, const double *diag, const double *v, double *out) {
    size_t w = net->w, h = net->h;
    for (size_t y = 0; y < h; y++)
        for (size_t x = 0; x < w; x++) {
            size_t i = y * w + x;
            if (!net->gas[i]) {
                out[i] = 0.0;
                continue;
            }
            double k[4];
            long nb[4];
            net_faces(net, x, y, k, nb);
            double s = diag[i] * v[i];
            for (int f = 0; f < 4; f++)
                if (nb[f] >= 0 && net->gas[nb[f]]) s -= k[f] * v[nb[f]];
            out[i] = s;
        }
}

typedef struct {
    const net_t *net;
    double *diag, *rhs, *heat, *slope;
    double u_bound, t_bound;
} system_t;

/* Jacobi PCG (heat.rs:413-447): >= 0 rounds, or -(rounds + 1) when a sum is not finite */
static int sys_solve(const system_t *s, double *x, int max_rounds) {
    const net_t *net = s->net;
    size_t n = net->w * net->h;
    double *q = (double *)calloc(n, sizeof(double)), *r = (double *)calloc(n, sizeof(double));
    double *z = (double *)calloc(n, sizeof(double)), *p = (double *)calloc(n, sizeof(double));
    net_matvec(net, s->diag, x, q);
    for (size_t i = 0; i < n; i++) {
        r[i] = net->gas[i] ? s->rhs[i] - q[i] : 0.0;
        z[i] = net->gas[i] ? r[i] / s->diag[i] : 0.0;
        p[i] = z[i];
    }
    double rz = 0;
    for (size_t i = 0; i < n; i++) rz += r[i] * z[i];
    int rounds = 0, ret;
    for (;;) {
        double e_heat = 0;
        for (size_t i = 0; i < n; i++)
            if (net->gas[i]) {
                double e = r[i] * s->slope[i] / s->diag[i];
                e_heat += s->heat[i] * e * e;
            }
        net_matvec(net, s->diag, p, q);
        double pq = 0;
        for (size_t i = 0; i < n; i++) pq += p[i] * q[i];
        if (!(isfinite(rz) && isfinite(e_heat) && isfinite(pq))) {
            ret = -(rounds + 1);
            break;
        }
        if ((rz <= s->u_bound && e_heat <= s->t_bound) || rounds == max_rounds) {
            ret = rounds;
            break;
        }
        double alpha = rz / pq;
        for (size_t i = 0; i < n; i++) {
            x[i] += alpha * p[i];
            r[i] -= alpha * q[i];
            z[i] = net->gas[i] ? r[i] / s->diag[i] : 0.0;
        }
        double rzn = 0;
        for (size_t i = 0; i < n; i++) rzn += r[i] * z[i];
        double beta = rzn / rz;
        for (size_t i = 0; i < n; i++) p[i] = z[i] + beta * p[i];
        rz = rzn;
        rounds++;
    }
    free(q);
    free(r);
    free(z);
    free(p);
    return ret;
}

void sro_flux_cells(const double *flux_x, const double *flux_y, size_t w, size_t h, double *fx, double *fy) {
    for (size_t y = 0; y < h; y++)
        for (size_t x = 0; x < w; x++) {
            fx[y * w + x] = 0.5 * (flux_x[y * (w + 1) + x] + flux_x[y * (w + 1) + x + 1]);
            fy[y * w + x] = 0.5 * (flux_y[y * w + x] + flux_y[(y + 1) * w + x]);
        }
}

int sro_heat_law(const sro_physics *ph, const double *planes, size_t w, size_t h, double *flux_x, double *flux_y) {
    net_t net;
    side_t *sides = network(ph, planes, w, h, u_of(ph, ph->t_floor), &net);
    size_t n = w * h;
    double *u = (double *)malloc(sizeof(double) * n);
    for (size_t i = 0; i < n; i++) u[i] = u_of(ph, sides[i].t);
    net_flux(ph, &net, u, flux_x, flux_y);
    free(u);
    free(sides);
    net_free(&net);
    return SRO_OK;
}

int sro_heat_first_system(const sro_physics *ph, const double *planes, size_t w, size_t h, double dt, double *own,
                          double *kx, double *ky, double *rhs, uint8_t *gas, double *u0o) {
    size_t n = w * h;
    double u_vac = u_of(ph, ph->t_floor);
    net_t net;
    side_t *sides = network(ph, planes, w, h, u_vac, &net);
    for (size_t i = 0; i < n; i++) {
        double t = sides[i].t, u = u_of(ph, t);
        if (net.gas[i]) {
            double heat = net.c[i] / dt, ci = heat * dt_du(ph, t), kv = k_vac(&net, i);
            own[i] = ci + kv;
            rhs[i] = ci * u + kv * u_vac;
        } else {
            own[i] = 1.0;
            rhs[i] = 0.0;
        }
        gas[i] = net.gas[i];
        if (u0o) u0o[i] = u;
    }
    memcpy(kx, net.kx, sizeof(double) * (w + 1) * h);
    memcpy(ky, net.ky, sizeof(double) * w * (h + 1));
    free(sides);
    net_free(&net);
    return SRO_OK;
}

static void commit_limited(const sro_physics *ph, double *planes, const net_t *net, const double *u, const double *t0,
                           double dt, double *flux_x, double *flux_y, double *emitted, sro_heat_report *rep) {
    size_t w = net->w, h = net->h, n = w * h;
    double *fxo = (double *)malloc(sizeof(double) * (w + 1) * h), *fyo = (double *)malloc(sizeof(double) * w * (h + 1));
    net_flux(ph, net, u, fxo, fyo);
    double *scale = (double *)malloc(sizeof(double) * n);
    double idx = 1.0 / ph->dx;
#define LEAVING(i, o)                                                                                                  \
    do {                                                                                                               \
        size_t _x = (i) % w, _y = (i) / w;                                                                             \
        (o)[0] = -fxo[_y * (w + 1) + _x];                                                                              \
        (o)[1] = fxo[_y * (w + 1) + _x + 1];                                                                           \
        (o)[2] = -fyo[_y * w + _x];                                                                                    \
        (o)[3] = fyo[(_y + 1) * w + _x];                                                                               \
    } while (0)
    for (size_t i = 0; i < n; i++) {
        if (!net->gas[i]) {
            scale[i] = 1.0;
            continue;
        }
        double o[4];
        LEAVING(i, o);
        double out = dt * idx * (fmax(o[0], 0.0) + fmax(o[1], 0.0) + fmax(o[2], 0.0) + fmax(o[3], 0.0));
        double held = net->c[i] * fmax(t0[i] - ph->t_floor, 0.0);
        scale[i] = out > held ? held / out : 1.0;
    }
    for (size_t y = 0; y < h; y++)
        for (size_t x = 0; x <= w; x++) {
            size_t k = y * (w + 1) + x;
            long from = fxo[k] > 0.0 ? (x > 0 ? (long)(y * w + x - 1) : -1) : (x < w ? (long)(y * w + x) : -1);
            fxo[k] *= from >= 0 ? scale[from] : 1.0;
        }
    for (size_t y = 0; y <= h; y++)
        for (size_t x = 0; x < w; x++) {
            size_t k = y * w + x;
            long from = fyo[k] > 0.0 ? (y > 0 ? (long)((y - 1) * w + x) : -1) : (y < h ? (long)(y * w + x) : -1);
            fyo[k] *= from >= 0 ? scale[from] : 1.0;
        }
    double radiated = 0;
    for (size_t i = 0; i < n; i++) {
        double em = 0;
        if (net->gas[i]) {
            double o[4], k[4];
            long nb[4];
            LEAVING(i, o);
            net_faces(net, i % w, i / w, k, nb);
            for (int f = 0; f < 4; f++)
                if (!(nb[f] >= 0 && net->gas[nb[f]])) em += dt * idx * o[f];
            planes[3 * n + i] -= dt * idx * (o[0] + o[1] + o[2] + o[3]);
        }
        if (emitted) emitted[i] = em;
        radiated += em;
    }
#undef LEAVING
    if (flux_x) memcpy(flux_x, fxo, sizeof(double) * (w + 1) * h);
    if (flux_y) memcpy(flux_y, fyo, sizeof(double) * w * (h + 1));
    rep->delta = 0.0;
    rep->converged = 0;
    rep->radiated = radiated;
    free(fxo);
    free(fyo);
    free(scale);
}

int sro_heat_step(const sro_physics *ph, double *planes, size_t w, size_t h, double dt, double *flux_x,
                  double *flux_y, double *emitted, sro_heat_report *rep) {
    size_t n = w * h;
    double u_vac = u_of(ph, ph->t_floor);
    net_t net;
    side_t *sides = network(ph, planes, w, h, u_vac, &net);
    double *t0 = (double *)malloc(sizeof(double) * n), *u0 = (double *)malloc(sizeof(double) * n);
    size_t ngas = 0;
    for (size_t i = 0; i < n; i++) {
        t0[i] = sides[i].t;
        u0[i] = u_of(ph, t0[i]);
        ngas += net.gas[i];
    }
    memset(rep, 0, sizeof *rep);
    rep->converged = 1;
    if (ph->pin_temperature || ngas == 0 || isnan(dt) || dt <= 0.0) {
        double *fxo = (double *)malloc(sizeof(double) * (w + 1) * h), *fyo = (double *)malloc(sizeof(double) * w * (h + 1));
        net_flux(ph, &net, u0, fxo, fyo);
        if (flux_x) memcpy(flux_x, fxo, sizeof(double) * (w + 1) * h);
        if (flux_y) memcpy(flux_y, fyo, sizeof(double) * w * (h + 1));
        if (emitted) memset(emitted, 0, sizeof(double) * n);
        free(fxo);
        free(fyo);
        free(t0);
        free(u0);
        free(sides);
        net_free(&net);
        return SRO_OK;
    }
    double heat0 = 0;
    double *ksum = (double *)calloc(n, sizeof(double)), *kvac = (double *)calloc(n, sizeof(double));
    for (size_t i = 0; i < n; i++)
        if (net.gas[i]) {
            heat0 += net.c[i] * t0[i];
            ksum[i] = k_sum(&net, i);
            kvac[i] = k_vac(&net, i);
        }
    double *t = (double *)malloc(sizeof(double) * n), *u = (double *)malloc(sizeof(double) * n);
    memcpy(t, t0, sizeof(double) * n);
    memcpy(u, u0, sizeof(double) * n);
    int passes = 0, rounds = 0, converged = 0;
    double *heat = (double *)calloc(n, sizeof(double)), *slope = (double *)calloc(n, sizeof(double));
    double *diag = (double *)calloc(n, sizeof(double)), *rhs = (double *)calloc(n, sizeof(double));
    double *x = (double *)calloc(n, sizeof(double));
    while (passes < MAX_PASSES && rounds < MAX_ROUNDS) {
        passes++;
        double eps2 = SOLVE_TOL * SOLVE_TOL, ub = 0, tb = 0;
#pragma omp parallel for schedule(static) reduction(+ : ub, tb)
        for (size_t i = 0; i < n; i++) {
            if (net.gas[i]) {
                heat[i] = net.c[i] / dt;
                slope[i] = dt_du(ph, t[i]);
                double ci = heat[i] * slope[i];
                diag[i] = ci + ksum[i];
                rhs[i] = ci * u[i] - heat[i] * (t[i] - t0[i]) + kvac[i] * u_vac;
                ub += ci * u[i] * u[i];
                tb += heat[i] * t[i] * t[i];
                x[i] = u[i];
            } else {
                heat[i] = slope[i] = 0.0;
                diag[i] = 1.0;
                rhs[i] = 0.0;
                x[i] = 0.0;
            }
        }
        system_t sys = {&net, diag, rhs, heat, slope, eps2 * ub, eps2 * tb};
        if (!(isfinite(sys.u_bound) && isfinite(sys.t_bound))) {
            passes--;
            break;
        }
        int r;
        if (g_mg_rtol > 0.0) {
            /* the same system in face form: own = diag - the gas-gas couplings (kvac stays in own) */
            double *own = (double *)malloc(sizeof(double) * n), rel = 0.0;
            for (size_t i = 0; i < n; i++) own[i] = net.gas[i] ? diag[i] - (ksum[i] - kvac[i]) : 1.0;
            int it = sro_fld_pcg(w, h, own, net.kx, net.ky, net.gas, rhs, x, g_mg_rtol, MAX_ROUNDS - rounds, 1, 0, 2,
                                 2, &rel);
            free(own);
            r = (it < 0 || !isfinite(rel)) ? (it < 0 ? it : -1) : it; /* a failed iteration k counts k - 1 rounds */
        } else {
            r = sys_solve(&sys, x, MAX_ROUNDS - rounds);
        }
        if (r < 0) {
            rounds += -r - 1;
            break;
        }
        rounds += r;
        double moved = 0;
        int every = 1;
#pragma omp parallel for schedule(static) reduction(+ : moved) reduction(& : every)
        for (size_t i = 0; i < n; i++) {
            if (!net.gas[i]) continue;
            double nx = t_of(ph, fmax(x[i], u_vac));
            nx = fmin(fmax(nx, 0.5 * t[i]), 2.0 * t[i]);
            double d = fabs(nx - t[i]);
            moved += net.c[i] * d;
            every &= d <= NEWTON_CELL * t[i];
            t[i] = nx;
            u[i] = u_of(ph, nx);
        }
        if (moved <= NEWTON_HEAT * heat0 && every) {
            converged = 1;
            break;
        }
    }
    rep->passes = passes;
    rep->rounds = rounds;
    if (!converged) {
        commit_limited(ph, planes, &net, u, t0, dt, flux_x, flux_y, emitted, rep);
    } else {
        double *fxo = (double *)malloc(sizeof(double) * (w + 1) * h), *fyo = (double *)malloc(sizeof(double) * w * (h + 1));
        net_flux(ph, &net, u, fxo, fyo);
        double radiated = 0, gained = 0, held = 0;
        for (size_t i = 0; i < n; i++) {
            double em = net.gas[i] ? dt * kvac[i] * (u[i] - u_vac) : 0.0;
            if (emitted) emitted[i] = em;
            radiated += em;
        }
        for (size_t i = 0; i < n; i++)
            if (net.gas[i]) {
                gained += net.c[i] * (t[i] - t0[i]);
                held += net.c[i] * t[i];
            }
        double delta = -(gained + radiated) / held;
        for (size_t i = 0; i < n; i++)
            if (net.gas[i]) planes[3 * n + i] += net.c[i] * ((1.0 + delta) * t[i] - t0[i]);
        if (flux_x) memcpy(flux_x, fxo, sizeof(double) * (w + 1) * h);
        if (flux_y) memcpy(flux_y, fyo, sizeof(double) * w * (h + 1));
        rep->delta = delta;
        rep->radiated = radiated;
        rep->converged = 1;
        free(fxo);
        free(fyo);
    }
    free(heat);
    free(slope);
    free(diag);
    free(rhs);
    free(x);
    free(t);
    free(u);
    free(ksum);
    free(kvac);
    free(t0);
    free(u0);
    free(sides);
    net_free(&net);
    return SRO_OK;
}
