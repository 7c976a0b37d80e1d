/* Oracle (test infrastructure only): P4 hydro in f64 — MUSCL-Hancock, MC limiter, HLL with Davis speeds.
 * Restates crates/sr-physics/src/reference/hydro.rs (module doc :1-96 is the specification):
 *   Prim::offset/mirrored/copied  hydro.rs:237-272
 *   mc / species_slopes / slope   hydro.rs:283-331 (HOT_FACE = 2)
 *   radiation_force / push_share  hydro.rs:371-410
 *   read / cons / prim / sound    hydro.rs:424-484
 *   flux / evolved / hll          hydro.rs:487-540
 *   wall / boundary               hydro.rs:559-582
 *   predicted                     hydro.rs:587-614
 *   dt                            hydro.rs:618-634
 *   step_held (no held cells)     hydro.rs:657-1025
 *   face_work                     hydro.rs:1048-1067
 * and the EOS it reads (crates/sr-physics/src/eos.rs:207-260; the cold pressure's terms from eos.c), plus the
 * composition (crates/sr-physics/src/registry.rs:146-154).
 * Held cells, sinks and the pull field are not restated (the path refuses a scene that starts at or above
 * Sigma_hold, floors.rs:200-260).
 * Each function is checked by the tests: closed forms, conservation, and the reference's own cases (the thin cell for
 * HOT_FACE, the Jeans adiabat for face_work, the radiation-force slab for push_share and the push's take-back); gravity's
 * kick take-back (hydro.rs:997-1001) by momentum conservation under a field given an external bias through a test
 * hook, which fails with the take-back left out or mis-signed.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "sro.h"

#define NS SRO_NS
#define NC SRO_COMPONENTS
#define HOT_FACE 2.0

/* species A and Z, assets/elements.json order: H He C O Ne Mg Si S Fe n */
static const double SP_A[NS] = {1, 4, 12, 16, 20, 24, 28, 32, 56, 1};
static const double SP_Z[NS] = {1, 2, 6, 8, 10, 12, 14, 16, 26, 0};

void sro_composition(const double *x, double *inv_mu, double *y_e, double *z_met) {
    double im = 0, ye = 0;
    for (int s = 0; s < NS; s++) {
        im += x[s] * (SP_Z[s] + 1.0) / SP_A[s];
        ye += x[s] * SP_Z[s] / SP_A[s];
    }
    if (inv_mu) *inv_mu = im;
    if (y_e) *y_e = ye;
    if (z_met) *z_met = 1.0 - x[0] - x[1] - x[9];
}

/* 1/mu of a cell: fixed, or the species' */
static double inv_mu_of(const sro_physics *ph, const double *x) {
    if (ph->mu > 0) return 1.0 / ph->mu;
    double im;
    sro_composition(x, &im, NULL, NULL);
    return im;
}

/* Y_e of a cell's species (registry.rs:146-154) */
static double y_e_of(const double *x) {
    double ye;
    sro_composition(x, NULL, &ye, NULL);
    return ye;
}

/* eps floor T_floor / (mu (gamma - 1)) */
static double eps_floor(const sro_physics *ph, double inv_mu) { return ph->t_floor * inv_mu / (ph->gamma - 1.0); }

typedef struct {
    double sigma, u, v, pi;
    double x[NS];
} prim_t;

typedef struct {
    double u[NC];
    prim_t p;
} face_t;

static prim_t offset(const prim_t *c, const prim_t *s, double f) {
    prim_t o;
    o.sigma = c->sigma + f * s->sigma;
    o.u = c->u + f * s->u;
    o.v = c->v + f * s->v;
    o.pi = c->pi + f * s->pi;
    for (int k = 0; k < NS; k++) o.x[k] = c->x[k] + f * s->x[k];
    return o;
}

static prim_t mirrored(prim_t p, int dir) {
    if (dir == 0)
        p.u = -p.u;
    else
        p.v = -p.v;
    return p;
}

static prim_t copied(prim_t p, int dir, int high) {
    double *n = dir == 0 ? &p.u : &p.v;
    *n = high ? fmax(*n, 0.0) : fmin(*n, 0.0);
    return p;
}

static double mc(double l, double c, double r) {
    double a = c - l, b = r - c;
    if (a * b <= 0.0) return 0.0;
    double m = fmin(fmin(fabs(0.5 * (a + b)), 2.0 * fabs(a)), 2.0 * fabs(b));
    return copysign(m, a);
}

/* the HOT_FACE bracket on (1) or off (0): off only to show a test that the bracket is what it checks */
static int g_hot_face = 1;
void sro_hydro_set_hot_face(int on) { g_hot_face = on; }
/* Test switch: 0 charges a face's negative work wholly to the two cells' energies, with no impulse between them (the
 * split before the pairs' relative motion paid first; a planted mistake for the Jeans adiabat test). 1 is the
 * reference's face_work. */
static int g_pair_work = 1;
void sro_hydro_set_pair_work(int on) { g_pair_work = on; }
/* Test hooks for gravity's kick take-back: a uniform field (bx, by) added to the half step's g, which the solve's
 * symmetric kernel never makes, so the source terms' net kick dt sum Sigma_src (g + b) is not zero; and the take-back
 * itself: 1 the reference's (hydro.rs:997-1001), 0 left out, -1 added with the wrong sign (planted mistakes). */
static double g_bias[2] = {0.0, 0.0};
static int g_kick_takeback = 1;
void sro_hydro_set_field_bias(double bx, double by) {
    g_bias[0] = bx;
    g_bias[1] = by;
}
void sro_hydro_set_kick_takeback(int mode) { g_kick_takeback = mode; }

static prim_t slope(const prim_t *l, const prim_t *c, const prim_t *r) {
    prim_t s;
    s.sigma = mc(l->sigma, c->sigma, r->sigma);
    s.u = mc(l->u, c->u, r->u);
    s.v = mc(l->v, c->v, r->v);
    s.pi = mc(l->pi, c->pi, r->pi);
    double central[NS], theta = 1.0;
    for (int i = 0; i < NS; i++) {
        central[i] = 0.5 * (r->x[i] - l->x[i]);
        if (central[i] != 0.0) theta = fmin(theta, mc(l->x[i], c->x[i], r->x[i]) / central[i]);
    }
    for (int i = 0; i < NS; i++) s.x[i] = theta * central[i];
    double hottest = fmax(fmax(l->pi / l->sigma, c->pi / c->sigma), r->pi / r->sigma);
    int hot = 0;
    for (int k = 0; k < 2; k++) {
        double f = k ? 0.5 : -0.5;
        double pi = c->pi + f * s.pi, bound = HOT_FACE * hottest * (c->sigma + f * s.sigma);
        /* hot when pi > bound or the comparison is unordered (NaN) */
        if (!(pi <= bound)) hot = 1;
    }
    if (hot && g_hot_face) {
        s.sigma = 0.0;
        s.pi = 0.0;
    }
    return s;
}

typedef struct {
    const sro_physics *ph;
} hy_t;

/* Pi = Pi_th + Pi_e + Pi_n (eos.rs:233-236), Pi_th = (gamma - 1) Sigma eps_th */
static double pi_of(const sro_physics *ph, double sigma, double eps, const double *x) {
    return (ph->gamma - 1.0) * sigma * eps + sro_pi_cold(ph, sigma, y_e_of(x), x[NS - 1]);
}

/* thermal energy recovered from E less the kinetic and the cold energy, floored (eos.rs:251-260) */
static double thermal(const sro_physics *ph, double E, double sigma, double mx, double my, double inv_mu,
                      const double *x, double *floor_added) {
    double kinetic = 0.5 * (mx * mx + my * my) / (sigma * sigma);
    double raw = E / sigma - kinetic - sro_eps_cold(ph, sigma, y_e_of(x), x[NS - 1]);
    double fl = eps_floor(ph, inv_mu);
    if (raw < fl) {
        if (floor_added) *floor_added = sigma * (fl - raw);
        return fl;
    }
    if (floor_added) *floor_added = 0.0;
    return raw;
}

static prim_t read_cell(const sro_physics *ph, const double *planes, size_t n, size_t i) {
    prim_t p;
    for (int s = 0; s < NS; s++) p.x[s] = planes[(4 + s) * n + i];
    double inv_mu = inv_mu_of(ph, p.x);
    double sigma = planes[i];
    if (sigma < ph->sigma_vac) {
        p.sigma = ph->sigma_floor;
        p.u = p.v = 0.0;
        p.pi = pi_of(ph, ph->sigma_floor, eps_floor(ph, inv_mu), p.x);
        return p;
    }
    double mx = planes[n + i], my = planes[2 * n + i];
    double eps = thermal(ph, planes[3 * n + i], sigma, mx, my, inv_mu, p.x, NULL);
    p.sigma = sigma;
    p.u = mx / sigma;
    p.v = my / sigma;
    p.pi = pi_of(ph, sigma, eps, p.x);
    return p;
}

static void cons(const sro_physics *ph, const prim_t *p, double *u) {
    double inv_mu = inv_mu_of(ph, p->x), y_e = y_e_of(p->x), x_n = p->x[NS - 1];
    double th = (p->pi - sro_pi_cold(ph, p->sigma, y_e, x_n)) / ((ph->gamma - 1.0) * p->sigma);
    double eps = fmax(th, eps_floor(ph, inv_mu));
    u[0] = p->sigma;
    u[1] = p->sigma * p->u;
    u[2] = p->sigma * p->v;
    u[3] = p->sigma * (eps + sro_eps_cold(ph, p->sigma, y_e, x_n)) + 0.5 * p->sigma * (p->u * p->u + p->v * p->v);
    for (int s = 0; s < NS; s++) u[4 + s] = p->sigma * p->x[s];
}

static int prim_of(const sro_physics *ph, const double *u, prim_t *p) {
    double sigma = u[0];
    if (isnan(sigma) || sigma <= 0.0) return 0;
    for (int k = 0; k < NC; k++)
        if (!isfinite(u[k])) return 0;
    for (int s = 0; s < NS; s++) p->x[s] = u[4 + s] / sigma;
    double inv_mu = inv_mu_of(ph, p->x);
    double raw = u[3] / sigma - 0.5 * (u[1] * u[1] + u[2] * u[2]) / (sigma * sigma) -
                 sro_eps_cold(ph, sigma, y_e_of(p->x), p->x[NS - 1]);
    if (isnan(raw) || raw <= 0.0) return 0;
    double eps = fmax(raw, eps_floor(ph, inv_mu));
    p->sigma = sigma;
    p->u = u[1] / sigma;
    p->v = u[2] / sigma;
    p->pi = pi_of(ph, sigma, eps, p->x);
    return 1;
}

static double sound(const sro_physics *ph, const prim_t *p) {
    double c2 = ph->gamma * p->pi / p->sigma;
    return sqrt(fmax(c2, 0.0));
}

static void flux(const prim_t *p, double energy, int dir, double *f) {
    double un = dir == 0 ? p->u : p->v, ut = dir == 0 ? p->v : p->u;
    double mass = p->sigma * un;
    f[0] = mass;
    f[1 + dir] = mass * un + p->pi;
    f[2 - dir] = mass * ut;
    f[3] = (energy + p->pi) * un;
    for (int s = 0; s < NS; s++) f[4 + s] = mass * p->x[s];
}

static face_t evolved(const sro_physics *ph, const prim_t *base, const double *delta) {
    face_t f;
    double u0[NC];
    cons(ph, base, u0);
    for (int k = 0; k < NC; k++) f.u[k] = u0[k] + delta[k];
    if (!prim_of(ph, f.u, &f.p)) {
        memcpy(f.u, u0, sizeof u0);
        f.p = *base;
    }
    return f;
}

static void hll(const sro_physics *ph, const face_t *l, const face_t *r, int dir, double *f) {
    double unl = dir == 0 ? l->p.u : l->p.v, unr = dir == 0 ? r->p.u : r->p.v;
    double cl = sound(ph, &l->p), cr = sound(ph, &r->p);
    double sl = fmin(unl - cl, unr - cr), sr = fmax(unl + cl, unr + cr);
    double fl[NC], fr[NC];
    flux(&l->p, l->u[3], dir, fl);
    flux(&r->p, r->u[3], dir, fr);
    if (sl >= 0.0) {
        memcpy(f, fl, sizeof fl);
    } else if (sr <= 0.0) {
        memcpy(f, fr, sizeof fr);
    } else {
        for (int k = 0; k < 4; k++) f[k] = (sr * fl[k] - sl * fr[k] + sl * sr * (r->u[k] - l->u[k])) / (sr - sl);
    }
    const prim_t *up = f[0] > 0.0 ? &l->p : &r->p;
    for (int s = 0; s < NS; s++) f[4 + s] = f[0] * up->x[s];
}

static void wall(const face_t *face, int dir, double *f) {
    memset(f, 0, sizeof(double) * NC);
    f[1 + dir] = face->p.pi;
}

/* predicted: the four face states [x-, x+, y-, y+] and the centre's (Sigma, Sigma u, Sigma v) half a step on */
static void predicted(const sro_physics *ph, const prim_t *grid, size_t pw, size_t gi, double half, const double *g,
                      face_t faces[4], double mid[3]) {
    const prim_t *c = &grid[gi];
    prim_t sx = slope(&grid[gi - 1], c, &grid[gi + 1]);
    prim_t sy = slope(&grid[gi - pw], c, &grid[gi + pw]);
    prim_t wf[4] = {offset(c, &sx, -0.5), offset(c, &sx, 0.5), offset(c, &sy, -0.5), offset(c, &sy, 0.5)};
    double uf[4][NC];
    for (int k = 0; k < 4; k++) cons(ph, &wf[k], uf[k]);
    double fxl[NC], fxr[NC], fyl[NC], fyr[NC];
    flux(&wf[0], uf[0][3], 0, fxl);
    flux(&wf[1], uf[1][3], 0, fxr);
    flux(&wf[2], uf[2][3], 1, fyl);
    flux(&wf[3], uf[3][3], 1, fyr);
    double hdx = half / ph->dx;
    double delta[NC];
    for (int k = 0; k < NC; k++) delta[k] = hdx * ((fxl[k] - fxr[k]) + (fyl[k] - fyr[k]));
    mid[0] = c->sigma + delta[0];
    mid[1] = c->sigma * c->u + delta[1];
    mid[2] = c->sigma * c->v + delta[2];
    if (!g) {
        for (int k = 0; k < 4; k++) faces[k] = evolved(ph, &wf[k], delta);
        return;
    }
    double gx = g[0], gy = g[1];
    mid[1] += half * c->sigma * gx;
    mid[2] += half * c->sigma * gy;
    for (int k = 0; k < 4; k++) {
        double d[NC];
        memcpy(d, delta, sizeof d);
        const prim_t *p = &wf[k];
        d[1] += half * p->sigma * gx;
        d[2] += half * p->sigma * gy;
        d[3] += half * p->sigma * ((p->u + 0.5 * half * gx) * gx + (p->v + 0.5 * half * gy) * gy);
        faces[k] = evolved(ph, p, d);
    }
}

double sro_hydro_dt(const sro_physics *ph, const double *planes, size_t w, size_t h, const double *gx,
                    const double *gy) {
    size_t n = w * h;
    double rate = 0.0, gmax = 0.0;
    for (size_t i = 0; i < n; i++) {
        prim_t p = read_cell(ph, planes, n, i);
        double c = sound(ph, &p);
        rate = fmax(rate, (fabs(p.u) + c) + (fabs(p.v) + c));
        if (gx) gmax = fmax(gmax, hypot(gx[i], gy[i]));
    }
    double dt = ph->cfl * ph->dx / rate;
    if (gmax > 0.0) dt = fmin(dt, ph->eta_g * sqrt(ph->dx / gmax));
    return dt;
}

/* gravity's source at one face (hydro.rs:1048-1067); d is (phi_low - phi_high) / dx */
static void face_work(double q, double sl, double sh, double pl, double ph_, double d, double dt, double out[3]) {
    double a = fmax(sl, 0.0), b = fmax(sh, 0.0);
    double w = a + b > 0.0 ? a / (a + b) : 0.5;
    double extra = dt * (q - 0.5 * (pl + ph_)) * d;
    double c = 0.0, kick = 0.0, el = 0.0, eh = 0.0;
    if (g_pair_work && extra < 0.0 && a > 0.0 && b > 0.0) {
        double ul = pl / a, uh = ph_ / b;
        double du = ul - uh;
        double inv = 1.0 / a + 1.0 / b;
        c = fmin(-extra, 0.25 * du * du / inv);
        if (c > 0.0) {
            double disc = sqrt(fmax(du * du - 2.0 * inv * c, 0.0));
            double sg = du > 0 ? 1.0 : (du < 0 ? -1.0 : (signbit(du) ? -1.0 : 1.0));
            kick = -2.0 * c / (du + sg * disc);
            el = ul * kick + 0.5 * kick * kick / a;
            eh = -uh * kick + 0.5 * kick * kick / b;
        }
    }
    double rest = extra + c;
    out[0] = kick;
    out[1] = 0.5 * dt * pl * d + w * rest + el;
    out[2] = 0.5 * dt * ph_ * d + (1.0 - w) * rest + eh;
}

/* kappa of a cell as the radiation force reads it (hydro.rs:371-387, heat.rs:216-219) */
static double kappa_cell(const sro_physics *ph, const double *x, double sigma, double t) {
    if (ph->opacity == SRO_OPACITY_KRAMERS) return ph->kappa0 * sigma * pow(t, -3.5);
    double dust = 0.0;
    if (ph->kappa_dust != 0.0 && t < ph->t_dust) {
        double zm;
        sro_composition(x, NULL, NULL, &zm);
        dust = ph->kappa_dust * zm;
    }
    return ph->kappa0 * (1.0 + x[0]) + dust;
}

static double push_share(const sro_physics *ph, const double *u, const double d[2]) {
    double d2 = d[0] * d[0] + d[1] * d[1];
    if (d2 == 0.0) return 1.0;
    double sigma = u[0];
    double x[NS];
    for (int s = 0; s < NS; s++) x[s] = u[4 + s] / sigma;
    double inv_mu = inv_mu_of(ph, x);
    double fa;
    double eps = thermal(ph, u[3], sigma, u[1], u[2], inv_mu, x, &fa);
    double raw = eps - fa / sigma;
    double a = sigma * (raw - eps_floor(ph, inv_mu));
    if (isnan(a) || a <= 0.0) return 0.0;
    double q = u[1] * d[0] + u[2] * d[1];
    double disc = sqrt(q * q + 2.0 * sigma * a * d2);
    double root = q >= 0.0 ? 2.0 * sigma * a / (q + disc) : (disc - q) / d2;
    return fmin(root, 1.0);
}

int sro_hydro_step(const sro_physics *ph, double *planes, size_t w, size_t h, double dt, const double *phi,
                   const double *gx, const double *gy, const double *rad_fx, const double *rad_fy, double *escaped,
                   double *sigma_mid_out) {
    const size_t G = SRO_GHOSTS;
    size_t n = w * h, pw = w + 2 * G, ph_ = h + 2 * G;
    prim_t *grid = (prim_t *)calloc(pw * ph_, sizeof(prim_t));
#define AT(x, y) (((y) + G) * pw + (x) + G)
    for (size_t y = 0; y < h; y++)
        for (size_t x = 0; x < w; x++) grid[AT(x, y)] = read_cell(ph, planes, n, y * w + x);
    if (ph->edge == SRO_EDGE_BOUNCE) {
        for (size_t y = 0; y < h; y++)
            for (size_t k = 1; k <= G; k++) {
                grid[(y + G) * pw + G - k] = mirrored(grid[AT(k - 1, y)], 0);
                grid[(y + G) * pw + G + w - 1 + k] = mirrored(grid[AT(w - k, y)], 0);
            }
        for (size_t px = 0; px < pw; px++)
            for (size_t k = 1; k <= G; k++) {
                grid[(G - k) * pw + px] = mirrored(grid[(G + k - 1) * pw + px], 1);
                grid[(G + h - 1 + k) * pw + px] = mirrored(grid[(G + h - k) * pw + px], 1);
            }
    } else {
        for (size_t y = 0; y < h; y++)
            for (size_t k = 1; k <= G; k++) {
                grid[(y + G) * pw + G - k] = copied(grid[AT(0, y)], 0, 0);
                grid[(y + G) * pw + G + w - 1 + k] = copied(grid[AT(w - 1, y)], 0, 1);
            }
        for (size_t px = 0; px < pw; px++)
            for (size_t k = 1; k <= G; k++) {
                grid[(G - k) * pw + px] = copied(grid[G * pw + px], 1, 0);
                grid[(G + h - 1 + k) * pw + px] = copied(grid[(G + h - 1) * pw + px], 1, 1);
            }
    }
    int grav = gx != NULL;
    double half = 0.5 * dt, dtdx = dt / ph->dx;
    face_t(*faces)[4] = (face_t(*)[4])malloc(sizeof(face_t) * 4 * n);
    double(*mids)[3] = (double(*)[3])malloc(sizeof(double) * 3 * n);
    /* gravity at a ghosted index: a ghost takes its edge cell's */
#pragma omp parallel for schedule(static)
    for (size_t y = 0; y < h; y++)
        for (size_t x = 0; x < w; x++) {
            size_t i = y * w + x;
            double g[2];
            if (grav) {
                g[0] = gx[i];
                g[1] = gy[i];
            }
            predicted(ph, grid, pw, AT(x, y), half, grav ? g : NULL, faces[i], mids[i]);
        }
    /* boundary flux of an edge cell on side s */
    double(*fx)[NC] = (double(*)[NC])calloc((w + 1) * h, sizeof(double[NC]));
    double(*fy)[NC] = (double(*)[NC])calloc(w * (h + 1), sizeof(double[NC]));
#pragma omp parallel for schedule(static)
    for (size_t y = 0; y < h; y++)
        for (size_t i = 0; i <= w; i++) {
            double *f = fx[y * (w + 1) + i];
            if (i == 0 || i == w) {
                int side = i == 0 ? 0 : 1;
                size_t cx = i == 0 ? 0 : w - 1;
                const face_t *own = &faces[y * w + cx][side];
                if (ph->edge == SRO_EDGE_BOUNCE) {
                    wall(own, 0, f);
                } else {
                    size_t cell = AT(cx, y), ghost = side == 0 ? cell - 1 : cell + 1;
                    size_t ex = cx;
                    double g[2];
                    if (grav) {
                        g[0] = gx[y * w + ex];
                        g[1] = gy[y * w + ex];
                    }
                    face_t gf[4];
                    double gm[3];
                    predicted(ph, grid, pw, ghost, half, grav ? g : NULL, gf, gm);
                    if (side == 1)
                        hll(ph, own, &gf[0], 0, f);
                    else
                        hll(ph, &gf[1], own, 0, f);
                    int inflow = side == 1 ? f[0] < 0.0 : f[0] > 0.0;
                    if (inflow) wall(own, 0, f);
                }
            } else {
                hll(ph, &faces[y * w + i - 1][1], &faces[y * w + i][0], 0, f);
            }
        }
#pragma omp parallel for schedule(static)
    for (size_t j = 0; j <= h; j++)
        for (size_t x = 0; x < w; x++) {
            double *f = fy[j * w + x];
            if (j == 0 || j == h) {
                int side = j == 0 ? 2 : 3;
                size_t cy = j == 0 ? 0 : h - 1;
                const face_t *own = &faces[cy * w + x][side];
                if (ph->edge == SRO_EDGE_BOUNCE) {
                    wall(own, 1, f);
                } else {
                    size_t cell = AT(x, cy), ghost = side == 2 ? cell - pw : cell + pw;
                    double g[2];
                    if (grav) {
                        g[0] = gx[cy * w + x];
                        g[1] = gy[cy * w + x];
                    }
                    face_t gf[4];
                    double gm[3];
                    predicted(ph, grid, pw, ghost, half, grav ? g : NULL, gf, gm);
                    if (side == 3)
                        hll(ph, own, &gf[2], 1, f);
                    else
                        hll(ph, &gf[3], own, 1, f);
                    int inflow = side == 3 ? f[0] < 0.0 : f[0] > 0.0;
                    if (inflow) wall(own, 1, f);
                }
            } else {
                hll(ph, &faces[(j - 1) * w + x][3], &faces[j * w + x][2], 1, f);
            }
        }
#define VAC(x, y) (planes[(y) * w + (x)] < ph->sigma_vac)
#define EMPTY(x, y)                                                                                                    \
    (VAC(x, y) && ((x) == 0 || VAC((x) - 1, y)) && ((x) + 1 == w || VAC((x) + 1, y)) && ((y) == 0 || VAC(x, (y) - 1)) && \
     ((y) + 1 == h || VAC(x, (y) + 1)))
    /* the half step's field (hydro.rs:821-856), without sinks: the given field is P2's own solve on Sigma^n, so
     * the "given - at_n" correction is exactly zero and the field is the solve on Sigma_mid alone */
    double *src = NULL, *fphi = NULL, *fgx = NULL, *fgy = NULL;
    if (grav) {
        src = (double *)malloc(sizeof(double) * n);
        for (size_t y = 0; y < h; y++)
            for (size_t x = 0; x < w; x++) {
                size_t i = y * w + x;
                double s;
                if (EMPTY(x, y)) {
                    s = 0.0;
                } else {
                    double start;
                    if (VAC(x, y)) {
                        double u[NC];
                        cons(ph, &grid[AT(x, y)], u);
                        start = u[0];
                    } else {
                        start = planes[i];
                    }
                    double div = (fx[y * (w + 1) + x + 1][0] - fx[y * (w + 1) + x][0]) + (fy[(y + 1) * w + x][0] - fy[y * w + x][0]);
                    s = 0.5 * (planes[i] + (start - dtdx * div));
                }
                src[i] = s >= ph->sigma_vac ? s : 0.0;
            }
        fphi = (double *)malloc(sizeof(double) * n);
        fgx = (double *)malloc(sizeof(double) * n);
        fgy = (double *)malloc(sizeof(double) * n);
        int e = sro_gravity_field(ph, src, w, h, fphi, fgx, fgy, NULL);
        if (e != SRO_OK) return e;
        if (g_bias[0] != 0.0 || g_bias[1] != 0.0)
            for (size_t i = 0; i < n; i++) {
                fgx[i] += g_bias[0];
                fgy[i] += g_bias[1];
            }
        if (sigma_mid_out)
            for (size_t i = 0; i < n; i++) sigma_mid_out[i] = src[i];
    }
    (void)phi;
    size_t nf = 2 * (w + h);
    double(*push)[2] = (double(*)[2])calloc(n, sizeof(double[2]));
    double kick_net[2] = {0, 0};
    long densest = -1;
    double densest_s = 0;
    double *next = (double *)malloc(sizeof(double) * NC * n);
    memcpy(next, planes, sizeof(double) * NC * n);
    for (size_t y = 0; y < h; y++)
        for (size_t x = 0; x < w; x++) {
            if (EMPTY(x, y)) continue;
            size_t i = y * w + x;
            double u[NC];
            if (VAC(x, y)) {
                cons(ph, &grid[AT(x, y)], u);
            } else {
                for (int k = 0; k < 4; k++) u[k] = planes[k * n + i];
                for (int s = 0; s < NS; s++) u[4 + s] = u[0] * planes[(4 + s) * n + i];
            }
            for (int k = 0; k < NC; k++) {
                double d = dtdx * ((fx[y * (w + 1) + x + 1][k] - fx[y * (w + 1) + x][k]) + (fy[(y + 1) * w + x][k] - fy[y * w + x][k]));
                u[k] -= d;
            }
            if (grav) {
                const double *m = mids[i];
                double kick[2] = {dt * src[i] * fgx[i], dt * src[i] * fgy[i]};
                kick_net[0] += dt * src[i] * fgx[i];
                kick_net[1] += dt * src[i] * fgy[i];
                if (src[i] > 0.0 && (densest < 0 || src[i] > densest_s)) {
                    densest = (long)i;
                    densest_s = src[i];
                }
                double p = fphi[i];
                double pr = p, pl = p, pd = p, pu = p;
                const double *mr = m, *ml = m, *md = m, *mu = m;
                if (x + 1 < w) { pr = fphi[i + 1]; mr = mids[i + 1]; }
                if (x > 0) { pl = fphi[i - 1]; ml = mids[i - 1]; }
                if (y + 1 < h) { pd = fphi[i + w]; md = mids[i + w]; }
                if (y > 0) { pu = fphi[i - w]; mu = mids[i - w]; }
                size_t kr = y * (w + 1) + x + 1, kl = y * (w + 1) + x, kd = (y + 1) * w + x, ku = y * w + x;
                double r[3], l[3], dn[3], up[3];
                double idx = 1.0 / ph->dx;
                face_work(fx[kr][0], m[0], mr[0], m[1], mr[1], (p - pr) * idx, dt, r);
                face_work(fx[kl][0], ml[0], m[0], ml[1], m[1], (pl - p) * idx, dt, l);
                face_work(fy[kd][0], m[0], md[0], m[2], md[2], (p - pd) * idx, dt, dn);
                face_work(fy[ku][0], mu[0], m[0], mu[2], m[2], (pu - p) * idx, dt, up);
                kick[0] += r[0] - l[0];
                kick[1] += dn[0] - up[0];
                u[1] += kick[0];
                u[2] += kick[1];
                u[3] += r[1] + l[2] + dn[1] + up[2];
            }
            if (rad_fx) {
                double sigma = planes[i];
                double a[2] = {0, 0};
                if (sigma >= ph->sigma_vac) {
                    double xs[NS];
                    for (int s = 0; s < NS; s++) xs[s] = planes[(4 + s) * n + i];
                    double inv_mu = inv_mu_of(ph, xs);
                    double t = 0.0;
                    if (ph->opacity == SRO_OPACITY_KRAMERS || (ph->kappa_dust != 0.0)) {
                        double eps = thermal(ph, planes[3 * n + i], sigma, planes[n + i], planes[2 * n + i], inv_mu,
                                             xs, NULL);
                        t = (ph->gamma - 1.0) * eps / inv_mu;
                    }
                    double kap = kappa_cell(ph, xs, sigma, t);
                    a[0] = kap * sigma * rad_fx[i] / ph->c_sb;
                    a[1] = kap * sigma * rad_fy[i] / ph->c_sb;
                }
                double d[2] = {dt * a[0], dt * a[1]};
                double s = push_share(ph, u, d);
                push[i][0] = s * d[0];
                push[i][1] = s * d[1];
                u[1] += push[i][0];
                u[2] += push[i][1];
            }
            if (escaped) {
                double phi_edge = grav ? fphi[i] : 0.0;
                struct {
                    int edge;
                    double sign;
                    const double *f;
                    size_t slot;
                } cr[4] = {{x == 0, -1.0, fx[y * (w + 1)], y},
                           {x + 1 == w, 1.0, fx[y * (w + 1) + w], h + y},
                           {y == 0, -1.0, fy[x], 2 * h + x},
                           {y + 1 == h, 1.0, fy[h * w + x], 2 * h + w + x}};
                for (int c = 0; c < 4; c++)
                    if (cr[c].edge) {
                        for (int k = 0; k < NC; k++) escaped[k * nf + cr[c].slot] += cr[c].sign * dtdx * cr[c].f[k];
                        escaped[3 * nf + cr[c].slot] += cr[c].sign * dtdx * cr[c].f[0] * phi_edge;
                    }
            }
            next[i] = u[0];
            next[n + i] = u[1];
            next[2 * n + i] = u[2];
            next[3 * n + i] = u[3];
            for (int s = 0; s < NS; s++) next[(4 + s) * n + i] = u[4 + s] / u[0];
        }
    /* the radiation force's net push taken back in proportion to each push (hydro.rs:986-996) */
    {
        double net0 = 0, net1 = 0, net2 = 0;
        for (size_t i = 0; i < n; i++) {
            net0 += push[i][0];
            net1 += push[i][1];
            net2 += sqrt(push[i][0] * push[i][0] + push[i][1] * push[i][1]);
        }
        if (net2 > 0.0)
            for (size_t i = 0; i < n; i++) {
                double m = sqrt(push[i][0] * push[i][0] + push[i][1] * push[i][1]);
                next[n + i] -= m * net0 / net2;
                next[2 * n + i] -= m * net1 / net2;
            }
    }
    /* gravity's net kick taken back from the densest free cell (hydro.rs:997-1001) */
    if (densest >= 0 && g_kick_takeback != 0) {
        next[n + densest] -= g_kick_takeback * kick_net[0];
        next[2 * n + densest] -= g_kick_takeback * kick_net[1];
    }
    memcpy(planes, next, sizeof(double) * NC * n);
    free(next);
    free(push);
    free(src);
    free(fphi);
    free(fgx);
    free(fgy);
    free(fx);
    free(fy);
    free(faces);
    free(mids);
    free(grid);
    return SRO_OK;
#undef AT
#undef VAC
#undef EMPTY
}
