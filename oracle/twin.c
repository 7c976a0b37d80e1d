/* Oracle (test infrastructure only): P8 floors and the twin's step driver in f64.
 * Restates crates/sr-physics/src/reference/floors.rs:45-175 (renormalised, Floors::cell, Floors::apply; Sigma_vac and
 * Sigma_floor read as the f32 the device holds, floors.rs:110; the cold energy of eos.rs in the vacuum reset and the
 * thermal floor) without the hold (floors.rs:200-260: no cell is held below Sigma_hold = 10, nor with the cold pressure
 * off), and crates/sr-physics/src/reference/mod.rs:129-297 (Twin::new, Twin::step: P1 dt, P2 gravity,
 * P4 hydro with the last P5 flux as the radiation force, P5 heat, P6 reactions at the step's dt, P8 floors, P9
 * dt_{n+1} with P2's g of the step's start state).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "sro.h"

#define NS SRO_NS
#define NC SRO_COMPONENTS

void sro_floors_apply(const sro_physics *ph, double *planes, size_t w, size_t h, double *vacuum_reset,
                      double *floor_added) {
    size_t n = w * h;
    double svac = (double)(float)ph->sigma_vac, sfloor = (double)(float)ph->sigma_floor;
    double vr = 0, fa = 0;
    for (size_t i = 0; i < n; i++) {
        double raw[NS], x[NS], sum = 0;
        int finite = 1;
        for (int s = 0; s < NS; s++) {
            raw[s] = planes[(4 + s) * n + i];
            if (!isfinite(raw[s])) finite = 0;
        }
        if (!finite) continue;
        for (int s = 0; s < NS; s++) {
            x[s] = fmax(raw[s], 0.0);
            sum += x[s];
        }
        if (sum > 0.0) {
            for (int s = 0; s < NS; s++) x[s] /= sum;
        } else {
            memset(x, 0, sizeof x);
            x[0] = 1.0;
        }
        double inv_mu, y_e;
        sro_composition(x, &inv_mu, &y_e, NULL);
        if (ph->mu > 0) inv_mu = 1.0 / ph->mu;
        double eps_fl = ph->t_floor * inv_mu / (ph->gamma - 1.0);
        double sigma = planes[i], E = planes[3 * n + i];
        if (sigma < svac) {
            double e = sfloor * (eps_fl + sro_eps_cold(ph, sfloor, y_e, x[NS - 1]));
            vr += sigma - sfloor;
            fa += e - E;
            planes[i] = sfloor;
            planes[n + i] = 0.0;
            planes[2 * n + i] = 0.0;
            planes[3 * n + i] = e;
        } else {
            double mx = planes[n + i], my = planes[2 * n + i];
            double rawe = E / sigma - 0.5 * (mx * mx + my * my) / (sigma * sigma) -
                          sro_eps_cold(ph, sigma, y_e, x[NS - 1]);
            if (rawe < eps_fl) {
                double add = sigma * (eps_fl - rawe);
                planes[3 * n + i] = E + add;
                fa += add;
            }
        }
        for (int s = 0; s < NS; s++) planes[(4 + s) * n + i] = x[s];
    }
    if (vacuum_reset) *vacuum_reset = vr;
    if (floor_added) *floor_added = fa;
}

struct sro_twin {
    sro_physics ph;
    size_t w, h;
    double *planes;
    double *gx, *gy, *phi;
    double *rfx, *rfy;
    double *escaped;
    int gravity, hydro, heat, have_flux;
    double dt, dt_next, sim_time;
    uint64_t steps;
    sro_heat_report last_heat;
    int react_on;
    sro_react_params rp;
    sro_react_pass last_react;
};

static void solve_gravity(sro_twin *t) {
    if (!t->gravity) return;
    sro_gravity_field(&t->ph, t->planes, t->w, t->h, t->phi, t->gx, t->gy, NULL);
}

sro_twin *sro_twin_new(const sro_physics *ph, size_t w, size_t h, const double *planes, int gravity, int hydro,
                       int heat) {
    sro_twin *t = (sro_twin *)calloc(1, sizeof *t);
    size_t n = w * h;
    t->ph = *ph;
    t->w = w;
    t->h = h;
    t->planes = (double *)malloc(sizeof(double) * NC * n);
    memcpy(t->planes, planes, sizeof(double) * NC * n);
    t->gx = (double *)calloc(n, sizeof(double));
    t->gy = (double *)calloc(n, sizeof(double));
    t->phi = (double *)calloc(n, sizeof(double));
    t->rfx = (double *)calloc(n, sizeof(double));
    t->rfy = (double *)calloc(n, sizeof(double));
    t->escaped = (double *)calloc(NC * 2 * (w + h), sizeof(double));
    t->gravity = gravity;
    t->hydro = hydro;
    t->heat = heat;
    solve_gravity(t);
    t->dt_next = sro_hydro_dt(&t->ph, t->planes, w, h, gravity ? t->gx : NULL, gravity ? t->gy : NULL);
    return t;
}

void sro_twin_free(sro_twin *t) {
    if (!t) return;
    free(t->planes);
    free(t->gx);
    free(t->gy);
    free(t->phi);
    free(t->rfx);
    free(t->rfy);
    free(t->escaped);
    free(t);
}

int sro_twin_step(sro_twin *t) {
    size_t w = t->w, h = t->h, n = w * h;
    double dt = t->dt_next;
    t->dt = dt;
    t->sim_time += dt;
    solve_gravity(t);
    if (t->hydro) {
        int e = sro_hydro_step(&t->ph, t->planes, w, h, dt, t->gravity ? t->phi : NULL, t->gravity ? t->gx : NULL,
                               t->gravity ? t->gy : NULL, t->have_flux ? t->rfx : NULL, t->have_flux ? t->rfy : NULL,
                               t->escaped, NULL);
        if (e != SRO_OK) return e;
    }
    if (t->heat) {
        double *fx = (double *)malloc(sizeof(double) * (w + 1) * h), *fy = (double *)malloc(sizeof(double) * w * (h + 1));
        sro_heat_step(&t->ph, t->planes, w, h, dt, fx, fy, NULL, &t->last_heat);
        sro_flux_cells(fx, fy, w, h, t->rfx, t->rfy);
        t->have_flux = 1;
        free(fx);
        free(fy);
    }
    if (t->react_on) { /* mod.rs:284-287 */
        int e = sro_react_step(&t->ph, &t->rp, t->planes, w, h, dt, NULL, NULL, NULL, &t->last_react);
        if (e != SRO_OK) return e;
    }
    sro_floors_apply(&t->ph, t->planes, w, h, NULL, NULL);
    t->dt_next = sro_hydro_dt(&t->ph, t->planes, w, h, t->gravity ? t->gx : NULL, t->gravity ? t->gy : NULL);
    t->steps++;
    (void)n;
    return SRO_OK;
}

void sro_twin_state(const sro_twin *t, double *planes) { memcpy(planes, t->planes, sizeof(double) * NC * t->w * t->h); }

void sro_twin_times(const sro_twin *t, double *dt, double *dt_next, double *sim_time, uint64_t *steps) {
    if (dt) *dt = t->dt;
    if (dt_next) *dt_next = t->dt_next;
    if (sim_time) *sim_time = t->sim_time;
    if (steps) *steps = t->steps;
}

void sro_twin_cap_dt_next(sro_twin *t, double cap) {
    if (cap > 0 && cap < t->dt_next) t->dt_next = cap;
}

void sro_twin_last_heat(const sro_twin *t, sro_heat_report *rep) { *rep = t->last_heat; }

/* P6 in the twin's step with these records (Twin::new's Burning, mod.rs:94-123); NULL switches it off */
void sro_twin_set_reactions(sro_twin *t, const sro_react_params *rp) {
    t->react_on = rp != NULL;
    if (rp) t->rp = *rp;
}

void sro_twin_last_react(const sro_twin *t, sro_react_pass *pass) { *pass = t->last_react; }

void sro_twin_escaped(const sro_twin *t, double *escaped) {
    memcpy(escaped, t->escaped, sizeof(double) * NC * 2 * (t->w + t->h));
}
