/* P6 burning: the oracle's f64 restatement of the reference twin (crates/sr-physics/src/reference/react.rs) and its
 * rate law (crates/sr-physics/src/rates.rs). Test infrastructure: only the tests, the smoke check and the bench's CPU
 * baseline call it.
 *
 * Per cell (react.rs:1-60): N_Fe first in closed form over the pass's dt, then the other enabled records sub-cycled by
 * linearised backward Euler, (I - hJ) dX = h S with S = sum_r nu_r omega_r and J its Jacobian in X (T, Sigma and Y_e
 * held), each sub-step sized by burn_dx on the fractions and burn_dt on T, checked and halved until it holds; the heat
 * of the mass each sub-step moved enters the energy and the record's group; the burn clock slows a sub-step whose
 * heating passes eps_th c_s / K_burn; the neutrino loss eps_nu h leaves the energy when the cooling is on; with T
 * pinned the energy follows the final composition at T0. A cell below Sigma_vac is not touched; a cell no record
 * converts is left bit for bit. */
#include <float.h>
#include <math.h>
#include <string.h>

#include "sro.h"

#define NS SRO_NS
#define NR SRO_REACT_RECORDS
#define ID_N 9                   /* react.rs:63 */
#define MAX_ATTEMPTS 1000000ULL  /* react.rs:72 */
#define MIN_H 1e-14              /* react.rs:74 */
#define LAND_ULPS 4.0            /* react.rs:76 */
#define X_N_STAR 0.5             /* react.rs:66 */

/* Z and A per species, assets/elements.json order H He C O Ne Mg Si S Fe n */
static const double SP_Z[NS] = {1, 2, 6, 8, 10, 12, 14, 16, 26, 0};
static const double SP_A[NS] = {1, 4, 12, 16, 20, 24, 28, 32, 56, 1};

/* ---------------------------------------------------------------------------------------------- the rate law */

/* ln of one Gamow-shaped term, ln w - (2/3) u - tau (e^{-u/3} - 1), u = ln(T/T_k), tau = 3 nu + 2 (rates.rs:44-48) */
static double term_ln(double t_k, double nu, double w, double t) {
    const double u = log(t / t_k);
    return log(w) - 2.0 / 3.0 * u - (3.0 * nu + 2.0) * expm1(-u / 3.0);
}

/* f(T) by the direct form, without the threshold (rates.rs:51-53) */
static double f_sum(const sro_react_record *r, double t) {
    double s = 0.0;
    for (int k = 0; k < r->n_terms; k++) s += exp(term_ln(r->t_k[k], r->nu[k], r->w[k], t));
    return s;
}

static double node_t(double t_thr, int i) { /* rates.rs:69-71 */
    return t_thr * exp(log(SRO_REACT_SPAN) * (double)i / (double)(SRO_REACT_TABLE - 1));
}

void sro_react_table(const sro_react_record *r, double *ln_f) { /* rates.rs:62-66 */
    if (r->n_terms == 0) return;
    for (int i = 0; i < SRO_REACT_TABLE; i++) ln_f[i] = log(f_sum(r, node_t(r->t_thr, i)));
}

/* f through the table: 0 below T_thr, log-log inside the span, the direct form above (rates.rs:98-111) */
static double f_of(const sro_react_record *r, const double *ln_f, double t, int table) {
    if (r->n_terms == 0) return 1.0;
    if (t < r->t_thr) return 0.0;
    if (!table) return f_sum(r, t);
    const double step = log(SRO_REACT_SPAN) / (double)(SRO_REACT_TABLE - 1);
    const double pos = log(t / r->t_thr) / step;
    if (pos >= (double)(SRO_REACT_TABLE - 1)) return f_sum(r, t);
    const int i = (int)floor(pos);
    const double frac = pos - (double)i;
    return exp(ln_f[i] + frac * (ln_f[i + 1] - ln_f[i]));
}

static double g_of(const sro_react_record *r, double y_e_sigma) { /* rates.rs:113-119 */
    if (!r->gated) return 1.0;
    return y_e_sigma > r->sigma_g ? pow(y_e_sigma / r->sigma_g - 1.0, r->gate_p) : 0.0;
}

double sro_react_omega(const sro_react_record *r, const double *ln_f, double sigma, const double *x, double t,
                       double y_e, int table) { /* rates.rs:121-131 */
    double orders = 1.0;
    for (int k = 0; k < r->n_in; k++) orders *= pow(x[r->in_species[k]], r->in_order[k]);
    return r->a_coef * pow(sigma, r->a_sigma) * orders * f_of(r, ln_f, t, table) * g_of(r, y_e * sigma);
}

/* ------------------------------------------------------------------------------------------------- the pass */

typedef struct {
    const sro_physics *ph;
    const sro_react_params *rp;
    double nu[NR][NS];                   /* outputs' shares minus inputs' (react.rs:188-195) */
    double ln_f[NR][SRO_REACT_TABLE];
} Ctx;

static void comp(const double *x, double *mu, double *y_e) {
    double inv_mu, ye, z_met;
    sro_composition(x, &inv_mu, &ye, &z_met);
    *mu = 1.0 / inv_mu;
    *y_e = ye;
}

static double omega(const Ctx *c, int r, double sigma, const double *x, double t, double y_e) {
    return sro_react_omega(&c->rp->rec[r], c->ln_f[r], sigma, x, t, y_e, 1);
}

/* eps_nu = A_nu (T / T_nu)^m for T >= T_nu with the cooling on, else 0 (react.rs:256-263) */
static double eps_nu(const Ctx *c, double t) {
    const sro_react_params *p = c->rp;
    return p->neutrinos && t >= p->t_nu ? p->a_nu * pow(t / p->t_nu, p->m_nu) : 0.0;
}

/* the burn clock's ratio k = qdot / q_max past q_max = eps_th sqrt(2 eps_th) / K_burn, else 1 (react.rs:451-462) */
static double stretch(const Ctx *c, double qdot, double t, double mu) {
    if (c->rp->pin_temperature || c->rp->burn_k <= 0.0) return 1.0;
    const double eps_th = t / mu;
    const double q_max = eps_th * sqrt(2.0 * eps_th) / c->rp->burn_k;
    return qdot > q_max ? qdot / q_max : 1.0;
}

/* N_Fe over dt in closed form, Sigma fixed: the mass moved from iron to neutrons (react.rs:470-513) */
static double nfe_exact(const Ctx *c, double sigma, double *x, double dt) {
    const sro_react_record *r = &c->rp->rec[c->rp->nfe];
    const int fe = r->in_species[0], n = r->out_species[0];
    const double y_fe = SP_Z[fe] / SP_A[fe];
    const double b = r->sigma_g / sigma;
    double others[NS], mu, a_y;
    memcpy(others, x, sizeof others);
    others[fe] = 0.0;
    comp(others, &mu, &a_y);
    const double cc = b - a_y;
    const double v = y_fe * x[fe];
    const double u = v - cc;
    if (!(u > 0.0 && v > 0.0) || dt <= 0.0) return 0.0;
    const double k = r->a_coef * pow(sigma, r->a_sigma) / b;
    const double w = cc >= 0.0 ? u : v, s = cc >= 0.0 ? cc : -cc;
    const double ks = k * s * dt;
    const double phi = ks > 0.0 ? -expm1(-ks) / s : k * dt;
    const double dy = u * v * phi / (1.0 + w * phi);
    double dm = fmin(dy / y_fe, x[fe]);
    const double left = w * exp(-ks) / (1.0 + w * phi);
    if (left <= LAND_ULPS * DBL_EPSILON * b) {
        if (cc >= 0.0) {
            x[fe] -= dm;
            x[n] += dm;
            double nudge = DBL_EPSILON * x[fe];
            for (int it = 0; it < 16; it++) {
                double m2, y_e;
                comp(x, &m2, &y_e);
                if (y_e * sigma <= r->sigma_g || x[fe] == 0.0) break;
                const double more = fmin(fmax((y_e * sigma - r->sigma_g) / (sigma * y_fe), nudge), x[fe]);
                x[fe] -= more;
                x[n] += more;
                dm += more;
                nudge *= 2.0;
            }
            return dm;
        }
        dm = x[fe];
    }
    x[fe] -= dm;
    x[n] += dm;
    return dm;
}

/* (I - h J) dX = h S by Gaussian elimination with partial pivoting, the last of equal pivots (react.rs:515-560) */
static void solve(const Ctx *c, const double grad[NR][NS], const double *s, double h, double *dx) {
    double a[NS][NS + 1];
    const int nr = c->rp->n_records;
    for (int i = 0; i < NS; i++) {
        for (int j = 0; j <= NS; j++) a[i][j] = 0.0;
        a[i][i] = 1.0;
        a[i][NS] = h * s[i];
        for (int r = 0; r < nr; r++) {
            if (c->nu[r][i] == 0.0) continue;
            for (int j = 0; j < NS; j++) a[i][j] -= h * c->nu[r][i] * grad[r][j];
        }
    }
    for (int col = 0; col < NS; col++) {
        int piv = col;
        for (int p = col + 1; p < NS; p++)
            if (fabs(a[p][col]) >= fabs(a[piv][col])) piv = p;
        if (piv != col)
            for (int j = 0; j <= NS; j++) {
                const double t = a[col][j];
                a[col][j] = a[piv][j];
                a[piv][j] = t;
            }
        const double d = a[col][col];
        if (d == 0.0) continue;
        for (int row = col + 1; row < NS; row++) {
            const double f = a[row][col] / d;
            if (f != 0.0)
                for (int j = col; j <= NS; j++) a[row][j] -= f * a[col][j];
        }
    }
    for (int i = NS - 1; i >= 0; i--) {
        double v = a[i][NS];
        for (int j = i + 1; j < NS; j++) v -= a[i][j] * dx[j];
        dx[i] = a[i][i] == 0.0 ? 0.0 : v / a[i][i];
    }
}

typedef struct {
    double x[NS], de, lost, booked[SRO_REACT_GROUPS], pinned_e;
    int nfe, pinned, unfinished, changed;
    uint64_t substeps, retried;
    double max_dx, max_dt_rel;
} Cell;

/* one cell over dt: x its fractions, e_tot0 its thermal plus cold energy per unit mass (react.rs:272-449) */
static void burn(const Ctx *c, double sigma, const double *x0, double e_tot0, double dt, Cell *cell) {
    const sro_physics *ph = c->ph;
    const sro_react_params *rp = c->rp;
    const int nr = rp->n_records, pin = rp->pin_temperature;
    double x[NS];
    memcpy(x, x0, sizeof x);
    double e_tot = e_tot0;
    memset(cell, 0, sizeof *cell);
    double mu0, y_e0;
    comp(x, &mu0, &y_e0);
    const double eps_th0 = fmax(e_tot0 - sro_eps_cold(ph, sigma, y_e0, x0[ID_N]), ph->t_floor / mu0);
    const double t0 = mu0 * eps_th0;
    if (rp->nfe >= 0) {
        const double w_nfe = omega(c, rp->nfe, sigma, x, t0, y_e0);
        cell->nfe = w_nfe > 0.0;
        if (cell->nfe) {
            const double dm = nfe_exact(c, sigma, x, dt);
            const double heat = rp->rec[rp->nfe].q * dm;
            if (!pin) e_tot += heat;
            cell->de += heat;
            cell->booked[rp->rec[rp->nfe].group] += heat;
            cell->changed |= dm > 0.0;
        }
    }
    double t_done = 0.0, h_try = INFINITY;
    uint64_t attempts = 0;
    while (t_done < dt) {
        double t_rem = dt - t_done;
        double mu, y_e;
        comp(x, &mu, &y_e);
        const double t =
            pin ? t0 : mu * fmax(e_tot - sro_eps_cold(ph, sigma, y_e, x[ID_N]), ph->t_floor / mu);
        double omegas[NR] = {0}, grad[NR][NS], s[NS] = {0}, qdot = 0.0;
        memset(grad, 0, sizeof grad);
        double enu = eps_nu(c, t);
        int any = 0;
        for (int r = 0; r < nr; r++) {
            if (r == rp->nfe) continue;
            const sro_react_record *rec = &rp->rec[r];
            const double w = omega(c, r, sigma, x, t, y_e);
            omegas[r] = w;
            cell->nfe |= rec->latch && w > 0.0;
            for (int k = 0; k < rec->n_in; k++) {
                const int j = rec->in_species[k];
                const double n = rec->in_order[k];
                if (n == 0.0) continue;
                double xj[NS];
                memcpy(xj, x, sizeof xj);
                xj[j] = 1.0;
                grad[r][j] += n * pow(x[j], n - 1.0) * omega(c, r, sigma, xj, t, y_e);
            }
            for (int i = 0; i < NS; i++) s[i] += c->nu[r][i] * w;
            qdot += rec->q * w;
            any |= w != 0.0;
        }
        if (enu == 0.0 && !any) break;
        const double k = stretch(c, qdot, t, mu);
        if (k != 1.0) {
            t_rem = t_rem / k;
            enu = enu * k;
        }
        double s_max = 0.0;
        for (int i = 0; i < NS; i++) s_max = fmax(s_max, fabs(s[i]));
        double h = t_rem;
        if (s_max > 0.0) h = fmin(h, rp->burn_dx / s_max);
        if (!pin) {
            double dinv_mu = 0.0;
            for (int i = 0; i < NS; i++) dinv_mu += s[i] * ((SP_Z[i] + 1.0) / SP_A[i]);
            const double dt_dt = mu * (qdot - enu) - t * mu * dinv_mu;
            if (dt_dt != 0.0) h = fmin(h, rp->burn_dt * t / fabs(dt_dt));
        }
        h = fmin(h, h_try);
        for (;;) {
            attempts++;
            if (attempts > MAX_ATTEMPTS) {
                cell->unfinished = 1;
                break;
            }
            const int last = h >= t_rem;
            const double h_use = fmin(h, t_rem);
            double dx[NS], x_new[NS];
            solve(c, (const double(*)[NS])grad, s, h_use, dx);
            double worst = 0.0;
            int negative = 0;
            for (int i = 0; i < NS; i++) {
                x_new[i] = x[i] + dx[i];
                worst = fmax(worst, fabs(dx[i]));
                if (x_new[i] < -1e-12) negative = 1;
            }
            double de = 0.0, by_group[SRO_REACT_GROUPS] = {0};
            for (int r = 0; r < nr; r++) {
                double lin = 0.0;
                for (int j = 0; j < NS; j++) lin += grad[r][j] * dx[j];
                const double xi = h_use * (omegas[r] + lin);
                de += rp->rec[r].q * xi;
                by_group[rp->rec[r].group] += rp->rec[r].q * xi;
            }
            const double lost = enu * h_use;
            double sum = 0.0;
            for (int i = 0; i < NS; i++) {
                x_new[i] = fmax(x_new[i], 0.0);
                sum += x_new[i];
            }
            for (int i = 0; i < NS; i++) x_new[i] /= sum;
            double mu_new, y_e_new, e_new, t_new;
            comp(x_new, &mu_new, &y_e_new);
            if (pin) {
                e_new = e_tot;
                t_new = t0;
            } else {
                e_new = e_tot + de - lost;
                t_new = mu_new * fmax(e_new - sro_eps_cold(ph, sigma, y_e_new, x_new[ID_N]), ph->t_floor / mu_new);
            }
            const double dt_rel = pin ? 0.0 : fabs(t_new - t) / t;
            const int bad = worst > rp->burn_dx * (1.0 + 1e-12) || dt_rel > rp->burn_dt * (1.0 + 1e-12) || negative;
            if (bad && h_use > MIN_H * dt) {
                h = 0.5 * h_use;
                cell->retried++;
                continue;
            }
            memcpy(x, x_new, sizeof x);
            e_tot = e_new;
            cell->de += de - lost;
            cell->lost += lost;
            for (int g = 0; g < SRO_REACT_GROUPS; g++) cell->booked[g] += by_group[g];
            cell->substeps++;
            cell->changed = 1;
            cell->max_dx = fmax(cell->max_dx, worst);
            cell->max_dt_rel = fmax(cell->max_dt_rel, dt_rel);
            t_done = last ? dt : (k == 1.0 ? t_done + h_use : t_done + h_use * k);
            h_try = fmax(2.0 * h_use, DBL_MIN);
            break;
        }
        if (cell->unfinished) break;
    }
    memcpy(cell->x, x, sizeof x);
    if (pin && cell->changed) {
        double mu, y_e;
        comp(x, &mu, &y_e);
        cell->pinned = 1;
        cell->pinned_e = t0 / mu + sro_eps_cold(ph, sigma, y_e, x[ID_N]);
    }
}

int sro_react_step(const sro_physics *ph, const sro_react_params *rp, double *planes, size_t w, size_t h, double dt,
                   double *nuclear, double *lost, double *s_nu, sro_react_pass *pass) {
    if (!ph || !rp || !planes || rp->n_records < 0 || rp->n_records > NR) return SRO_ERR_ARG;
    static Ctx c; /* large tables: one pass at a time (the oracle is single-threaded test code) */
    memset(&c, 0, sizeof c);
    c.ph = ph;
    c.rp = rp;
    for (int r = 0; r < rp->n_records; r++) {
        const sro_react_record *rec = &rp->rec[r];
        for (int k = 0; k < rec->n_out; k++) c.nu[r][rec->out_species[k]] += rec->out_share[k];
        for (int k = 0; k < rec->n_in; k++) c.nu[r][rec->in_species[k]] -= rec->in_share[k];
        sro_react_table(rec, c.ln_f[r]);
    }
    const size_t n = w * h;
    sro_react_pass ps;
    memset(&ps, 0, sizeof ps);
    if (nuclear) memset(nuclear, 0, sizeof(double) * SRO_REACT_GROUPS * n);
    if (lost) memset(lost, 0, sizeof(double) * n);
    if (s_nu) memset(s_nu, 0, sizeof(double) * n);
    for (size_t i = 0; i < n; i++) {
        const double sigma = planes[i];
        if (sigma < ph->sigma_vac) continue;
        double x[NS];
        for (int s = 0; s < NS; s++) x[s] = planes[(4 + s) * n + i];
        const double mx = planes[n + i], my = planes[2 * n + i];
        const double kinetic = 0.5 * (mx * mx + my * my) / (sigma * sigma);
        const double e_tot = planes[3 * n + i] / sigma - kinetic;
        Cell cell;
        burn(&c, sigma, x, e_tot, dt, &cell);
        ps.substeps += cell.substeps;
        ps.retried += cell.retried;
        if (cell.substeps > ps.max_substeps) ps.max_substeps = cell.substeps;
        ps.max_dx = fmax(ps.max_dx, cell.max_dx);
        ps.max_dt_rel = fmax(ps.max_dt_rel, cell.max_dt_rel);
        ps.unfinished += (uint64_t)cell.unfinished;
        ps.latch |= cell.nfe;
        if (!cell.changed) continue;
        for (int s = 0; s < NS; s++) planes[(4 + s) * n + i] = cell.x[s];
        planes[3 * n + i] = cell.pinned ? sigma * (cell.pinned_e + kinetic) : planes[3 * n + i] + sigma * cell.de;
        for (int g = 0; g < SRO_REACT_GROUPS; g++) {
            if (nuclear) nuclear[g * n + i] = sigma * cell.booked[g];
            ps.nuclear[g] += sigma * cell.booked[g];
        }
        if (lost) lost[i] = sigma * cell.lost;
        ps.neutrino_lost += sigma * cell.lost;
        if (s_nu && cell.x[ID_N] >= X_N_STAR && dt > 0.0) s_nu[i] = sigma * cell.lost / dt;
    }
    if (pass) *pass = ps;
    return SRO_OK;
}
