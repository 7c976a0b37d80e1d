// P6 burning on the GPU (crates/sr-physics/src/reference/react.rs; the rate law crates/sr-physics/src/rates.rs): one
// thread per cell in FP64, each cell burning alone over the pass's dt. N_Fe first in closed form, then the other
// enabled records sub-cycled by linearised backward Euler, (I - hJ) dX = h S, each sub-step sized by burn_dx on the
// fractions and burn_dt on T, checked, halved until it holds; the heat of the mass each sub-step moved into the energy
// and the record's group; the burn clock slowing a sub-step whose heating passes eps_th c_s / K_burn; the neutrino loss
// eps_nu h with the cooling on; with T pinned the energy follows the final composition at T0. A cell below Sigma_vac is
// not touched and a cell no record converts is left bit for bit. The state's fractions and energy are the world's f32
// planes, Sigma its FP64 plane; the cold energy is read from FP64 tables; every expression and sum order is the
// oracle's, so the cells differ from it only by the device's f64 elementary functions and the f32 planes.
#include <cfloat>
#include <cmath>
#include <cstring>

#include "common.hpp"
#include "react.hpp"

namespace sr {

#define NS REACT_NS
#define NR SR_REACT_RECORDS
#define NG SR_REACT_GROUPS
#define ID_N 9
#define X_N_STAR 0.5

__constant__ double r_z[NS] = {1, 2, 6, 8, 10, 12, 14, 16, 26, 0};
__constant__ double r_a[NS] = {1, 4, 12, 16, 20, 24, 28, 32, 56, 1};
// (Z + 1) / A and Z / A per species
__constant__ double r_zp1_a[NS] = {2.0, 3.0 / 4.0, 7.0 / 12.0, 9.0 / 16.0, 11.0 / 20.0, 13.0 / 24.0, 15.0 / 28.0,
                                   17.0 / 32.0, 27.0 / 56.0, 1.0};
__constant__ double r_z_a[NS] = {1.0, 0.5, 0.5, 0.5, 0.5, 0.5, 0.5, 0.5, 26.0 / 56.0, 0.0};

// ln of one Gamow-shaped term and f's direct form (rates.rs:44-53), on either side
__host__ __device__ inline double term_ln(double t_k, double nu, double w, double t) {
    const double u = log(t / t_k);
    return log(w) - 2.0 / 3.0 * u - (3.0 * nu + 2.0) * expm1(-u / 3.0);
}

__host__ __device__ inline double f_sum(const sr_react_record &r, double t) {
    double s = 0.0;
    for (int k = 0; k < r.n_terms; k++) s += exp(term_ln(r.t_k[k], r.nu[k], r.w[k], t));
    return s;
}

int react_check(const sr_react_params *p) {
    if (!p) return SR_OK;
    if (p->n_records < 0 || p->n_records > NR) return SR_ERR_ARGUMENT;
    if (p->nfe < -1 || p->nfe >= p->n_records) return SR_ERR_ARGUMENT;
    if (!(p->burn_dx > 0.0) || !(p->burn_dt > 0.0) || !(p->burn_k >= 0.0)) return SR_ERR_ARGUMENT;
    if (p->neutrinos && !(p->t_nu > 0.0)) return SR_ERR_ARGUMENT;
    for (int r = 0; r < p->n_records; r++) {
        const sr_react_record &c = p->rec[r];
        if (c.n_in < 1 || c.n_in > 4 || c.n_out < 0 || c.n_out > 4 || c.group < 0 || c.group >= NG) return SR_ERR_ARGUMENT;
        if (c.n_terms < 0 || c.n_terms > 4 || (c.n_terms > 0 && !(c.t_thr > 0.0))) return SR_ERR_ARGUMENT;
        for (int k = 0; k < c.n_terms; k++)
            if (!(c.t_k[k] > 0.0) || !(c.w[k] > 0.0)) return SR_ERR_ARGUMENT;
        for (int k = 0; k < c.n_in; k++)
            if (c.in_species[k] < 0 || c.in_species[k] >= NS || !(c.in_order[k] > 0.0)) return SR_ERR_ARGUMENT;
        for (int k = 0; k < c.n_out; k++)
            if (c.out_species[k] < 0 || c.out_species[k] >= NS) return SR_ERR_ARGUMENT;
        if (c.gated && !(c.sigma_g > 0.0)) return SR_ERR_ARGUMENT;
    }
    if (p->nfe >= 0) {  // the closed form's shape (react.rs:205-212)
        const sr_react_record &c = p->rec[p->nfe];
        if (c.n_terms != 0 || c.n_in != 1 || c.n_out != 1 || c.in_order[0] != 1.0 || !c.gated || c.gate_p != 1.0 ||
            c.out_species[0] != ID_N)
            return SR_ERR_ARGUMENT;
    }
    return SR_OK;
}

void react_build(const sr_react_params *p, const sr_physics &ph, ReactDev *out) {
    std::memset(out, 0, sizeof *out);
    out->P = *p;
    for (int r = 0; r < p->n_records; r++) {
        const sr_react_record &c = p->rec[r];
        for (int k = 0; k < c.n_out; k++) out->nu[r][c.out_species[k]] += c.out_share[k];
        for (int k = 0; k < c.n_in; k++) out->nu[r][c.in_species[k]] -= c.in_share[k];
        if (c.n_terms > 0)  // rates.rs:62-71: log-spaced nodes from T_thr over the span
            for (int i = 0; i < REACT_TABLE; i++)
                out->ln_f[r][i] =
                    std::log(f_sum(c, c.t_thr * std::exp(std::log(2.5) * (double)i / (double)(REACT_TABLE - 1))));
    }
    out->sigma_vac = ph.sigma_vac;
    out->t_floor = ph.t_floor;
    const double k[2][2] = {{ph.k1e, ph.k2e}, {ph.k1n, ph.k2n}};
    for (int q = 0; q < 2; q++) {
        out->cold_pair[q] = k[q][0] > 0.0 && k[q][1] > 0.0;
        if (out->cold_pair[q]) cold_table(k[q][0], k[q][1], out->cold_lnu[q]);
    }
    out->cold_ln_min = std::log(1e-8);
    out->cold_step = (std::log(1e8) - std::log(1e-8)) / (COLD_POINTS - 1);
}

// -------------------------------------------------------------------------------------------------- the device side

// 1/mu and Y_e (registry.rs:146-154): sums of X (Z + 1) / A and X Z / A in species order
__device__ inline void comp(const double *x, double *mu, double *y_e) {
    double im = 0.0, ye = 0.0;
#pragma unroll
    for (int s = 0; s < NS; s++) {
        im += x[s] * r_zp1_a[s];
        ye += x[s] * r_z_a[s];
    }
    *mu = 1.0 / im;
    *y_e = ye;
}

// u(x) of one pair from its FP64 table: linear in (ln x, ln u), the K1 x^2 power law below the first node
__device__ inline double cold_u_d(const ReactDev &D, const double *lnu, double x) {
    if (x <= 0.0) return 0.0;
    const double pos = (log(x) - D.cold_ln_min) / D.cold_step;
    if (pos < 0.0) return exp(lnu[0]) * (x / 1e-8) * (x / 1e-8);
    int i = (int)floor(pos);
    if (i > COLD_POINTS - 2) i = COLD_POINTS - 2;
    const double frac = pos - (double)i;
    return exp(lnu[i] + frac * (lnu[i + 1] - lnu[i]));
}

__device__ inline double eps_cold_d(const ReactDev &D, double sigma, double y_e, double x_n) {
    double u = 0.0;
    if (D.cold_pair[0]) u += cold_u_d(D, D.cold_lnu[0], y_e * sigma);
    if (D.cold_pair[1]) u += cold_u_d(D, D.cold_lnu[1], x_n * sigma);
    return u / sigma;
}

// x^n for the whole orders 0..3 by products, pow() otherwise
__device__ __forceinline__ double ipow(double x, double n) {
    if (n == 1.0) return x;
    if (n == 2.0) return x * x;
    if (n == 3.0) return x * x * x;
    if (n == 0.0) return 1.0;
    return pow(x, n);
}

// f(T) through the table (rates.rs:98-118): 0 below T_thr, the direct sum past the table's span
__device__ inline double temp_factor(const ReactDev &D, int r, double t) {
    const sr_react_record &c = D.P.rec[r];
    if (c.n_terms == 0) return 1.0;
    if (t < c.t_thr) return 0.0;
    const double step = log(2.5) / (double)(REACT_TABLE - 1);
    const double pos = log(t / c.t_thr) / step;
    if (pos >= (double)(REACT_TABLE - 1)) return f_sum(c, t);
    const int i = (int)floor(pos);
    const double frac = pos - (double)i;
    return exp(D.ln_f[r][i] + frac * (D.ln_f[r][i + 1] - D.ln_f[r][i]));
}

// the gate g(Y_e Sigma) (rates.rs:119-125)
__device__ inline double gate(const sr_react_record &c, double sigma, double y_e) {
    if (!c.gated) return 1.0;
    const double yes = y_e * sigma;
    return yes > c.sigma_g ? pow(yes / c.sigma_g - 1.0, c.gate_p) : 0.0;
}

// omega = A Sigma^a prod X^n f(T) g(Y_e Sigma) (rates.rs:121-131)
__device__ inline double omega(const ReactDev &D, int r, double sigma, const double *x, double t, double y_e) {
    const sr_react_record &c = D.P.rec[r];
    const double f = temp_factor(D, r, t);
    double orders = 1.0;
    for (int k = 0; k < c.n_in; k++) orders *= ipow(x[c.in_species[k]], c.in_order[k]);
    const double g = gate(c, sigma, y_e);
    return c.a_coef * ipow(sigma, c.a_sigma) * orders * f * g;
}

// N_Fe over dt in closed form (react.rs:470-513)
__device__ double nfe_exact(const ReactDev &D, double sigma, double *x, double dt) {
    const sr_react_record &c = D.P.rec[D.P.nfe];
    const int fe = c.in_species[0], n = c.out_species[0];
    const double y_fe = r_z_a[fe];
    const double b = c.sigma_g / sigma;
    double others[NS], mu, a_y;
    for (int s = 0; s < NS; s++) others[s] = x[s];
    others[fe] = 0.0;
    comp(others, &mu, &a_y);
    const double cc = b - a_y;
    const double v = y_fe * x[fe];
    const double u = v - cc;
    if (!(u > 0.0 && v > 0.0) || dt <= 0.0) return 0.0;
    const double k = c.a_coef * ipow(sigma, c.a_sigma) / b;
    const double w = cc >= 0.0 ? u : v, s = cc >= 0.0 ? cc : -cc;
    const double ks = k * s * dt;
    const double phi = ks > 0.0 ? -expm1(-ks) / s : k * dt;
    const double dy = u * v * phi / (1.0 + w * phi);
    double dm = fmin(dy / y_fe, x[fe]);
    const double left = w * exp(-ks) / (1.0 + w * phi);
    if (left <= 4.0 * DBL_EPSILON * b) {
        if (cc >= 0.0) {
            x[fe] -= dm;
            x[n] += dm;
            double nudge = DBL_EPSILON * x[fe];
            for (int it = 0; it < 16; it++) {
                double m2, y_e;
                comp(x, &m2, &y_e);
                if (y_e * sigma <= c.sigma_g || x[fe] == 0.0) break;
                const double more = fmin(fmax((y_e * sigma - c.sigma_g) / (sigma * y_fe), nudge), x[fe]);
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

// (I - h J) dX = h S, Gaussian elimination with partial pivoting, the last of equal pivots (react.rs:515-560)
__device__ void solve(const ReactDev &D, const double (*grad)[NS], const double *s, double h, double *dx) {
    double a[NS][NS + 1];
    const int nr = D.P.n_records;
    for (int i = 0; i < NS; i++) {
        for (int j = 0; j <= NS; j++) a[i][j] = 0.0;
        a[i][i] = 1.0;
        a[i][NS] = h * s[i];
        for (int r = 0; r < nr; r++) {
            if (D.nu[r][i] == 0.0) continue;
            for (int j = 0; j < NS; j++) a[i][j] -= h * D.nu[r][i] * grad[r][j];
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

struct CellOut {
    double x[NS], de, lost, booked[NG], pinned_e;
    int nfe, pinned, unfinished, changed;
    double substeps, retried, max_dx, max_dt_rel;
};

// one cell over dt (react.rs:272-449)
__device__ void burn(const ReactDev &D, double sigma, const double *x0, double e_tot0, double dt, CellOut &cell) {
    const sr_react_params &P = D.P;
    const int nr = P.n_records, pin = P.pin_temperature;
    double x[NS];
    for (int s = 0; s < NS; s++) x[s] = x0[s];
    double e_tot = e_tot0;
    cell.de = cell.lost = cell.pinned_e = cell.substeps = cell.retried = cell.max_dx = cell.max_dt_rel = 0.0;
    for (int g = 0; g < NG; g++) cell.booked[g] = 0.0;
    cell.nfe = cell.pinned = cell.unfinished = cell.changed = 0;
    double mu0, y_e0;
    comp(x, &mu0, &y_e0);
    const double eps_th0 = fmax(e_tot0 - eps_cold_d(D, sigma, y_e0, x0[ID_N]), D.t_floor / mu0);
    const double t0 = mu0 * eps_th0;
    if (P.nfe >= 0) {
        const double w_nfe = omega(D, P.nfe, sigma, x, t0, y_e0);
        cell.nfe = w_nfe > 0.0;
        if (cell.nfe) {
            const double dm = nfe_exact(D, sigma, x, dt);
            const double heat = P.rec[P.nfe].q * dm;
            if (!pin) e_tot += heat;
            cell.de += heat;
            cell.booked[P.rec[P.nfe].group] += heat;
            cell.changed |= dm > 0.0;
        }
    }
    double t_done = 0.0, h_try = HUGE_VAL;
    unsigned long long attempts = 0;
    while (t_done < dt) {
        double t_rem = dt - t_done;
        double mu, y_e;
        comp(x, &mu, &y_e);
        const double t = pin ? t0 : mu * fmax(e_tot - eps_cold_d(D, sigma, y_e, x[ID_N]), D.t_floor / mu);
        double omegas[NR], grad[NR][NS], s[NS], qdot = 0.0;
        for (int i = 0; i < NS; i++) s[i] = 0.0;
        double enu = P.neutrinos && t >= P.t_nu ? P.a_nu * pow(t / P.t_nu, P.m_nu) : 0.0;
        int any = 0;
        for (int r = 0; r < nr; r++) {
            omegas[r] = 0.0;
            for (int j = 0; j < NS; j++) grad[r][j] = 0.0;
            if (r == P.nfe) continue;
            const sr_react_record &c = P.rec[r];
            // omega at X and, for each input j, at X with X_j = 1 (whose factor 1^n is exactly 1): f, g and A Sigma^a
            // once, each input's power once, the products in omega()'s order
            const double f = temp_factor(D, r, t), g = gate(c, sigma, y_e);
            const double ap = c.a_coef * ipow(sigma, c.a_sigma);
            const int n_in = c.n_in;
            double pw[4];
#pragma unroll
            for (int k = 0; k < 4; k++) pw[k] = k < n_in ? ipow(x[c.in_species[k]], c.in_order[k]) : 1.0;
            double orders = 1.0;
#pragma unroll
            for (int k = 0; k < 4; k++)
                if (k < n_in) orders *= pw[k];
            const double w = ap * orders * f * g;
            omegas[r] = w;
            cell.nfe |= c.latch && w > 0.0;
#pragma unroll
            for (int k = 0; k < 4; k++) {
                if (k >= n_in) break;
                const int j = c.in_species[k];
                const double n = c.in_order[k];
                if (n == 0.0) continue;
                double oj = 1.0;
#pragma unroll
                for (int q = 0; q < 4; q++)
                    if (q < n_in) oj *= c.in_species[q] == j ? 1.0 : pw[q];
                grad[r][j] += n * ipow(x[j], n - 1.0) * (ap * oj * f * g);
            }
            for (int i = 0; i < NS; i++) s[i] += D.nu[r][i] * w;
            qdot += c.q * w;
            any |= w != 0.0;
        }
        if (enu == 0.0 && !any) break;
        double k = 1.0;  // the burn clock (react.rs:451-462)
        if (!pin && P.burn_k > 0.0) {
            const double eps_th = t / mu;
            const double q_max = eps_th * sqrt(2.0 * eps_th) / P.burn_k;
            if (qdot > q_max) k = qdot / q_max;
        }
        if (k != 1.0) {
            t_rem = t_rem / k;
            enu = enu * k;
        }
        double s_max = 0.0;
        for (int i = 0; i < NS; i++) s_max = fmax(s_max, fabs(s[i]));
        double h = t_rem;
        if (s_max > 0.0) h = fmin(h, P.burn_dx / s_max);
        if (!pin) {
            double dinv_mu = 0.0;
            for (int i = 0; i < NS; i++) dinv_mu += s[i] * r_zp1_a[i];
            const double dt_dt = mu * (qdot - enu) - t * mu * dinv_mu;
            if (dt_dt != 0.0) h = fmin(h, P.burn_dt * t / fabs(dt_dt));
        }
        h = fmin(h, h_try);
        for (;;) {
            attempts++;
            if (attempts > 1000000ull) {
                cell.unfinished = 1;
                break;
            }
            const int last = h >= t_rem;
            const double h_use = fmin(h, t_rem);
            double dx[NS], x_new[NS];
            solve(D, grad, s, h_use, dx);
            double worst = 0.0;
            int negative = 0;
            for (int i = 0; i < NS; i++) {
                x_new[i] = x[i] + dx[i];
                worst = fmax(worst, fabs(dx[i]));
                if (x_new[i] < -1e-12) negative = 1;
            }
            double de = 0.0, by_group[NG];
            for (int g = 0; g < NG; g++) by_group[g] = 0.0;
            for (int r = 0; r < nr; r++) {
                double lin = 0.0;
                for (int j = 0; j < NS; j++) lin += grad[r][j] * dx[j];
                const double xi = h_use * (omegas[r] + lin);
                de += P.rec[r].q * xi;
                by_group[P.rec[r].group] += P.rec[r].q * xi;
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
                t_new = mu_new * fmax(e_new - eps_cold_d(D, sigma, y_e_new, x_new[ID_N]), D.t_floor / mu_new);
            }
            const double dt_rel = pin ? 0.0 : fabs(t_new - t) / t;
            const int bad = worst > P.burn_dx * (1.0 + 1e-12) || dt_rel > P.burn_dt * (1.0 + 1e-12) || negative;
            if (bad && h_use > 1e-14 * dt) {
                h = 0.5 * h_use;
                cell.retried += 1.0;
                continue;
            }
            for (int i = 0; i < NS; i++) x[i] = x_new[i];
            e_tot = e_new;
            cell.de += de - lost;
            cell.lost += lost;
            for (int g = 0; g < NG; g++) cell.booked[g] += by_group[g];
            cell.substeps += 1.0;
            cell.changed = 1;
            cell.max_dx = fmax(cell.max_dx, worst);
            cell.max_dt_rel = fmax(cell.max_dt_rel, dt_rel);
            t_done = last ? dt : (k == 1.0 ? t_done + h_use : t_done + h_use * k);
            h_try = fmax(2.0 * h_use, DBL_MIN);
            break;
        }
        if (cell.unfinished) break;
    }
    for (int s = 0; s < NS; s++) cell.x[s] = x[s];
    if (pin && cell.changed) {
        double mu, y_e;
        comp(x, &mu, &y_e);
        cell.pinned = 1;
        cell.pinned_e = t0 / mu + eps_cold_d(D, sigma, y_e, x[ID_N]);
    }
}

__global__ void __launch_bounds__(REACT_NT) k_react(const ReactDev *__restrict__ Dp, HydroState st, int n, double dt_arg,
                                                   const double *__restrict__ scal, double *__restrict__ nuclear,
                                                   double *__restrict__ lost, double *__restrict__ snu, double *psum,
                                                   double *pmax) {
    const ReactDev &D = *Dp;
    const double dt = dt_arg >= 0.0 ? dt_arg : scal[0];
    double vs[REACT_SUMS], vm[REACT_MAXES];
    for (int k = 0; k < REACT_SUMS; k++) vs[k] = 0.0;
    for (int k = 0; k < REACT_MAXES; k++) vm[k] = 0.0;
    for (int i = blockIdx.x * REACT_NT + threadIdx.x; i < n; i += gridDim.x * REACT_NT) {
        for (int g = 0; g < NG; g++) nuclear[(size_t)g * n + i] = 0.0;
        lost[i] = 0.0;
        snu[i] = 0.0;
        const double sigma = st.sigma[i];
        if (sigma < D.sigma_vac) continue;
        double x[NS];
        for (int s = 0; s < NS; s++) x[s] = (double)st.x[(size_t)s * n + i];
        const double mx = st.mx[i], my = st.my[i], e = st.e[i];
        const double kinetic = 0.5 * (mx * mx + my * my) / (sigma * sigma);
        const double e_tot = e / sigma - kinetic;
        CellOut c;
        burn(D, sigma, x, e_tot, dt, c);
        vs[0] += c.substeps;
        vs[1] += c.retried;
        vs[2] += (double)c.unfinished;
        vm[0] = fmax(vm[0], c.substeps);
        vm[1] = fmax(vm[1], c.max_dx);
        vm[2] = fmax(vm[2], c.max_dt_rel);
        vm[3] = fmax(vm[3], (double)c.nfe);
        if (!c.changed) continue;
        for (int s = 0; s < NS; s++) st.x[(size_t)s * n + i] = (float)c.x[s];
        st.e[i] = (float)(c.pinned ? sigma * (c.pinned_e + kinetic) : e + sigma * c.de);
        for (int g = 0; g < NG; g++) {
            const double b = sigma * c.booked[g];
            nuclear[(size_t)g * n + i] = b;
            vs[3 + g] += b;
        }
        lost[i] = sigma * c.lost;
        vs[3 + NG] += sigma * c.lost;
        if (c.x[ID_N] >= X_N_STAR && dt > 0.0) snu[i] = sigma * c.lost / dt;
    }
    block_partials<REACT_SUMS>(vs, psum, gridDim.x);
    block_partials<REACT_MAXES, 1>(vm, pmax, gridDim.x);
}

int react_launch(const ReactDev *D, const HydroState &st, int n, double dt, const double *scal, double *nuclear,
                 double *lost, double *snu, double *psum, double *pmax, int *nblocks, hipStream_t s) {
    const int nb = std::max(1, std::min(REACT_BLOCKS, (n + REACT_NT - 1) / REACT_NT));
    hipLaunchKernelGGL(k_react, dim3(nb), dim3(REACT_NT), 0, s, D, st, n, dt, scal, nuclear, lost, snu, psum, pmax);
    if (nblocks) *nblocks = nb;
    SR_HIP(hipGetLastError());
    return SR_OK;
}

} // namespace sr
