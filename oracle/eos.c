/* Oracle (test infrastructure only): the cold pressure of the equation of state, restating
 * crates/sr-physics/src/eos.rs:24-140, :239-262.
 *
 * P(x; K1, K2) = 1 / (1 / (K1 x^2) + 1 / (K2 x^(3/2))) with x = Y_e Sigma (electrons) or X_n Sigma (neutron matter);
 * its energy per area u(x) = x int_0^x P(s) / s^2 ds, tabulated on 512 points log-spaced over [1e-8, 1e8] by
 * Simpson's rule in t = sqrt(s) (64 sub-intervals per segment, 512 for the first) and read back by linear
 * interpolation in (ln x, ln u), the K1 x^2 power law below the table; a pair with a K of 0 is off. The closed form
 * u = x 2 (K2^2 / K1) (y - ln(1 + y)), y = (K1 / K2) sqrt x, is the table's check (eos.rs:52-68), never its source.
 */
#include <math.h>
#include <string.h>

#include "sro.h"

#define COLD_POINTS 512
#define COLD_QUAD 64
static const double COLD_X_MIN = 1e-8, COLD_X_MAX = 1e8;

double sro_cold_pressure(double x, double k1, double k2) {
    if (x <= 0.0 || k1 <= 0.0 || k2 <= 0.0) return 0.0;
    return 1.0 / (1.0 / (k1 * x * x) + 1.0 / (k2 * pow(x, 1.5)));
}

/* int of P(s) / s^2 ds over s in [t0^2, t1^2], as int of P(t^2) / t^4 2t dt by Simpson's rule with n (even) steps */
static double quad_segment(double k1, double k2, double t0, double t1, int n) {
    double h = (t1 - t0) / n, sum = 0.0;
    for (int i = 0; i <= n; i++) {
        double t = t0 + h * i, f = 0.0;
        if (t > 0.0) {
            double s = t * t;
            f = sro_cold_pressure(s, k1, k2) / (s * s) * 2.0 * t;
        }
        double wgt = (i == 0 || i == n) ? 1.0 : ((i % 2 == 1) ? 4.0 : 2.0);
        sum += wgt * f;
    }
    return sum * h / 3.0;
}

double sro_cold_energy_exact(double x, double k1, double k2) {
    if (x <= 0.0 || k1 <= 0.0 || k2 <= 0.0) return 0.0;
    double y = k1 / k2 * sqrt(x), g;
    if (y < 1e-3) {
        g = 0.0;
        for (int k = 2; k < 10; k++) g += ((k % 2 == 0) ? 1.0 : -1.0) * pow(y, k) / k;
    } else {
        g = y - log1p(y);
    }
    return x * 2.0 * k2 * k2 / k1 * g;
}

double sro_cold_node_x(int i) {
    double frac = (double)i / (COLD_POINTS - 1);
    return exp(log(COLD_X_MIN) + frac * (log(COLD_X_MAX) - log(COLD_X_MIN)));
}

void sro_cold_table(double k1, double k2, double *ln_u) {
    double t_prev = 0.0, j = 0.0;
    for (int i = 0; i < COLD_POINTS; i++) {
        double xi = sro_cold_node_x(i), t = sqrt(xi);
        j += quad_segment(k1, k2, t_prev, t, i == 0 ? COLD_QUAD * 8 : COLD_QUAD);
        ln_u[i] = log(xi * j);
        t_prev = t;
    }
}

double sro_cold_table_u(const double *ln_u, double x) {
    if (x <= 0.0) return 0.0;
    double step = (log(COLD_X_MAX) - log(COLD_X_MIN)) / (COLD_POINTS - 1);
    double pos = (log(x) - log(COLD_X_MIN)) / step;
    if (pos < 0.0) return exp(ln_u[0]) * (x / COLD_X_MIN) * (x / COLD_X_MIN);
    int i = (int)floor(pos);
    if (i > COLD_POINTS - 2) i = COLD_POINTS - 2;
    double frac = pos - i;
    return exp(ln_u[i] + frac * (ln_u[i + 1] - ln_u[i]));
}

/* the tables of the two pairs, built when the constants change (the oracle runs one physics at a time) */
static struct {
    double k1, k2;
    int on;
    double ln_u[COLD_POINTS];
} g_pair[2];

static const double *pair_table(int which, double k1, double k2) {
    if (!(k1 > 0.0 && k2 > 0.0)) return NULL;
    if (!g_pair[which].on || g_pair[which].k1 != k1 || g_pair[which].k2 != k2) {
        sro_cold_table(k1, k2, g_pair[which].ln_u);
        g_pair[which].k1 = k1;
        g_pair[which].k2 = k2;
        g_pair[which].on = 1;
    }
    return g_pair[which].ln_u;
}

int sro_cold_on(const sro_physics *ph) {
    return (ph->k1e > 0.0 && ph->k2e > 0.0) || (ph->k1n > 0.0 && ph->k2n > 0.0);
}

double sro_pi_cold(const sro_physics *ph, double sigma, double y_e, double x_n) {
    return sro_cold_pressure(y_e * sigma, ph->k1e, ph->k2e) + sro_cold_pressure(x_n * sigma, ph->k1n, ph->k2n);
}

double sro_eps_cold(const sro_physics *ph, double sigma, double y_e, double x_n) {
    const double *te = pair_table(0, ph->k1e, ph->k2e), *tn = pair_table(1, ph->k1n, ph->k2n);
    double u = 0.0;
    if (te) u += sro_cold_table_u(te, y_e * sigma);
    if (tn) u += sro_cold_table_u(tn, x_n * sigma);
    return u / sigma;
}
