/* Oracle (test infrastructure only): P3 sinks in f64, restating crates/sr-physics/src/reference/sinks.rs:1-429
 * (r_acc, disk, Sinks::form with the mass test, take, accrete, the softened pull and the pair pull, the kick-drift-kick,
 * merge, step) and crates/sr-physics/src/reference/grav.rs:79-96 (add_sinks: the sinks' softened potential and its exact
 * pull added to a field). m_tov_sb, the sheet's neutron-matter ceiling (sinks.rs:104-115 over sheet.rs), is not restated:
 * the mass test reads it from the caller (infinite: no sink from mass, as the reference's with the neutrons' cold
 * pressure off). The gas's eps_th is the EOS's (eos.rs:249-260) with this library's gamma and fixed-mu generalisations.
 */
#include <math.h>
#include <string.h>

#include "sro.h"

#define NS SRO_NS
#define ID_N 9
#define R_MIN 0.5   /* sinks.rs:48 */
#define R_CORE 4.0  /* sinks.rs:55-57 */
#define R_Y_E 3.0
#define CORE_Y_E 0.40

double sro_r_acc(double g, double m, double c_sb) { return fmax(2.0 * g * m / (c_sb * c_sb), 2.0); }

static double r_acc_of(const sro_sinks *k, int i) { return sro_r_acc(k->g, k->s[i].m, k->c_sb); }

/* fractions as the floors renormalise them (floors.rs:45-60): negatives to 0, then to a sum of 1 (pure H if none) */
static void renormalised(const double *raw, double *x) {
    double sum = 0.0;
    for (int s = 0; s < NS; s++) {
        x[s] = fmax(raw[s], 0.0);
        sum += x[s];
    }
    if (sum > 0.0) {
        for (int s = 0; s < NS; s++) x[s] /= sum;
    } else {
        memset(x, 0, sizeof(double) * NS);
        x[0] = 1.0;
    }
}

static void cell_x(const double *planes, size_t n, size_t i, double *x) {
    double raw[NS];
    for (int s = 0; s < NS; s++) raw[s] = planes[(4 + s) * n + i];
    renormalised(raw, x);
}

/* the gas of cell i as the criteria read it (sinks.rs:186-194): its velocity and eps_th */
static void gas(const sro_physics *ph, const double *planes, size_t n, size_t i, double *ux, double *uy,
                double *eps_th) {
    const double sigma = planes[i], mx = planes[n + i], my = planes[2 * n + i], e = planes[3 * n + i];
    double x[NS], inv_mu, y_e;
    cell_x(planes, n, i, x);
    sro_composition(x, &inv_mu, &y_e, NULL);
    if (ph->mu > 0) inv_mu = 1.0 / ph->mu;
    const double kinetic = 0.5 * (mx * mx + my * my) / (sigma * sigma);
    const double raw = e / sigma - kinetic - sro_eps_cold(ph, sigma, y_e, x[ID_N]);
    const double floor = ph->t_floor * inv_mu / (ph->gamma - 1.0);
    *eps_th = raw < floor ? floor : raw;
    *ux = mx / sigma;
    *uy = my / sigma;
}

/* the velocity as the divergence reads it: at rest below Sigma_vac (sinks.rs:196-203) */
static void velocity(const sro_sinks *k, const double *planes, size_t n, size_t i, double *u) {
    const double sigma = planes[i];
    if (sigma < k->sigma_vac) {
        u[0] = u[1] = 0.0;
        return;
    }
    u[0] = planes[n + i] / sigma;
    u[1] = planes[2 * n + i] / sigma;
}

/* the share 1 - f of cell i, f = Sigma_floor / Sigma (sinks.rs:205-219) */
static void take(sro_sinks *k, double *planes, size_t n, size_t i, double *dm, double *dp) {
    const double sigma = planes[i], mx = planes[n + i], my = planes[2 * n + i], e = planes[3 * n + i];
    const double f = k->sigma_floor / sigma, taken = 1.0 - k->sigma_floor / sigma;
    const double thermal_cold = e - 0.5 * (mx * mx + my * my) / sigma;
    k->swallowed += taken * thermal_cold;
    for (int c = 0; c < 4; c++) planes[c * n + i] *= f;
    *dm = taken * sigma;
    dp[0] = taken * mx;
    dp[1] = taken * my;
}

/* the cells of the disk r about (x, y), centre to centre, clipped to the world, row-major (sinks.rs:117-125) */
static int disk(long x, long y, double r, size_t w, size_t h, size_t *out) {
    const long kk = (long)r;
    int c = 0;
    for (long dy = -kk; dy <= kk; dy++)
        for (long dx = -kk; dx <= kk; dx++) {
            const long cx = x + dx, cy = y + dy;
            if (cx >= 0 && cy >= 0 && cx < (long)w && cy < (long)h && (double)(dx * dx + dy * dy) <= r * r)
                out[c++] = (size_t)cy * w + (size_t)cx;
        }
    return c;
}

static int peak(const double *planes, size_t w, size_t x, size_t y) {
    const double sigma = planes[y * w + x];
    for (size_t ny = y - 1; ny <= y + 1; ny++)
        for (size_t nx = x - 1; nx <= x + 1; nx++)
            if (!(sigma >= planes[ny * w + nx])) return 0;
    return 1;
}

/* the mass test (sinks.rs:221-236) */
static int core(const sro_sinks *k, const double *planes, size_t w, size_t h, size_t x, size_t y) {
    const size_t n = w * h;
    size_t cells[81];
    int c = disk((long)x, (long)y, R_CORE, w, h, cells);
    double mass = 0.0;
    for (int j = 0; j < c; j++) mass += planes[cells[j]];
    if (mass < k->m_tov) return 0;
    double m = 0.0, ye = 0.0;
    c = disk((long)x, (long)y, R_Y_E, w, h, cells);
    for (int j = 0; j < c; j++) {
        double xs[NS], inv_mu, y_e;
        cell_x(planes, n, cells[j], xs);
        sro_composition(xs, &inv_mu, &y_e, NULL);
        m += planes[cells[j]];
        ye += planes[cells[j]] * y_e;
    }
    return ye <= CORE_Y_E * m;
}

int sro_sinks_form(sro_sinks *k, const sro_physics *ph, double *planes, size_t w, size_t h, const double *phi,
                   int *fx, int *fy) {
    *fx = *fy = -1;
    if (k->count >= SRO_MAX_SINKS || w < 3 || h < 3) return 0;
    const size_t n = w * h;
    int have = 0;
    double best_s = 0.0;
    size_t bx = 0, by = 0;
    for (size_t y = 1; y + 1 < h; y++)
        for (size_t x = 1; x + 1 < w; x++) {
            const size_t i = y * w + x;
            const double sigma = planes[i];
            if (have && sigma <= best_s) continue;
            if (sigma < k->sigma_bh && !peak(planes, w, x, y)) continue;
            double ul[2], ur[2], ud[2], uu[2];
            velocity(k, planes, n, i + 1, ur);
            velocity(k, planes, n, i - 1, ul);
            velocity(k, planes, n, i + w, uu);
            velocity(k, planes, n, i - w, ud);
            const double div = (ur[0] - ul[0]) / 2.0 + (uu[1] - ud[1]) / 2.0;
            if (div >= 0.0) continue;
            int lowest = 1;
            for (size_t ny = y - 1; ny <= y + 1 && lowest; ny++)
                for (size_t nx = x - 1; nx <= x + 1; nx++)
                    if (!(phi[i] <= phi[ny * w + nx])) {
                        lowest = 0;
                        break;
                    }
            if (!lowest) continue;
            double ux, uy, eps_th;
            gas(ph, planes, n, i, &ux, &uy, &eps_th);
            if (0.5 * (ux * ux + uy * uy) + eps_th + phi[i] >= 0.0) continue;
            if (sigma < k->sigma_bh && !core(k, planes, w, h, x, y)) continue;
            have = 1;
            best_s = sigma;
            bx = x;
            by = y;
        }
    if (!have) return 0;
    size_t cells[81];
    int c;
    if (peak(planes, w, bx, by) && core(k, planes, w, h, bx, by)) {
        size_t all[81];
        const int ca = disk((long)bx, (long)by, R_CORE, w, h, all);
        c = 0;
        for (int j = 0; j < ca; j++)
            if (planes[all[j]] >= k->sigma_vac && planes[all[j]] > k->sigma_floor) cells[c++] = all[j];
    } else {
        cells[0] = by * w + bx;
        c = 1;
    }
    double m = 0.0, p[2] = {0.0, 0.0};
    for (int j = 0; j < c; j++) {
        double dm, dp[2];
        take(k, planes, n, cells[j], &dm, dp);
        m += dm;
        p[0] += dp[0];
        p[1] += dp[1];
    }
    sro_sink *s = &k->s[k->count++];
    s->x = (double)bx + 0.5;
    s->y = (double)by + 0.5;
    s->vx = p[0] / m;
    s->vy = p[1] / m;
    s->m = m;
    *fx = (int)bx;
    *fy = (int)by;
    return 1;
}

double sro_sinks_accrete(sro_sinks *k, const sro_physics *ph, double *planes, size_t w, size_t h) {
    const size_t n = w * h;
    double total = 0.0;
    for (int q = 0; q < k->count; q++) {
        const sro_sink s = k->s[q];
        const double ra = r_acc_of(k, q);
        const size_t x0 = (size_t)fmax(floor(s.x - ra), 0.0);
        size_t x1 = (size_t)fmax(ceil(s.x + ra), 0.0);
        if (x1 > w - 1) x1 = w - 1;
        const size_t y0 = (size_t)fmax(floor(s.y - ra), 0.0);
        size_t y1 = (size_t)fmax(ceil(s.y + ra), 0.0);
        if (y1 > h - 1) y1 = h - 1;
        double dm = 0.0, dp[2] = {0.0, 0.0};
        for (size_t y = y0; y <= y1; y++)
            for (size_t x = x0; x <= x1; x++) {
                const size_t i = y * w + x;
                const double cx = (double)x + 0.5 - s.x, cy = (double)y + 0.5 - s.y;
                const double r = hypot(cx, cy);
                if (r > ra || planes[i] < k->sigma_vac || planes[i] <= k->sigma_floor) continue;
                double ux, uy, eps_th;
                gas(ph, planes, n, i, &ux, &uy, &eps_th);
                const double rx = ux - s.vx, ry = uy - s.vy;
                if (0.5 * (rx * rx + ry * ry) + eps_th - k->g * s.m / fmax(r, R_MIN) >= 0.0) continue;
                double m, p[2];
                take(k, planes, n, i, &m, p);
                dm += m;
                dp[0] += p[0];
                dp[1] += p[1];
            }
        const double m = s.m + dm;
        k->s[q].m = m;
        k->s[q].vx = (s.m * s.vx + dp[0]) / m;
        k->s[q].vy = (s.m * s.vy + dp[1]) / m;
        total += dm;
    }
    return total;
}

/* the pull on sink i at the positions at (sinks.rs:322-356): the gas's, cell by cell row-major, then the other sinks' */
static void acceleration(const sro_sinks *k, int i, const double (*at)[2], const double *planes, size_t w, size_t h,
                         double *a) {
    const double ra = r_acc_of(k, i);
    a[0] = a[1] = 0.0;
    for (size_t c = 0; c < w * h; c++) {
        const double sigma = planes[c];
        if (sigma < k->sigma_vac) continue;
        const double dx = (double)(c % w) + 0.5 - at[i][0], dy = (double)(c / w) + 0.5 - at[i][1];
        const double soft = dx * dx + dy * dy + ra * ra;
        const double pull = k->g * sigma / (soft * sqrt(soft));
        a[0] += pull * dx;
        a[1] += pull * dy;
    }
    double b[2] = {0.0, 0.0};
    for (int j = 0; j < k->count; j++) {
        if (j == i) continue;
        const double dx = at[j][0] - at[i][0], dy = at[j][1] - at[i][1];
        const double rp = fmax(r_acc_of(k, i), r_acc_of(k, j));
        const double soft = dx * dx + dy * dy + rp * rp;
        const double pull = k->g * k->s[j].m / (soft * sqrt(soft));
        b[0] += pull * dx;
        b[1] += pull * dy;
    }
    a[0] += b[0];
    a[1] += b[1];
}

void sro_sinks_pull(const sro_sinks *k, const double *planes, size_t w, size_t h, double *a) {
    double at[SRO_MAX_SINKS][2];
    for (int i = 0; i < k->count; i++) {
        at[i][0] = k->s[i].x;
        at[i][1] = k->s[i].y;
    }
    for (int i = 0; i < k->count; i++) acceleration(k, i, (const double(*)[2])at, planes, w, h, a + 2 * i);
}

void sro_sinks_advance(sro_sinks *k, const double *planes, size_t w, size_t h, double dt) {
    const int cnt = k->count;
    double at[SRO_MAX_SINKS][2], a[SRO_MAX_SINKS][2];
    for (int i = 0; i < cnt; i++) {
        at[i][0] = k->s[i].x;
        at[i][1] = k->s[i].y;
    }
    for (int i = 0; i < cnt; i++) acceleration(k, i, (const double(*)[2])at, planes, w, h, a[i]);
    for (int i = 0; i < cnt; i++) {
        sro_sink *s = &k->s[i];
        s->vx += 0.5 * dt * a[i][0];
        s->vy += 0.5 * dt * a[i][1];
        s->x += dt * s->vx;
        s->y += dt * s->vy;
    }
    for (int i = 0; i < cnt; i++) {
        at[i][0] = k->s[i].x;
        at[i][1] = k->s[i].y;
    }
    for (int i = 0; i < cnt; i++) acceleration(k, i, (const double(*)[2])at, planes, w, h, a[i]);
    for (int i = 0; i < cnt; i++) {
        k->s[i].vx += 0.5 * dt * a[i][0];
        k->s[i].vy += 0.5 * dt * a[i][1];
    }
}

int sro_sinks_merge(sro_sinks *k) {
    int lost = 0;
again:
    for (int i = 0; i < k->count; i++)
        for (int j = i + 1; j < k->count; j++) {
            const sro_sink a = k->s[i], b = k->s[j];
            if (hypot(a.x - b.x, a.y - b.y) < fmax(r_acc_of(k, i), r_acc_of(k, j))) {
                const double m = a.m + b.m;
                k->s[i].x = (a.m * a.x + b.m * b.x) / m;
                k->s[i].y = (a.m * a.y + b.m * b.y) / m;
                k->s[i].vx = (a.m * a.vx + b.m * b.vx) / m;
                k->s[i].vy = (a.m * a.vy + b.m * b.vy) / m;
                k->s[i].m = m;
                for (int q = j; q + 1 < k->count; q++) k->s[q] = k->s[q + 1];
                k->count--;
                lost++;
                goto again;
            }
        }
    return lost;
}

void sro_sinks_add_field(const sro_sinks *k, double *phi, double *gx, double *gy, size_t w, size_t h) {
    for (int q = 0; q < k->count; q++) {
        const sro_sink s = k->s[q];
        const double ra = r_acc_of(k, q);
        for (size_t y = 0; y < h; y++)
            for (size_t x = 0; x < w; x++) {
                const double dx = (double)x + 0.5 - s.x, dy = (double)y + 0.5 - s.y;
                const double soft = dx * dx + dy * dy + ra * ra;
                const size_t i = y * w + x;
                if (phi) phi[i] -= k->g * s.m / sqrt(soft);
                const double pull = k->g * s.m / (soft * sqrt(soft));
                if (gx) gx[i] -= pull * dx;
                if (gy) gy[i] -= pull * dy;
            }
    }
}

int sro_sinks_step(sro_sinks *k, const sro_physics *ph, double *planes, size_t w, size_t h, const double *phi_gas,
                   double dt, double *phi_work, sro_sinks_report *rep) {
    memset(rep, 0, sizeof *rep);
    memcpy(phi_work, phi_gas, sizeof(double) * w * h);
    sro_sinks_add_field(k, phi_work, NULL, NULL, w, h);
    sro_sinks_form(k, ph, planes, w, h, phi_work, &rep->formed_x, &rep->formed_y);
    rep->accreted = sro_sinks_accrete(k, ph, planes, w, h);
    sro_sinks_advance(k, planes, w, h, dt);
    rep->merged = sro_sinks_merge(k);
    return SRO_OK;
}
