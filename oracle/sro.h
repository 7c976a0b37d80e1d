/* The oracle: a CPU f64 restatement of the reference's twin for the grid-physics path.
 *
 * TEST INFRASTRUCTURE ONLY. Nothing in the shipped library links, loads or calls this code; only the test suite,
 * the smoke check and the bench's CPU-baseline leg use it, and only as the checker.
 *
 * Each function cites the reference file:line it restates (paths relative to the reference repository root).
 * Conventions follow the reference: channel-major planes, row-major cells, index = y * w + x, y down.
 *
 * Generalisations over the reference (each reduces to the reference exactly at its defaults gamma = 2, dx = 1,
 * opacity = reference law, mu from the composition):
 *   - adiabatic index gamma: Pi_th = (gamma-1) Sigma eps, T = mu (gamma-1) eps, c^2 = gamma Pi / Sigma,
 *     heat capacity C = Sigma / (mu (gamma-1));
 *   - cell size dx: flux divergences / dx, gradients / dx, sheet kernel K = -G dx / r;
 *   - opacity law: reference kappa0 (1 + X_H) (+ dust) or Kramers kappa0 Sigma T^-3.5;
 *   - mu: from the species (reference) or fixed.
 * Not restated (absent from the workloads): the cold (Fermi) pressure tables (K pairs off), held cells, sinks,
 * player edits, burning.
 */
#ifndef SRO_H
#define SRO_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SRO_NS 10
#define SRO_COMPONENTS (4 + SRO_NS)
#define SRO_GHOSTS 2
#define SRO_MIN_PADDED 32

enum { SRO_OK = 0, SRO_ERR_POW2 = -1, SRO_ERR_BOX = -2, SRO_ERR_LEN = -3, SRO_ERR_ARG = -4 };
enum { SRO_EDGE_BOUNCE = 0, SRO_EDGE_LEAVE = 1 };
enum { SRO_OPACITY_REFERENCE = 0, SRO_OPACITY_KRAMERS = 1 };

typedef struct {
    double gamma;
    double G;
    double dx;
    double sigma_vac, sigma_floor, t_floor;
    double cfl, eta_g;
    double a_r, c_sb, kappa0, k_cond, kappa_dust, t_dust;
    int opacity;
    double mu;          /* > 0: fixed mean molecular weight; <= 0: from the species fractions */
    int edge;
    /* heat test flags (reference heat.rs:62-74) */
    int pin_temperature;
    double constant_diffusivity; /* > 0 switches it on */
    int flux_limiter;
    double max_padded;  /* FFT cap per padded axis (reference 2048) */
    /* cold pressure pairs (eos.rs:24-30): electrons (k1e, k2e) on Y_e Sigma, neutron matter (k1n, k2n) on X_n Sigma; a
     * pair with a K of 0 is off (the reference's shipped 21.3 / 134 and 2.67 / 58.2) */
    double k1e, k2e, k1n, k2n;
} sro_physics;

/* ---------------- cold pressure (reference crates/sr-physics/src/eos.rs) ---------------- */
double sro_cold_pressure(double x, double k1, double k2);
double sro_cold_energy_exact(double x, double k1, double k2);
double sro_cold_node_x(int i);
void sro_cold_table(double k1, double k2, double *ln_u); /* 512 entries */
double sro_cold_table_u(const double *ln_u, double x);
int sro_cold_on(const sro_physics *ph);
double sro_pi_cold(const sro_physics *ph, double sigma, double y_e, double x_n);
double sro_eps_cold(const sro_physics *ph, double sigma, double y_e, double x_n);
/* test switch: the hydro's HOT_FACE bracket on (default) or off */
void sro_hydro_set_hot_face(int on);
void sro_hydro_set_pair_work(int on);
/* test hooks: a uniform field added to the half step's g; gravity's kick take-back 1 (on), 0 (off), -1 (mis-signed) */
void sro_hydro_set_field_bias(double bx, double by);
void sro_hydro_set_kick_takeback(int mode);

/* ---------------- FFT and convolution (reference crates/sr-physics/src/reference/fft.rs) ---------------- */
int sro_padded_size(size_t len, size_t max_padded, size_t *out);
/* in-place complex transform of n (power of two) interleaved (re, im) doubles */
int sro_fft(double *data, size_t n, int inverse);
int sro_fft2(double *data, size_t nx, size_t ny, int inverse);
/* zero-padded convolution with kernel(dx, dy, ctx); out is (w+2 ring) x (h+2 ring) */
typedef double (*sro_kernel_fn)(long dx, long dy, void *ctx);
int sro_convolve_ringed(const double *field, size_t w, size_t h, size_t ring, sro_kernel_fn k, void *ctx,
                        size_t max_padded, double *out);

/* ---------------- gravity (reference crates/sr-physics/src/reference/grav.rs) ---------------- */
#define SRO_SELF_CELL 3.525494348078172
double sro_grav_kernel(double G, double dx, long i, long j);
int sro_gravity_field(const sro_physics *ph, const double *sigma, size_t w, size_t h, double *phi, double *gx,
                      double *gy, double *phi_ringed /* may be NULL */);

/* ---------------- EOS + composition (reference eos.rs, registry.rs:146-154) ---------------- */
void sro_composition(const double *x, double *inv_mu, double *y_e, double *z_met);

/* ---------------- hydro (reference crates/sr-physics/src/reference/hydro.rs) ---------------- */
double sro_hydro_dt(const sro_physics *ph, const double *planes, size_t w, size_t h, const double *gx,
                    const double *gy);
/* one step in place; gx/gy/phi may be NULL (no gravity), fx/fy may be NULL (no radiation force).
 * escaped: COMPONENTS * 2 (w + h) doubles accumulated (may be NULL). */
int sro_hydro_step(const sro_physics *ph, double *planes, size_t w, size_t h, double dt, const double *phi,
                   const double *gx, const double *gy, const double *rad_fx, const double *rad_fy, double *escaped,
                   double *sigma_mid /* optional out */);

/* ---------------- heat / FLD (reference crates/sr-physics/src/reference/heat.rs) ---------------- */
typedef struct {
    int passes, rounds, converged;
    double delta, radiated;
} sro_heat_report;
/* one implicit P5 step in place; flux_x (w+1)h and flux_y w(h+1) outputs (may be NULL), emitted w*h (may be NULL) */
/* the twin's heat solver: 0 the reference's Jacobi PCG rule, > 0 multigrid PCG to that relative residual */
void sro_heat_set_mg(double rtol);
int sro_heat_step(const sro_physics *ph, double *planes, size_t w, size_t h, double dt, double *flux_x,
                  double *flux_y, double *emitted, sro_heat_report *rep);
/* the law on the start state (heat.rs:307) */
int sro_heat_law(const sro_physics *ph, const double *planes, size_t w, size_t h, double *flux_x, double *flux_y);
/* the first Newton pass's frozen system (own = c_i + sum of vacuum/edge faces, kx, ky face coefficients, rhs, gas);
 * own/rhs/gas are w*h, kx (w+1)h, ky w(h+1). */
int sro_heat_first_system(const sro_physics *ph, const double *planes, size_t w, size_t h, double dt, double *own,
                          double *kx, double *ky, double *rhs, uint8_t *gas, double *u0);
void sro_flux_cells(const double *flux_x, const double *flux_y, size_t w, size_t h, double *fx, double *fy);

/* ---------------- FLD linear solve model (f64) ---------------- */
/* A v = own v + sum_gas_faces k (v_i - v_j); non-gas rows are identity (gas == NULL: all gas). */
void sro_fld_matvec(size_t w, size_t h, const double *own, const double *kx, const double *ky, const uint8_t *gas,
                    const double *v, double *out);
/* precond 0: Jacobi, 1: multigrid V(nu1, nu2) with Galerkin 2x2 aggregation (the same operator the device applies);
 * stops at ||b - A x|| <= tol ||b||; returns iterations or < 0 on failure. */
int sro_fld_pcg(size_t w, size_t h, const double *own, const double *kx, const double *ky, const uint8_t *gas,
                const double *b, double *x, double tol, int max_iter, int precond, int levels, int nu1, int nu2,
                double *final_rel);
/* one application z = M^-1 r of the multigrid preconditioner (f64 model) */
void sro_mg_set_coarse_scale(double omega); /* the coarse correction's scale for both calls; 1 by default */
int sro_mg_apply(size_t w, size_t h, const double *own, const double *kx, const double *ky, const uint8_t *gas,
                 const double *r, double *z, int levels, int nu1, int nu2);

/* ---------------- floors (reference crates/sr-physics/src/reference/floors.rs) ---------------- */
void sro_floors_apply(const sro_physics *ph, double *planes, size_t w, size_t h, double *vacuum_reset,
                      double *floor_added);

/* ---------------- the twin's step driver (reference crates/sr-physics/src/reference/mod.rs) ---------------- */
typedef struct sro_twin sro_twin;
sro_twin *sro_twin_new(const sro_physics *ph, size_t w, size_t h, const double *planes, int gravity, int hydro,
                       int heat);
void sro_twin_free(sro_twin *t);
int sro_twin_step(sro_twin *t);
void sro_twin_state(const sro_twin *t, double *planes);
void sro_twin_times(const sro_twin *t, double *dt, double *dt_next, double *sim_time, uint64_t *steps);
void sro_twin_cap_dt_next(sro_twin *t, double cap);
void sro_twin_last_heat(const sro_twin *t, sro_heat_report *rep);
void sro_twin_escaped(const sro_twin *t, double *escaped);

/* ---------------- P6 burning (reference crates/sr-physics/src/reference/react.rs, rates.rs) ---------------- */
#define SRO_REACT_RECORDS 16
#define SRO_REACT_GROUPS 7 /* registry.rs:460 GROUPS: H He C Ne O Si N */
#define SRO_REACT_TABLE 256
#define SRO_REACT_SPAN 2.5
/* One enabled record with its physics.json keys resolved (registry.rs:485-535, rates.rs:31-60): inputs (species,
 * share, order), outputs (species, share), its group, q = q_ratio Q_H, the law's A, a, Gamow terms (T_k, nu, w),
 * T_thr (0: f = 1), the gate (Sigma_g, p) and whether it is N_Fe (its rate sets the latch). */
typedef struct {
    int n_in, n_out;
    int in_species[4];
    double in_share[4], in_order[4];
    int out_species[4];
    double out_share[4];
    int group;
    double q, a_coef, a_sigma;
    int n_terms;
    double t_k[4], nu[4], w[4];
    double t_thr;
    int gated;
    double sigma_g, gate_p;
    int latch;
} sro_react_record;
/* The pass's constants: the records, burn_dx, burn_dt, K_burn, the cooling's A_nu, T_nu, m (physics.json), the
 * scene's neutrinos switch and pin_temperature, and the record N_Fe's closed form steps (-1: none). */
typedef struct {
    int n_records;
    sro_react_record rec[SRO_REACT_RECORDS];
    double burn_dx, burn_dt, burn_k, a_nu, t_nu, m_nu;
    int neutrinos, pin_temperature, nfe;
} sro_react_params;
typedef struct {
    double nuclear[SRO_REACT_GROUPS]; /* energy per area released, summed over the world */
    double neutrino_lost;
    uint64_t substeps, retried, max_substeps, unfinished;
    double max_dx, max_dt_rel;
    int latch;
} sro_react_pass;
/* ln f at the table's nodes of one record (SRO_REACT_TABLE values; none without terms) */
void sro_react_table(const sro_react_record *r, double *ln_f);
/* one record's rate omega at (Sigma, X, T, Y_e), f through the table (table != 0) or by the direct form */
double sro_react_omega(const sro_react_record *r, const double *ln_f, double sigma, const double *x, double t,
                       double y_e, int table);
/* P6 over dt on channel-major planes (react.rs:563-603): per-cell energy booked by group (7 n, may be NULL),
 * neutrino loss (n), S_nu (n) */
int sro_react_step(const sro_physics *ph, const sro_react_params *rp, double *planes, size_t w, size_t h, double dt,
                   double *nuclear, double *lost, double *s_nu, sro_react_pass *pass);
/* ---------------- P3 sinks (reference crates/sr-physics/src/reference/sinks.rs, grav.rs:79-96) ---------------- */
#define SRO_MAX_SINKS 16
typedef struct {
    double x, y, vx, vy, m; /* position (cells, cell (i, j) centred at (i + 1/2, j + 1/2)), velocity, mass */
} sro_sink;
/* the sinks, what they swallowed, and the constants P3 reads: G, Sigma_BH, Sigma_vac, Sigma_floor, c_sb, m_tov_sb */
typedef struct {
    int count;
    sro_sink s[SRO_MAX_SINKS];
    double swallowed;
    double g, sigma_bh, sigma_vac, sigma_floor, c_sb, m_tov;
} sro_sinks;
typedef struct {
    int formed_x, formed_y; /* -1 when none formed */
    double accreted;
    int merged;
} sro_sinks_report;
double sro_r_acc(double g, double m, double c_sb);
int sro_sinks_form(sro_sinks *k, const sro_physics *ph, double *planes, size_t w, size_t h, const double *phi,
                   int *fx, int *fy);
double sro_sinks_accrete(sro_sinks *k, const sro_physics *ph, double *planes, size_t w, size_t h);
void sro_sinks_pull(const sro_sinks *k, const double *planes, size_t w, size_t h, double *a);
void sro_sinks_advance(sro_sinks *k, const double *planes, size_t w, size_t h, double dt);
int sro_sinks_merge(sro_sinks *k);
/* the sinks' softened potential and exact pull added to a field (any of the three may be NULL) */
void sro_sinks_add_field(const sro_sinks *k, double *phi, double *gx, double *gy, size_t w, size_t h);
/* one P3 step (sinks.rs:401-417): formation against phi_gas with the sinks' potential added (phi_work, w h values,
 * receives it), accretion, the kick-drift-kick in the gas accretion left, merging */
int sro_sinks_step(sro_sinks *k, const sro_physics *ph, double *planes, size_t w, size_t h, const double *phi_gas,
                   double dt, double *phi_work, sro_sinks_report *rep);

/* P6 in the twin's step, between P5 and P8 (NULL: off) */
void sro_twin_set_reactions(sro_twin *t, const sro_react_params *rp);
void sro_twin_last_react(const sro_twin *t, sro_react_pass *pass);

/* ---------------- the reference's RNG (reference fft.rs:292-300) ---------------- */
void sro_lcg_fill(uint64_t seed, size_t count, double *out); /* uniform [0, 1) */

#ifdef __cplusplus
}
#endif
#endif
