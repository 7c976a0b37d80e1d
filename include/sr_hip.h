/* sr_hip.h — the C ABI of the MI355X-native grid-physics path (P2 FFT gravity, P4 MUSCL-Hancock/HLL hydro,
 * P5 implicit flux-limited diffusion by MG-PCG) of a sandbox_reactions-style engine.
 *
 * The reference engine has no plugin registry: a pass is a Rust type contributing dispatch lists to `Step`, and the
 * operator API is `Step`'s methods (crates/sr-engine/src/step/mod.rs:120-169, :367-527, :691-777). This header is
 * the C surface a HIP backend crate would bind through `extern "C"`; every entry point names the reference
 * interface it replaces. Plain pointers and sizes only. All functions return SR_OK (0) or a negative SR_ERR_* code
 * and leave a message for sr_last_error(); none aborts across the boundary.
 *
 * Ownership: the caller owns host arrays and passes their lengths; a world owns its device memory until
 * sr_destroy. Functions taking `*_dev` pointers read and write device memory the caller owns, on the caller's
 * stream (a hipStream_t passed as void*, NULL = the default stream).
 * Threading: a world is confined to one host thread; it runs on its own hipStream. sr_read_* calls block.
 */
#ifndef SR_HIP_H
#define SR_HIP_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes: the reference's named refusals (crates/sr-engine/src/step/fft.rs:35-51, state.rs:93-100,
 * gpu.rs:45-80) as integers */
#define SR_OK 0
#define SR_ERR_WORLD_SIZE (-1)  /* StateError::WorldSize: an axis not a multiple of 8 or outside [8, 4096] */
#define SR_ERR_BOX_AXIS (-2)    /* FftError::BoxAxis / BoxOutOfRange: a padded FFT axis past the cap */
#define SR_ERR_LENGTH (-3)      /* FftError::LengthMismatch: a buffer whose length is not what the call needs */
#define SR_ERR_DEVICE (-4)      /* a HIP runtime error (the reference's device error, exit 3) */
#define SR_ERR_ARGUMENT (-5)    /* a null handle or pointer, or a parameter out of range */
#define SR_ERR_NOT_POW2 (-6)    /* FftError::NotPowerOfTwo */
#define SR_ERR_NON_FINITE (-7)  /* the non-finite guard tripped (the reference's exit 5) */

#define SR_COMPONENTS 14 /* Sigma, Sigma u, Sigma v, E, then 10 species fractions (state.rs:26-46) */
#define SR_SPECIES 10

enum sr_edge { SR_EDGE_BOUNCE = 0, SR_EDGE_LEAVE = 1 };
enum sr_opacity { SR_OPACITY_REFERENCE = 0, SR_OPACITY_KRAMERS = 1 };
enum sr_precond { SR_PRECOND_JACOBI = 0, SR_PRECOND_MULTIGRID = 1, SR_PRECOND_LEARNED = 2 };
enum sr_stop { SR_STOP_REFERENCE = 0, SR_STOP_RELATIVE = 1 };
enum sr_tile { SR_TILE_16 = 0, SR_TILE_32 = 1, SR_TILE_64 = 2 }; /* smoother tiles: 16x16, 32x32, 64x32 */
/* the fused smoothers of the last two tile-smoothed levels: separate launches, or one launch per direction with 8x8
 * coarse / 16x16 fine tiles (small) or 16x16 / 32x32 (large, nu1 and nu2 <= 2) */
enum sr_fuse { SR_FUSE_OFF = 0, SR_FUSE_SMALL = 1, SR_FUSE_LARGE = 2 };
enum sr_field { SR_FIELD_PHI = 0, SR_FIELD_GX = 1, SR_FIELD_GY = 2, SR_FIELD_FLUX_X = 3, SR_FIELD_FLUX_Y = 4,
                SR_FIELD_SIGMA_MID = 5, SR_FIELD_PHI_MID = 6 };

/* The constants the path reads (assets/physics.json keys, plus the generalisations the workloads need: gamma,
 * dx, G, the opacity law and a fixed mu). Defaults via sr_physics_defaults() are the reference's. */
typedef struct {
    double gamma;       /* adiabatic index; the reference hard-wires 2 (crates/sr-physics/src/eos.rs:207-247) */
    double G;           /* gravitational constant (reference/grav.rs:19) */
    double dx;          /* cell side; the reference's cells are 1 */
    double sigma_vac, sigma_floor, t_floor;
    double cfl, eta_g;
    double a_r, c_sb, kappa0, k_cond, kappa_dust, t_dust;
    int opacity;        /* enum sr_opacity */
    double mu;          /* > 0: fixed mean molecular weight; <= 0: from the species (registry.rs:146-154) */
    int edge;           /* enum sr_edge */
    int pin_temperature;         /* heat test flags (reference/heat.rs:62-74) */
    double constant_diffusivity; /* > 0 switches it on */
    int flux_limiter;
    int max_padded;     /* FFT cap per padded axis; the reference's is 2048 (step/fft.rs:26-63), the library's default 4096 */
    double k1e, k2e;    /* the electrons' cold pressure pair on Y_e Sigma (eos.rs:24-30; shipped 21.3, 134) */
    double k1n, k2n;    /* neutron matter's on X_n Sigma (shipped 2.67, 58.2); a pair of zeros is off (the default
                         * here, as the reference's Sod and heat scenes set it); one K of a pair alone is refused */
} sr_physics;

typedef struct {
    int gravity, hydro, heat; /* scene switches (reference/mod.rs:47-55) */
} sr_switches;

/* FLD solver settings (the heat pass's stopping rules and multigrid shape) */
typedef struct {
    int stop;              /* enum sr_stop: the reference's two energy-weighted measures, or ||b-Ax||/||b|| */
    double solve_tol;      /* the reference GPU's 1e-5 (heat.wgsl:355-360) or a relative residual such as 1e-8 */
    double newton_heat;    /* Newton pass stop on sum C|dT| / sum C T0 (reference GPU 1e-6) */
    double newton_cell;    /* Newton pass stop per cell |dT| / T (reference GPU 1e-3) */
    int max_passes;        /* reference 128 */
    int max_rounds;        /* reference 4096 */
    int levels;            /* multigrid levels, 0 = down to 4x4 */
    int nu1, nu2;          /* red-black Gauss-Seidel sweeps before / after the coarse correction */
    double forcing;        /* 0: every Newton pass solved to solve_tol (the reference's rule). > 0 with
                            * SR_STOP_RELATIVE: inexact Newton, a pass solved to forcing x the previous pass's
                            * relative heat change (at most 1e-2, at least solve_tol); the pass that meets the Newton
                            * criteria after a loose solve is followed by one solved to solve_tol */
    double coarse_scale;   /* omega, the scale of every coarse-grid correction: 1 (or 0) is the reference's plain
                            * prolongation; in (0, 2) the V-cycle stays an SPD preconditioner, and about 1.2 cuts the
                            * iterations of the stiff lognormal systems by a tenth */
} sr_heat_settings;

typedef struct {
    double dt, dt_next, sim_time; /* the step's dt, the next step's, sim time in f64 */
    uint64_t steps;
    int non_finite;               /* 1 when the guard has found a non-finite value since the upload */
    int nf_x, nf_y, nf_channel;   /* its cell (the lowest cell index of that step's findings) and channel (0..13) */
    uint64_t nf_step;             /* the index of the step whose P9 found it (0 also for the upload's own scan), counted
                                   * from the upload; the first finding is kept (step/dt.rs:40-47, NonFinite) */
} sr_dt_report;

typedef struct {
    double dt;
    int passes, rounds, converged;
    double delta, radiated;
    double final_residual;        /* ||b - A x|| / ||b|| of the last pass, FP64 */
} sr_heat_report;

typedef struct {
    uint64_t steps;
    double step_ms;               /* mean wall ms per step over the timed steps */
    double pass_ms[10];           /* per pass, the reference's Pass::ORDER slots P0..P9 (step/mod.rs:200-219) */
} sr_step_cost;

/* What the steps since the upload (or the last sr_read_work with reset) did, counted on the host as the launches are
 * made, and the HBM bytes their algorithm moves (each operand plane once per kernel, f32 4 B, f64 8 B): the
 * numerator of the full step's roofline. No counterpart in the reference (its read_cost gives times only). */
typedef struct {
    uint64_t steps, hydro_steps, convolutions, heat_steps, newton_passes;
    uint64_t pcg_iterations, precond_applications, true_residuals, hierarchy_builds;
    double algorithmic_bytes;
    double bytes_p2, bytes_p4, bytes_p5, bytes_p89;  /* the same by pass */
} sr_step_work;

/* How a heat-pass or standalone solve decides when to stop (sr_fld_set_solve_mode, sr_set_solve_mode): on the host
 * from partial sums read back every few iterations (SR_SOLVE_HOST, the default; the learned preconditioner always
 * does), or on the device, each PCG iteration opened by a head kernel that reads the last one's residual and decides
 * (the FP64 true residual and restart, or the next inner iteration, or done), the later kernels of the iteration
 * skipping themselves once it is done; the iterations launched one by one (SR_SOLVE_DEVICE) or replayed as a
 * captured hipGraph of two (SR_SOLVE_GRAPH). The two device modes give the same bits. SR_SOLVE_HOST looks ahead: at
 * a check the last rate says will not end the inner loop, the next iteration's preconditioner application is queued
 * before the host waits for the partials, so the device does not idle meanwhile (when the check ends the loop after
 * all, that application goes unused); SR_SOLVE_HOST_WAIT is the same loop waiting at every check on an idle device.
 * The two host modes give the same bits. A world replays its step's fixed parts (P1..P4 with the take-backs, P8/P9)
 * as captured graphs in every mode (with the clocks off). */
enum { SR_SOLVE_HOST = 0, SR_SOLVE_DEVICE = 1, SR_SOLVE_GRAPH = 2, SR_SOLVE_HOST_WAIT = 3 };

typedef struct sr_world sr_world;
typedef struct sr_gravity sr_gravity;
typedef struct sr_fld sr_fld;
typedef struct sr_net sr_net;

/* ------------------------------------------------------------------------------------------------ errors / device */
const char *sr_last_error(void);
int sr_device_info(char *name, size_t name_len, int *cu_count, size_t *hbm_bytes);
void sr_physics_defaults(sr_physics *ph);
void sr_heat_settings_defaults(sr_heat_settings *hs);

/* ------------------------------------------------------------------------------------------------ P2 gravity / FFT */
/* Fft::transform (crates/sr-engine/src/step/fft.rs:116-134): an in-place complex 2D transform of an nx x ny
 * row-major interleaved (re, im) f32 array in device memory; forward unnormalised, inverse divided by nx*ny. */
int sr_fft2(float *data_dev, int nx, int ny, int inverse, void *stream);
/* sr_fft2 without allocating: no counterpart in the reference. For callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops). The twiddle table lives in a library-owned device array of 8192
 * complex f32 values (64 KB, held by each GPU the library runs on until the process ends): axes longer than 8192 are
 * refused (SR_ERR_LENGTH). Every call on one GPU shares that array, so calls that overlap on one GPU (on different
 * streams) race. Returns once the work is enqueued on `stream`, without waiting for it; results bitwise equal to
 * sr_fft2's. */
int sr_fft2_noalloc(float *data_dev, int nx, int ny, int inverse, void *stream);

/* Convolution::new + Gravity::new (step/fft.rs:136-182, step/grav.rs:133-200; twin reference/grav.rs:99-121):
 * a gravity solver for a w x h box, its kernel's spectrum built once (f64 on the host, stored f32). */
int sr_gravity_create(int w, int h, double G, double dx, double sigma_vac, int max_padded, sr_gravity **out);
/* Gravity::field (reference/grav.rs:115-121): phi on the box grown by one ring ((w+2) x (h+2)), and g = -grad phi
 * over the box (gx, gy may be NULL). sigma_dev is the w x h surface density; cells below sigma_vac are no source. */
int sr_gravity_field(sr_gravity *g, const float *sigma_dev, float *phi_ringed_dev, float *gx_dev, float *gy_dev,
                     void *stream);
void sr_gravity_destroy(sr_gravity *g);

/* ------------------------------------------------------------------------------------------------ P5 standalone FLD */
/* The heat pass's linear system (step/heat.rs:609-673 heat_first_system / apply_heat_precond): a solver for the
 * 5-point SPD operator (A v)_i = own_i v_i + sum_f k_f (v_i - v_j) over a w x h grid, face coefficients kx
 * ((w+1) x h, face x between cells x-1 and x) and ky (w x (h+1)), couplings to non-gas cells and the edge already
 * folded into own. gas_dev (w x h bytes, may be NULL: all gas) marks the rows in the system. */
int sr_fld_create(int w, int h, const sr_heat_settings *hs, sr_fld **out);
int sr_fld_set_system(sr_fld *f, const float *own_dev, const float *kx_dev, const float *ky_dev,
                      const uint8_t *gas_dev, void *stream);
/* Solves A x = b by preconditioned flexible CG with FP64 x, r and dots, FP32 preconditioner, to
 * ||b - A x||_2 <= tol ||b||_2 (FP64 true residual); x_dev holds the start and receives the answer. */
int sr_fld_solve(sr_fld *f, const double *b_dev, double *x_dev, double tol, int max_iter, int precond,
                 int *iterations, double *final_rel, void *stream);
/* z = M^-1 r for one preconditioner application (symmetry / positivity tests) */
int sr_fld_apply_precond(sr_fld *f, const float *r_dev, float *z_dev, int precond, void *stream);
/* out = A v in FP64 */
int sr_fld_matvec(sr_fld *f, const double *v_dev, double *out_dev, void *stream);
int sr_fld_levels(sr_fld *f, int *levels);
/* the solve's stopping decisions: SR_SOLVE_HOST, SR_SOLVE_DEVICE or SR_SOLVE_GRAPH (no counterpart in the reference:
 * its heat rounds are chained dispatches whose sizes the GPU writes, step/mod.rs:129-138) */
int sr_fld_set_solve_mode(sr_fld *f, int mode);
/* Tile shapes of the smoothers of the levels above the single-block ones (SR_TILE_16 16x16, SR_TILE_32 32x32,
 * SR_TILE_64 64x32, the last for nu1, nu2 <= 2 only): the project's autotuner times every shape on the first system
 * built and keeps the fastest per level below level 0 and smoother; sr_fld_set_tiles fixes them instead. Level 0
 * keeps 32x32 (its dot partials are counted per tile). No result bit depends on the shapes (exact red-black
 * Gauss-Seidel in every tiling). No counterpart in the reference, whose tiles are fixed by WebGPU's limits
 * (heat.wgsl:1401-1445: 16x16 with 4 RB pairs). */
int sr_fld_tiles(sr_fld *f, int *pre, int *post, int max_levels, int *levels);
int sr_fld_set_tiles(sr_fld *f, const int *pre, const int *post, int levels);
/* The last two tile-smoothed levels' pre-smoothers in one launch and their post-smoothers in another, when at least
 * three levels are tile-smoothed (level 0 always apart): each block redoes the finer level's smoothing over the cells
 * its coarser tile's halo reaches, the two levels' launches being latency-bound. The same bits as separate launches.
 * The autotuner picks per direction (SR_FUSE_OFF, _SMALL or _LARGE) on the first system built; sr_fld_set_fusion
 * fixes the choice (for both it stays off when fewer levels are tile-smoothed). No counterpart in the reference (its
 * V-cycle dispatches each level's passes, heat.wgsl:1401-1445). */
int sr_fld_fusion(sr_fld *f, int *pre, int *post);
int sr_fld_set_fusion(sr_fld *f, int pre, int post);
/* The single-block V-cycle's smallest levels (from the first of at most 16 x 16 cells, when one or two levels follow
 * it and the last is solved by its dense inverse) on one wave, neighbours by cross-lane shuffles and no block
 * barrier (on, the default), or on the whole block from LDS (off): the same values either way. No counterpart in the
 * reference (its one coarse workgroup, heat.wgsl:1485-1526). */
int sr_fld_set_coarse_wave(sr_fld *f, int on);
/* Host-decided solves: each PCG iteration's update (x += alpha p, r -= alpha q, ||r||^2 partials) folded into the next
 * preconditioner application's level-0 pre-smoother (on, the default, where level 0 has 32 x 32 tiles apart from the
 * fused pair and at most 2^20 + 2^18 cells: on larger grids the update streams faster on its own), or launched on
 * its own (off). x, r and z take the same values either way; the ||r||^2 partials are
 * summed per tile when folded. No counterpart in the reference (its update is a round of its chained dispatches,
 * heat.wgsl:605-877). */
int sr_fld_set_update_fusion(sr_fld *f, int on);
void sr_fld_destroy(sr_fld *f);
/* The learned preconditioner (SR_PRECOND_LEARNED): z = V r + N(r - A V r), V the multigrid V-cycle and N the
 * network. With fallback_iterations > 0 the solve continues with plain multigrid from where it stands once the
 * learned preconditioner has used that many iterations, or earlier when after 4 iterations it converges more slowly
 * per iteration than the last plain multigrid solve of the same system; 0: never. No counterpart in the reference (its set_heat_precond, step/heat.rs:595, offers
 * Jacobi and multigrid); n may be NULL to detach. Grid sides must be multiples of 4. */
int sr_fld_set_network(sr_fld *f, sr_net *n, int fallback_iterations);
/* sr_fld_solve, also reporting the iterations made with the learned preconditioner and whether it gave way to plain
 * multigrid (after fallback_iterations, or when after 4 iterations it converges more slowly per iteration than the
 * last plain multigrid solve of the same system did) */
int sr_fld_solve_stats(sr_fld *f, const double *b_dev, double *x_dev, double tol, int max_iter, int precond,
                       int *iterations, double *final_rel, int *learned_iterations, int *fell_back, void *stream);

/* ----------------------------------------------------------------------------- the learned preconditioner's network */
/* A 3-level U-Net (16/32/64 channels, 3x3 convolutions, BF16 activations and weights, FP32 accumulation on the
 * matrix cores) mapping (residual, log coupling, diagonal share) to a correction, with its backward pass and Adam on
 * the device. No counterpart in the reference. Parameters are FP32 in the order of the layers e1a, e1, e2a, e2,
 * e3a, e3, d2a, d2, d1a, d1 (each weights (3, 3, cin, cout) then biases), then the 1x1 output's 16 weights and bias. */
int sr_net_create(int max_w, int max_h, sr_net **out);
int sr_net_param_count(sr_net *n, size_t *count);
int sr_net_set_params(sr_net *n, const float *params, size_t count); /* host array; resets Adam's moments */
int sr_net_get_params(sr_net *n, float *params, size_t count);
int sr_net_get_grads(sr_net *n, float *grads, size_t count);        /* of the last sr_net_loss_grad */
/* z = N(r) scaled to the system's units, for a w x h system (own, kx, ky as sr_fld_set_system's, f32 device) */
int sr_net_apply(sr_net *n, int w, int h, const float *own_dev, const float *kx_dev, const float *ky_dev,
                 const float *r_dev, float *z_dev, void *stream);
/* the energy-norm loss (z - e)^T A (z - e) / e^T A e of z = N(r) against e = A^-1 r, and every parameter's gradient */
int sr_net_loss_grad(sr_net *n, int w, int h, const float *own_dev, const float *kx_dev, const float *ky_dev,
                     const float *r_dev, const float *e_dev, double *loss, void *stream);
int sr_net_adam(sr_net *n, double lr, void *stream);   /* one Adam step with the last gradients */
/* one activation of the last forward pass as f32 (names e1a e1 p1 e2a e2 p2 e3a e3 d2a d2 d1a d1 x), channels-last */
int sr_net_activation(sr_net *n, const char *name, float *out, size_t count);
void sr_net_destroy(sr_net *n);

/* ------------------------------------------------------------------------------------------------ the world / step */
/* Step::with_heat (step/mod.rs:415-527) + State::new (state.rs:139): refuses an axis not a multiple of 8 in
 * [8, 4096] (StateError::WorldSize) and a padded FFT axis past max_padded (FftError::BoxAxis). */
int sr_world_create(int w, int h, const sr_physics *ph, const sr_switches *sw, const sr_heat_settings *hs,
                    sr_world **out);
/* State::upload (state.rs:251): 14 canonical f32 planes, channel-major, row-major; n = 14 w h. Resets the clock
 * and solves P2 on the state for the first dt (reference/mod.rs:173-175). */
int sr_upload(sr_world *wd, const float *planes, size_t n);
/* State::readback (state.rs:273) */
int sr_readback(sr_world *wd, float *planes, size_t n);
/* The step's records a state dump carries beyond the cells (crates/sr-engine/src/dump.rs:8-20, :55-70 `Resume`),
 * which load_dump restores so that a resumed run goes on bit for bit: dt_n, dt_{n+1} and the sim time. */
typedef struct {
    double dt, dt_next, sim_time;
} sr_resume_records;
/* load_dump's restore after State::upload (dump.rs:8-20; Step::set_heat_rate and side_flux, step/heat.rs:567-598):
 * the clock from `rec`; with heat on and `flux` given, F's two planes (F_x then F_y, 2 w h f32, flux_n = 2 w h) that
 * the next P4 pushes the gas with (NULL: none, as after an upload); and `sigma` (w h f64, sigma_n = w h, may be NULL)
 * in place of the uploaded Sigma plane: the library carries Sigma in FP64 and a dump's f32 plane is that value
 * rounded, so resuming bit for bit needs the FP64 plane (sr_readback_sigma). The rest of the step (phi, g, the heat
 * pass's planes) is rebuilt by the next step from these. The step count stays the upload's (0). */
int sr_resume(sr_world *wd, const sr_resume_records *rec, const float *flux, size_t flux_n, const double *sigma,
              size_t sigma_n);
/* Sigma's FP64 plane (w h values), the state's own precision */
int sr_readback_sigma(sr_world *wd, double *sigma, size_t n);
/* Step::run (step/mod.rs:703): `steps` steps, P1 -> P2 -> P4 -> P5 -> P8 -> P9 (reference/mod.rs:264-297) */
int sr_run(sr_world *wd, uint32_t steps);
/* A batch of independent worlds, each `steps` steps as sr_run would take them (every world's result is the one it
 * reaches alone). No counterpart in the reference, where a Step drives one world (step/mod.rs:367-527, :703-718) and
 * headless runs one scene (headless.rs:1087-1260): the batched-ensemble workload of many uneven scenes. The worlds
 * are shared out over `threads` host threads (1..256), largest first onto the least-loaded; each world is used by one
 * thread during the call, so no world may appear twice or be used elsewhere meanwhile. The first error stops the
 * batch and is returned. */
int sr_ensemble_run(sr_world **worlds, int count, uint32_t steps, int threads);
int sr_read_dt(sr_world *wd, sr_dt_report *out);              /* Step::read_dt (step/mod.rs:691) */
int sr_read_heat(sr_world *wd, sr_heat_report *out);          /* read_heat (step/heat.rs:499) */
int sr_read_field(sr_world *wd, int field, float *out, size_t n); /* side fields: phi, g, F (step/heat.rs:549) */
/* FP64 totals of the 14 conserved quantities (Sigma, Sigma u, Sigma v, E, each species' mass) and what has left
 * through the edges (reference/hydro.rs:210-224, :954-976) */
int sr_read_totals(sr_world *wd, double *totals, double *escaped);
int sr_set_precond(sr_world *wd, int precond);                /* set_heat_precond (step/heat.rs:595) */
/* heat_first_system (step/heat.rs:604-640): the system of the last step's first Newton pass (its iterate at the start
 * temperatures T0, rebuilt from the planes and faces the step leaves, with its multigrid hierarchy), level 0 read back
 * in sr_fld_set_system's form: own (w h; couplings to non-gas cells and the edge folded in, 1 off the gas), kx
 * ((w+1) h) and ky (w (h+1)) the gas-gas couplings, gas (w h bytes, 1 gas). n = w h. Host arrays (any may be NULL).
 * Needs a step run with heat since the upload (SR_ERR_ARGUMENT otherwise). */
int sr_heat_first_system(sr_world *wd, float *own, float *kx, float *ky, uint8_t *gas, size_t n);
/* apply_heat_precond (step/heat.rs:645-673): z = M^-1 r by the world's preconditioner (sr_set_precond) on the system
 * sr_heat_first_system last built; r and z host arrays of w h f32, z 0 off the gas. */
int sr_apply_heat_precond(sr_world *wd, const float *r, float *z, size_t n);
/* the network SR_PRECOND_LEARNED uses in the world's heat pass (no counterpart in the reference) */
int sr_set_network(sr_world *wd, sr_net *n, int fallback_iterations);
/* the heat pass's solves: SR_SOLVE_HOST (the default), SR_SOLVE_DEVICE or SR_SOLVE_GRAPH */
int sr_set_solve_mode(sr_world *wd, int mode);
/* the heat pass's fused smoother pair, as sr_fld_set_fusion */
int sr_set_fusion(sr_world *wd, int pre, int post);
/* the heat pass's update folded into the next pre-smoother, as sr_fld_set_update_fusion */
int sr_set_update_fusion(sr_world *wd, int on);
/* P6 burning (crates/sr-physics/src/reference/react.rs; the rate law crates/sr-physics/src/rates.rs; the records
 * assets/reactions.json resolved against physics.json, registry.rs:485-535, :680-753). One enabled record: its inputs
 * (species, share, order), outputs (species, share), group (H He C Ne O Si N, registry.rs:460), q = q_ratio Q_H, the
 * law omega = A Sigma^a prod X_i^n_i f(T) g(Y_e Sigma) with f's Gamow-shaped terms (T_k, nu, w) above T_thr (no terms:
 * f = 1) and the gate ((Y_e Sigma / Sigma_g - 1)^p above Sigma_g), and whether its rate sets the collapse latch. The
 * species are assets/elements.json's order: H He C O Ne Mg Si S Fe n. */
#define SR_REACT_RECORDS 16
#define SR_REACT_GROUPS 7
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
} sr_react_record;
/* The pass: the enabled records in the registry's order, burn_dx, burn_dt, K_burn, the cooling's A_nu, T_nu, m, the
 * scene's neutrinos switch (the cooling) and pin_temperature, and the record stepped in closed form (N_Fe: one input
 * of order 1 into neutrons, no temperature law, p = 1; -1 none; react.rs:197-219). */
typedef struct {
    int n_records;
    sr_react_record rec[SR_REACT_RECORDS];
    double burn_dx, burn_dt, burn_k, a_nu, t_nu, m_nu;
    int neutrinos, pin_temperature, nfe;
} sr_react_params;
/* What the last pass did (react.rs:120-157's Pass, summed over the world): energy per area booked per group, the
 * neutrino loss, the accepted and retried sub-steps, the most in one cell, the cells that ran out of attempts, the
 * largest change of a fraction and of T in one accepted sub-step, and the latch (an N_Fe rate > 0 at a pass's start) */
typedef struct {
    double nuclear[SR_REACT_GROUPS];
    double neutrino_lost;
    uint64_t substeps, retried, max_substeps, unfinished;
    double max_dx, max_dt_rel;
    int latch;
} sr_react_report;
/* P6 in the world's step, between P5 and P8 (reference/mod.rs:263-297), with these records (NULL: off, the default);
 * the scene's `reactions` switch (Step::with_reactions, step/mod.rs:427-440) */
int sr_set_reactions(sr_world *wd, const sr_react_params *p);
/* P6 alone over dt on the world's state (the reference's rig, tests/gpu/burn.rs:120-235) */
int sr_react_apply(sr_world *wd, double dt);
int sr_read_react(sr_world *wd, sr_react_report *out);
/* the last pass's per-cell energy booked per group (7 w h, group-major), neutrino loss and S_nu (w h each) */
int sr_read_react_planes(sr_world *wd, double *nuclear, double *lost, double *s_nu, size_t n);

/* P3 sinks (reference/sinks.rs; Step::with_sinks, step/sinks.rs): point masses in cells (cell (i, j) centred at
 * (i + 1/2, j + 1/2)), at most SR_MAX_SINKS. Each step with P3 runs, after P2 and before P4: formation at the densest
 * cell meeting all four criteria (Sigma >= Sigma_BH, or a local Sigma maximum whose disk r <= 4 holds m_tov_sb with a
 * disk r <= 3 at Y_e <= 0.40; converging; the potential's minimum among its 8 neighbours; bound; the world's outer
 * ring never forms; one a step), accretion of the bound gas within each sink's r_acc = max(2, 2GM/c_sb^2) in index
 * order (the share 1 - Sigma_floor / Sigma of a cell's Sigma, momentum and E; its thermal and cold energy booked
 * `swallowed`), the kick-drift-kick by the gas's softened pull and the pairs' (softened by the larger r_acc), then
 * merging of pairs closer than the larger r_acc, the lowest pair first. P2 adds the sinks' softened potential to phi
 * and their exact pull to g (reference/grav.rs:79-96), and the half step's field the same part (hydro.rs:836-846). */
#define SR_MAX_SINKS 16
typedef struct {
    double x, y, vx, vy, m;
} sr_sink;
/* Sigma_BH, and m_tov_sb: the sheet's neutron-matter ceiling for the run's physics, which this library does not
 * compute (sinks.rs:104-115): the caller gives it (INFINITY: no sink from mass, the reference's own with the
 * neutrons' cold pressure off) */
typedef struct {
    double sigma_bh, m_tov;
} sr_sinks_params;
/* the last pass: the cell a sink formed at (-1, -1: none), the mass accreted, the sinks lost to merging; and the
 * thermal and cold energy swallowed since the sinks were last set */
typedef struct {
    int count, formed_x, formed_y, merged;
    double accreted, swallowed;
} sr_sinks_report;
/* P3 in the world's step (NULL: off, the default); the world's dx must be 1 (the sinks' law is in cells). Setting it
 * clears the sinks. */
int sr_set_sinks(sr_world *wd, const sr_sinks_params *p);
/* the sinks placed as given (a dump's header, a test's), count <= SR_MAX_SINKS; swallowed restarts at 0 */
int sr_put_sinks(sr_world *wd, const sr_sink *sinks, int count);
/* the sinks (up to max of them into out) and the last pass's report (Step::read_sinks, step/sinks.rs) */
int sr_read_sinks(sr_world *wd, sr_sink *out, int max, sr_sinks_report *rep);
/* P3 alone over dt on the world's state, against the potential phi (w h values, the cells' phi; NULL: the world's own
 * from its last P2, the sinks' term included): the reference's rig with phi uploaded by hand (tests/gpu/sink.rs) */
int sr_sinks_apply(sr_world *wd, double dt, const float *phi, size_t n);

/* The active box (step/boxfit.rs:1-14, BoxMode and BoxReport :27-55): the bounding box of the cells with Sigma >=
 * Sigma_vac, grown by 8 cells, rounded out to multiples of 8, clamped to the world (empty, all 0, with no such cell),
 * its FFT size per axis the smallest power of two >= 2 x the box's, at least 32. It is fit at the upload (before
 * step 0's dt) and after P9 at step indices = 0 (mod 16); P8's floors and P9's dt, wave-speed maximum and non-finite
 * guard run over its cells only, so vacuum outside it keeps its uploaded values. SR_BOX_WHOLE_WORLD (the default
 * here; the twin's world) holds it at the world on the same schedule; SR_BOX_FIT is the reference's Step::new
 * (step/mod.rs:368). Set before sr_upload. Hydro, gravity and the heat pass leave cells outside the box as they are
 * either way (every cell there and its neighbours are below Sigma_vac) and run over the world: gravity's FFT is the
 * world's, where the reference sizes it to the box (the same field up to round-off). */
enum { SR_BOX_WHOLE_WORLD = 0, SR_BOX_FIT = 1 };
typedef struct {
    uint32_t x0, y0, w, h; /* origin and size in cells, multiples of 8; all 0 when empty */
    uint32_t fft[2];       /* the FFT size per axis */
    uint32_t from_step;    /* the first step that runs inside this box (0: the upload's fit) */
    uint32_t fits;         /* fits written since the upload, its own included */
    uint32_t tiles[3];     /* the reference's indirect sizes over the box: 8 x 8 tiles, then 256-cell groups */
    uint32_t cells[3];
} sr_box_report;
int sr_set_box_mode(sr_world *wd, int mode);
int sr_read_box(sr_world *wd, sr_box_report *out);
/* load_dump's restore of the box (Step::restore, step/latch.rs:496-505): the record (x0, y0, w, h, fft_w, fft_h,
 * from_step, fits) a dump holds and the step index the refit schedule and the guard count from; after sr_resume */
int sr_resume_box(sr_world *wd, const uint32_t *record, uint64_t step);
int sr_set_dt_cap(sr_world *wd, double cap);                  /* Twin::cap_dt_next (reference/mod.rs:225) */
int sr_set_timing(sr_world *wd, int on);                      /* Step::set_timing (step/mod.rs:722) */
int sr_read_cost(sr_world *wd, sr_step_cost *out);            /* Step::read_cost (step/mod.rs:753) */
/* the work counters above; reset != 0 zeroes them after reading */
int sr_read_work(sr_world *wd, sr_step_work *out, int reset);
void sr_destroy(sr_world *wd);

/* Timing of a kernel over the steps run with timing on (since sr_set_timing switched it on): launches, mean device ms
 * per launch (HIP events on the world's stream around each launch), and the algorithmic HBM bytes of one launch.
 * `kernel`: "hydro_update", "hydro_mid" (P4), "mg_pre0", "mg_post0" (the level-0 smoothers), "mg_coarse" (the
 * single-block V-cycle), "fld_matvec", "fld_update" (the PCG iteration's two kernels), "fld_update_pre0" (the update
 * folded into the next level-0 pre-smoother, sr_set_update_fusion). No counterpart in the reference (its read_cost
 * times passes). */
int sr_kernel_time(sr_world *wd, const char *kernel, uint64_t *launches, double *mean_ms, double *bytes_per_launch);

#ifdef __cplusplus
}
#endif
#endif
