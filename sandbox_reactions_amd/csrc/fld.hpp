// P5 linear solver: FP32 multigrid-preconditioned flexible CG inside FP64 iterative refinement.
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

#include "../../include/sr_hip.h"

namespace sr {

class Net;
struct CoarseSet;
struct PendingUpdate;
struct SolveCtl;

// How a solve decides when to stop: on the host from read-back partials (the learned preconditioner always), or on
// the device by each slot's head, the slots launched one by one or replayed as a captured graph of two
enum { SOLVE_HOST = SR_SOLVE_HOST, SOLVE_DEVICE = SR_SOLVE_DEVICE, SOLVE_GRAPH = SR_SOLVE_GRAPH,
       SOLVE_HOST_WAIT = SR_SOLVE_HOST_WAIT };

// tile shapes of the level smoothers
enum { TILE_16 = SR_TILE_16, TILE_32 = SR_TILE_32, TILE_64 = SR_TILE_64 };
// the last two tile-smoothed levels' smoothers: separate launches, or fused with small or large tiles
enum { FUSE_OFF = SR_FUSE_OFF, FUSE_SMALL = SR_FUSE_SMALL, FUSE_LARGE = SR_FUSE_LARGE };

// One level of the hierarchy: (A v)_i = own_i v_i + sum_f a_f (v_i - v_j); ax is (w+1) x h (face x between cells
// x-1 and x, 0 at the edges), ay is w x (h+1).
struct FldLevel {
    int w = 0, h = 0;
    float *own = nullptr, *ax = nullptr, *ay = nullptr, *z = nullptr, *b = nullptr, *zp = nullptr;
    // the tile smoothers' per-cell record {own, left face, up face, 1 / (own + all four faces)}, levels handled
    // by the tile kernels only
    float4 *cf = nullptr;
};

struct FldStats {
    int iterations = 0, refinements = 0;
    double rel = 0.0;      // ||b - A x|| / ||b|| in FP64
    bool converged = false;
    bool finite = true;
    bool fell_back = false;  // the learned preconditioner gave way to plain multigrid
    int learned_iterations = 0;
};

// The reference's energy-weighted stopping rule (crates/sr-physics/src/reference/heat.rs:424-431): r^T D^-1 r <=
// u_bound and sum heat_i (r_i slope_i / D_i)^2 <= t_bound, D the diagonal own + sum of couplings.
struct RefStop {
    const float *heat = nullptr, *slope = nullptr;
    double u_bound = 0.0, t_bound = 0.0;
};

// launches counted as they are made (the work counters behind sr_read_work)
struct FldWork {
    unsigned long long iterations = 0, precond = 0, true_residuals = 0, builds_faces = 0, builds_own = 0;
};

class FldSolver {
  public:
    ~FldSolver();
    // omega scales the coarse-grid correction (1: plain prolongation; SPD for 0 < omega < 2)
    int init(int w, int h, int levels, int nu1, int nu2, double omega = 1.0);
    // level 0's operator from device arrays (own, couplings ax/ay, gas mask for the coarsening); builds the
    // Galerkin hierarchy. own/ax/ay are copied.
    int set_system(const float *own, const float *ax, const float *ay, const uint8_t *gas, hipStream_t s);
    // the operator is level 0's own arrays, filled in place by a caller kernel; then rebuild() the coarse levels
    float *own0() { return lv[0].own; }
    float *ax0() { return lv[0].ax; }
    float *ay0() { return lv[0].ay; }
    uint8_t *gas0() { return gas; }
    int rebuild(bool faces, hipStream_t s);
    // A x = b to ||b - Ax|| <= tol ||b|| (or the reference rule when ref != nullptr); x holds the start
    // `expect`: the iterations a similar solve took (the same Newton pass of the last step), the device-decided
    // solve's first burst of slots; 0 when unknown
    int solve(const double *b, double *x, double tol, int max_iter, int precond, const RefStop *ref, FldStats *st,
              hipStream_t s, int expect = 0);
    int apply_precond(const float *r, float *z, int precond, hipStream_t s);
    void set_solve_mode(int m) { solve_mode = m; }
    int get_solve_mode() const { return solve_mode; }
    // the network of the learned preconditioner (z = V r + N(r - A V r)); after `fallback` iterations without
    // convergence (0: never) the solve continues with plain multigrid
    int set_network(Net *n, int fallback);
    int matvec64(const double *v, double *out, hipStream_t s);
    // tile shapes per tile-smoothed level (the levels above the single-block ones): read, or fixed by the caller
    // (which turns the autotuner off). Level 0's post-smoother keeps 32x32: its dot partials are counted per tile.
    int tiles(int *pre, int *post, int max_levels, int *levels) const;
    int set_tiles(const int *pre, const int *post, int levels);
    // the fused pair (levels coarse_from - 2 >= 1 and coarse_from - 1, one launch per direction; no result bit depends
    // on it): the autotuner's choice per direction, or the caller's
    int fusion(int *pre, int *post) const;
    int set_fusion(int pre, int post);
    // the single-block V-cycle's smallest levels on one wave with no block barrier (on) or on the whole block from
    // LDS (off); the same values either way
    void set_update_fusion(bool on) { fuse_update = on; }
    bool update_fusion() const { return fuse_update; }
    void set_coarse_wave(bool on) {
        coarse_wave = on;
        tiles_version++;
    }
    int nlevels() const { return (int)lv.size(); }
    int width() const { return w0; }
    int height() const { return h0; }
    double bytes_per_iteration() const;
    // algorithmic HBM bytes of the counted work (level 0 and the hierarchy; per fine cell as the kernels' comments
    // state)
    double work_bytes(const FldWork &w) const;
    FldWork work;
    // kernel clocks (timing on): event pairs recorded on the stream around every launch of the level-0 pre- and
    // post-smoother, the single-block V-cycle, the matvec and the update, without waiting; summed per kernel when read
    // (clock), so timing does not stall the solve
    enum { CLK_PRE0, CLK_POST0, CLK_COARSE, CLK_MATVEC, CLK_UPDATE, CLK_UPDATE_PRE0, CLK_N };
    void set_timing(bool on) { timing = on; }
    int clock(int kernel, double *ms, unsigned long long *launches);
    void reset_clocks();
    // algorithmic HBM bytes of one launch of a clocked kernel
    double clock_bytes(int kernel) const;
    bool timing = false;

  private:
    int w0 = 0, h0 = 0, nu1 = 2, nu2 = 2, coarse_from = 0;
    int head = 0, ltx = 0;  // the single-block V-cycle's register-resident levels (0 or 2), log2 of its layout's width
    float omega = 1.0f;
    int nbe = 1, ntiles0 = 1, nbd_max = 1; // blocks of the elementwise kernels, level-0 tiles
    // the tile shape (TILE_16, TILE_32, TILE_64) of each tile-smoothed level's pre- and post-smoother: the autotuner's
    // choice at the first rebuild, or the caller's (set_tiles)
    std::vector<int> pre_tile, post_tile;
    bool tuned = false, autotune = true;
    int fuse_pre = FUSE_SMALL, fuse_post = FUSE_SMALL;
    bool coarse_wave = true;  // the single-block V-cycle's levels of at most 16 x 16 cells on one wave
    // host-decided solves: each PCG update folded into the next V-cycle's level-0 pre-smoother (needs 32 x 32 tiles
    // at level 0 apart from the fused pair)
    bool fuse_update = true;
    bool update_fusable() const;
    bool fuse_fixed = false;
    // the finer level of the fused pair, or -1 when fewer than three levels are tile-smoothed (level 0 stays apart)
    int pair_level() const { return coarse_from >= 3 ? coarse_from - 2 : -1; }
    int launch_pre_pair(int a, const float *ba, hipStream_t s, const int *gate = nullptr);
    int launch_post_pair(int a, hipStream_t s, const int *gate = nullptr);
    int tune(hipStream_t s);
    int launch_pre_level(int l, const float *b, hipStream_t s, const int *gate = nullptr);
    int launch_post_level(int l, const float *b, float *zout, const float *zold, double *part, int *nblocks,
                          hipStream_t s, const int *gate = nullptr);
    std::vector<FldLevel> lv;
    uint8_t *gas = nullptr;
    // level-0 vectors
    float *r32 = nullptr, *r32b = nullptr, *za = nullptr, *zb = nullptr, *p = nullptr, *q = nullptr;
    float *pack = nullptr; // the single-block levels' own/ax/ay back to back, the coarsest's inverse, 1/diag
    int pack_floats = 0;
    // per-block partial sums (summed by the kernel that consumes them, or on the host)
    double *part_true = nullptr, *part_pq = nullptr, *part_rr = nullptr, *part_rr_t = nullptr, *part_dots = nullptr;
    double *host_part = nullptr; // pinned
    // the device-decided solve: two copies of its state (by slot parity), the head's ticket counter, a pinned copy for
    // the host's checks; the captured pair of slots and what it was captured for
    int solve_mode = SOLVE_HOST;
    SolveCtl *ctl = nullptr, *host_ctl = nullptr;
    unsigned *ticket = nullptr;
    hipGraphExec_t slot_exec = nullptr;
    struct GraphKey {
        const double *b = nullptr;
        const double *x = nullptr;
        const float *heat = nullptr, *slope = nullptr;
        int precond = -1, tiles = -1;
        hipStream_t s = nullptr;
    } slot_key;
    int tiles_version = 0;
    int last_iterations = 0;  // of the last device-decided solve: where the next one first looks
    int launch_slot(int parity, const double *b, double *x, const RefStop *ref, int precond, hipStream_t s);
    int solve_slots(const double *b, double *x, double tol, int max_iter, int precond, const RefStop *ref,
                    FldStats *st, hipStream_t s, int expect);
    std::vector<hipEvent_t> evs;   // pairs: start, end
    std::vector<int> ev_kernel;    // the kernel each recorded pair times
    size_t ev_used = 0;
    double clock_ms[CLK_N] = {};
    unsigned long long clock_n[CLK_N] = {};
    int tick(int kernel, hipStream_t s);  // records a pair's start when timing; returns its slot (or -1)
    int tock(int slot, hipStream_t s);
    int fold_clocks();
    Net *net = nullptr;
    int fallback_it = 0;
    double plain_rate = 0.0;  // ln(residual reduction) per iteration of the last converged plain multigrid solve
    // `upd`: the pending update folded into level 0's pre-smoother (r is then the new residual's buffer); `stage`:
    // 0 the whole V-cycle, 1 its first launch only, 2 the rest after it
    int vcycle(const float *r, float *z, const float *zold, double *dots, int *nbd, hipStream_t s,
               const int *gate = nullptr, const PendingUpdate *upd = nullptr, int stage = 0);
    CoarseSet coarse_set() const;
    int learned(const float *r, float *z, const float *zold, double *dots, hipStream_t s);
    int read_partials(const double *part, int k, int nb, double *out, hipStream_t s);
    int copy_partials(const double *part, int k, int nb, hipStream_t s);
    int wait_partials(int k, int nb, double *out);
    hipEvent_t ev_part = nullptr;
    // the hierarchy's rebuild captured per variant (without, with the faces) and the layout it was captured for
    hipGraphExec_t rebuild_exec[2] = {nullptr, nullptr};
    int rebuild_key[2] = {-1, -1};
};

} // namespace sr
