// P5 heat pass: the implicit FLD step's kernels and Newton driver.
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

#include "../../include/sr_hip.h"
#include "cold.hpp"
#include "fld.hpp"

namespace sr {

struct HeatParams {
    int w, h;
    double dx;
    float inv_dx;
    double gm1;             // gamma - 1
    double t_floor;
    float sigma_vac, sigma_floor;
    float fixed_imu;        // 1/mu when fixed, else 0
    double a_r, c_sb, k_cond;
    float kappa0, kappa_dust, t_dust;
    int opacity;            // 0 reference, 1 Kramers
    int cd;                 // constant_diffusivity on
    double chi;
    int limiter;
    int pinned;
    ColdEos cold;           // the cold energy comes off E before the temperature is read
};

struct HeatPlanes {
    uint8_t *gas = nullptr;
    double *t0 = nullptr, *t = nullptr; // the start temperature and the Newton iterate, FP64
    float *kappa = nullptr, *imu = nullptr, *cap = nullptr, *sig = nullptr;
    float *kx = nullptr, *ky = nullptr, *kvac = nullptr, *heat = nullptr, *slope = nullptr, *scale = nullptr;
};

// the state planes the pass reads (E is written)
struct HydroLike {
    const double *sigma;
    const float *mx, *my;
    float *e;
    const float *x;
};

// the heat pass's launches, counted as they are made (the work counters behind sr_read_work)
struct HeatWork {
    unsigned long long steps = 0, passes = 0, commits = 0, limited = 0;
};

class HeatPass {
  public:
    ~HeatPass();
    int init(const HeatParams &p, const sr_heat_settings &hs);
    // one P5 step on the state (E in place); writes the cell-centred flux the next P4 reads
    int step(const HydroLike &st, const double *scal, double dt_host, float *fx, float *fy, int precond,
             hipStream_t s);
    // the last step's first Newton pass's system (its iterate at T0) rebuilt into the solver's level 0 and hierarchy
    // from the planes and faces the step left (clobbers the last iterate, which no later pass reads)
    int first_system(const double *scal, hipStream_t s);
    sr_heat_report report{};
    HeatWork work;
    // algorithmic HBM bytes of the counted work outside the linear solver
    double work_bytes(const HeatWork &w) const;
    FldSolver solver;
    HeatParams P{};
    sr_heat_settings hs{};

  private:
    int read_sums(int k, double *out, hipStream_t s);
    // the partials' copy queued with an event (copy_sums), work queued behind it, then the host's wait for the copy
    // alone and the sums in block order (wait_sums)
    // into one of two pinned slots (4 nb doubles each), so a second copy can be in flight
    int copy_sums(int k, hipStream_t s, int slot = 0);
    int wait_sums(int k, double *out, int slot = 0);
    hipEvent_t ev_sums[2] = {nullptr, nullptr};
    HeatPlanes H;
    std::vector<int> pass_rounds;  // the last step's solve iterations per Newton pass: the next step's first bursts
    double *rhs = nullptr, *xs = nullptr, *part = nullptr;
    double *host = nullptr;
    int nb = 1;
};

} // namespace sr
