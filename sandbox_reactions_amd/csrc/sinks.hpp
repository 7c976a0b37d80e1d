// P3 sinks on the device: the sinks' state, their part of P2's field and the pass's launches.
#pragma once
#include <hip/hip_runtime.h>

#include "../../include/sr_hip.h"
#include "cold.hpp"
#include "hydro.hpp"

namespace sr {

// The sinks as the device holds them (up to SR_MAX_SINKS, slots 0..count-1) and what the last P3 did; a pass reads
// and writes it on the device only, so a step with P3 needs no host round trip.
struct SinkDev {
    int count;
    int formed_x, formed_y, merged;  // -1, -1 when none formed
    double accreted, swallowed;      // this pass's accretion; the thermal and cold energy taken since the sinks were set
    double x[SR_MAX_SINKS], y[SR_MAX_SINKS], vx[SR_MAX_SINKS], vy[SR_MAX_SINKS], m[SR_MAX_SINKS];
};

// What P3 reads besides the state: the world's size and constants, Sigma_BH and m_tov_sb
struct SinkParams {
    int w, h;
    double g, sigma_bh, sigma_vac, sigma_floor, c_sb, m_tov, t_floor, gm1, fixed_imu;
    ColdEos cold;
};

// Per-world device buffers of the pass
struct SinkBufs {
    SinkDev *dev;
    double *cand_s;  // per tile: the best candidate's Sigma (-1: none)
    int *cand_i;     // per tile: its cell
    double *pull;    // per block and sink: the gas's pull on the sink, x and y
    float *ps;       // the sinks' potential on the ringed grid, (w + 2) (h + 2)
    float *gsx, *gsy;  // their exact pull per cell
    int tiles, pull_blocks;
};

int sinks_alloc_sizes(int w, int h, int *tiles, int *pull_blocks);
// P2's part: the sinks' potential added to phi (ringed) and their exact pull to g_x, g_y; the same terms kept in the
// planes the half step's field adds (zero with no sink)
int sinks_field(const SinkParams &K, const SinkBufs &B, float *phi, float *gx, float *gy, hipStream_t s);
// formation against the potential phi (ringed), accretion, the kick-drift-kick, merging; dt < 0: the step's, scal[0]
int sinks_pass(const SinkParams &K, const SinkBufs &B, const HydroState &st, const float *phi, double dt,
               const double *scal, hipStream_t s);

} // namespace sr
