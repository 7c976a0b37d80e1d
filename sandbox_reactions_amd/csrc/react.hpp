// P6 burning: the records and constants on the device, the pass's per-cell kernel.
#pragma once
#include <hip/hip_runtime.h>

#include "../../include/sr_hip.h"
#include "cold.hpp"
#include "hydro.hpp"

namespace sr {

#define REACT_NS 10
#define REACT_TABLE 256

// Everything the pass reads besides the state, in device memory: the records and constants as given, each record's
// nu (outputs' shares minus inputs'), its ln f table, Sigma_vac, T_floor and the cold energy's tables in FP64 (the
// electrons' and the neutrons' ln u at the nodes, a pair's flag 0 when it is off; ln x_min and the node spacing in
// ln x), so that T = mu (e - eps_cold) carries no f32 rounding into the rates
struct ReactDev {
    sr_react_params P;
    double nu[SR_REACT_RECORDS][REACT_NS];
    double ln_f[SR_REACT_RECORDS][REACT_TABLE];
    double sigma_vac, t_floor;
    double cold_lnu[2][COLD_POINTS];
    int cold_pair[2];
    double cold_ln_min, cold_step;
};

// the host's checks of the records and the device copy's build (tables in FP64 on the host) for the world's physics
int react_check(const sr_react_params *p);
void react_build(const sr_react_params *p, const sr_physics &ph, ReactDev *out);

// P6 over dt (dt < 0: the step's dt, scal[0]) on the world's state; per cell the energy booked per group (7 n), the
// neutrino loss and S_nu (n each); per block partials: sums (substeps, retried, unfinished, nuclear x 7, lost) into
// psum[11][nb], maxima (max sub-steps, max dX, max dT/T, latch) into pmax[4][nb]
constexpr int REACT_NT = 64, REACT_BLOCKS = 4096, REACT_SUMS = 11, REACT_MAXES = 4;
int react_launch(const ReactDev *D, const HydroState &st, int n, double dt, const double *scal, double *nuclear,
                 double *lost, double *snu, double *psum, double *pmax, int *nblocks, hipStream_t s);

} // namespace sr
