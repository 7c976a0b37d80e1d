// P4 hydro kernels' host-side types.
#pragma once
#include <hip/hip_runtime.h>

#include "cold.hpp"

namespace sr {

enum { HYDRO_MID = 0, HYDRO_UPDATE = 1 };

// Constants of the scheme in the device's f32 (gamma, the floors, the opacity law).
struct HydroParams {
    int w, h;
    float gamma, gm1;      // gamma, gamma - 1
    float t_floor_g;       // T_floor / (gamma - 1): eps_floor = t_floor_g / mu
    float sigma_vac, sigma_floor;
    float fixed_imu;       // 1/mu when fixed, else 0 (from the species)
    int edge;              // 0 bounce, 1 leave
    float inv_dx;
    float kappa0, c_sb, kappa_dust, t_dust;
    int opacity;           // 0 reference, 1 Kramers
    ColdEos cold;          // the cold pressure pairs (off: on == 0)
};

// The state on the device: Sigma in FP64, the momenta and E in FP32, ten species fractions (planes back to back).
struct HydroState {
    double *sigma;
    float *mx, *my, *e, *x;
};

struct HydroPartials {
    double *sums;              // 5 per block: net kick x, y; net push x, y; sum |push|
    unsigned long long *keys;  // per block: the densest free cell's key (Sigma_mid bits, ~index)
};

// The sinks' part of the half step's field (reference/hydro.rs:836-846): their softened potential on the ringed grid of
// phi_mid ((w + 2) x (h + 2)) and their exact pull per cell; all null without sinks
struct SinkPlanes {
    const float *ps, *gsx, *gsy;
};

// `sinks` (UPDATE only; null or all-null planes: none): the kick takes the sinks' exact pull beside the gas's g_mid,
// the net kick taken back is the gas's own, and the faces' work reads phi_mid with the sinks' potential
int hydro_launch(int mode, const HydroParams &P, const HydroState &in, const HydroState &out, const double *scal,
                 const float *gx, const float *gy, const float *sigma_mid, const float *phi_mid, const float *rfx,
                 const float *rfy, float *sigma_mid_out, float *pushmag, double *escaped, const HydroPartials &part,
                 hipStream_t s, const SinkPlanes *sinks = nullptr);
int hydro_takeback(const HydroParams &P, const HydroState &st, const float *pushmag, const HydroPartials &part,
                   double *takeback, bool radiation, hipStream_t s);
int hydro_blocks(int w, int h);

} // namespace sr
