// Error reporting, device queries and defaults of the C ABI.
#include <hip/hip_runtime.h>

#include <cstring>
#include <string>

#include "../../include/sr_hip.h"

namespace sr {

static thread_local std::string g_last_error;

void set_error(const std::string &msg) { g_last_error = msg; }

int fail(int code, const std::string &msg) {
    g_last_error = msg;
    return code;
}

} // namespace sr

extern "C" const char *sr_last_error(void) { return sr::g_last_error.c_str(); }

extern "C" int sr_device_info(char *name, size_t name_len, int *cu_count, size_t *hbm_bytes) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return sr::fail(SR_ERR_DEVICE, "no HIP device");
    hipDeviceProp_t p;
    if (hipGetDeviceProperties(&p, dev) != hipSuccess) return sr::fail(SR_ERR_DEVICE, "hipGetDeviceProperties");
    if (name && name_len) {
        std::strncpy(name, p.gcnArchName, name_len - 1);
        name[name_len - 1] = 0;
    }
    if (cu_count) *cu_count = p.multiProcessorCount;
    if (hbm_bytes) *hbm_bytes = p.totalGlobalMem;
    return SR_OK;
}

// The reference's assets/physics.json values the path reads, gamma = 2, dx = 1, G = 1, the reference opacity.
extern "C" void sr_physics_defaults(sr_physics *ph) {
    if (!ph) return;
    std::memset(ph, 0, sizeof *ph);
    ph->gamma = 2.0;
    ph->G = 1.0;
    ph->dx = 1.0;
    ph->sigma_vac = 1e-6;
    ph->sigma_floor = 1e-8;
    ph->t_floor = 0.05;
    ph->cfl = 0.4;
    ph->eta_g = 0.3;
    ph->a_r = 6.25e-7;
    ph->c_sb = 12800.0;
    ph->kappa0 = 1.0;
    ph->k_cond = 0.0;
    ph->kappa_dust = 0.0;
    ph->t_dust = 0.0;
    ph->opacity = SR_OPACITY_REFERENCE;
    ph->mu = 0.0;
    ph->edge = SR_EDGE_LEAVE;
    ph->pin_temperature = 0;
    ph->constant_diffusivity = 0.0;
    ph->flux_limiter = 1;
    ph->max_padded = 4096;
}

// The reference GPU's heat tolerances (crates/sr-engine/shaders/heat.wgsl:355-360) and ceilings (heat.rs:55-56),
// V(2,2) multigrid down to 4x4.
extern "C" void sr_heat_settings_defaults(sr_heat_settings *hs) {
    if (!hs) return;
    std::memset(hs, 0, sizeof *hs);
    hs->stop = SR_STOP_REFERENCE;
    hs->solve_tol = 1e-5;
    hs->newton_heat = 1e-6;
    hs->newton_cell = 1e-3;
    hs->max_passes = 128;
    hs->max_rounds = 4096;
    hs->levels = 0;
    hs->nu1 = 2;
    hs->nu2 = 2;
    hs->coarse_scale = 1.0;
}
