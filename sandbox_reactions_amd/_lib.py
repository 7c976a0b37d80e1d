"""ctypes declarations of include/sr_hip.h (the library is libsrhip.so in this directory)."""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsrhip.so")

SR_OK = 0
ERRORS = {
    -1: "WorldSize", -2: "BoxAxis", -3: "LengthMismatch", -4: "Device", -5: "Argument", -6: "NotPowerOfTwo",
    -7: "NonFinite",
}
SR_ERR_WORLD_SIZE, SR_ERR_BOX_AXIS, SR_ERR_LENGTH, SR_ERR_DEVICE, SR_ERR_ARGUMENT, SR_ERR_NOT_POW2, \
    SR_ERR_NON_FINITE = -1, -2, -3, -4, -5, -6, -7

EDGE_BOUNCE, EDGE_LEAVE = 0, 1
OPACITY_REFERENCE, OPACITY_KRAMERS = 0, 1
PRECOND_JACOBI, PRECOND_MULTIGRID, PRECOND_LEARNED = 0, 1, 2
STOP_REFERENCE, STOP_RELATIVE = 0, 1
TILE_16, TILE_32, TILE_64 = 0, 1, 2
SOLVE_HOST, SOLVE_DEVICE, SOLVE_GRAPH, SOLVE_HOST_WAIT = 0, 1, 2, 3
FUSE_OFF, FUSE_SMALL, FUSE_LARGE = 0, 1, 2
BOX_WHOLE_WORLD, BOX_FIT = 0, 1
FIELD_PHI, FIELD_GX, FIELD_GY, FIELD_FLUX_X, FIELD_FLUX_Y, FIELD_SIGMA_MID, FIELD_PHI_MID = range(7)
COMPONENTS = 14


class Physics(C.Structure):
    _fields_ = [
        ("gamma", C.c_double), ("G", C.c_double), ("dx", C.c_double),
        ("sigma_vac", C.c_double), ("sigma_floor", C.c_double), ("t_floor", C.c_double),
        ("cfl", C.c_double), ("eta_g", C.c_double),
        ("a_r", C.c_double), ("c_sb", C.c_double), ("kappa0", C.c_double), ("k_cond", C.c_double),
        ("kappa_dust", C.c_double), ("t_dust", C.c_double),
        ("opacity", C.c_int), ("mu", C.c_double), ("edge", C.c_int),
        ("pin_temperature", C.c_int), ("constant_diffusivity", C.c_double), ("flux_limiter", C.c_int),
        ("max_padded", C.c_int),
        ("k1e", C.c_double), ("k2e", C.c_double), ("k1n", C.c_double), ("k2n", C.c_double),
    ]


class Switches(C.Structure):
    _fields_ = [("gravity", C.c_int), ("hydro", C.c_int), ("heat", C.c_int)]


class HeatSettings(C.Structure):
    _fields_ = [("stop", C.c_int), ("solve_tol", C.c_double), ("newton_heat", C.c_double),
                ("newton_cell", C.c_double), ("max_passes", C.c_int), ("max_rounds", C.c_int),
                ("levels", C.c_int), ("nu1", C.c_int), ("nu2", C.c_int), ("forcing", C.c_double),
                ("coarse_scale", C.c_double)]


class DtReport(C.Structure):
    _fields_ = [("dt", C.c_double), ("dt_next", C.c_double), ("sim_time", C.c_double), ("steps", C.c_uint64),
                ("non_finite", C.c_int), ("nf_x", C.c_int), ("nf_y", C.c_int), ("nf_channel", C.c_int),
                ("nf_step", C.c_uint64)]


class HeatReport(C.Structure):
    _fields_ = [("dt", C.c_double), ("passes", C.c_int), ("rounds", C.c_int), ("converged", C.c_int),
                ("delta", C.c_double), ("radiated", C.c_double), ("final_residual", C.c_double)]


class StepCost(C.Structure):
    _fields_ = [("steps", C.c_uint64), ("step_ms", C.c_double), ("pass_ms", C.c_double * 10)]


class StepWork(C.Structure):
    _fields_ = [(k, C.c_uint64) for k in ("steps", "hydro_steps", "convolutions", "heat_steps", "newton_passes",
                                          "pcg_iterations", "precond_applications", "true_residuals",
                                          "hierarchy_builds")] + \
               [(k, C.c_double) for k in ("algorithmic_bytes", "bytes_p2", "bytes_p4", "bytes_p5", "bytes_p89")]


class ResumeRecords(C.Structure):
    _fields_ = [("dt", C.c_double), ("dt_next", C.c_double), ("sim_time", C.c_double)]


REACT_RECORDS, REACT_GROUPS = 16, 7


class ReactRecord(C.Structure):
    """sr_react_record: one enabled record with its physics keys resolved."""
    _fields_ = [("n_in", C.c_int), ("n_out", C.c_int), ("in_species", C.c_int * 4), ("in_share", C.c_double * 4),
                ("in_order", C.c_double * 4), ("out_species", C.c_int * 4), ("out_share", C.c_double * 4),
                ("group", C.c_int), ("q", C.c_double), ("a_coef", C.c_double), ("a_sigma", C.c_double),
                ("n_terms", C.c_int), ("t_k", C.c_double * 4), ("nu", C.c_double * 4), ("w", C.c_double * 4),
                ("t_thr", C.c_double), ("gated", C.c_int), ("sigma_g", C.c_double), ("gate_p", C.c_double),
                ("latch", C.c_int)]


class ReactParams(C.Structure):
    """sr_react_params: the records and the pass's constants."""
    _fields_ = [("n_records", C.c_int), ("rec", ReactRecord * REACT_RECORDS), ("burn_dx", C.c_double),
                ("burn_dt", C.c_double), ("burn_k", C.c_double), ("a_nu", C.c_double), ("t_nu", C.c_double),
                ("m_nu", C.c_double), ("neutrinos", C.c_int), ("pin_temperature", C.c_int), ("nfe", C.c_int)]


class ReactReport(C.Structure):
    """sr_react_report: what the last P6 pass did, summed over the world."""
    _fields_ = [("nuclear", C.c_double * REACT_GROUPS), ("neutrino_lost", C.c_double), ("substeps", C.c_uint64),
                ("retried", C.c_uint64), ("max_substeps", C.c_uint64), ("unfinished", C.c_uint64),
                ("max_dx", C.c_double), ("max_dt_rel", C.c_double), ("latch", C.c_int)]


MAX_SINKS = 16


class Sink(C.Structure):
    """sr_sink: a sink's position (cells), velocity and mass."""
    _fields_ = [("x", C.c_double), ("y", C.c_double), ("vx", C.c_double), ("vy", C.c_double), ("m", C.c_double)]


class SinksParams(C.Structure):
    """sr_sinks_params: Sigma_BH and m_tov_sb (infinite: no sink from mass)."""
    _fields_ = [("sigma_bh", C.c_double), ("m_tov", C.c_double)]


class SinksReport(C.Structure):
    """sr_sinks_report: the sinks' count, the last pass's formation, merges and accretion, what they swallowed."""
    _fields_ = [("count", C.c_int), ("formed_x", C.c_int), ("formed_y", C.c_int), ("merged", C.c_int),
                ("accreted", C.c_double), ("swallowed", C.c_double)]


class BoxReport(C.Structure):
    _fields_ = [("x0", C.c_uint32), ("y0", C.c_uint32), ("w", C.c_uint32), ("h", C.c_uint32),
                ("fft", C.c_uint32 * 2), ("from_step", C.c_uint32), ("fits", C.c_uint32),
                ("tiles", C.c_uint32 * 3), ("cells", C.c_uint32 * 3)]


# (name, restype, argtypes): every entry point include/sr_hip.h declares
_vp, _i, _d, _sz, _u32 = C.c_void_p, C.c_int, C.c_double, C.c_size_t, C.c_uint32
_pp = C.POINTER(C.c_void_p)
SIGNATURES = [
    ("sr_last_error", C.c_char_p, []),
    ("sr_device_info", _i, [C.c_char_p, _sz, C.POINTER(_i), C.POINTER(_sz)]),
    ("sr_physics_defaults", None, [C.POINTER(Physics)]),
    ("sr_heat_settings_defaults", None, [C.POINTER(HeatSettings)]),
    ("sr_fft2", _i, [_vp, _i, _i, _i, _vp]),
    ("sr_fft2_noalloc", _i, [_vp, _i, _i, _i, _vp]),
    ("sr_gravity_create", _i, [_i, _i, _d, _d, _d, _i, _pp]),
    ("sr_gravity_field", _i, [_vp, _vp, _vp, _vp, _vp, _vp]),
    ("sr_gravity_destroy", None, [_vp]),
    ("sr_fld_create", _i, [_i, _i, C.POINTER(HeatSettings), _pp]),
    ("sr_fld_set_system", _i, [_vp, _vp, _vp, _vp, _vp, _vp]),
    ("sr_fld_solve", _i, [_vp, _vp, _vp, _d, _i, _i, C.POINTER(_i), C.POINTER(_d), _vp]),
    ("sr_fld_apply_precond", _i, [_vp, _vp, _vp, _i, _vp]),
    ("sr_fld_matvec", _i, [_vp, _vp, _vp, _vp]),
    ("sr_fld_levels", _i, [_vp, C.POINTER(_i)]),
    ("sr_fld_destroy", None, [_vp]),
    ("sr_world_create", _i, [_i, _i, C.POINTER(Physics), C.POINTER(Switches), C.POINTER(HeatSettings), _pp]),
    ("sr_upload", _i, [_vp, _vp, _sz]),
    ("sr_readback", _i, [_vp, _vp, _sz]),
    ("sr_resume", _i, [_vp, C.POINTER(ResumeRecords), _vp, _sz, _vp, _sz]),
    ("sr_readback_sigma", _i, [_vp, _vp, _sz]),
    ("sr_heat_first_system", _i, [_vp, _vp, _vp, _vp, _vp, _sz]),
    ("sr_apply_heat_precond", _i, [_vp, _vp, _vp, _sz]),
    ("sr_run", _i, [_vp, _u32]),
    ("sr_fld_tiles", _i, [_vp, C.POINTER(_i), C.POINTER(_i), _i, C.POINTER(_i)]),
    ("sr_fld_set_tiles", _i, [_vp, C.POINTER(_i), C.POINTER(_i), _i]),
    ("sr_ensemble_run", _i, [C.POINTER(_vp), _i, _u32, _i]),
    ("sr_read_dt", _i, [_vp, C.POINTER(DtReport)]),
    ("sr_read_heat", _i, [_vp, C.POINTER(HeatReport)]),
    ("sr_read_field", _i, [_vp, _i, _vp, _sz]),
    ("sr_read_totals", _i, [_vp, _vp, _vp]),
    ("sr_set_precond", _i, [_vp, _i]),
    ("sr_set_dt_cap", _i, [_vp, _d]),
    ("sr_set_timing", _i, [_vp, _i]),
    ("sr_read_cost", _i, [_vp, C.POINTER(StepCost)]),
    ("sr_read_work", _i, [_vp, C.POINTER(StepWork), _i]),
    ("sr_set_solve_mode", _i, [_vp, _i]),
    ("sr_fld_set_solve_mode", _i, [_vp, _i]),
    ("sr_fld_fusion", _i, [_vp, C.POINTER(_i), C.POINTER(_i)]),
    ("sr_fld_set_fusion", _i, [_vp, _i, _i]),
    ("sr_fld_set_coarse_wave", _i, [_vp, _i]),
    ("sr_fld_set_update_fusion", _i, [_vp, _i]),
    ("sr_set_update_fusion", _i, [_vp, _i]),
    ("sr_set_fusion", _i, [_vp, _i, _i]),
    ("sr_set_box_mode", _i, [_vp, _i]),
    ("sr_set_sinks", _i, [_vp, C.POINTER(SinksParams)]),
    ("sr_put_sinks", _i, [_vp, C.POINTER(Sink), C.c_int]),
    ("sr_read_sinks", _i, [_vp, C.POINTER(Sink), C.c_int, C.POINTER(SinksReport)]),
    ("sr_sinks_apply", _i, [_vp, _d, _vp, _sz]),
    ("sr_set_reactions", _i, [_vp, C.POINTER(ReactParams)]),
    ("sr_react_apply", _i, [_vp, _d]),
    ("sr_read_react", _i, [_vp, C.POINTER(ReactReport)]),
    ("sr_read_react_planes", _i, [_vp, _vp, _vp, _vp, _sz]),
    ("sr_read_box", _i, [_vp, C.POINTER(BoxReport)]),
    ("sr_resume_box", _i, [_vp, C.POINTER(C.c_uint32), C.c_uint64]),
    ("sr_destroy", None, [_vp]),
    ("sr_kernel_time", _i, [_vp, C.c_char_p, C.POINTER(C.c_uint64), C.POINTER(_d), C.POINTER(_d)]),
    ("sr_fld_set_network", _i, [_vp, _vp, _i]),
    ("sr_fld_solve_stats", _i, [_vp, _vp, _vp, _d, _i, _i, C.POINTER(_i), C.POINTER(_d), C.POINTER(_i),
                                C.POINTER(_i), _vp]),
    ("sr_set_network", _i, [_vp, _vp, _i]),
    ("sr_net_create", _i, [_i, _i, _pp]),
    ("sr_net_param_count", _i, [_vp, C.POINTER(_sz)]),
    ("sr_net_set_params", _i, [_vp, _vp, _sz]),
    ("sr_net_get_params", _i, [_vp, _vp, _sz]),
    ("sr_net_get_grads", _i, [_vp, _vp, _sz]),
    ("sr_net_apply", _i, [_vp, _i, _i, _vp, _vp, _vp, _vp, _vp, _vp]),
    ("sr_net_loss_grad", _i, [_vp, _i, _i, _vp, _vp, _vp, _vp, _vp, C.POINTER(_d), _vp]),
    ("sr_net_adam", _i, [_vp, _d, _vp]),
    ("sr_net_activation", _i, [_vp, C.c_char_p, _vp, _sz]),
    ("sr_net_destroy", None, [_vp]),
]


def load(path: str = LIB_PATH) -> C.CDLL:
    if not os.path.exists(path):
        raise OSError(f"{path} is not built: run `python -m sandbox_reactions_amd.build` first")
    lib = C.CDLL(path)
    missing = []
    for name, res, args in SIGNATURES:
        try:
            fn = getattr(lib, name)
        except AttributeError:
            missing.append(name)
            continue
        fn.restype = res
        fn.argtypes = args
    lib.missing_symbols = missing
    return lib
