"""sandbox_reactions_amd — the MI355X-native grid-physics path of a sandbox_reactions-style engine.

The product is libsrhip.so (hand-written HIP for gfx950 behind the C ABI in include/sr_hip.h). This package is the
host-side mirror of the reference's step API over that ABI (crates/sr-engine/src/step/mod.rs `Step`,
crates/sr-engine/src/state.rs `State`, crates/sr-physics/src/reference/grav.rs `Gravity`): names, argument meaning
and refusals follow the reference. There is no CPU fallback: every call goes to the HIP library, and a missing
library or device raises.
"""
from __future__ import annotations

import ctypes as C
from typing import Optional

import numpy as np

from . import _lib
from ._lib import (BOX_FIT, BOX_WHOLE_WORLD, COMPONENTS, EDGE_BOUNCE, EDGE_LEAVE, FIELD_FLUX_X, FIELD_FLUX_Y, FIELD_GX,
                   FIELD_GY, FIELD_PHI, FIELD_PHI_MID, FIELD_SIGMA_MID, FUSE_LARGE, FUSE_OFF, FUSE_SMALL,
                   OPACITY_KRAMERS, OPACITY_REFERENCE, PRECOND_JACOBI, PRECOND_LEARNED, PRECOND_MULTIGRID, SOLVE_DEVICE,
                   SOLVE_GRAPH, SOLVE_HOST, SOLVE_HOST_WAIT, STOP_REFERENCE, STOP_RELATIVE, TILE_16, TILE_32, TILE_64,
                   DtReport, HeatReport, HeatSettings, Physics, StepCost, Switches)

_LIB = None


class SrError(RuntimeError):
    """A refusal or failure of the HIP library: `code` is the SR_ERR_* value, `kind` its reference name."""

    def __init__(self, code: int, message: str):
        self.code = code
        self.kind = _lib.ERRORS.get(code, "Unknown")
        super().__init__(f"{self.kind} ({code}): {message}")


def lib():
    global _LIB
    if _LIB is None:
        _LIB = _lib.load()
    return _LIB


def _check(code: int):
    if code != _lib.SR_OK:
        msg = lib().sr_last_error()
        raise SrError(code, msg.decode() if msg else "")


def require_device():
    """Raises unless a HIP device is visible and the library is loaded (the product never runs without both)."""
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("no HIP device visible: the sr-hip path runs only on a GPU")
    lib()


def device_info():
    name = C.create_string_buffer(64)
    cu = C.c_int()
    mem = C.c_size_t()
    _check(lib().sr_device_info(name, 64, C.byref(cu), C.byref(mem)))
    return name.value.decode(), cu.value, mem.value


def physics(**over) -> Physics:
    """The reference's constants (assets/physics.json; gamma 2, dx 1, G 1), overridden by keyword."""
    ph = Physics()
    lib().sr_physics_defaults(C.byref(ph))
    for k, v in over.items():
        if not hasattr(ph, k):
            raise KeyError(f"unknown physics key {k}")
        setattr(ph, k, v)
    return ph


def heat_settings(**over) -> HeatSettings:
    hs = HeatSettings()
    lib().sr_heat_settings_defaults(C.byref(hs))
    for k, v in over.items():
        if not hasattr(hs, k):
            raise KeyError(f"unknown heat setting {k}")
        setattr(hs, k, v)
    return hs


def _ptr(t) -> Optional[int]:
    if t is None:
        return None
    return t.data_ptr()


def _stream(stream=None):
    import torch
    s = stream if stream is not None else torch.cuda.current_stream()
    return C.c_void_p(s.cuda_stream)


def fft2(data, inverse: bool = False, stream=None):
    """Fft::transform (step/fft.rs:116-134) on a complex64 (ny, nx) CUDA tensor, in place."""
    import torch
    if data.dtype != torch.complex64 or not data.is_contiguous():
        raise SrError(_lib.SR_ERR_ARGUMENT, "fft2 wants a contiguous complex64 tensor")
    ny, nx = data.shape
    _check(lib().sr_fft2(C.c_void_p(data.data_ptr()), nx, ny, int(inverse), _stream(stream)))
    return data


class Gravity:
    """P2 for one w x h box (reference/grav.rs:99-121; GPU step/grav.rs:133-376): phi on the box grown by one ring
    and g = -grad phi, from a float32 CUDA tensor of Sigma."""

    def __init__(self, w: int, h: int, G: float = 1.0, dx: float = 1.0, sigma_vac: float = 1e-6,
                 max_padded: int = 4096):
        self.w, self.h = w, h
        p = C.c_void_p()
        _check(lib().sr_gravity_create(w, h, G, dx, sigma_vac, max_padded, C.byref(p)))
        self._p = p

    def field(self, sigma, stream=None):
        import torch
        if sigma.shape != (self.h, self.w) or sigma.dtype != torch.float32 or not sigma.is_contiguous():
            raise SrError(_lib.SR_ERR_LENGTH, f"sigma must be a contiguous float32 ({self.h}, {self.w}) tensor")
        phi = torch.empty((self.h + 2, self.w + 2), dtype=torch.float32, device=sigma.device)
        gx = torch.empty_like(sigma)
        gy = torch.empty_like(sigma)
        _check(lib().sr_gravity_field(self._p, C.c_void_p(sigma.data_ptr()), C.c_void_p(phi.data_ptr()),
                                      C.c_void_p(gx.data_ptr()), C.c_void_p(gy.data_ptr()), _stream(stream)))
        return phi, gx, gy

    def __del__(self):
        if getattr(self, "_p", None) and _LIB is not None:
            _LIB.sr_gravity_destroy(self._p)
            self._p = None


from .fld import Fld  # noqa: E402
from .net import Net  # noqa: E402
from .step import World, ensemble_run  # noqa: E402

__all__ = ["SrError", "Gravity", "World", "ensemble_run", "Fld", "fft2", "physics", "heat_settings", "require_device", "lib",
           "device_info", "COMPONENTS", "EDGE_BOUNCE", "EDGE_LEAVE", "OPACITY_KRAMERS", "OPACITY_REFERENCE",
           "PRECOND_JACOBI", "PRECOND_MULTIGRID", "PRECOND_LEARNED", "STOP_REFERENCE", "STOP_RELATIVE",
           "FIELD_PHI", "FIELD_GX", "FIELD_GY", "FIELD_FLUX_X", "FIELD_FLUX_Y", "FIELD_SIGMA_MID",
           "FIELD_PHI_MID", "DtReport", "HeatReport", "HeatSettings", "Physics", "StepCost", "Switches"]
