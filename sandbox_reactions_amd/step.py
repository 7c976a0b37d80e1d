"""The world: the reference's Step + State surface over the C ABI (crates/sr-engine/src/step/mod.rs:367-777,
crates/sr-engine/src/state.rs:139-280; the twin's driver crates/sr-physics/src/reference/mod.rs:129-297)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import (MAX_SINKS, BoxReport, DtReport, HeatReport, ReactReport, Sink, SinksParams, SinksReport, StepCost,
                   StepWork, Switches)


class World:
    """A w x h world on the GPU. `upload` takes the 14 canonical planes (channel-major, row-major, f32), `run` steps
    P1 -> P2 -> P4 -> P5 -> P8 -> P9 (the passes switched on), `readback` returns the planes."""

    def __init__(self, w: int, h: int, physics=None, gravity=True, hydro=True, heat=False, heat_settings=None,
                 precond=_lib.PRECOND_MULTIGRID):
        from . import _check, lib, physics as mkphys, heat_settings as mkhs
        self.w, self.h = w, h
        self._ph = physics if physics is not None else mkphys()
        self._sw = Switches(int(gravity), int(hydro), int(heat))
        self._hs = heat_settings if heat_settings is not None else mkhs()
        p = C.c_void_p()
        _check(lib().sr_world_create(w, h, C.byref(self._ph), C.byref(self._sw), C.byref(self._hs), C.byref(p)))
        self._p = p
        if precond != _lib.PRECOND_MULTIGRID:
            self.set_precond(precond)

    def _c(self, code):
        from . import _check
        _check(code)

    @property
    def lib(self):
        from . import lib
        return lib()

    def upload(self, planes: np.ndarray):
        a = np.ascontiguousarray(planes, dtype=np.float32)
        if a.shape != (_lib.COMPONENTS, self.h, self.w):
            from . import SrError
            raise SrError(_lib.SR_ERR_LENGTH, f"planes must be (14, {self.h}, {self.w}), got {a.shape}")
        self._c(self.lib.sr_upload(self._p, a.ctypes.data, a.size))

    def readback(self) -> np.ndarray:
        out = np.empty((_lib.COMPONENTS, self.h, self.w), np.float32)
        self._c(self.lib.sr_readback(self._p, out.ctypes.data, out.size))
        return out

    def resume(self, dt: float, dt_next: float, sim_time: float, flux: np.ndarray | None = None,
               sigma: np.ndarray | None = None):
        """load_dump's restore after `upload` (crates/sr-engine/src/dump.rs:8-20): the clock, F's two planes
        ((2, h, w) f32, the next P4's push) and Sigma's FP64 plane ((h, w)), each optional but the clock."""
        rec = _lib.ResumeRecords(dt, dt_next, sim_time)
        f = None if flux is None else np.ascontiguousarray(flux, dtype=np.float32)
        s = None if sigma is None else np.ascontiguousarray(sigma, dtype=np.float64)
        if f is not None and f.shape != (2, self.h, self.w):
            from . import SrError
            raise SrError(_lib.SR_ERR_LENGTH, f"flux must be (2, {self.h}, {self.w}), got {f.shape}")
        if s is not None and s.shape != (self.h, self.w):
            from . import SrError
            raise SrError(_lib.SR_ERR_LENGTH, f"sigma must be ({self.h}, {self.w}), got {s.shape}")
        self._c(self.lib.sr_resume(self._p, C.byref(rec), None if f is None else f.ctypes.data,
                                   0 if f is None else f.size, None if s is None else s.ctypes.data,
                                   0 if s is None else s.size))

    def set_reactions(self, params):
        """P6 in the step with these records (sandbox_reactions_amd.react.params(...)); None switches it off."""
        self._react = params
        self._c(self.lib.sr_set_reactions(self._p, None if params is None else C.byref(params)))

    def react_apply(self, dt: float):
        """P6 alone over dt on the current state."""
        self._c(self.lib.sr_react_apply(self._p, dt))

    def read_react(self) -> ReactReport:
        r = ReactReport()
        self._c(self.lib.sr_read_react(self._p, C.byref(r)))
        return r

    def read_react_planes(self):
        """The last pass's (energy booked per group (7, h, w), neutrino loss (h, w), S_nu (h, w))."""
        nuc = np.empty((7, self.h, self.w))
        lost, snu = np.empty((self.h, self.w)), np.empty((self.h, self.w))
        self._c(self.lib.sr_read_react_planes(self._p, nuc.ctypes.data, lost.ctypes.data, snu.ctypes.data,
                                              self.w * self.h))
        return nuc, lost, snu

    def set_sinks(self, sigma_bh: float = 6000.0, m_tov: float = float("inf"), on: bool = True):
        """P3 in the step with Sigma_BH and m_tov_sb (infinite: no sink from mass); on=False switches it off. Clears
        the sinks."""
        p = SinksParams(sigma_bh, m_tov)
        self._c(self.lib.sr_set_sinks(self._p, C.byref(p) if on else None))

    def put_sinks(self, sinks):
        """The sinks placed as given: (x, y, vx, vy, m) each, at most 16."""
        arr = (Sink * MAX_SINKS)(*[Sink(*map(float, s)) for s in sinks][:MAX_SINKS])
        self._c(self.lib.sr_put_sinks(self._p, arr, len(sinks)))

    def read_sinks(self):
        """((x, y, vx, vy, m) per sink, the last pass's SinksReport)."""
        arr = (Sink * MAX_SINKS)()
        rep = SinksReport()
        self._c(self.lib.sr_read_sinks(self._p, arr, MAX_SINKS, C.byref(rep)))
        return [(s.x, s.y, s.vx, s.vy, s.m) for s in arr[:rep.count]], rep

    def sinks_apply(self, dt: float, phi=None):
        """P3 alone over dt against phi ((h, w) f32 values the cells' potential; None: the world's own)."""
        if phi is None:
            self._c(self.lib.sr_sinks_apply(self._p, dt, None, 0))
            return
        a = np.ascontiguousarray(phi, dtype=np.float32)
        self._c(self.lib.sr_sinks_apply(self._p, dt, a.ctypes.data, a.size))

    def set_box_mode(self, mode: int):
        """BOX_WHOLE_WORLD (the default) or BOX_FIT (the reference's Step::new): P8 and P9 over the active box. Before
        upload."""
        self._c(self.lib.sr_set_box_mode(self._p, mode))

    def read_box(self) -> BoxReport:
        """The active box's record and the reference's indirect sizes over it (BoxReport, step/boxfit.rs:38-55)."""
        r = BoxReport()
        self._c(self.lib.sr_read_box(self._p, C.byref(r)))
        return r

    def resume_box(self, record, step: int):
        """The box record a dump holds (8 words) and the step index the refit schedule counts from; after resume."""
        rec = (C.c_uint32 * 8)(*[int(v) for v in record])
        if len(record) != 8:
            from . import SrError
            raise SrError(_lib.SR_ERR_ARGUMENT, "the box record has 8 words")
        self._c(self.lib.sr_resume_box(self._p, rec, int(step)))

    def readback_sigma(self) -> np.ndarray:
        """Sigma's FP64 plane, (h, w)."""
        out = np.empty((self.h, self.w), np.float64)
        self._c(self.lib.sr_readback_sigma(self._p, out.ctypes.data, out.size))
        return out

    def heat_first_system(self):
        """heat_first_system (step/heat.rs:604-640): (own, kx, ky, gas) of the last step's first Newton pass, shaped
        (h, w), (h, w + 1), (h + 1, w), (h, w) as sr_fld_set_system takes them."""
        own = np.empty((self.h, self.w), np.float32)
        kx = np.empty((self.h, self.w + 1), np.float32)
        ky = np.empty((self.h + 1, self.w), np.float32)
        gas = np.empty((self.h, self.w), np.uint8)
        self._c(self.lib.sr_heat_first_system(self._p, own.ctypes.data, kx.ctypes.data, ky.ctypes.data,
                                              gas.ctypes.data, own.size))
        return own, kx, ky, gas

    def apply_heat_precond(self, r: np.ndarray) -> np.ndarray:
        """apply_heat_precond (step/heat.rs:645-673): M^-1 r on the system heat_first_system built, 0 off the gas."""
        a = np.ascontiguousarray(r, dtype=np.float32)
        if a.shape != (self.h, self.w):
            from . import SrError
            raise SrError(_lib.SR_ERR_LENGTH, f"r must be ({self.h}, {self.w}), got {a.shape}")
        z = np.empty_like(a)
        self._c(self.lib.sr_apply_heat_precond(self._p, a.ctypes.data, z.ctypes.data, a.size))
        return z

    def run(self, steps: int = 1):
        self._c(self.lib.sr_run(self._p, steps))

    def read_dt(self) -> DtReport:
        r = DtReport()
        self._c(self.lib.sr_read_dt(self._p, C.byref(r)))
        return r

    def read_heat(self) -> HeatReport:
        r = HeatReport()
        self._c(self.lib.sr_read_heat(self._p, C.byref(r)))
        return r

    def read_field(self, field: int) -> np.ndarray:
        out = np.empty((self.h, self.w), np.float32)
        self._c(self.lib.sr_read_field(self._p, field, out.ctypes.data, out.size))
        return out

    def totals(self):
        t = np.zeros(14)
        e = np.zeros(14)
        self._c(self.lib.sr_read_totals(self._p, t.ctypes.data, e.ctypes.data))
        return t, e

    def set_precond(self, precond: int):
        self._c(self.lib.sr_set_precond(self._p, precond))

    def set_network(self, net, fallback_iterations: int = 0):
        """The network the learned preconditioner uses in the heat pass (None detaches it)."""
        self._net = net
        self._c(self.lib.sr_set_network(self._p, net._p if net is not None else None, fallback_iterations))

    def set_solve_mode(self, mode: int):
        """SOLVE_HOST, SOLVE_DEVICE or SOLVE_GRAPH: how the heat pass's solves decide when to stop."""
        self._c(self.lib.sr_set_solve_mode(self._p, mode))

    def set_fusion(self, pre: int, post: int):
        """FUSE_OFF, FUSE_SMALL or FUSE_LARGE for the heat solver's last two tile-smoothed levels (as Fld.set_fusion)."""
        self._c(self.lib.sr_set_fusion(self._p, pre, post))

    def set_update_fusion(self, on: bool):
        """The heat solver's update folded into the next pre-smoother (True, the default) or on its own."""
        self._c(self.lib.sr_set_update_fusion(self._p, int(on)))

    def set_dt_cap(self, cap: float):
        self._c(self.lib.sr_set_dt_cap(self._p, cap))

    def set_timing(self, on: bool = True):
        self._c(self.lib.sr_set_timing(self._p, int(on)))

    def read_cost(self) -> StepCost:
        c = StepCost()
        self._c(self.lib.sr_read_cost(self._p, C.byref(c)))
        return c

    def read_work(self, reset: bool = False) -> StepWork:
        """The launches' counts and the algorithmic HBM bytes of the steps since the upload or the last reset."""
        w = StepWork()
        self._c(self.lib.sr_read_work(self._p, C.byref(w), int(reset)))
        return w

    def kernel_time(self, name: str):
        n = C.c_uint64()
        ms = C.c_double()
        b = C.c_double()
        self._c(self.lib.sr_kernel_time(self._p, name.encode(), C.byref(n), C.byref(ms), C.byref(b)))
        return n.value, ms.value, b.value

    def close(self):
        if getattr(self, "_p", None):
            self.lib.sr_destroy(self._p)
            self._p = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def ensemble_run(worlds, steps: int = 1, threads: int = 8):
    """Steps every world of `worlds` (distinct World objects) `steps` times, the batch shared out over `threads` host
    threads (sr_ensemble_run); each world ends where it would stepped alone."""
    from . import _check, lib
    ws = list(worlds)
    arr = (C.c_void_p * len(ws))(*[w._p.value for w in ws])
    _check(lib().sr_ensemble_run(arr, len(ws), steps, threads))
