"""The heat pass's linear solver on its own (step/heat.rs:609-673: heat_first_system / apply_heat_precond): FP32
multigrid-preconditioned flexible CG inside FP64 iterative refinement, on device tensors."""
from __future__ import annotations

import ctypes as C

from . import _lib


class Fld:
    """Solver for the 5-point SPD system own_i v_i + sum_f k_f (v_i - v_j) on a w x h grid (float32 own, kx
    (h, w+1), ky (h+1, w) couplings on CUDA tensors; gas uint8 (h, w) or None)."""

    def __init__(self, w: int, h: int, levels: int = 0, nu1: int = 2, nu2: int = 2, coarse_scale: float = 1.0):
        from . import _check, heat_settings, lib
        self.w, self.h = w, h
        hs = heat_settings(levels=levels, nu1=nu1, nu2=nu2, coarse_scale=coarse_scale)
        p = C.c_void_p()
        self._lib = lib()  # held for the destructor, which may run while the interpreter shuts down
        _check(self._lib.sr_fld_create(w, h, C.byref(hs), C.byref(p)))
        self._p = p

    @staticmethod
    def _s(stream):
        import torch
        s = stream if stream is not None else torch.cuda.current_stream()
        return C.c_void_p(s.cuda_stream)

    def set_system(self, own, kx, ky, gas=None, stream=None):
        from . import _check, lib
        for t in (own, kx, ky):
            assert t.is_contiguous()
        _check(lib().sr_fld_set_system(self._p, C.c_void_p(own.data_ptr()), C.c_void_p(kx.data_ptr()),
                                       C.c_void_p(ky.data_ptr()), C.c_void_p(gas.data_ptr()) if gas is not None
                                       else None, self._s(stream)))

    def solve(self, b, x, tol=1e-8, max_iter=10000, precond=_lib.PRECOND_MULTIGRID, stream=None):
        """Solves in place into x (float64 CUDA tensor); returns (iterations, ||b - Ax|| / ||b||)."""
        from . import _check, lib
        it = C.c_int()
        rel = C.c_double()
        _check(lib().sr_fld_solve(self._p, C.c_void_p(b.data_ptr()), C.c_void_p(x.data_ptr()), tol, max_iter,
                                  precond, C.byref(it), C.byref(rel), self._s(stream)))
        return it.value, rel.value

    def solve_stats(self, b, x, tol=1e-8, max_iter=10000, precond=_lib.PRECOND_MULTIGRID, stream=None):
        """As solve(): (iterations, relative residual, iterations with the learned preconditioner, whether it gave
        way to plain multigrid)."""
        from . import _check, lib
        it, rel, li, fb = C.c_int(), C.c_double(), C.c_int(), C.c_int()
        _check(lib().sr_fld_solve_stats(self._p, C.c_void_p(b.data_ptr()), C.c_void_p(x.data_ptr()), tol,
                                        max_iter, precond, C.byref(it), C.byref(rel), C.byref(li), C.byref(fb),
                                        self._s(stream)))
        return it.value, rel.value, li.value, bool(fb.value)

    def set_solve_mode(self, mode: int):
        """SOLVE_HOST (the default), SOLVE_DEVICE or SOLVE_GRAPH: how a solve decides when to stop."""
        from . import _check, lib
        _check(lib().sr_fld_set_solve_mode(self._p, mode))

    def set_network(self, net, fallback_iterations: int = 0):
        """The network of the learned preconditioner (None detaches it)."""
        from . import _check, lib
        self._net = net  # keep it alive while attached
        _check(lib().sr_fld_set_network(self._p, net._p if net is not None else None, fallback_iterations))

    def apply_precond(self, r, z, precond=_lib.PRECOND_MULTIGRID, stream=None):
        from . import _check, lib
        _check(lib().sr_fld_apply_precond(self._p, C.c_void_p(r.data_ptr()), C.c_void_p(z.data_ptr()), precond,
                                          self._s(stream)))

    def matvec(self, v, out, stream=None):
        from . import _check, lib
        _check(lib().sr_fld_matvec(self._p, C.c_void_p(v.data_ptr()), C.c_void_p(out.data_ptr()), self._s(stream)))

    @property
    def levels(self) -> int:
        from . import _check, lib
        n = C.c_int()
        _check(lib().sr_fld_levels(self._p, C.byref(n)))
        return n.value

    def tiles(self):
        """([pre shape per level], [post shape per level]) of the tile-smoothed levels (TILE_16, TILE_32, TILE_64):
        the autotuner's choice once a system is set, or set_tiles'."""
        from . import _check, lib
        pre, post, n = (C.c_int * 32)(), (C.c_int * 32)(), C.c_int()
        _check(lib().sr_fld_tiles(self._p, pre, post, 32, C.byref(n)))
        return list(pre[:n.value]), list(post[:n.value])

    def set_tiles(self, pre, post):
        from . import SrError, _check, lib
        if len(pre) != len(post):
            raise SrError(_lib.SR_ERR_ARGUMENT, "set_tiles: one pre and one post shape per level")
        k = len(pre)
        _check(lib().sr_fld_set_tiles(self._p, (C.c_int * max(1, k))(*pre), (C.c_int * max(1, k))(*post), k))

    def fusion(self):
        """(pre, post): FUSE_OFF, FUSE_SMALL or FUSE_LARGE for the last two tile-smoothed levels' smoothers, the
        autotuner's choice once a system is set, or set_fusion's; FUSE_OFF for both with fewer than three levels
        tile-smoothed."""
        from . import _check, lib
        pre, post = C.c_int(), C.c_int()
        _check(lib().sr_fld_fusion(self._p, C.byref(pre), C.byref(post)))
        return pre.value, post.value

    def set_fusion(self, pre: int, post: int):
        from . import _check, lib
        _check(lib().sr_fld_set_fusion(self._p, pre, post))

    def set_update_fusion(self, on: bool):
        """Each PCG update folded into the next V-cycle's level-0 pre-smoother (True, the default) or on its own."""
        from . import _check, lib
        _check(lib().sr_fld_set_update_fusion(self._p, int(on)))

    def set_coarse_wave(self, on: bool):
        """The single-block V-cycle's smallest levels on one wave (True, the default) or on the whole block."""
        from . import _check, lib
        _check(lib().sr_fld_set_coarse_wave(self._p, int(on)))

    def close(self):
        """Release the solver's device memory now (the destructor does it otherwise)."""
        if getattr(self, "_p", None):
            self._lib.sr_fld_destroy(self._p)
            self._p = None

    def __del__(self):
        self.close()
