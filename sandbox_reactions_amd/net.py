"""The learned FLD preconditioner's network on the device (sr_net_* of include/sr_hip.h): a 3-level U-Net with
16/32/64 channels and 3x3 convolutions on BF16 activations, its backward pass and Adam. No counterpart in the
reference, whose preconditioners are Jacobi and multigrid (crates/sr-engine/src/step/heat.rs:595)."""
from __future__ import annotations

import ctypes as C

import numpy as np

# (name, cin, cout) in parameter order; each layer's weights (3, 3, cin, cout) then biases, then the 1x1 output
LAYERS = [("e1a", 3, 16), ("e1", 16, 16), ("e2a", 16, 32), ("e2", 32, 32), ("e3a", 32, 64), ("e3", 64, 64),
          ("d2a", 96, 32), ("d2", 32, 32), ("d1a", 48, 16), ("d1", 16, 16)]
ACTIVATIONS = {"x": (8, 1), "e1a": (16, 1), "e1": (16, 1), "p1": (16, 2), "e2a": (32, 2), "e2": (32, 2),
               "p2": (32, 4), "e3a": (64, 4), "e3": (64, 4), "d2a": (32, 2), "d2": (32, 2), "d1a": (16, 1),
               "d1": (16, 1)}


def flatten(params: dict) -> np.ndarray:
    """A {name.w, name.b, out.w, out.b} dict (the layout of oracle/unet.py) as the library's flat f32 vector."""
    parts = []
    for name, _, _ in LAYERS:
        parts += [np.asarray(params[name + ".w"]).ravel(), np.asarray(params[name + ".b"]).ravel()]
    parts += [np.asarray(params["out.w"]).ravel(), np.asarray(params["out.b"]).ravel()]
    return np.concatenate(parts).astype(np.float32)


def unflatten(v: np.ndarray) -> dict:
    out, o = {}, 0
    for name, cin, cout in LAYERS:
        k = 9 * cin * cout
        out[name + ".w"] = v[o:o + k].reshape(3, 3, cin, cout)
        o += k
        out[name + ".b"] = v[o:o + cout]
        o += cout
    out["out.w"] = v[o:o + 16]
    out["out.b"] = v[o + 16:o + 17]
    return out


def init_params(seed: int = 0) -> dict:
    """He-normal weights, zero biases, a small output layer (the untrained network's correction starts near 0)."""
    rng = np.random.default_rng(seed)
    p = {}
    for name, cin, cout in LAYERS:
        p[name + ".w"] = (rng.normal(size=(3, 3, cin, cout)) * np.sqrt(2.0 / (9 * cin))).astype(np.float32)
        p[name + ".b"] = np.zeros(cout, np.float32)
    p["out.w"] = (rng.normal(size=16) * 0.01).astype(np.float32)
    p["out.b"] = np.zeros(1, np.float32)
    return p


class Net:
    def __init__(self, max_w: int, max_h: int, params: dict | None = None):
        from . import _check, lib
        p = C.c_void_p()
        self._lib = lib()  # held for the destructor, which may run while the interpreter shuts down
        _check(self._lib.sr_net_create(max_w, max_h, C.byref(p)))
        self._p = p
        n = C.c_size_t()
        _check(lib().sr_net_param_count(self._p, C.byref(n)))
        self.nparams = n.value
        self.set_params(params if params is not None else init_params())

    @staticmethod
    def _s(stream):
        import torch
        s = stream if stream is not None else torch.cuda.current_stream()
        return C.c_void_p(s.cuda_stream)

    def set_params(self, params):
        from . import _check, lib
        v = flatten(params) if isinstance(params, dict) else np.ascontiguousarray(params, np.float32)
        _check(lib().sr_net_set_params(self._p, v.ctypes.data, v.size))

    def params(self) -> np.ndarray:
        from . import _check, lib
        v = np.empty(self.nparams, np.float32)
        _check(lib().sr_net_get_params(self._p, v.ctypes.data, v.size))
        return v

    def grads(self) -> np.ndarray:
        from . import _check, lib
        v = np.empty(self.nparams, np.float32)
        _check(lib().sr_net_get_grads(self._p, v.ctypes.data, v.size))
        return v

    def apply(self, own, kx, ky, r, z, stream=None):
        from . import _check, lib
        h, w = own.shape
        _check(lib().sr_net_apply(self._p, w, h, *(C.c_void_p(t.data_ptr()) for t in (own, kx, ky, r, z)),
                                  self._s(stream)))

    def loss_grad(self, own, kx, ky, r, e, stream=None) -> float:
        from . import _check, lib
        h, w = own.shape
        loss = C.c_double()
        _check(lib().sr_net_loss_grad(self._p, w, h, *(C.c_void_p(t.data_ptr()) for t in (own, kx, ky, r, e)),
                                      C.byref(loss), self._s(stream)))
        return loss.value

    def adam(self, lr: float, stream=None):
        from . import _check, lib
        _check(lib().sr_net_adam(self._p, lr, self._s(stream)))

    def activation(self, name: str, w: int, h: int) -> np.ndarray:
        from . import _check, lib
        c, f = ACTIVATIONS[name]
        out = np.empty((h // f, w // f, c), np.float32)
        _check(lib().sr_net_activation(self._p, name.encode(), out.ctypes.data, out.size))
        return out

    def __del__(self):
        if getattr(self, "_p", None):
            self._lib.sr_net_destroy(self._p)
            self._p = None


def _time_main():
    """python -m sandbox_reactions_amd.net --time N [--reps R]: the committed network on the N^2 benchmark FLD system,
    R applications timed with HIP events (the kernels a counter run attributes to the network)."""
    import argparse
    import sys
    import torch
    import sandbox_reactions_amd as sr
    from . import scenes, train
    ap = argparse.ArgumentParser()
    ap.add_argument("--time", type=int, default=2048)
    ap.add_argument("--reps", type=int, default=10)
    a = ap.parse_args()
    sr.require_device()
    n = a.time
    c, kx, ky, _, _ = scenes.fld_system(n)
    t = lambda v: torch.tensor(np.ascontiguousarray(v, np.float32), device="cuda")
    own, tkx, tky = t(c), t(kx), t(ky)
    net = Net(n, n, train.load())
    r = torch.randn((n, n), device="cuda")
    z = torch.empty_like(r)
    net.apply(own, tkx, tky, r, z)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(a.reps):
        net.apply(own, tkx, tky, r, z)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / a.reps
    flops = 2.0 * n * n * (sum(9 * ci * co * f for (_, ci, co), f in
                               zip(LAYERS, (1, 1, .25, .25, 1 / 16, 1 / 16, .25, .25, 1, 1))) + 16)
    print(f"network {n}^2: {ms:.3f} ms per application, {flops / ms / 1e9:.1f} TFLOP/s "
          f"({flops / ms / 1e9 / 2500.0 * 100:.1f} % of 2.5 PFLOP/s dense BF16)", file=sys.stderr, flush=True)


if __name__ == "__main__":
    _time_main()
