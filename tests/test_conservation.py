"""Conservation over a run (crates/sr-engine/tests/gpu/conservation.rs's invariants, on the passes this path carries):
a leave world whose gas touches two edges (a disk cut by the left edge, a block in the top-right corner), gravity and
hydro on, 60 steps. Mass: grid + escaped constant to 1e-6 of its start (the floors' vacuum resets, which the path does
not book, stay far below it); momentum: grid + escaped constant to 1e-5 of the largest sum of |Sigma u| the run has
had (the reference's bound: f32's floor with every route booked; gravity's net kick is taken back from the densest
free cell each step, reference/hydro.rs:997-1001, so no force error moves it). The twin (f64) is held to the same
mass bound (its drift, 1.6e-7, is the floors' refill of the floor gas that leaves through the edges) and to 1e-7 on
momentum."""
import json

import numpy as np
import pytest

import oracle
from helpers import both
from sandbox_reactions_amd import scene_json

SCENE = {"version": 1, "world": {"width": 64, "height": 64, "edge": "leave"},
         "switches": {"gravity": True, "hydro": True, "heat": False, "reactions": False, "neutrinos": False,
                      "sinks": False},
         "objects": [{"kind": "disk", "x": 8, "y": 20, "a": 14, "sigma_c": 1.0, "t": 0.5, "composition": {"H": 1.0}},
                     {"kind": "rect", "x0": 40, "y0": 50, "w": 24, "h": 14, "sigma": 0.5, "t": 0.5,
                      "composition": {"H": 1.0}}],
         "overrides": {"k1e": 0, "k2e": 0, "k1n": 0, "k2n": 0}}
STEPS = 60


def scene():
    s = scene_json.parse(json.dumps(SCENE))
    kw = s.physics_kwargs()
    return s, kw, s.planes(kw["gamma"]).astype(np.float32).astype(np.float64)


def run_invariants(step, state_sums):
    """(worst mass drift over its start, worst |momentum drift| over the running peak of sum |Sigma u|)."""
    m0, p0, _ = state_sums()
    worst_m = worst_p = peak = 0.0
    for _ in range(STEPS):
        step()
        m, p, mag = state_sums()
        peak = max(peak, mag)
        worst_m = max(worst_m, abs(m / m0 - 1.0))
        worst_p = max(worst_p, float(np.hypot(*(p - p0))) / peak)
    print(f"mass drift {worst_m:.2e}, momentum drift {worst_p:.2e} of the peak sum |Sigma u| {peak:.2f}", flush=True)
    return worst_m, worst_p, peak


def test_the_twin_conserves_mass_and_momentum():
    s, kw, st = scene()
    op, _ = both(**kw)
    twin = oracle.Twin(op, st, gravity=True, hydro=True, heat=False)

    def sums():
        q, e = twin.state(), twin.escaped()
        return (q[0].sum() + e[0].sum(), np.array([q[1].sum() + e[1].sum(), q[2].sum() + e[2].sum()]),
                float(np.hypot(q[1], q[2]).sum()))

    m, p, peak = run_invariants(twin.step, sums)
    assert peak > 10.0
    assert m <= 1e-6 and p <= 1e-7


@pytest.mark.gpu
def test_the_gpu_conserves_mass_and_momentum(gpu_lib):
    s, kw, st = scene()
    world = scene_json.world(s)
    world.upload(st.astype(np.float32))

    def sums():
        t, e = world.totals()
        q = world.readback().astype(np.float64)
        return t[0] + e[0], np.array([t[1] + e[1], t[2] + e[2]]), float(np.hypot(q[1], q[2]).sum())

    m, p, peak = run_invariants(lambda: world.run(1), sums)
    assert peak > 10.0
    assert m <= 1e-6 and p <= 1e-5
