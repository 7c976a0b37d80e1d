"""The world's two edges (crates/sr-engine/tests/gpu/edges.rs), graded as the reference grades them.

Bounce: scenes/edge_pulse.json, a 256 x 64 world of gas at rest (Sigma 1, Pi 1) holding a right-going acoustic pulse
32 cells wide (u = v0 = 0.002, one simple wave, Mach 1.4e-3). It reflects off the right wall; at t = 200 the world's
x momentum over the pulse slab's mass is within 2 % of -v0; the walls' mass flux is exactly 0, and their pressure
impulse is booked as escaped momentum and nothing else: escaped mass, energy and species exactly 0, the momentum grid
+ escaped constant to 1e-12 (twin) / 1e-6 (GPU) of the walls' gross booked impulse, and the reversal 2 M v0 in the
walls' net x booking within 2 %.

Leave: a wind, the 128 x 64 world moving at (0.6, 0.4) out through the right and bottom sides with a helium block
riding it (Sigma + 2, Pi + 0.02), to t = 40, by when over a third of the mass has gone. The grid's mass loss equals
the booked escaped mass (the twin to its f64 reach, 1e-12; the GPU to 1e-6 of the escaped mass), and no face's booked
mass ever falls in a step (the twin, every face; the GPU, the total per step). The GPU's escaped totals are the twin's
within 1e-3."""
import json
import os

import numpy as np
import pytest

import oracle
from helpers import both
from sandbox_reactions_amd import scene_json

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "scenes")
V0, SLAB, T_PULSE, T_WIND = 0.002, (64, 96), 200.0, 40.0
WIND = {"version": 1, "world": {"width": 128, "height": 64, "edge": "leave"},
        "switches": {"gravity": False, "hydro": True, "heat": False, "reactions": False, "neutrinos": False,
                     "sinks": False},
        "objects": [{"kind": "rect", "x0": 0, "y0": 0, "w": 128, "h": 64, "sigma": 1.0, "pressure": 1.0,
                     "composition": {"H": 1.0}, "v": [0.6, 0.4]},
                    {"kind": "rect", "x0": 88, "y0": 24, "w": 24, "h": 16, "sigma": 2.0, "pressure": 0.02,
                     "composition": {"He": 1.0}, "v": [0.6, 0.4]}],
        "overrides": {"k1e": 0, "k2e": 0, "k1n": 0, "k2n": 0}}


def load(s):
    kw = s.physics_kwargs()
    return kw, s.planes(kw["gamma"]).astype(np.float32).astype(np.float64)


def to_time(until, step, sim_time, cap, each=None):
    steps = 0
    while sim_time() < until:
        cap(until - sim_time())
        step()
        steps += 1
        if each:
            each()
        assert steps < 20000
    return steps


def grade_pulse(start, mom, esc_net, gross, tol_mom):
    """mom: the grid's (x, y) momentum at t = 200; esc_net: the walls' net (x, y) booking; gross: their gross impulse."""
    slab_mass = start[0][:, SLAB[0]:SLAB[1]].sum()
    v = mom[0] / slab_mass
    drift = np.hypot(mom[0] + esc_net[0] - start[1].sum(), mom[1] + esc_net[1] - start[2].sum())
    reversal = 2.0 * slab_mass * V0
    print(f"v at t = 200: {v / V0:+.4f} v0; walls' net x booking {esc_net[0] / reversal:.4f} of 2 M v0; momentum "
          f"drift {drift / gross:.2e} of the impulse {gross:.4f}", flush=True)
    assert abs(v / -V0 - 1.0) <= 0.02
    assert abs(esc_net[0] / reversal - 1.0) <= 0.02
    assert drift <= tol_mom * gross


def test_the_twin_reflects_the_pulse_and_books_the_walls():
    s = scene_json.load(os.path.join(HERE, "edge_pulse.json"))
    kw, st = load(s)
    op, _ = both(**kw)
    twin = oracle.Twin(op, st, gravity=False, hydro=True, heat=False)
    to_time(T_PULSE, twin.step, lambda: twin.times()[2], twin.cap_dt_next)
    e = twin.escaped()
    assert np.all(e[0] == 0) and np.all(e[3] == 0) and np.all(e[4:] == 0)
    end = twin.state()
    assert abs(end[0].sum() - st[0].sum()) <= 1e-12 * st[0].sum()
    gross = np.abs(e[1]).sum() + np.abs(e[2]).sum()
    grade_pulse(st, (end[1].sum(), end[2].sum()), (e[1].sum(), e[2].sum()), gross, 1e-12)


def test_the_twin_lets_the_wind_out_and_books_it():
    s = scene_json.parse(json.dumps(WIND))
    kw, st = load(s)
    op, _ = both(**kw)
    twin = oracle.Twin(op, st, gravity=False, hydro=True, heat=False)
    last = [twin.escaped()[0].copy()]

    def each():
        e = twin.escaped()[0]
        assert np.all(e >= last[0] - 1e-15 * np.abs(last[0]).max()), "a face's booked mass fell"
        last[0] = e.copy()

    to_time(T_WIND, twin.step, lambda: twin.times()[2], twin.cap_dt_next, each)
    e, end = twin.escaped(), twin.state()
    lost = st[0].sum() - end[0].sum()
    print(f"wind: lost {lost:.6f} of {st[0].sum():.1f}, booked {e[0].sum():.6f}", flush=True)
    assert lost > st[0].sum() / 3.0
    assert abs(lost - e[0].sum()) <= 1e-12 * e[0].sum()


@pytest.mark.gpu
def test_the_gpu_reflects_the_pulse_and_books_the_walls(gpu_lib):
    s = scene_json.load(os.path.join(HERE, "edge_pulse.json"))
    kw, st = load(s)
    world = scene_json.world(s)
    world.upload(st.astype(np.float32))
    n = to_time(T_PULSE, lambda: world.run(1), lambda: world.read_dt().sim_time, world.set_dt_cap)
    print(f"edge pulse: {n} steps", flush=True)
    t, e = world.totals()
    assert e[0] == 0 and e[3] == 0 and np.all(e[4:] == 0)
    assert abs(t[0] - st[0].sum()) <= 1e-12 * st[0].sum()
    # the library books each component's net; the walls' gross impulse is at least the net reversal, which bounds it
    reversal = 2.0 * st[0][:, SLAB[0]:SLAB[1]].sum() * V0
    grade_pulse(st, (t[1], t[2]), (e[1], e[2]), reversal, 1e-6)


@pytest.mark.gpu
def test_the_gpu_lets_the_wind_out_and_books_it(gpu_lib):
    s = scene_json.parse(json.dumps(WIND))
    kw, st = load(s)
    op, _ = both(**kw)
    world = scene_json.world(s)
    world.upload(st.astype(np.float32))
    twin = oracle.Twin(op, st, gravity=False, hydro=True, heat=False)
    last = [0.0]

    def step():
        dt = world.read_dt().dt_next
        twin.cap_dt_next(dt)
        twin.step()
        world.set_dt_cap(dt)
        world.run(1)
        booked = world.totals()[1][0]
        assert booked >= last[0], "the booked mass fell"
        last[0] = booked

    n = to_time(T_WIND, step, lambda: world.read_dt().sim_time, lambda c: None)
    t, e = world.totals()
    lost = st[0].sum() - t[0]
    te = twin.escaped().sum(1)
    print(f"wind on the GPU: {n} steps, lost {lost:.6f}, booked {e[0]:.6f}, twin {te[0]:.6f}", flush=True)
    assert abs(lost - e[0]) <= 1e-6 * e[0]
    for k in (0, 1, 2, 3):
        assert abs(e[k] - te[k]) <= 1e-3 * np.abs(te[k]), (k, e[k], te[k])
