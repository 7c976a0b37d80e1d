"""The reference's own scenes (tests/golden/scenes, loaded by sandbox_reactions_amd.scene_json) on the GPU against
the f64 twin, each on the same dt: the Sod tube, the reference disk under self-gravity, a cold collapsing disk, the
edge pulse against a bounce wall, the hot disk radiating through a thin rim, and the constant-diffusivity pulse."""
import os

import numpy as np
import pytest

import oracle
from helpers import both, rel_l1
from sandbox_reactions_amd import scene_json

pytestmark = pytest.mark.gpu
HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "scenes")


def run(name, steps, tight=False):
    import sandbox_reactions_amd as sr
    s = scene_json.load(os.path.join(HERE, name))
    kw = s.physics_kwargs()
    op, gp = both(**kw)
    hs = sr.heat_settings(stop=1, solve_tol=1e-10, newton_heat=1e-9, newton_cell=1e-6) if tight else None
    world = scene_json.world(s, heat_settings=hs)
    start = s.planes(kw["gamma"]).astype(np.float64)
    twin = oracle.Twin(op, start, gravity=s.switches["gravity"], hydro=s.switches["hydro"], heat=s.switches["heat"])
    for _ in range(steps):
        dt = world.read_dt().dt_next
        twin.cap_dt_next(dt)
        twin.step()
        world.set_dt_cap(dt)
        world.run(1)
    return world.readback(), twin.state(), world


@pytest.mark.parametrize("name, steps, bar", [("sod.json", 1, 2e-6), ("ref_sod.json", 5, 1e-5),
                                              ("edge_pulse.json", 10, 1e-5), ("ref_disk.json", 5, 1e-5),
                                              ("cold_disk.json", 5, 1e-5)])
def test_hydro_and_gravity_scenes(gpu_lib, name, steps, bar):
    print(f"scene {name}, {steps} steps", flush=True)
    g, t, _ = run(name, steps)
    assert rel_l1(g[0], t[0]) <= bar, (name, rel_l1(g[0], t[0]))
    assert rel_l1(g[3], t[3]) <= 10 * bar, (name, rel_l1(g[3], t[3]))


@pytest.mark.parametrize("name", ["heat_rim.json", "heat_pulse.json", "heat_hot.json"])
def test_heat_scenes(gpu_lib, name):
    print(f"scene {name}", flush=True)
    g, t, world = run(name, 1, tight=True)
    assert world.read_heat().converged
    assert rel_l1(g[3], t[3]) <= 5e-5, (name, rel_l1(g[3], t[3]))


def run_scene(s, steps):
    import sandbox_reactions_amd as sr
    kw = s.physics_kwargs()
    op, gp = both(**kw)
    world = scene_json.world(s)
    start = s.planes(kw["gamma"]).astype(np.float64)
    twin = oracle.Twin(op, start, gravity=s.switches["gravity"], hydro=s.switches["hydro"], heat=s.switches["heat"])
    for _ in range(steps):
        dt = world.read_dt().dt_next
        twin.cap_dt_next(dt)
        twin.step()
        world.set_dt_cap(dt)
        world.run(1)
    return world.readback(), twin.state(), world


PASSES_ON = {"gravity": True, "hydro": True, "heat": True, "reactions": False, "neutrinos": False, "sinks": False}


@pytest.mark.parametrize("case", ["asym_blob", "neutron_disk"])
def test_cold_pressure_scenes(gpu_lib, case):
    """The shipped cold pressure on (eos.rs: electrons and neutron matter), gravity, hydro and heat: asym_blob.json
    with its passes the path carries, and a disk of half hydrogen, half neutrons beside a `pressure` rect of helium
    (both pairs, and a pressure object's cold share). Five steps on the twin's dt."""
    import json
    print(f"cold pressure scene {case}", flush=True)
    if case == "asym_blob":
        with open(os.path.join(HERE, "asym_blob.json")) as f:
            d = json.load(f)
    else:
        d = {"version": 1, "world": {"width": 64, "height": 64, "edge": "bounce"},
             "objects": [{"kind": "disk", "x": 30, "y": 33, "a": 12, "sigma_c": 2.0, "t": 1.0,
                          "composition": {"H": 0.5, "n": 0.5}},
                         {"kind": "rect", "x0": 44, "y0": 10, "w": 10, "h": 8, "sigma": 0.8, "pressure": 30.0,
                          "composition": {"He": 1.0}}]}
    d["switches"] = PASSES_ON
    s = scene_json.parse(json.dumps(d))
    assert s.cold_on()
    g, t, world = run_scene(s, 5)
    assert rel_l1(g[0], t[0]) <= 1e-5, (case, rel_l1(g[0], t[0]))
    assert rel_l1(g[3], t[3]) <= 1e-4, (case, rel_l1(g[3], t[3]))
    assert world.read_heat().converged
