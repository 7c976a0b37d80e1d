"""The scene loader (sandbox_reactions_amd/scene_json.py) on the reference's own scene files (copied under
tests/golden/scenes as input data), against the planes the reference's Scene::planes defines
(crates/sr-engine/src/scene.rs:428-497): closed-form cell values, the Maclaurin disk's mass, overlap sums, the floor
state, and the loader's refusals."""
import json
import math
import os

import numpy as np
import pytest

from sandbox_reactions_amd import scene_json, scenes

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "scenes")


def load(name):
    return scene_json.load(os.path.join(HERE, name))


def test_sod_is_the_shock_tube():
    s = load("sod.json")
    assert (s.width, s.height, s.edge) == (400, 64, 0)
    assert s.switches["hydro"] and not s.switches["gravity"] and not s.switches["heat"]
    p = s.planes()
    ref = scenes.sod_scene(400, 64).astype(np.float32)
    assert np.array_equal(p, ref)


def test_hot_helium_slab_energy():
    s = load("heat_hot.json")
    p = s.planes()
    assert np.all(p[0] == np.float32(30.0))
    assert np.allclose(p[3], 30.0 * 18800.0 * 0.75, rtol=1e-7)  # gamma 2: E = Sigma T / mu, 1/mu(He) = 3/4
    assert np.all(p[5] == 1.0) and np.all(p[4] == 0.0)


def test_maclaurin_disk_samples_and_mass():
    s = load("ref_disk.json")
    p = s.planes().astype(np.float64)
    a, sc = 24.0, 1.0
    yy, xx = np.mgrid[0:64, 0:64]
    r2 = ((xx + 0.5 - 32) ** 2 + (yy + 0.5 - 32) ** 2) / a ** 2
    inside = r2 < 1.0
    assert np.allclose(p[0][inside], sc * np.sqrt(1.0 - r2[inside]), rtol=1e-6)
    assert np.all(p[0][~inside] == np.float32(1e-8))
    # the continuous disk holds 2 pi a^2 sigma_c / 3; cell-centre sampling is within 1 % of it at a = 24
    m = p[0][inside].sum()
    assert abs(m - 2.0 * math.pi * a * a * sc / 3.0) <= 1e-2 * m
    im = 0.75 * 2.0 + 0.25 * 0.75
    assert np.allclose(p[3][inside], p[0][inside] * 0.05 * im, rtol=1e-6)
    assert np.allclose(p[4][inside], 0.75) and np.allclose(p[5][inside], 0.25)


def test_overlapping_objects_add():
    s = load("heat_rim.json")
    p = s.planes().astype(np.float64)
    c = 32, 32
    sig_disk = 1.0  # the disk's centre cell centre is 0.5 cells off: r^2 / a^2 = 0.5 / 400
    sd = sig_disk * math.sqrt(1.0 - 0.5 / 400.0)
    assert p[0][c] == pytest.approx(2e-6 + sd, rel=1e-6)
    assert p[3][c] == pytest.approx(2e-6 * 5.0 * 2.0 + sd * 100.0 * 2.0, rel=1e-6)
    assert p[0][0, 0] == pytest.approx(2e-6, rel=1e-6)


def test_uncovered_cells_hold_the_floor():
    p = scene_json.parse(json.dumps(base())).planes().astype(np.float64)
    rest = p[:, 8:, 8:]
    assert np.all(rest[0] == np.float32(1e-8))
    assert np.allclose(rest[3], 1e-8 * 0.05 * 2.0, rtol=1e-6)  # Sigma_floor T_floor / mu(H), gamma 2
    assert np.all(rest[4] == 1.0) and np.all(rest[1] == 0.0)
    assert np.all(p[0, :8, :8] == 1.0)


def test_hot_spot_in_a_cool_sheet():
    s = load("heat_spot.json")
    p = s.planes().astype(np.float64)
    assert np.allclose(p[3, 28:36, 28:36], 1e-4 * 2.0 * 2.0, rtol=1e-6)
    assert np.allclose(p[3, :28, :], 1e-4 * 0.05 * 2.0, rtol=1e-6)


def test_test_settings_and_constants():
    s = load("heat_pulse.json")
    kw = s.physics_kwargs()
    assert kw["constant_diffusivity"] == 0.5 and kw["edge"] == 0 and kw["flux_limiter"] == 1
    assert kw["sigma_floor"] == 1e-8 and kw["c_sb"] == 12800.0


OFF = {"reactions": False, "neutrinos": False, "sinks": False}


def base():  # an 8 x 8 rect of hydrogen in a 64 x 64 world, cold pressure off
    return {"version": 1, "world": {"width": 64, "height": 64}, "overrides": {k: 0 for k in scene_json.COLD},
            "objects": [{"kind": "rect", "x0": 0, "y0": 0, "w": 8, "h": 8, "sigma": 1.0, "t": 1.0,
                         "composition": {"H": 1.0}}]}


@pytest.mark.parametrize("edit, message", [
    (lambda d: d["objects"][0].update(pressure=1.0), "not both"),
    (lambda d: d["objects"][0].update(composition={"H": 0.9}), "sum to"),
    (lambda d: d["objects"][0].update(colour=1), "unknown key"),
    (lambda d: d["overrides"].update(k1e=21.3), "pair is on or off"),
    (lambda d: (d.pop("overrides"), d.update(switches=OFF), d["objects"][0].update(sigma=12.0)), "sigma_hold"),
    (lambda d: (d.pop("overrides"), d.update(switches=OFF), d["objects"][0].pop("t"),
                d["objects"][0].update(pressure=1e-3)), "below the cold pressure"),
    (lambda d: d["objects"].append({"kind": "preset", "preset": "sun", "x": 1, "y": 1}), "calibration"),
    (lambda d: d.update(switches={"reactions": True, "neutrinos": True, "sinks": False}), "neutrinos"),
    (lambda d: (d.update(switches={"reactions": False, "neutrinos": False, "sinks": True}),
                d["overrides"].update(k1n=3.0, k2n=50.0)), "neutron pair other than the shipped"),
    (lambda d: d["world"].update(width=60), "world size"),
    (lambda d: d["overrides"].update(warp=1.0), "unknown key"),
    (lambda d: d.update(version=2), "version"),
])
def test_refusals(edit, message):
    d = base()
    edit(d)
    with pytest.raises(scene_json.SceneError, match=message):
        s = scene_json.parse(json.dumps(d))
        s.check_path()
        s.planes()


def test_every_fixture_parses():
    """Every fixture, asym_blob.json with the shipped cold pressure on."""
    assert load("asym_blob.json").cold_on() and not load("sod.json").cold_on()
    for name in sorted(os.listdir(HERE)):
        if name.endswith(".json"):
            s = load(name)
            p = s.planes()
            assert p.shape == (14, s.height, s.width) and np.all(np.isfinite(p))
            frac = p[4:].sum(axis=0)
            assert np.allclose(frac, 1.0, atol=1e-6)


def test_cold_energy_laid_down_with_the_gas():
    """With the shipped pairs on (no overrides): a `t` object's cells hold Sigma (T / (mu (gamma - 1)) + eps_cold),
    a `pressure` object's eps_th = (Pi - Pi_e - Pi_n) / ((gamma - 1) Sigma), the floor cells the floor state with its
    cold energy (scene.rs:446-470); eps_cold and Pi_cold from the oracle's f64 restatement."""
    import oracle
    d = base()
    d.pop("overrides")
    d["objects"][0].update(composition={"H": 0.6, "He": 0.3, "n": 0.1}, t=2.0)
    d["objects"].append({"kind": "rect", "x0": 20, "y0": 20, "w": 4, "h": 4, "sigma": 0.5, "pressure": 10.0,
                         "composition": {"H": 1.0}})
    s = scene_json.parse(json.dumps(d))
    p = s.planes().astype(np.float64)
    ph = oracle.reference_physics(**oracle.SHIPPED_COLD)
    ye = 0.6 * 1.0 + 0.3 * 0.5
    inv_mu = 0.6 * 2.0 + 0.3 * 0.75 + 0.1 * 1.0
    want = 1.0 * (2.0 * inv_mu + oracle.eps_cold(ph, 1.0, ye, 0.1))
    assert abs(p[3][3, 3] / want - 1.0) < 1e-6
    eth = (10.0 - oracle.pi_cold(ph, 0.5, 1.0, 0.0)) / 0.5
    assert abs(p[3][21, 21] / (0.5 * (eth + oracle.eps_cold(ph, 0.5, 1.0, 0.0))) - 1.0) < 1e-6
    sf = 1e-8
    assert abs(p[3][40, 40] / (sf * (0.05 * 2.0 + oracle.eps_cold(ph, sf, 1.0, 0.0))) - 1.0) < 1e-6
    kw = s.physics_kwargs()
    assert (kw["k1e"], kw["k2e"], kw["k1n"], kw["k2n"]) == (21.3, 134.0, 2.67, 58.2)


def test_host_cold_eos_is_the_oracles():
    import oracle
    from sandbox_reactions_amd import cold
    for k1, k2 in ((21.3, 134.0), (2.67, 58.2)):
        assert np.max(np.abs(cold.table(k1, k2) - oracle.cold_table(k1, k2))) < 1e-12
        xs = np.exp(np.linspace(np.log(1e-11), np.log(1e9), 500))
        u = cold.energy(xs, k1, k2)
        t = oracle.cold_table(k1, k2)
        assert np.max(np.abs(u / np.array([oracle.cold_table_u(t, x) for x in xs]) - 1.0)) < 1e-12
        assert np.max(np.abs(cold.pressure(xs, k1, k2) / np.array([oracle.cold_pressure(x, k1, k2) for x in xs])
                             - 1.0)) < 1e-13


def test_a_resized_world_keeps_the_objects_about_its_centre():
    """--world over a scene shifts every object by half the change in size (scene.rs:428-435): a disk at the
    world's centre stays at the new centre, its cells the same values, and a rect keeps its offset from the
    centre."""
    s = load("ref_disk.json")
    base = s.planes().astype(np.float64)
    s.resize(128, 96)
    p = s.planes().astype(np.float64)
    assert p.shape == (14, 96, 128)
    yy, xx = np.mgrid[0:96, 0:128]
    m = p[0] - 1e-8
    cx, cy = (m * (xx + 0.5)).sum() / m.sum(), (m * (yy + 0.5)).sum() / m.sum()
    assert abs(cx - 64.0) < 1e-9 and abs(cy - 48.0) < 1e-9
    assert np.array_equal(p[:, 16:80, 32:96], base)
    r = load("sod.json")  # 400 x 64, the left half's rect from x = 0
    r.resize(416, 80)
    q = r.planes()
    assert np.array_equal(q[:, 8:72, 8:408], load("sod.json").planes())
    r.resize(400, 64)
    assert np.array_equal(r.planes(), load("sod.json").planes())
    with pytest.raises(scene_json.SceneError, match="world size"):
        r.resize(100, 64)


def test_the_burning_scene_carries_p6():
    """ref_burn.json (heat and reactions on, burn_k 10): every shipped record with burn_k from the scene's overrides,
    the cooling off with the neutrinos switch; a hot start at Sigma 30 >= sigma_hold is not held (eps_th far above
    eta_hold eps_cold), while a cold one is refused; reactions off gives no parameters."""
    from sandbox_reactions_amd import react
    s = load("ref_burn.json")
    s.check_path()
    p = s.react_params()
    assert p.n_records == len(react.RECORDS) and p.burn_k == 10.0 and p.neutrinos == 0 and p.pin_temperature == 0
    assert p.burn_dx == react.CONSTANTS["burn_dx"]
    s.planes()
    d = json.load(open(os.path.join(HERE, "ref_burn.json")))
    d["objects"][0]["t"] = 1.0
    with pytest.raises(scene_json.SceneError, match="starts held"):
        scene_json.parse(json.dumps(d)).planes()
    d["switches"]["reactions"] = False
    d["objects"][0]["t"] = 100.0
    assert scene_json.parse(json.dumps(d)).react_params() is None


def test_the_sink_scene_carries_p3():
    """sink_cloud.json (gravity, hydro and sinks on, the cold pressure off by its overrides): P3 with Sigma_BH 6000 and
    no sink from mass; with the shipped neutron pair m_tov_sb is the reference's 4,853.17; sinks off gives none."""
    s = load("sink_cloud.json")
    s.check_path()
    assert s.sinks_params() == (6000.0, math.inf)
    d = json.load(open(os.path.join(HERE, "sink_cloud.json")))
    d["overrides"] = {"sigma_bh": 5000.0}
    assert scene_json.parse(json.dumps(d)).sinks_params() == (5000.0, scene_json.M_TOV_SHIPPED)
    d["switches"]["sinks"] = False
    assert scene_json.parse(json.dumps(d)).sinks_params() is None
