"""P8 floors, P9's dt and the non-finite guard, on the cases of the reference's tests (crates/sr-engine/tests/gpu/floors.rs
and dt.rs), graded by what fixes them independently of the oracle's code:

- floors (floors.rs:109-175): fractions clipped at 0 and renormalised, pure hydrogen when none is positive; a cell below
  Sigma_vac becomes the floor state (Sigma_floor at rest, E = Sigma_floor T_floor / (mu (gamma - 1))); a cell whose
  thermal energy is below the floor's is raised to it exactly, its momentum and Sigma untouched; every other cell
  untouched; the booked vacuum reset and floor energy equal the mass and energy the pass changed. 1/mu of each pure
  species is the fully ionised gas's (Z + 1)/A, written here as numbers;
- dt (hydro.rs:618-634): on fields built from chosen (u, v, c), dt = C / max(|u| + |v| + 2c) from the chosen values,
  and C / 25 where a planted cell has u = 8, v = -5, c = 6; the gravity term eta_g sqrt(dx / |g|max) where it wins; an
  empty world steps on its floor gas;
- the guard (step/dt.rs:40-47, headless.rs:620-625): the first non-finite value by cell, then channel, with its step.

The CPU cases grade the oracle; the GPU cases run the library's own floors + dt pass (a world with every other pass
off) against the same grades and the oracle. The reference's tolerances: f32 arithmetic 2e-6 relative, dt 1e-6."""
import os
import subprocess
import sys

import numpy as np
import pytest

import oracle
from helpers import REFERENCE, both

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
C_CFL, ETA_G = 0.4, 0.3
# 1/mu = (Z + 1)/A for H, He, C, O, Ne, Mg, Si, S, Fe, n (fully ionised; a free neutron is one particle of mass 1)
INV_MU = np.array([2.0, 3.0 / 4.0, 7.0 / 12.0, 9.0 / 16.0, 11.0 / 20.0, 13.0 / 24.0, 15.0 / 28.0, 17.0 / 32.0,
                   27.0 / 56.0, 1.0])
T_FLOOR = REFERENCE["t_floor"]
SVAC = float(np.float32(REFERENCE["sigma_vac"]))
SFLOOR = float(np.float32(REFERENCE["sigma_floor"]))


def floors_field(w=64, h=64, seed=19):
    """Vacuum (Sigma 0, 1e-7, negative), cold (thermal energy below the floor's, moving), dense cold and hot cells, by
    cell index mod 4; fractions normalised, unnormalised (sum 2.5), of mixed signs or all <= 0, by (index / 4) mod 4."""
    rng = np.random.default_rng(seed)
    n = w * h
    p = np.zeros((14, n))
    kind = np.arange(n) % 4
    frac = (np.arange(n) // 4) % 4
    raw = rng.random((10, n))
    x = raw / raw.sum(0)
    x = np.where(frac == 1, 2.5 * x, x)
    x = np.where(frac == 2, x - 0.05, x)
    x = np.where(frac == 3, -np.abs(x), x)
    p[4:] = x
    eps_fl = T_FLOOR * (INV_MU[:, None] * renormalised(x)).sum(0)  # gamma = 2
    sigma = np.select([kind == 0, kind == 1, kind == 2], [rng.choice([0.0, 1e-7, -1e-3], n),
                                                         rng.uniform(0.1, 2.0, n), rng.uniform(20.0, 200.0, n)],
                      rng.uniform(0.5, 2.0, n))
    u, v = rng.uniform(-2, 2, n), rng.uniform(-2, 2, n)
    eps = np.where(kind == 3, eps_fl * rng.uniform(1.5, 50.0, n), eps_fl * rng.uniform(0.0, 0.9, n))
    p[0], p[1], p[2] = sigma, sigma * u, sigma * v
    p[3] = sigma * (eps + 0.5 * (u * u + v * v))
    return p.reshape(14, h, w).astype(np.float32).astype(np.float64)


def renormalised(x):
    """Fractions clipped at 0 over their sum; pure hydrogen where none is positive (floors.rs:45-60)."""
    xc = np.clip(x, 0.0, None)
    s = xc.sum(0)
    xn = xc / np.where(s > 0, s, 1.0)
    xn[0] = np.where(s > 0, xn[0], 1.0)
    return xn


def expected_floors(p, gamma=2.0):
    """The floors' result from their rules, in f64 from the f32 inputs."""
    q = p.copy()
    xn = renormalised(p[4:])
    inv_mu = (INV_MU[:, None, None] * xn).sum(0)
    eps_fl = T_FLOOR * inv_mu / (gamma - 1.0)
    vac = p[0] < SVAC
    q[0] = np.where(vac, SFLOOR, p[0])
    q[1] = np.where(vac, 0.0, p[1])
    q[2] = np.where(vac, 0.0, p[2])
    s = np.where(vac, 1.0, p[0])
    ke = 0.5 * (p[1] ** 2 + p[2] ** 2) / s
    eth = p[3] / s - ke / s
    raised = ~vac & (eth < eps_fl)
    q[3] = np.where(vac, SFLOOR * eps_fl, np.where(raised, p[0] * eps_fl + ke, p[3]))
    q[4:] = xn
    return q, vac, raised


@pytest.mark.parametrize("gamma", [2.0, 5.0 / 3.0])
def test_the_floors_follow_their_rules_in_the_oracle(gamma):
    p = floors_field()
    ph, _ = both(gamma=gamma)
    out, vr, fa = oracle.floors_apply(ph, p)
    want, vac, raised = expected_floors(p, gamma)
    assert vac.sum() > 500 and raised.sum() > 1000 and (~vac & ~raised).sum() > 1000
    assert np.all(out[0] == want[0]) and np.all(out[1:3] == want[1:3])
    assert np.allclose(out[3], want[3], rtol=1e-12, atol=0)
    assert np.allclose(out[4:], want[4:], rtol=0, atol=1e-15)
    untouched = ~vac & ~raised
    assert np.all(out[:4, untouched] == p[:4, untouched])
    # the booked terms: the mass the reset removed, the energy the floors added
    assert abs(vr - (p[0].sum() - out[0].sum())) <= 1e-12 * np.abs(p[0]).sum()
    assert abs(fa - (out[3].sum() - p[3].sum())) <= 1e-12 * np.abs(out[3]).sum()


def test_a_fixed_mu_replaces_the_species():
    """physics.json's mu > 0 fixes 1/mu for every cell: the vacuum's E is Sigma_floor T_floor / mu whatever its mix."""
    p = floors_field()
    ph, _ = both(mu=0.6)
    out = oracle.floors_apply(ph, p)[0]
    vac = p[0] < SVAC
    assert np.allclose(out[3, vac], SFLOOR * T_FLOOR / 0.6, rtol=1e-12, atol=0)


def known_field(w, h, seed, peak):
    """The reference's dt field (dt.rs:86-112): Sigma in [0.5, 2), (u, v) in [-2, 2)^2, c in [0.5, 3), fractions drawn and
    normalised; every 97th cell vacuum (Sigma 0, x_H 1); the cell at `peak` moving u = 8, v = -5 with c = 6. Returns the
    f32 planes and the rate |u| + |v| + 2c of each cell from the chosen values."""
    rng = np.random.default_rng(seed)
    n = w * h
    p = np.zeros((14, n))
    sigma = rng.uniform(0.5, 2.0, n)
    u, v, c = rng.uniform(-2, 2, n), rng.uniform(-2, 2, n), rng.uniform(0.5, 3.0, n)
    if peak is not None:
        u[peak], v[peak], c[peak] = 8.0, -5.0, 6.0
    x = rng.random((10, n))
    x /= x.sum(0)
    p[0], p[1], p[2] = sigma, sigma * u, sigma * v
    p[3] = sigma * (0.5 * c * c + 0.5 * (u * u + v * v))
    p[4:] = x
    rate = np.abs(u) + np.abs(v) + 2 * c
    vac = (np.arange(n) % 97 == 13) & (np.arange(n) != (peak if peak is not None else -1))
    p[:, vac] = 0.0
    p[4, vac] = 1.0
    rate[vac] = 2.0 * np.sqrt(2.0 * T_FLOOR * 2.0)  # floor gas of hydrogen at rest
    return p.reshape(14, h, w).astype(np.float32).astype(np.float64), rate


CASES = [((64, 64), 0), ((72, 88), 72 * 88 - 1), ((600, 400), 123457), ((600, 400), 0), ((64, 64), None),
         ((72, 88), None)]


@pytest.mark.parametrize("size, peak", CASES)
def test_the_oracle_dt_is_the_cfl_formula(size, peak):
    (w, h) = size
    p, rate = known_field(w, h, 0x5EED0020 + w + (peak or 0), peak)
    ph, _ = both()
    dt = oracle.hydro_dt(ph, p)
    assert abs(dt / (C_CFL / rate.max()) - 1) <= 1e-6, (dt, C_CFL / rate.max())
    if peak is not None:
        assert abs(dt / (C_CFL / 25.0) - 1) <= 1e-6


def test_the_gravity_term_bounds_dt_where_it_wins():
    p, rate = known_field(64, 64, 5, None)
    ph, _ = both()
    gx = np.zeros((64, 64))
    gy = np.zeros((64, 64))
    gx[10, 20], gy[10, 20] = 300.0, -400.0  # |g| = 500
    dt = oracle.hydro_dt(ph, p, gx, gy)
    want = min(C_CFL / rate.max(), ETA_G * np.sqrt(1.0 / 500.0))
    assert ETA_G * np.sqrt(1.0 / 500.0) < C_CFL / rate.max()
    assert abs(dt / want - 1) <= 1e-12


# ---- the library's floors + dt pass on the GPU ----

def bare_world(w, h, **over):
    import sandbox_reactions_amd as sr
    _, gp = both(**over)
    return sr.World(w, h, physics=gp, gravity=False, hydro=False, heat=False)


@pytest.mark.gpu
def test_the_gpu_floors_follow_their_rules_and_the_oracle(gpu_lib):
    print("floors 64x64 field", flush=True)
    p = floors_field()
    world = bare_world(64, 64)
    world.upload(p)
    world.run(1)
    g = world.readback().astype(np.float64)
    want, vac, raised = expected_floors(p)
    ph, _ = both()
    twin = oracle.floors_apply(ph, p)[0]
    untouched = ~vac & ~raised
    assert np.all(g[:4, untouched] == p[:4, untouched])
    assert np.all(g[0, vac] == SFLOOR) and np.all(g[1:3, vac] == 0)
    assert np.all(g[1:3, raised] == p[1:3, raised]) and np.all(g[0, raised] == p[0, raised])
    for ref in (want, twin):
        assert np.max(np.abs(g[3] - ref[3]) / np.abs(ref[3])) <= 2e-6
        assert np.max(np.abs(g[4:] - ref[4:])) <= 2e-6
    # the raised cells' thermal energy lands on the floor's
    s = p[0, raised]
    eth = g[3, raised] / s - 0.5 * (g[1, raised] ** 2 + g[2, raised] ** 2) / s ** 2
    eps_fl = T_FLOOR * (INV_MU[:, None] * renormalised(p[4:, raised])).sum(0)
    assert np.max(np.abs(eth - eps_fl) / (eps_fl + 0.5 * (g[1, raised] ** 2 + g[2, raised] ** 2) / s ** 2)) <= 4e-6
    # P9 on the floored state
    assert abs(world.read_dt().dt_next / oracle.hydro_dt(ph, g) - 1) <= 1e-6


@pytest.mark.gpu
@pytest.mark.parametrize("size, peak", CASES)
def test_the_gpu_dt_is_the_cfl_formula(gpu_lib, size, peak):
    (w, h) = size
    print(f"dt field {w}x{h} peak {peak}", flush=True)
    p, rate = known_field(w, h, 0x5EED0020 + w + (peak or 0), peak)
    world = bare_world(w, h)
    world.upload(p)
    world.run(1)
    r = world.read_dt()
    ph, _ = both()
    assert abs(r.dt / (C_CFL / rate.max()) - 1) <= 1e-6, (r.dt, C_CFL / rate.max())
    assert abs(r.dt_next / oracle.hydro_dt(ph, world.readback().astype(np.float64)) - 1) <= 1e-6
    assert (abs(r.dt / (C_CFL / 25.0) - 1) <= 1e-6) == (peak is not None)
    assert r.steps == 1 and not r.non_finite


@pytest.mark.gpu
def test_an_empty_world_steps_on_its_floor_gas(gpu_lib):
    """dt.rs:210-236's world with no gas, its vacuum iron: every cell is read as floor gas at rest, so dt = C / 2c with
    c^2 = 2 T_floor / mu; the clock is the f64 sum of the steps. The floor gas here is the cell's own (iron, 1/mu =
    27/56), as the twin reads it (reference/hydro.rs:424-435); the reference's engine steps an empty world on
    hydrogen's, a consequence of its active-box fit (an empty box), which this library does not carry."""
    p = np.zeros((14, 64, 64))
    p[4 + 8] = 1.0
    world = bare_world(64, 64)
    world.upload(p)
    world.run(10)
    r = world.read_dt()
    c_fe = np.sqrt(2.0 * T_FLOOR * INV_MU[8])
    assert abs(r.dt / (C_CFL / (2 * c_fe)) - 1) <= 1e-6 and r.dt_next == r.dt
    t = 0.0
    for _ in range(10):
        t += r.dt
    assert r.sim_time == t


@pytest.mark.gpu
def test_the_guard_names_the_first_non_finite_value(gpu_lib):
    """dt.rs:262-300: a NaN energy at (37, 21) before step 0 is named with step 0; two values planted into a clean
    run's state (Inf mom_y at (5, 60), NaN energy at (50, 2), the lower cell index) name (50, 2) and energy, and the
    first finding is kept over later steps; a NaN x_C is named and its cell's fractions are left as planted."""
    w = h = 64
    p, _ = known_field(w, h, 3, None)
    q = p.copy()
    q[3, 21, 37] = np.nan
    world = bare_world(w, h)
    world.upload(q)
    world.run(5)
    r = world.read_dt()
    assert (r.non_finite, r.nf_step, r.nf_x, r.nf_y, r.nf_channel) == (1, 0, 37, 21, 3)
    world.upload(p)
    world.run(70)
    assert not world.read_dt().non_finite
    planes = world.readback()
    planes[2, 60, 5] = np.inf
    planes[3, 2, 50] = np.nan
    world.upload(planes)
    world.run(3)
    world.run(2)
    r = world.read_dt()
    assert (r.non_finite, r.nf_step, r.nf_x, r.nf_y, r.nf_channel) == (1, 0, 50, 2, 3)
    world.run(80)
    r2 = world.read_dt()
    assert (r2.non_finite, r2.nf_step, r2.nf_x, r2.nf_y, r2.nf_channel) == (1, 0, 50, 2, 3) and r2.steps == 85
    q = p.copy()
    q[4 + 2, 10, 20] = np.nan
    world.upload(q)
    world.run(1)
    r = world.read_dt()
    assert (r.non_finite, r.nf_step, r.nf_x, r.nf_y, r.nf_channel) == (1, 0, 20, 10, 6)
    g = world.readback()
    assert np.array_equal(g[4:, 10, 20].view(np.uint32), q[4:, 10, 20].astype(np.float32).view(np.uint32))


@pytest.mark.gpu
def test_headless_exits_5_with_the_reference_line(gpu_lib, tmp_path):
    """dt.rs:302-369's scene (every physics off, a 10 x 10 hydrogen rect) run 3 steps to a dump; a NaN energy planted at
    (25, 25) in it; resumed for 5 steps: exit 5 and the reference's line, naming step 3 (the first step after the
    resume, which this library's guard scans; the reference's scans only at its schedule and names step 7); the clean
    dump resumed the same way: exit 0."""
    import json
    from sandbox_reactions_amd import dump
    scene = tmp_path / "block.json"
    scene.write_text(json.dumps({"version": 1, "world": {"width": 64, "height": 64, "edge": "leave"},
                                 "switches": {"gravity": False, "hydro": False, "heat": False, "reactions": False,
                                              "neutrinos": False, "sinks": False},
                                 "objects": [{"kind": "rect", "x0": 20, "y0": 20, "w": 10, "h": 10, "sigma": 1.0,
                                              "t": 2.0, "composition": {"H": 1.0}, "v": [0, 0]}]}))

    def headless(*args):
        r = subprocess.run([sys.executable, "-m", "sandbox_reactions_amd.headless", "--scene", str(scene), *args],
                           cwd=ROOT, capture_output=True, text=True, timeout=300)
        return r.returncode, r.stdout.strip().split("\n")

    code, lines = headless("--steps", "3", "--out", str(tmp_path / "start"))
    assert code == 0, lines
    d = dump.read(str(tmp_path / "start" / "dumps" / "s_3.bin"))
    assert d["step"] == 3 and np.all(np.isfinite(d["planes"]))
    clean = tmp_path / "clean.bin"
    dump.write(str(clean), d["planes"], step=3, sim_time=d["sim_time"], dt=d["dt"], dt_next=d["dt_next"])
    planes = d["planes"].copy()
    planes[3, 25, 25] = np.nan
    nan = tmp_path / "nan.bin"
    dump.write(str(nan), planes, step=3, sim_time=d["sim_time"], dt=d["dt"], dt_next=d["dt_next"])
    code, lines = headless("--load-dump", str(clean), "--steps", "5")
    assert code == 0 and lines[-1].startswith("SR-HEADLESS DONE steps=5 "), lines
    code, lines = headless("--load-dump", str(nan), "--steps", "5")
    assert code == 5, lines
    assert lines[-1] == "SR-ERROR non-finite value in the state at step 3, cell (25, 25), channel energy", lines
