"""The cold collapse (crates/sr-engine/tests/gpu/collapse.rs): gravity's source in the flow graded by a closed form. The
reference's scenes/cold_disk.json is a cold Maclaurin disk Sigma_c sqrt(1 - r^2/a^2), Sigma_c = 1, a = 40, on 128 x 128,
gravity and hydro on, the cold pressure off; it falls in homologously, every radius scaled by cos^2 eta at
t/t_ff = (2/pi)(eta + sin(2 eta)/2), t_ff = sqrt(a / G Sigma_c)/2, so the half-mass radius reads r50(t)/r50(0) =
0.8368 at 0.5 t_ff and 0.5279 at 0.8 t_ff; the reference grades each within 3 % on its twin and on its GPU.

Here the oracle's twin (FFT gravity each step, dt with its gravity term, the last step to each read time cut to land
on it) and the GPU world (gravity + hydro, the same cut through its dt cap) are graded the same way, and the GPU's
ratio is held to the twin's."""
import math
import os

import numpy as np
import pytest

import oracle
from helpers import both, mass_radius
from sandbox_reactions_amd import scene_json

SCENE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "scenes", "cold_disk.json")
A, SIGMA_C = 40.0, 1.0
READS = (0.5, 0.8)
TOLERANCE = 0.03


def t_ff():
    return 0.5 * math.sqrt(A / SIGMA_C)


def homologous(tau):
    """cos^2 eta with (2/pi)(eta + sin(2 eta)/2) = tau, eta by bisection on [0, pi/2] (collapse.rs:66-78)."""
    lo, hi = 0.0, 0.5 * math.pi
    for _ in range(100):
        mid = 0.5 * (lo + hi)
        if (2.0 / math.pi) * (mid + 0.5 * math.sin(2.0 * mid)) < tau:
            lo = mid
        else:
            hi = mid
    return math.cos(0.5 * (lo + hi)) ** 2


def start():
    s = scene_json.load(SCENE)
    kw = s.physics_kwargs()
    return s, kw, s.planes(kw["gamma"]).astype(np.float64)


def test_the_closed_form_gives_the_reference_numbers():
    """0.8368 at 0.5 t_ff, 0.5279 at 0.8 t_ff (cut to four digits), t_ff = 3.162, the ends 1 and 0 (collapse.rs:175-184);
    and the curve against a numerical integration of the edge's equation of motion, which fixes it independently.
    Inside a Maclaurin disk of radius R the pull is g = (pi^2 G Sigma_c / 2R) r, linear in r, so a pressureless disk
    stays a Maclaurin disk; its mass 2 pi Sigma_c R^2 / 3 is conserved, so Sigma_c = Sigma_c0 a^2 / R^2 and the edge obeys
    R'' = -k / R^2, k = pi^2 G Sigma_c0 a^2 / 2 (whose fall from rest takes (pi/2) sqrt(a^3 / 2k) = t_ff)."""
    assert abs(homologous(0.5) - 0.8368) < 1e-4 and abs(homologous(0.8) - 0.5279) < 1e-4
    assert abs(t_ff() - 3.162) < 5e-4
    assert homologous(1.0) < 1e-9 and abs(homologous(0.0) - 1.0) < 1e-12
    k = 0.5 * math.pi ** 2 * SIGMA_C * A ** 2
    r, v, t, dt = A, 0.0, 0.0, 1e-4
    for tau in READS:
        until = tau * t_ff()
        while t < until - 1e-12:
            h = min(dt, until - t)
            # velocity Verlet on R'' = -k/R^2
            v += 0.5 * h * (-k / r ** 2)
            r += h * v
            v += 0.5 * h * (-k / r ** 2)
            t += h
        assert abs(r / A - homologous(tau)) < 1e-5, (tau, r / A, homologous(tau))


def test_the_scene_is_the_cold_maclaurin_disk():
    """Sigma = Sigma_c sqrt(1 - r^2/a^2) about (64, 64), at rest, hydrogen, E = Sigma T_floor / mu (mu = 1/2); gravity and
    hydro on, the rest off; only the four cold-pressure constants overridden, to 0 (collapse.rs:146-173)."""
    s, kw, st = start()
    assert s.switches["gravity"] and s.switches["hydro"]
    assert not (s.switches["heat"] or s.switches["reactions"] or s.switches["neutrinos"] or s.switches["sinks"])
    assert sorted(s.overrides) == ["k1e", "k1n", "k2e", "k2n"] and not s.cold_on()
    h, w = st.shape[1:]
    y, x = np.mgrid[0:h, 0:w]
    r = np.hypot(x + 0.5 - 64.0, y + 0.5 - 64.0)
    inside = r < A
    sigma = np.where(inside, SIGMA_C * np.sqrt(np.clip(1.0 - r * r / (A * A), 0.0, None)), 0.0)
    assert np.all(np.abs(st[0][inside] - sigma[inside]) <= 1e-6)
    assert np.all(st[1][inside] == 0) and np.all(st[2][inside] == 0)
    e = st[0][inside] * kw["t_floor"] * 2.0
    assert np.all(np.abs(st[3][inside] - e) <= 1e-6 * e)
    mass = st[0][inside].sum()
    exact = 2.0 * math.pi * SIGMA_C * A * A / 3.0
    assert abs(mass / exact - 1.0) < 2.0 * math.sqrt(2.0) * A ** -1.5


def test_the_half_mass_radius_of_the_sampled_disk():
    """r50 of the scene's own disk against the continuum's: M(<r) ~ 1 - (1 - r^2/a^2)^(3/2), so r50 = a sqrt(1 - 2^(-2/3))
    = 0.6083 a; the grid's sampling moves it by well under a cell. A wrong fraction or a lost centre moves it by cells."""
    _, _, st = start()
    r50, _ = mass_radius(st[0], fraction=0.5)
    assert abs(r50 - A * math.sqrt(1.0 - 2.0 ** (-2.0 / 3.0))) < 0.1, r50
    r90, _ = mass_radius(st[0], fraction=0.9)
    assert abs(r90 - A * math.sqrt(1.0 - 0.1 ** (2.0 / 3.0))) < 0.1, r90


def grade(name, reads):
    for tau_read, (t, ratio) in zip(READS, reads):
        tau = t / t_ff()
        expected = homologous(tau)
        err = ratio / expected - 1.0
        print(f"{name}: t = {t:.5f} ({tau:.5f} t_ff), r50/r50(0) = {ratio:.5f}, homologous {expected:.5f}, "
              f"error {100 * err:+.3f} %", flush=True)
        assert abs(tau - tau_read) < 0.05
        assert abs(err) <= TOLERANCE, (name, tau, ratio, expected)


def run_collapse(step, sim_time, cap, sigma):
    """Steps until each read time, the last step capped to land on it; (sim time, r50/r50(0)) per read."""
    r0, _ = mass_radius(sigma(), fraction=0.5)
    reads, steps = [], 0
    for tau in READS:
        until = tau * t_ff()
        while sim_time() < until:
            cap(until - sim_time())
            step()
            steps += 1
            assert steps < 2000
        reads.append((sim_time(), mass_radius(sigma(), fraction=0.5)[0] / r0))
    return reads


def twin_reads():
    s, kw, st = start()
    op, _ = both(**kw)
    twin = oracle.Twin(op, st, gravity=True, hydro=True, heat=False)
    return run_collapse(twin.step, lambda: twin.times()[2], twin.cap_dt_next, lambda: twin.state()[0])


def test_the_twin_collapses_homologously():
    grade("twin", twin_reads())


@pytest.mark.gpu
def test_the_gpu_collapses_homologously(gpu_lib):
    """The GPU world (gravity + hydro, the scene's leave edge) graded the same way, and its ratios within 0.5 % of the
    twin's on the same read times."""
    s, kw, _ = start()
    world = scene_json.world(s)
    reads = run_collapse(lambda: world.run(1), lambda: world.read_dt().sim_time, world.set_dt_cap,
                         lambda: world.readback()[0])
    grade("gpu", reads)
    for (tg, rg), (tt, rt) in zip(reads, twin_reads()):
        assert abs(tg / tt - 1) < 1e-9 and abs(rg / rt - 1) < 5e-3, (tg, tt, rg, rt)
