"""The radiation force in the flow (crates/sr-engine/tests/gpu/radforce.rs): f = kappa Sigma F / c_sb from the previous
step's flux F is P4's momentum source, E unchanged by it; the push's work is capped at the heat above the floor
(push_share, reference/hydro.rs:389-410), and the grid's net push is taken back from each pushed cell in proportion to
its push (hydro.rs:986-996).

The scene (radforce.rs:5-10): 64 x 64, bounce, hydrogen at rest, the cold pressure off; a medium of Sigma 1 at T_m and a
slab (columns 24..39, every row) of Sigma 0.1 at T_s, with Pi = 2 on both sides at (T_m, T_s) = (1, 10). The oracle is
graded from the reference's words, not the scheme: the forced step against the force-free step from the same state
at the same dt, so what the force adds is, per cell, s d - |s d| N / S with d = dt kappa Sigma F / c_sb (kappa =
kappa_0 (1 + X_H)), s the largest share <= 1 whose work (|p0 + s d|^2 - |p0|^2) / 2 Sigma is at most
Sigma (eps_th - T_floor / mu) of the free step's cell (0 when that is not positive), found here by bisection, and N, S
the net and gross of the pushes s d; Sigma, E and every species are the free step's bit for bit. Three F fields
written by hand (the reference writes one by hand too): uncapped (T 1 / 10, a weak F), capped at the slab's edges (T
0.06 / 1000, a strong F), and a net push (F's left half halved).

On the GPU the force comes from the world's own heat pass: the scene with heat on and the temperatures pinned (P5
computes F and leaves E alone, the reference's set-up) is stepped against the twin on the same dt."""
import numpy as np
import pytest

import oracle
from helpers import both, rel_l1

N = 64
DT = 0.01


def slab_state(t_medium, t_slab):
    x = np.arange(N)[None, :].repeat(N, 0)
    slab = (x >= 24) & (x < 40)
    p = np.zeros((14, N, N))
    p[0] = np.where(slab, 0.1, 1.0)
    p[3] = p[0] * 2.0 * np.where(slab, t_slab, t_medium)  # E = Sigma T / mu / (gamma - 1), hydrogen, gamma 2, at rest
    p[4] = 1.0
    return p.astype(np.float32).astype(np.float64)


def f_field(f0, left_half=1.0):
    """F pointing away from the slab's centre line along x, a weak sinusoid along y."""
    x = np.arange(N) + 0.5
    fx = f0 * np.tanh((x - 32.0) / 4.0)[None, :].repeat(N, 0)
    fx[:, :32] *= left_half
    fy = 0.1 * f0 * np.sin(2 * np.pi * (np.arange(N) + 0.5) / N)[:, None].repeat(N, 1)
    return fx, fy


def expected_gain(ph, start, free, fx, fy):
    """Per cell, s d - |s d| N / S from the reference's words (radforce.rs:20-30, :35-36)."""
    kap = ph.kappa0 * (1.0 + start[4])
    d = np.stack([DT * kap * start[0] * fx / ph.c_sb, DT * kap * start[0] * fy / ph.c_sb])
    sig, p0 = free[0], free[1:3]
    eth = free[3] / sig - 0.5 * (p0 ** 2).sum(0) / sig ** 2
    heat = sig * (eth - ph.t_floor * 2.0)  # T_floor / mu of hydrogen at gamma = 2

    def over(s):
        q = p0 + s * d
        return ((q ** 2).sum(0) - (p0 ** 2).sum(0)) / (2.0 * sig) - heat

    lo, hi = np.zeros((N, N)), np.ones((N, N))
    whole = over(hi) <= 0
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        ok = over(mid) <= 0
        lo, hi = np.where(ok, mid, lo), np.where(ok, hi, mid)
    s = np.where(heat <= 0, 0.0, np.where(whole, 1.0, lo))
    push = s * d
    net = push.sum((1, 2))
    gross = np.hypot(push[0], push[1]).sum()
    return push - np.hypot(push[0], push[1])[None] * net[:, None, None] / gross, s, d


@pytest.mark.parametrize("case", ["uncapped", "capped", "net_push"])
def test_the_force_is_the_capped_push_less_the_net(case):
    t_m, t_s, f0, left = {"uncapped": (1.0, 10.0, 6400.0, 1.0), "capped": (0.06, 1000.0, 640000.0, 1.0),
                          "net_push": (1.0, 10.0, 6400.0, 0.5)}[case]
    op, _ = both(edge=0)
    st = slab_state(t_m, t_s)
    fx, fy = f_field(f0, left)
    free = oracle.hydro_step(op, st, DT)
    forced = oracle.hydro_step(op, st, DT, radiation=(fx, fy))
    for c in [0, 3] + list(range(4, 14)):
        assert np.array_equal(forced[c], free[c]), c
    want, s, d = expected_gain(op, st, free, fx, fy)
    gain = forced[1:3] - free[1:3]
    assert np.max(np.abs(gain - want)) <= 1e-12 * np.abs(d).max()
    capped = int((s < 1).sum())
    if case == "capped":
        assert capped > 1000 and s.min() < 0.5
    else:
        assert capped == 0
    moved = forced[1:3].sum((1, 2)) - free[1:3].sum((1, 2))
    assert np.all(np.abs(moved) <= 1e-12 * np.abs(d).sum())
    if case == "net_push":
        net = (s * d).sum((1, 2))
        assert net[0] > 1e-3 * np.abs(d[0]).sum()  # a real net push, taken back


@pytest.mark.gpu
@pytest.mark.parametrize("temps", [(1.0, 10.0), (0.06, 1000.0)])
def test_the_gpu_force_from_its_heat_pass_matches_the_twin(gpu_lib, temps):
    """The slab with hydro and heat on, temperatures pinned (radforce.rs:7-8), 5 steps on the twin's dt: momenta within
    1e-5 (rel-L1) of the twin's, Sigma 1e-6, and E as the twin's (the force moves no energy)."""
    import sandbox_reactions_amd as sr
    print(f"radiation force slab {temps}", flush=True)
    op, gp = both(edge=0, pin_temperature=1)
    st = slab_state(*temps)
    world = sr.World(N, N, physics=gp, gravity=False, hydro=True, heat=True)
    world.upload(st.astype(np.float32))
    twin = oracle.Twin(op, st, gravity=False, hydro=True, heat=True)
    for _ in range(5):
        dt = world.read_dt().dt_next
        twin.cap_dt_next(dt)
        twin.step()
        world.set_dt_cap(dt)
        world.run(1)
    g, t = world.readback().astype(np.float64), twin.state()
    assert rel_l1(g[0], t[0]) <= 1e-6
    mom_g, mom_t = np.hypot(g[1], g[2]), np.hypot(t[1], t[2])
    assert np.abs(mom_t).sum() > 0
    assert np.abs(g[1:3] - t[1:3]).sum() <= 1e-5 * mom_t.sum(), np.abs(g[1:3] - t[1:3]).sum() / mom_t.sum()
    assert rel_l1(g[3], t[3]) <= 1e-5
