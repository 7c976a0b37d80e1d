"""The radiating slab of the reference's heat tests (crates/sr-engine/tests/gpu/heat.rs:457-502, :783-811): a hot hydrogen
slab (Sigma 1, T 30) at the world's left edge beside vacuum, 64 x 64, heat only, the cold pressure off. These pin the
heat pass's Newton passes on a radiating scene by what the result must satisfy, computed here from the pass's own
output:

- one pass of dt = 2 solves backward Euler's nonlinear balance in every gas cell with the flux it kept,
  |C (T - T0) + dt sum_f F_f| <= 1e-8 (C T0 + dt sum_f |F_f|), C = Sigma / mu (gamma = 2), in at least 3 Newton passes
  (the radiative law is nonlinear), and the grid loses what it books as radiated (1e-9);
- over 12 passes of dt = 2 the grid loses what it books (1e-6), the emitted heat sits exactly on the slab's surface
  ring (x = 0, x = 31, y = 16, y = 47: 124 cells) and nowhere else, and no vacuum cell changes.

On the GPU the world steps the same scene at its own dt (P9's, ~0.018), against the twin on that dt."""
import json

import numpy as np
import pytest

import oracle
from helpers import both, rel_l1
from sandbox_reactions_amd import scene_json

INV_MU = 2.0  # hydrogen


def rect_world(w=64, h=64, rect=(0, 16, 32, 32), sigma=1.0, t=30.0):
    """heat.rs:85-95's inline scene: heat only, bounce, the cold pressure off, one rect of hydrogen."""
    x0, y0, rw, rh = rect
    d = {"version": 1, "world": {"width": w, "height": h, "edge": "bounce"}, "slowdown": False,
         "switches": {"gravity": False, "hydro": False, "heat": True, "reactions": False, "neutrinos": False,
                      "sinks": False},
         "objects": [{"kind": "rect", "x0": x0, "y0": y0, "w": rw, "h": rh, "sigma": sigma, "t": t,
                      "composition": {"H": 1.0}, "v": [0, 0]}],
         "overrides": {"k1e": 0, "k2e": 0, "k1n": 0, "k2n": 0}}
    return scene_json.parse(json.dumps(d))


def start():
    s = rect_world()
    kw = s.physics_kwargs()
    return s, kw, s.planes(kw["gamma"]).astype(np.float32).astype(np.float64)


def temps(p):
    return p[3] / p[0] / INV_MU


def ring():
    y, x = np.mgrid[0:64, 0:64]
    slab = (x < 32) & (y >= 16) & (y < 48)
    return slab, slab & ((x == 0) | (x == 31) | (y == 16) | (y == 47))


def test_one_pass_solves_the_nonlinear_balance_cell_by_cell():
    s, kw, st = start()
    op, _ = both(**kw)
    dt = 2.0
    p, rep, fx, fy, em = oracle.heat_step(op, st, dt)
    t0, t1 = temps(st), temps(p)
    out = fx[:, 1:] - fx[:, :-1] + fy[1:, :] - fy[:-1, :]
    mag = np.abs(fx[:, 1:]) + np.abs(fx[:, :-1]) + np.abs(fy[1:, :]) + np.abs(fy[:-1, :])
    c = st[0] * INV_MU
    gas = st[0] >= kw["sigma_vac"]
    worst = np.max(np.abs(c * (t1 - t0) + dt * out)[gas] / (c * t0 + dt * mag)[gas])
    lost = st[3].sum() - p[3].sum()
    print(f"{rep.passes} Newton passes, {rep.rounds} rounds, delta {rep.delta:.1e}; worst balance {worst:.2e}; "
          f"lost {lost:.6f} of radiated {rep.radiated:.6f}", flush=True)
    assert rep.converged and rep.passes >= 3
    assert worst <= 1e-8
    assert rep.radiated > 0 and abs(lost - rep.radiated) <= 1e-9 * rep.radiated


def test_twelve_passes_book_what_the_grid_lost_on_the_slab_ring():
    s, kw, st = start()
    op, _ = both(**kw)
    slab, surface = ring()
    slab_energy = st[3][slab].sum()
    p = st.copy()
    radiated, emitted, passes = 0.0, np.zeros((64, 64)), 0
    for _ in range(12):
        p, rep, _, _, em = oracle.heat_step(op, p, 2.0)
        assert rep.converged
        radiated += rep.radiated
        emitted += em
        passes = max(passes, rep.passes)
    lost = st[3].sum() - p[3].sum()
    assert radiated / slab_energy > 0.01 and passes >= 3
    assert abs(lost - radiated) <= 1e-6 * radiated
    assert abs(emitted.sum() - radiated) <= 1e-12 * radiated
    assert np.all(emitted[surface] > 0) and np.all(emitted[~surface] == 0) and surface.sum() == 124
    assert np.array_equal(p[:, ~slab], st[:, ~slab])


@pytest.mark.gpu
def test_the_gpu_slab_books_what_it_lost_and_matches_the_twin(gpu_lib):
    """12 steps of the scene on the GPU at P9's dt, the twin on the same dt: the radiated heat the steps book equals the
    energy the grid lost (1e-5) and the twin's (1e-4), E within 1e-5 of the twin's, and no vacuum cell changes."""
    import sandbox_reactions_amd as sr
    s, kw, st = start()
    op, _ = both(**kw)
    hs = sr.heat_settings(stop=1, solve_tol=1e-10, newton_heat=1e-9, newton_cell=1e-6)
    world = scene_json.world(s, heat_settings=hs)
    world.upload(st.astype(np.float32))
    twin = oracle.Twin(op, st, gravity=False, hydro=False, heat=True)
    e0 = world.totals()[0][3]
    radiated = twin_radiated = 0.0
    for k in range(12):
        dt = world.read_dt().dt_next
        twin.cap_dt_next(dt)
        twin.step()
        world.set_dt_cap(dt)
        world.run(1)
        r = world.read_heat()
        assert r.converged
        radiated += r.radiated
        twin_radiated += twin.last_heat().radiated
    g = world.readback().astype(np.float64)
    lost = e0 - world.totals()[0][3]
    slab, _ = ring()
    print(f"GPU: lost {lost:.6f}, radiated {radiated:.6f}, twin {twin_radiated:.6f}", flush=True)
    assert abs(lost - radiated) <= 1e-5 * radiated
    assert abs(radiated - twin_radiated) <= 1e-4 * twin_radiated
    assert rel_l1(g[3], twin.state()[3]) <= 1e-5
    assert np.array_equal(g[:, ~slab], st.astype(np.float32)[:, ~slab])


def scalar_newton(c, dt, kvac, t0, a_r, uvac):
    """Newton's method on one cell's backward-Euler balance F(u) = (C/dt)(T(u) - T0) + k_vac (u - u_vac), u = a_r T^4,
    from u(T0): u' = u - F(u) / F'(u), each new T clamped to [T/2, 2T], stopped by the reference's constants (the heat
    moved at most 1e-12 of C T0 and the cell at most 1e-10 of its T, heat.rs:55-60). Returns (steps, T, clamped)."""
    t, clamped = t0, False
    for k in range(1, 129):
        u = a_r * t ** 4
        f = c / dt * (t - t0) + kvac * (u - uvac)
        un = u - f / (c / dt / (4.0 * a_r * t ** 3) + kvac)
        tn = (max(un, uvac) / a_r) ** 0.25
        clamped |= not (0.5 * t <= tn <= 2.0 * t)
        tn = min(max(tn, 0.5 * t), 2.0 * t)
        d = abs(tn - t)
        t_last, t = t, tn
        if c * d <= 1e-12 * c * t0 and d <= 1e-10 * t_last:
            return k, t, clamped
    return None, t, clamped


@pytest.mark.parametrize("dt, clamps", [(1e-3, False), (2.0, True), (1e5, True)])
def test_one_cell_newton_passes_are_newtons_method(dt, clamps):
    """A single gas cell (Sigma 1, T 30, hydrogen) in an 8 x 8 vacuum is a scalar equation, so the heat pass's Newton passes are
    Newton's method on it and its linear solves are exact: the pass count is the number of Newton steps the
    reference's stopping rule takes (the clamp to [T/2, 2T] engaged in the two longer steps), the committed T is the balance's
    root found by bracketing (scipy's brentq), and the radiated heat is dt k_vac (u - u_vac) at that root. k_vac and C
    come from the first pass's system (own = C/dt dT/du + k_vac, rhs = C/dt dT/du u0 + k_vac u_vac)."""
    from scipy.optimize import brentq
    s = rect_world(w=8, h=8, rect=(3, 3, 1, 1))
    kw = s.physics_kwargs()
    op, _ = both(**kw)
    st = s.planes(kw["gamma"]).astype(np.float64)
    a_r, uvac = op.a_r, op.a_r * op.t_floor ** 4
    t0 = temps(st)[3, 3]
    c = st[0, 3, 3] * INV_MU / (kw["gamma"] - 1.0)
    own, _, _, rhs, gas, u0 = oracle.heat_first_system(op, st, dt)
    assert gas.sum() == 1 and gas[3, 3]
    ci = c / dt / (4.0 * a_r * t0 ** 3)
    kvac = own[3, 3] - ci
    assert kvac > 0 and rhs[3, 3] == pytest.approx(ci * u0[3, 3] + kvac * uvac, rel=1e-13)
    steps, t_newton, clamped = scalar_newton(c, dt, kvac, t0, a_r, uvac)
    assert steps is not None and clamped == clamps
    root = brentq(lambda t: c / dt * (t - t0) + kvac * (a_r * t ** 4 - uvac), op.t_floor, t0, xtol=1e-15, rtol=1e-15)
    p, rep, _, _, em = oracle.heat_step(op, st, dt)
    print(f"dt {dt}: {rep.passes} passes, Newton's method {steps} steps; T {temps(p)[3, 3]:.12f}, root {root:.12f}",
          flush=True)
    assert rep.converged and rep.passes == steps
    assert temps(p)[3, 3] == pytest.approx(root, rel=1e-12) and t_newton == pytest.approx(root, rel=1e-12)
    assert rep.radiated == pytest.approx(dt * kvac * (a_r * root ** 4 - uvac), rel=1e-10)
    assert em[3, 3] == rep.radiated and np.array_equal(np.delete(p.reshape(14, -1), 27, axis=1),
                                                       np.delete(st.reshape(14, -1), 27, axis=1))


@pytest.mark.gpu
@pytest.mark.parametrize("dt_cap", [1e-3, 1e-2])
def test_the_gpu_one_cell_reaches_the_root_in_newtons_passes_or_more(gpu_lib, dt_cap):
    """The same one-cell scene stepped by the world (heat only, the twin's Newton constants: 1e-12 of the heat and 1e-10
    of a cell's T) at a capped dt, each pass's solve to 1e-12 of its right-hand side: the cell's T at the balance's root
    (f32), and at least the passes Newton's method takes (it converges quadratically, no iteration does better). The
    world's system matrix is FP32, so its last passes converge linearly at these constants: at dt 1e-3 it takes 22
    passes where Newton's method takes 5, as a scalar model with the diagonal rounded to f32 each pass does too. The
    workloads stop at 1e-3 of a cell's T, far above where the rounding binds. (Both caps sit below P9's dt, ~0.018.)"""
    import sandbox_reactions_amd as sr
    from scipy.optimize import brentq
    s = rect_world(w=8, h=8, rect=(3, 3, 1, 1))
    kw = s.physics_kwargs()
    op, _ = both(**kw)
    st = s.planes(kw["gamma"]).astype(np.float32).astype(np.float64)
    hs = sr.heat_settings(stop=sr.STOP_RELATIVE, solve_tol=1e-12, newton_heat=1e-12, newton_cell=1e-10)
    world = scene_json.world(s, heat_settings=hs)
    world.upload(st.astype(np.float32))
    world.set_dt_cap(dt_cap)
    world.run(1)
    r = world.read_heat()
    dt = r.dt
    a_r, uvac = op.a_r, op.a_r * op.t_floor ** 4
    t0 = temps(st)[3, 3]
    c = st[0, 3, 3] * INV_MU / (kw["gamma"] - 1.0)
    own = oracle.heat_first_system(op, st, dt)[0]
    kvac = own[3, 3] - c / dt / (4.0 * a_r * t0 ** 3)
    steps, _, _ = scalar_newton(c, dt, kvac, t0, a_r, uvac)
    root = brentq(lambda t: c / dt * (t - t0) + kvac * (a_r * t ** 4 - uvac), op.t_floor, t0, xtol=1e-15, rtol=1e-15)
    g = world.readback().astype(np.float64)
    print(f"dt {dt}: GPU {r.passes} passes, Newton's method {steps}; T {temps(g)[3, 3]:.7f}, root {root:.7f}",
          flush=True)
    assert dt == pytest.approx(dt_cap, rel=1e-12)
    assert r.converged and r.passes >= steps
    assert temps(g)[3, 3] == pytest.approx(root, rel=2e-6)
