"""P5 on the GPU against the oracle twin: the heat pass alone (the reference's heat scopes, crates/sr-engine/tests/
gpu/heat.rs: a hot slab beside vacuum, a pulse under constant diffusivity) and the full step with it."""
import numpy as np
import pytest

import oracle
from helpers import both, f32, rel_l1
from sandbox_reactions_amd import scenes

pytestmark = pytest.mark.gpu


def tight(sr, **over):
    d = dict(stop=1, solve_tol=1e-10, newton_heat=1e-9, newton_cell=1e-6)
    d.update(over)
    return sr.heat_settings(**d)


def hot_slab(n=64, sigma=30.0, t=18800.0, he=True):
    """A 32 x 32 slab of hot helium (scenes/heat_hot.json's gas) in vacuum: light leaves through its rim."""
    p = np.zeros((14, n, n))
    p[0] = 1e-9
    p[3] = 1e-9
    p[4] = 1.0
    sl = slice(n // 4, n // 4 + n // 2)
    p[0][sl, sl] = sigma
    if he:
        p[4][sl, sl] = 0.0
        p[5][sl, sl] = 1.0
    inv_mu = 0.75 if he else 2.0
    p[3][sl, sl] = sigma * t * inv_mu  # gamma 2: E = Sigma eps, eps = T / mu
    return p


def heat_pair(p, hs_over=None, **phys):
    """One heat-only step on the GPU at the dt P9 gives the uploaded state, and the oracle's P5 at the same dt."""
    import sandbox_reactions_amd as sr
    op, gp = both(**phys)
    _, h, w = p.shape
    start = f32(p)
    world = sr.World(w, h, physics=gp, gravity=False, hydro=False, heat=True,
                     heat_settings=tight(sr, **(hs_over or {})))
    world.upload(start.astype(np.float32))
    dt = world.read_dt().dt_next
    world.run(1)
    assert world.read_dt().dt == dt
    planes, rep, fxf, fyf, em = oracle.heat_step(op, start, dt)
    return planes, rep, fxf, fyf, world


def test_hot_slab_beside_vacuum(gpu_lib):
    print("hot slab", flush=True)
    p = hot_slab()
    planes, rep, fxf, fyf, world = heat_pair(p)
    assert rep.converged and rep.passes >= 2
    g = world.readback()
    r = world.read_heat()
    assert r.converged and r.passes >= 2
    assert rel_l1(g[3], planes[3]) <= 1e-5, rel_l1(g[3], planes[3])
    assert abs(r.radiated - rep.radiated) <= 1e-4 * rep.radiated
    fx, fy = oracle.flux_cells(fxf, fyf)
    gfx = world.read_field(gpu_lib.FIELD_FLUX_X)
    scale = np.abs(fx).max()
    assert np.abs(gfx - fx).max() <= 1e-4 * scale


def test_constant_diffusivity_pulse(gpu_lib):
    print("pulse", flush=True)
    n = 64
    p = np.zeros((14, n, n))
    p[0] = 1.0
    p[4] = 1.0
    yy, xx = np.mgrid[0:n, 0:n]
    T = 0.05 + 2.0 * np.exp(-((xx - n / 2 + 0.5) ** 2 + (yy - n / 2 + 0.5) ** 2) / (2 * 4.0 ** 2))
    p[3] = T * 2.0  # gamma 2, 1/mu = 2 (hydrogen): E = Sigma T / mu
    planes, rep, _, _, world = heat_pair(p, constant_diffusivity=0.2, edge=0)
    g = world.readback()
    assert rel_l1(g[3], planes[3]) <= 1e-5


@pytest.mark.parametrize("opacity", [0, 1])
def test_full_step_with_heat_matches_the_twin(gpu_lib, opacity):
    import sandbox_reactions_amd as sr
    print(f"full step opacity {opacity}", flush=True)
    n = 64
    p = scenes.turbulent_scene(n)
    kw = dict(scenes.config_physics_kwargs(n), opacity=opacity)
    op, gp = both(**kw)
    start = f32(p)
    twin = oracle.Twin(op, start, gravity=True, hydro=True, heat=True)
    world = sr.World(n, n, physics=gp, gravity=True, hydro=True, heat=True, heat_settings=tight(sr))
    world.upload(start.astype(np.float32))
    for _ in range(3):
        dtg = world.read_dt().dt_next
        twin.cap_dt_next(dtg)
        twin.step()
        world.set_dt_cap(dtg)
        world.run(1)
    g, t = world.readback(), twin.state()
    for c in range(4):
        assert rel_l1(g[c], t[c]) <= 5e-5, (c, rel_l1(g[c], t[c]))
    r = world.read_heat()
    assert r.converged


def test_full_step_with_heat_on_a_ragged_world(gpu_lib):
    """The full step with gravity and heat on a 120 x 72 box cut from the turbulent scene (no side a multiple of any
    tile; the padded FFT 256 x 256): the same bars against the twin as the square case."""
    import sandbox_reactions_amd as sr
    w, h = 120, 72
    print(f"ragged full step {w}x{h}", flush=True)
    p = scenes.turbulent_scene(128)[:, :h, :w].copy()
    kw = dict(scenes.config_physics_kwargs(w))
    op, gp = both(**kw)
    start = f32(p)
    twin = oracle.Twin(op, start, gravity=True, hydro=True, heat=True)
    world = sr.World(w, h, physics=gp, gravity=True, hydro=True, heat=True, heat_settings=tight(sr, coarse_scale=1.2))
    world.upload(start.astype(np.float32))
    for _ in range(3):
        dtg = world.read_dt().dt_next
        twin.cap_dt_next(dtg)
        twin.step()
        world.set_dt_cap(dtg)
        world.run(1)
    g, t = world.readback(), twin.state()
    for c in range(4):
        assert rel_l1(g[c], t[c]) <= 5e-5, (c, rel_l1(g[c], t[c]))
    assert world.read_heat().converged


def test_inexact_newton_agrees_with_exact_newton(gpu_lib):
    """Loose early Newton passes (forcing 0.1) against every pass solved to 1e-8: the same state to the Newton
    tolerance, the last pass still to 1e-8, and fewer solver rounds."""
    import sandbox_reactions_amd as sr
    print("inexact Newton 256", flush=True)
    n = 256
    kw = scenes.config_physics_kwargs(n)
    planes = scenes.turbulent_scene(n).astype(np.float32)
    out, rounds = [], []
    for forcing in (0.0, 0.1):
        hs = sr.heat_settings(stop=sr.STOP_RELATIVE, solve_tol=1e-8, newton_heat=1e-6, newton_cell=1e-3,
                              forcing=forcing)
        world = sr.World(n, n, physics=sr.physics(**kw), gravity=True, hydro=True, heat=True, heat_settings=hs)
        world.upload(planes)
        world.set_dt_cap(world.read_dt().dt_next)
        world.run(1)
        r = world.read_heat()
        assert r.converged and r.final_residual <= 1e-8
        out.append(world.readback())
        rounds.append(r.rounds)
    print(f"rounds: every pass to 1e-8 {rounds[0]}, inexact {rounds[1]}", flush=True)
    assert rel_l1(out[1][3], out[0][3]) <= 1e-5
    assert rel_l1(out[1][0], out[0][0]) <= 1e-7
    assert rounds[1] < rounds[0]


def test_ceiling_commit_moves_no_more_heat_than_the_cells_hold(gpu_lib):
    """The reference's ceiling cut (heat_ceiling.bin) with the round ceiling forced from the host (64 rounds): the
    step commits by limited fluxes with delta 0 and keeps the reference's bounds to f32's 1e-5."""
    import sandbox_reactions_amd as sr
    from test_dump import assert_bounded, ceiling_cut
    print("ceiling cut 64x64", flush=True)
    before, oph, dt = ceiling_cut()
    _, gp = both(a_r=4e-5, c_sb=200.0)
    hs = sr.heat_settings(max_rounds=64)
    world = sr.World(64, 64, physics=gp, gravity=False, hydro=False, heat=True, heat_settings=hs)
    world.upload(before.astype(np.float32))
    world.set_dt_cap(dt)
    world.run(1)
    r = world.read_heat()
    assert not r.converged and r.delta == 0.0
    after = world.readback().astype(np.float64)
    assert_bounded(oph, before, after, r.radiated, 1e-5, 2.0 * float(np.finfo(np.float32).eps))
