"""P4 (and P2 with it) on the GPU against the oracle twin, through the World C ABI.

Bars: density rel-L1 <= 1e-5 per step on smooth scenes (the FP32 hydro against the FP64 twin); the reference's G-REF
bands (Sigma, E rel-L1 <= 1e-3, crates/sr-engine/tests/gpu/cpu_gpu.rs:30-36) after many steps; total mass conserved
to FP64 round-off; bit-for-bit reruns (the reference's determinism rule, docs/agent/architecture.md:103-108).
"""
import numpy as np
import pytest

import oracle
from helpers import both, f32, rel_l1
from sandbox_reactions_amd import scenes

pytestmark = pytest.mark.gpu


def run_pair(planes, steps, gravity, hydro=True, heat=False, hs=None, **phys):
    import sandbox_reactions_amd as sr
    op, gp = both(**phys)
    _, h, w = planes.shape
    start = f32(planes)
    twin = oracle.Twin(op, start, gravity=gravity, hydro=hydro, heat=heat)
    world = sr.World(w, h, physics=gp, gravity=gravity, hydro=hydro, heat=heat, heat_settings=hs)
    world.upload(start.astype(np.float32))
    d0 = world.read_dt()
    _, tdn, _, _ = twin.times()
    assert abs(d0.dt_next - tdn) <= 1e-5 * tdn, (d0.dt_next, tdn)
    for _ in range(steps):
        dt_gpu = world.read_dt().dt_next
        twin.cap_dt_next(dt_gpu)   # the twin steps on the GPU's dt (G-REF, reference/mod.rs:225-228)
        twin.step()
        world.set_dt_cap(dt_gpu)
        world.run(1)
    return twin, world


@pytest.mark.parametrize("edge", [0, 1])
def test_sod_one_step_matches_the_twin(gpu_lib, edge):
    print(f"sod edge {edge}", flush=True)
    p = scenes.sod_scene(400, 8)
    twin, world = run_pair(p, 1, gravity=False, edge=edge)
    g, t = world.readback(), twin.state()
    for c in (0, 3):
        assert rel_l1(g[c], t[c]) <= 2e-6, (c, rel_l1(g[c], t[c]))
    # momenta per cell against the step's pressure impulse dt/dx max(Pi) (the y momentum is 0 in exact arithmetic:
    # the f32 rounding of the wall's pressure against the interior flux leaves ~1 ulp of that impulse in the wall rows)
    impulse = world.read_dt().dt * 1.0
    for c in (1, 2):
        assert np.max(np.abs(g[c] - t[c])) <= 1e-6 * impulse, (c, np.max(np.abs(g[c] - t[c])) / impulse)


def test_sod_many_steps_stays_on_the_twin(gpu_lib):
    print("sod 60 steps", flush=True)
    p = scenes.sod_scene(400, 8)
    twin, world = run_pair(p, 60, gravity=False, edge=0)
    g, t = world.readback(), twin.state()
    assert rel_l1(g[0], t[0]) <= 1e-3 and rel_l1(g[3], t[3]) <= 1e-3
    dt, _, st, steps = twin.times()
    r = world.read_dt()
    # the twin steps on min(its dt, the GPU's): the clocks agree to the f32 wave speeds (G-REF allows 5e-4)
    assert steps == r.steps == 60 and abs(r.sim_time - st) <= 1e-6 * st


def test_blob_with_gravity_one_step_density_within_1e5(gpu_lib):
    print("blob 128 gravity", flush=True)
    n = 128
    p = scenes.blob_scene(n)
    twin, world = run_pair(p, 1, gravity=True, **scenes.config_physics_kwargs(n))
    g, t = world.readback(), twin.state()
    assert rel_l1(g[0], t[0]) <= 1e-5
    assert rel_l1(g[3], t[3]) <= 1e-5
    mom = np.hypot(t[1], t[2]).sum()
    assert np.abs(g[1] - t[1]).sum() / mom <= 1e-4 and np.abs(g[2] - t[2]).sum() / mom <= 1e-4


def test_turbulent_scene_with_gravity_five_steps(gpu_lib):
    print("turbulent 128 gravity 5 steps", flush=True)
    n = 128
    p = scenes.turbulent_scene(n)
    twin, world = run_pair(p, 5, gravity=True, **scenes.config_physics_kwargs(n))
    g, t = world.readback(), twin.state()
    for c in range(4):
        assert rel_l1(g[c], t[c]) <= 2e-5, (c, rel_l1(g[c], t[c]))
    for s in range(10):
        assert np.max(np.abs(g[4 + s] - t[4 + s])) <= 1e-5


def test_mixed_species_and_vacuum_match_the_twin(gpu_lib):
    """A composition-weighted mu, vacuum cells and both edge modes (the floor state, empty cells, species slopes)."""
    print("species + vacuum", flush=True)
    n = 64
    rng = np.random.default_rng(4)
    p = scenes.blob_scene(n)
    x = rng.uniform(0.0, 1.0, (10, n, n))
    x[9] = 0.0
    x /= x.sum(0)
    p[4:] = x
    p[0][:, :10] = 1e-9            # a vacuum band
    p[1][:, :10] = p[2][:, :10] = 0.0
    p[3][:, :10] = 1e-9
    twin, world = run_pair(p, 3, gravity=True, gamma=2.0, mu=0.0, dx=1.0 / n, edge=1)
    g, t = world.readback(), twin.state()
    for c in range(4):
        assert rel_l1(g[c], t[c]) <= 2e-5, (c, rel_l1(g[c], t[c]))
    assert np.max(np.abs(g[4:] - t[4:])) <= 2e-5


def test_mass_is_conserved_to_fp64_roundoff_and_reruns_are_bitwise(gpu_lib):
    import sandbox_reactions_amd as sr
    print("conservation", flush=True)
    n = 128
    p = scenes.turbulent_scene(n)
    ph = sr.physics(**dict(scenes.config_physics_kwargs(n), edge=0))
    outs = []
    for _ in range(2):
        world = sr.World(n, n, physics=ph, gravity=True)
        world.upload(p.astype(np.float32))
        t0, _ = world.totals()
        world.run(20)
        t1, esc = world.totals()
        outs.append(world.readback())
        # bounce edges: nothing leaves, so the FP64 mass is constant to round-off
        assert abs(t1[0] - t0[0]) <= 1e-13 * t0[0], (t1[0] - t0[0]) / t0[0]
        assert esc[0] == 0.0
    assert np.array_equal(outs[0], outs[1])


def test_leave_edges_book_what_escapes(gpu_lib):
    import sandbox_reactions_amd as sr
    print("escaped booking", flush=True)
    n = 64
    p = scenes.turbulent_scene(n)
    ph = sr.physics(**dict(scenes.config_physics_kwargs(n), edge=1))
    world = sr.World(n, n, physics=ph, gravity=False)
    world.upload(p.astype(np.float32))
    t0, _ = world.totals()
    world.run(10)
    t1, esc = world.totals()
    assert esc[0] > 0.0
    assert abs((t1[0] + esc[0]) - t0[0]) <= 1e-13 * t0[0]


def test_mirrored_sod_is_the_mirror_image(gpu_lib):
    import sandbox_reactions_amd as sr
    print("mirror", flush=True)
    p = scenes.sod_scene(400, 8)
    q = p[:, :, ::-1].copy()
    q[1] = -q[1]
    ph = sr.physics(edge=0)
    a = sr.World(400, 8, physics=ph, gravity=False)
    b = sr.World(400, 8, physics=ph, gravity=False)
    a.upload(p.astype(np.float32))
    b.upload(q.astype(np.float32))
    a.run(30)
    b.run(30)
    ra, rb = a.readback(), b.readback()
    # the same arithmetic in mirrored order: equal up to f32 rounding of reordered sums
    assert np.allclose(ra[0], rb[0][:, ::-1], rtol=1e-5, atol=1e-7)
    assert np.allclose(ra[1], -rb[1][:, ::-1], rtol=1e-5, atol=1e-6)


def test_world_refusals(gpu_lib):
    import sandbox_reactions_amd as sr
    with pytest.raises(sr.SrError) as e:
        sr.World(100, 64)
    assert e.value.kind == "WorldSize"
    w = sr.World(64, 64, gravity=False)
    with pytest.raises(sr.SrError) as e:
        w.upload(np.zeros((14, 32, 64), np.float32))
    assert e.value.kind == "LengthMismatch"
