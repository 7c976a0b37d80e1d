"""The benchmark's own workload at its full size on the GPU, graded by properties that do not depend on size: the
turbulent 1024^2 world (lognormal rho sigma_ln 1.0 and T 0.5, Mach 0.5, gamma 5/3, Kramers, leave edges) stepped with
gravity, hydro and the FLD heat pass under the benchmark's settings (relative stop 1e-8, inexact Newton with forcing
0.1, coarse scale 1.2) for 40 steps:

- every heat pass converged, its last solve to the relative residual 1e-8;
- the guard never trips and every value stays finite;
- mass, grid + escaped, constant to FP64 round-off (1e-12: no cell of this world goes below Sigma_vac, so the floors
  never reset one), and every species' mass to 1e-9 of it (the fractions are f32 planes, renormalised each step, so
  a species' total carries their rounding: 7e-11 measured);
- the run is deterministic: the same start stepped twice for 10 steps gives the same state to the last bit.

The oracle cannot run this size in seconds; the same step is held to the twin at 128^2 and 256^2 by the other GPU
tests."""
import numpy as np
import pytest

from sandbox_reactions_amd import scenes

pytestmark = pytest.mark.gpu
N = 1024


def world_of(sr, planes):
    kw = scenes.config_physics_kwargs(N)
    hs = sr.heat_settings(stop=sr.STOP_RELATIVE, solve_tol=1e-8, newton_heat=1e-6, newton_cell=1e-3, forcing=0.1,
                          coarse_scale=1.2)
    w = sr.World(N, N, physics=sr.physics(**kw), gravity=True, hydro=True, heat=True, heat_settings=hs)
    w.upload(planes)
    return w


def test_the_benchmark_world_keeps_its_invariants_over_40_steps(gpu_lib):
    import sandbox_reactions_amd as sr
    print("turbulent 1024^2, 40 steps", flush=True)
    planes = scenes.turbulent_scene(N).astype(np.float32)
    assert planes[0].min() >= 1e-3  # far above Sigma_vac: the floors reset nothing
    world = world_of(sr, planes)
    t0, e0 = world.totals()
    worst_res = 0.0
    for k in range(40):
        world.run(1)
        h = world.read_heat()
        assert h.converged, (k, h.passes, h.rounds)
        worst_res = max(worst_res, h.final_residual)
        assert not world.read_dt().non_finite, k
    t1, e1 = world.totals()
    g = world.readback()
    print(f"worst final residual {worst_res:.2e}; mass drift {((t1[0] + e1[0]) / t0[0] - 1):.2e}; escaped "
          f"{e1[0] / t0[0]:.2e} of the mass", flush=True)
    assert worst_res <= 1e-8
    assert np.all(np.isfinite(g))
    assert abs((t1[0] + e1[0]) - (t0[0] + e0[0])) <= 1e-12 * t0[0]
    for s in range(4, 14):
        assert abs((t1[s] + e1[s]) - (t0[s] + e0[s])) <= 1e-9 * t0[0], s


def test_the_benchmark_world_reruns_bit_for_bit(gpu_lib):
    import sandbox_reactions_amd as sr
    print("turbulent 1024^2, 10 steps twice", flush=True)
    planes = scenes.turbulent_scene(N).astype(np.float32)
    outs = []
    for _ in range(2):
        world = world_of(sr, planes)
        world.run(10)
        outs.append((world.readback(), world.read_dt().sim_time, world.read_heat().rounds))
        world.close()
    assert np.array_equal(outs[0][0].view(np.uint32), outs[1][0].view(np.uint32))
    assert outs[0][1] == outs[1][1] and outs[0][2] == outs[1][2]
