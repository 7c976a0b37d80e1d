"""The batched ensemble: independent worlds of uneven sizes stepped together (sr_ensemble_run). The scene generator's
draws on the CPU; on the GPU, every world of a batch ends bit for bit where the same world ends stepped alone (the
kernels reduce in fixed orders, so sharing the device with other worlds changes nothing), and the refusals."""
import numpy as np
import pytest

from sandbox_reactions_amd import scenes


def test_ensemble_scenes_are_seeded_and_drawn_from_the_sizes():
    a = scenes.ensemble_scenes(count=12, sides=(16, 32, 64))
    b = scenes.ensemble_scenes(count=12, sides=(16, 32, 64))
    assert [s for s, _ in a] == [s for s, _ in b]
    assert all(np.array_equal(p, q) for (_, p), (_, q) in zip(a, b))
    assert {s for s, _ in a} <= {16, 32, 64} and len({s for s, _ in a}) > 1
    for s, p in a:
        assert p.shape == (14, s, s) and p.dtype == np.float32
        rho = p[0].astype(np.float64)
        assert np.all(rho > 0) and rho.max() >= 0.1 * 0.9  # at least one blob of peak >= 0.1 above the floor
        assert np.all(p[1] == 0) and np.all(p[2] == 0) and np.all(p[4] == 1)
        T = p[3] / rho * (2.0 / 3.0)  # E = rho T / (mu (gamma - 1)), mu 1, gamma 5/3
        assert np.all(np.abs(T - 1.0) <= 0.0101)


def test_the_default_ensemble_is_the_64_scene_workload():
    e = scenes.ensemble_scenes()
    sides = [s for s, _ in e]
    assert len(e) == 64 and set(sides) == set(scenes.ENSEMBLE_SIDES)
    assert 15e6 < sum(s * s for s in sides) < 35e6  # about 64 x 348k cells expected


def make_world(sr, side, planes):
    hs = sr.heat_settings(stop=sr.STOP_RELATIVE, solve_tol=1e-8, forcing=0.1, coarse_scale=1.2)
    w = sr.World(side, side, physics=sr.physics(**scenes.config_physics_kwargs(side)), gravity=True, hydro=True,
                 heat=True, heat_settings=hs)
    w.upload(planes)
    return w


@pytest.mark.gpu
def test_an_ensemble_ends_where_each_world_ends_alone(gpu_lib):
    sr = gpu_lib
    batch = scenes.ensemble_scenes(count=5, sides=(32, 48, 64))
    print("ensemble of sides", [s for s, _ in batch], flush=True)
    together = [make_world(sr, s, p) for s, p in batch]
    alone = [make_world(sr, s, p) for s, p in batch]
    sr.ensemble_run(together, 3, threads=3)
    for w in alone:
        w.run(3)
    for (s, _), a, b in zip(batch, together, alone):
        ra, rb = a.read_dt(), b.read_dt()
        assert ra.steps == rb.steps == 3 and ra.sim_time == rb.sim_time and not ra.non_finite
        assert np.array_equal(a.readback(), b.readback()), s


@pytest.mark.gpu
def test_ensemble_refusals(gpu_lib):
    sr = gpu_lib
    (s, p), = scenes.ensemble_scenes(count=1, sides=(32,))
    w = make_world(sr, s, p)
    with pytest.raises(sr.SrError):
        sr.ensemble_run([w, w], 1)
    with pytest.raises(sr.SrError):
        sr.ensemble_run([], 1)
    with pytest.raises(sr.SrError):
        sr.ensemble_run([w], 1, threads=0)
    sr.ensemble_run([w], 1, threads=4)
    assert w.read_dt().steps == 1
