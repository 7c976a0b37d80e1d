"""The world's work counters (sr_read_work) and kernel clocks (sr_kernel_time) against what the steps did: counts
equal to the heat reports' passes and rounds, the clocks' launch counts equal to the counters, the algorithmic bytes
of the passes that have closed forms per cell, and the reset."""
import numpy as np
import pytest

from sandbox_reactions_amd import scenes

pytestmark = pytest.mark.gpu


def make(sr, n):
    kw = scenes.config_physics_kwargs(n)
    hs = sr.heat_settings(stop=sr.STOP_RELATIVE, solve_tol=1e-8, forcing=0.1, coarse_scale=1.2)
    w = sr.World(n, n, physics=sr.physics(**kw), gravity=True, hydro=True, heat=True, heat_settings=hs)
    w.upload(scenes.turbulent_scene(n).astype(np.float32))
    return w


def test_work_counters_follow_the_steps(gpu_lib):
    sr = gpu_lib
    n = 128
    print(f"work counters {n}^2", flush=True)
    w = make(sr, n)
    reports = []
    for _ in range(2):
        w.run(1)
        reports.append(w.read_heat())
    k = w.read_work()
    assert k.steps == 2 and k.hydro_steps == 2 and k.convolutions == 4 and k.heat_steps == 2
    assert k.newton_passes == sum(r.passes for r in reports)
    assert k.pcg_iterations == sum(r.rounds for r in reports)
    # one application opens each inner iteration (an inner loop ends on an update, before the next application); a
    # true residual per refinement and one to end each solve
    assert k.precond_applications == k.pcg_iterations and k.true_residuals >= k.newton_passes
    assert k.hierarchy_builds == k.newton_passes
    cells = n * n
    # P4: the mass-flux pass 72 B and the update 136 B per cell; from the second step on the update also reads the
    # last heat pass's flux and |push| (12 B); P8/P9 128 B per cell
    assert k.bytes_p4 == cells * (2 * (72 + 136) + 12)
    assert k.bytes_p89 == cells * 128 * 2
    assert k.bytes_p2 > 0 and k.bytes_p5 > 0
    assert k.algorithmic_bytes == pytest.approx(k.bytes_p2 + k.bytes_p4 + k.bytes_p5 + k.bytes_p89, rel=1e-15)
    again = w.read_work(reset=True)
    assert again.steps == 2
    z = w.read_work()
    assert z.steps == 0 and z.pcg_iterations == 0 and z.algorithmic_bytes == 0.0


def test_kernel_clocks_count_every_launch(gpu_lib):
    sr = gpu_lib
    n = 256
    print(f"kernel clocks {n}^2", flush=True)
    w = make(sr, n)
    w.run(1)
    w.read_work(reset=True)
    w.set_timing(True)
    w.run(2)
    k = w.read_work()
    clocks = {name: w.kernel_time(name) for name in ("hydro_update", "hydro_mid", "mg_pre0", "mg_post0",
                                                       "mg_coarse", "fld_matvec", "fld_update", "fld_update_pre0")}
    w.set_timing(False)
    for name, (launches, mean_ms, nbytes) in clocks.items():
        if name != "fld_update":  # every update is folded into the next pre-smoother here
            assert launches > 0 and 0.0 < mean_ms < 50.0 and nbytes > 0, name
    assert clocks["hydro_update"][0] == 2 and clocks["hydro_mid"][0] == 2
    # host-decided: every launch is counted work; each update once, on its own or folded into the next pre-smoother
    # (the fold at a solve's last check starts an application that the check ends)
    for name in ("mg_post0", "mg_coarse"):
        assert clocks[name][0] == k.precond_applications, name
    assert clocks["fld_matvec"][0] == clocks["fld_update"][0] + clocks["fld_update_pre0"][0] == k.pcg_iterations
    assert clocks["mg_pre0"][0] + clocks["fld_update_pre0"][0] >= k.precond_applications
    assert clocks["fld_update_pre0"][2] == 49 * n * n
    # device-decided, launched one by one with the clocks on: every slot's launches are clocked, the ones that skip
    # themselves past a solve's end included, and each inner iteration has exactly one application
    w.read_work(reset=True)
    w.set_solve_mode(sr.SOLVE_DEVICE)
    w.set_timing(True)
    w.run(1)
    k = w.read_work()
    clocks = {name: w.kernel_time(name) for name in ("mg_pre0", "mg_post0", "mg_coarse", "fld_matvec", "fld_update")}
    w.set_timing(False)
    assert k.precond_applications == k.pcg_iterations > 0
    for name in clocks:
        assert clocks[name][0] >= k.pcg_iterations, name
    assert len({v[0] for v in clocks.values()}) == 1
    assert clocks["mg_pre0"][2] == 21 * n * n and clocks["fld_update"][2] == 32 * n * n
    with pytest.raises(sr.SrError):
        w.kernel_time("no_such_kernel")


def test_the_fused_smoother_pair_in_the_world(gpu_lib):
    """The heat pass with the last two tile-smoothed levels' smoothers fused (either tiling) or launched apart: the
    same state to the last bit after two full steps, host- and device-decided. 512^2 tile-smooths three levels."""
    sr = gpu_lib
    n = 512
    print(f"fused pair in the world {n}^2", flush=True)
    for mode in (sr.SOLVE_HOST, sr.SOLVE_GRAPH):
        states = []
        for form in ((sr.FUSE_OFF, sr.FUSE_OFF), (sr.FUSE_SMALL, sr.FUSE_SMALL), (sr.FUSE_LARGE, sr.FUSE_LARGE)):
            w = make(sr, n)
            w.set_solve_mode(mode)
            w.set_fusion(*form)
            w.run(2)
            states.append((w.readback(), w.read_dt().sim_time))
            w.close()
        for s, t in states[1:]:
            assert np.array_equal(s.view(np.uint32), states[0][0].view(np.uint32)) and t == states[0][1], mode
    with pytest.raises(sr.SrError):
        make(sr, 16).set_fusion(0, 5)


def test_host_and_device_decided_solves(gpu_lib):
    """The heat pass's solves decided on the host (partials read back, with and without the lookahead), on the device
    (slots launched one by one) and by the captured graph of slots: the two device modes give the same bits after
    three steps, and so do the two host modes; all converge every pass, and the host-decided run (which stops only at
    its checks, so a few iterations later) agrees with them within the solves' tolerance in E."""
    sr = gpu_lib
    n = 128
    print(f"solve modes {n}^2", flush=True)
    out = {}
    for mode in (sr.SOLVE_HOST, sr.SOLVE_HOST_WAIT, sr.SOLVE_DEVICE, sr.SOLVE_GRAPH):
        w = make(sr, n)
        w.set_solve_mode(mode)
        reps = []
        for _ in range(3):
            w.run(1)
            reps.append(w.read_heat())
        assert all(r.converged and r.final_residual <= 1e-8 for r in reps), mode
        out[mode] = (w.readback(), [r.rounds for r in reps], w.read_dt().sim_time)
        w.close()
    d, g, h, hw = out[sr.SOLVE_DEVICE], out[sr.SOLVE_GRAPH], out[sr.SOLVE_HOST], out[sr.SOLVE_HOST_WAIT]
    assert np.array_equal(d[0].view(np.uint32), g[0].view(np.uint32)) and d[1] == g[1] and d[2] == g[2]
    # the host's lookahead queues preconditioner applications that a check may leave unused: the same bits
    assert np.array_equal(h[0].view(np.uint32), hw[0].view(np.uint32)) and h[1] == hw[1] and h[2] == hw[2]
    print(f"rounds per step: host {h[1]}, device {d[1]}", flush=True)
    e_h, e_d = h[0][3].astype(np.float64), d[0][3].astype(np.float64)
    assert np.abs(e_h - e_d).sum() / np.abs(e_h).sum() <= 1e-6
    with pytest.raises(sr.SrError):
        make(sr, 16).set_solve_mode(7)


def test_the_lookahead_changes_no_solve_bit(gpu_lib):
    """The standalone stiff solve at 512^2 with the host's lookahead and waiting at every check: the same iterations
    and the same FP64 answer to the last bit."""
    import torch
    from sandbox_reactions_amd import scenes
    sr = gpu_lib
    n = 512
    print(f"lookahead {n}^2", flush=True)
    c, kx, ky, b, _ = scenes.fld_system(n)
    t = lambda a: torch.tensor(a.astype(np.float32), device="cuda").contiguous()
    bt = torch.tensor(b, device="cuda")
    res = []
    for mode in (sr.SOLVE_HOST, sr.SOLVE_HOST_WAIT):
        f = sr.Fld(n, n, coarse_scale=1.2)
        f.set_solve_mode(mode)
        f.set_system(t(c), t(kx), t(ky))
        x = torch.zeros((n, n), dtype=torch.float64, device="cuda")
        it, rel = f.solve(bt, x, tol=1e-8)
        assert rel <= 1e-8
        res.append((it, x.cpu().numpy()))
    assert res[0][0] == res[1][0] and np.array_equal(res[0][1].view(np.uint64), res[1][1].view(np.uint64))


def test_the_folded_update_changes_no_bit(gpu_lib):
    """The heat pass with each PCG update folded into the next level-0 pre-smoother and with it launched apart: the same
    state to the last bit after two full steps at 512^2, and the standalone stiff solve at 512^2 in the same iterations
    to the same FP64 answer."""
    import torch
    sr = gpu_lib
    n = 512
    print(f"folded update {n}^2", flush=True)
    states = []
    for on in (True, False):
        w = make(sr, n)
        w.set_update_fusion(on)
        w.run(2)
        states.append((w.readback(), w.read_dt().sim_time, w.read_work().pcg_iterations))
        w.close()
    assert np.array_equal(states[0][0].view(np.uint32), states[1][0].view(np.uint32))
    assert states[0][1:] == states[1][1:]
    c, kx, ky, b, _ = scenes.fld_system(n)
    t = lambda a: torch.tensor(a.astype(np.float32), device="cuda").contiguous()
    bt = torch.tensor(b, device="cuda")
    res = []
    for on in (True, False):
        f = sr.Fld(n, n, coarse_scale=1.2)
        f.set_update_fusion(on)
        f.set_system(t(c), t(kx), t(ky))
        x = torch.zeros((n, n), dtype=torch.float64, device="cuda")
        it, rel = f.solve(bt, x, tol=1e-8)
        assert rel <= 1e-8
        res.append((it, x.cpu().numpy()))
    assert res[0][0] == res[1][0] and np.array_equal(res[0][1].view(np.uint64), res[1][1].view(np.uint64))
