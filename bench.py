"""Benchmark of the grid-physics path on one MI355X.

The headline workload is the full hydro + gravity + FLD step at 1024^2 (a lognormal density with sigma_ln 1.0, a
lognormal temperature with sigma_ln 0.5, random velocities at rms Mach 0.5, box L = 1, gamma = 5/3, Kramers opacity,
FLD Newton passes solved by MG-PCG, inexactly before the last one and to a relative residual of 1e-8 on it). `value`
is Mcell-steps/s of that step with the state resident in HBM. Beside it: the stiff FLD solve alone at 2048^2 (time to
1e-8 relative residual by plain MG-PCG and by the learned preconditioner), the roofline of the step's dominant
HBM-bound kernel (the FLD's level-0 pre-smoother; the hydro update beside it) from HIP events over the timed region,
and the oracle (the f64 CPU twin restated in C) timed on a bounded sample of the same workload on one host core.

    python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

HBM_PEAK_GBS = 8000.0
N_STEP = 1024
N_FLD = 2048


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--fld-n", type=int, default=N_FLD)
    ap.add_argument("--step-n", type=int, default=N_STEP)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fld", action="store_true", help="skip the standalone FLD solve (counter runs)")
    ap.add_argument("--no-ensemble", action="store_true", help="skip the 64-scene ensemble")
    ap.add_argument("--ensemble-threads", type=int, default=16)
    return ap.parse_args()


def log(msg):
    print(msg, file=sys.stderr, flush=True)


# the coarse-grid correction's scale for the multigrid V-cycle (1 is the reference's plain prolongation): 1.2 cuts the
# iterations of these stiff lognormal systems by about a tenth (the f64 model: 119 -> 107 at 1024^2, 152 -> 135 at
# 2048^2)
COARSE_SCALE = 1.2


def full_step(sr, scenes, n, steps, warmup):
    import torch
    kw = scenes.config_physics_kwargs(n)
    ph = sr.physics(**kw)
    hs = sr.heat_settings(stop=sr.STOP_RELATIVE, solve_tol=1e-8, newton_heat=1e-6, newton_cell=1e-3, forcing=0.1,
                          coarse_scale=COARSE_SCALE)
    planes = scenes.turbulent_scene(n).astype(np.float32)
    world = sr.World(n, n, physics=ph, gravity=True, hydro=True, heat=True, heat_settings=hs)
    world.upload(planes)
    log(f"full step {n}^2: warmup {warmup}")
    world.run(warmup)
    world.set_timing(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    world.run(steps)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    cost = world.read_cost()
    kernels = {k: world.kernel_time(k) for k in ("mg_pre0", "hydro_update")}
    heat = world.read_heat()
    log(f"full step: {elapsed / steps * 1e3:.3f} ms/step; " +
        "; ".join(f"{k} {v[1]:.4f} ms x {v[0]}" for k, v in kernels.items()) +
        f"; last heat passes {heat.passes} rounds {heat.rounds}")
    return dict(elapsed=elapsed, cost=cost, kernels=kernels, heat=heat, world=world)


def fld_solve(sr, scenes, n):
    """The stiff FLD system alone: plain MG-PCG, then the learned preconditioner (the committed trained network)
    with its fallback to plain multigrid, then the learned preconditioner without fallback (one solve, for the
    record), and one network application timed alone."""
    import torch
    from sandbox_reactions_amd import train
    log(f"FLD system {n}^2: building")
    c, kx, ky, b, dt = scenes.fld_system(n)
    t = lambda a: torch.tensor(a.astype(np.float32), device="cuda").contiguous()
    own, tkx, tky = t(c), t(kx), t(ky)
    f = sr.Fld(n, n, coarse_scale=COARSE_SCALE)
    f.set_system(own, tkx, tky)
    bt = torch.tensor(b, device="cuda")

    def timed(precond, reps):
        out = []
        for rep in range(reps):
            x = torch.zeros((n, n), dtype=torch.float64, device="cuda")
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            it, rel, li, fb = f.solve_stats(bt, x, tol=1e-8, precond=precond)
            torch.cuda.synchronize()
            out.append((time.perf_counter() - t0, it, rel, li, fb))
        return out

    plain = timed(sr.PRECOND_MULTIGRID, 4)[1:]
    ms = float(np.median([p[0] for p in plain]) * 1e3)
    res = dict(ms=ms, iterations=int(plain[0][1]), rel=float(max(p[2] for p in plain)), levels=f.levels,
               tiles=f.tiles())
    log(f"FLD solve {n}^2: {ms:.2f} ms, {plain[0][1]} iterations, rel {res['rel']:.2e}, {f.levels} levels")
    if not os.path.exists(train.DEFAULT_WEIGHTS):
        log("no trained network: the learned preconditioner is not measured")
        return res
    net = sr.Net(n, n, train.load())
    f.set_network(net, fallback_iterations=plain[0][1])
    learned = timed(sr.PRECOND_LEARNED, 3)[1:]
    f.set_network(net, fallback_iterations=0)
    pure = timed(sr.PRECOND_LEARNED, 1)[0] if plain[0][1] > 0 else None
    r = torch.randn((n, n), device="cuda")
    z = torch.empty_like(r)
    net.apply(own, tkx, tky, r, z)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(5):
        net.apply(own, tkx, tky, r, z)
    e1.record()
    torch.cuda.synchronize()
    apply_ms = e0.elapsed_time(e1) / 5
    lms = float(np.median([p[0] for p in learned]) * 1e3)
    res.update(learned_ms=lms, learned_iterations=int(learned[0][1]), learned_net_iterations=int(learned[0][3]),
               learned_fell_back=bool(learned[0][4]), learned_rel=float(max(p[2] for p in learned)),
               net_apply_ms=apply_ms,
               learned_without_fallback={"ms": pure[0] * 1e3, "iterations": pure[1], "rel": pure[2]}
               if pure else None)
    log(f"FLD learned {n}^2: {lms:.2f} ms, {learned[0][1]} iterations ({learned[0][3]} with the network, "
        f"fell back {learned[0][4]}); without fallback {pure[1]} iterations in {pure[0] * 1e3:.1f} ms; "
        f"network apply {apply_ms:.3f} ms")
    return res


def copy_bandwidth(mib=512, reps=10):
    """Device copies of a known size (libsrcalib.so: one float per lane, one float4 per lane) and the runtime's
    own copy: the achievable HBM rates on this box (read + write bytes / s), and the calibration kernels of the
    FETCH_SIZE / WRITE_SIZE counters."""
    import ctypes
    import torch
    from sandbox_reactions_amd import build
    cal = ctypes.CDLL(build.CALIB_LIB)
    cal.sr_calib_copy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int, ctypes.c_void_p]
    src = torch.ones(mib * 2**20 // 4, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    out = {}
    for name, fn in [("copy_float", lambda: cal.sr_calib_copy(src.data_ptr(), dst.data_ptr(), src.numel(), 1, stream)),
                     ("copy_float4", lambda: cal.sr_calib_copy(src.data_ptr(), dst.data_ptr(), src.numel(), 4, stream)),
                     ("copy_runtime", lambda: dst.copy_(src))]:
        fn()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        out[name] = round(2 * src.numel() * 4 / (e0.elapsed_time(e1) / reps * 1e-3) / 1e9, 1)
    return out


def event_bracket_ms(reps=512):
    """The mean time a pair of timing events reads around an empty kernel that follows another (libsrcalib.so): the
    part of an event-timed kernel mean that a profiler's kernel trace does not count."""
    import ctypes
    import torch
    from sandbox_reactions_amd import build
    cal = ctypes.CDLL(build.CALIB_LIB)
    cal.sr_calib_event_bracket.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.POINTER(ctypes.c_double)]
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    ms = ctypes.c_double()
    cal.sr_calib_event_bracket(16, stream, ctypes.byref(ms))
    if cal.sr_calib_event_bracket(reps, stream, ctypes.byref(ms)) != 0:
        return None
    return round(ms.value, 5)


def measured_traffic(kernel, n, grid=None):
    """HBM bytes per launch of `kernel` (launched with `grid` threads, when given) at n^2 from the committed counter
    summary (rocprofv3 FETCH_SIZE and WRITE_SIZE runs, each scaled by what the counter reads for a copy of known
    size), or None."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", f"traffic_full_step_{n}.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        t = json.load(f)
    k = t.get("by_grid", {}).get(f"{kernel}|{grid}") if grid is not None else t.get("kernels", {}).get(kernel)
    return None if k is None else k.get("traffic_bytes")


def ensemble(sr, scenes, threads, steps=2):
    """The batched ensemble: 64 independent scenes of sides 128..1024 (log-uniform), blobs on a lognormal
    background, the full step with the same heat settings as the single world, all worlds stepped together by
    sr_ensemble_run over `threads` host threads; one more step with one thread for comparison."""
    import torch
    batch = scenes.ensemble_scenes()
    cells = sum(s * s for s, _ in batch)
    log(f"ensemble: {len(batch)} scenes, {cells / 1e6:.1f} Mcells; creating")
    hs = sr.heat_settings(stop=sr.STOP_RELATIVE, solve_tol=1e-8, newton_heat=1e-6, newton_cell=1e-3, forcing=0.1,
                          coarse_scale=COARSE_SCALE)
    worlds = []
    for side, planes in batch:
        w = sr.World(side, side, physics=sr.physics(**scenes.config_physics_kwargs(side)), gravity=True, hydro=True,
                      heat=True, heat_settings=hs)
        w.upload(planes)
        worlds.append(w)
    del batch
    sr.ensemble_run(worlds, 1, threads)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    sr.ensemble_run(worlds, steps, threads)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    t1 = time.perf_counter()
    sr.ensemble_run(worlds, 1, 1)
    torch.cuda.synchronize()
    serial = time.perf_counter() - t1
    bad = sum(1 for w in worlds if w.read_dt().non_finite)
    res = {"workload": "64 scenes, sides 128..1024 log-uniform, 1-3 Gaussian blobs (peak rho 0.1-1, sigma "
                       "0.03-0.15 L) on a lognormal background (sigma_ln 1.0), full hydro+gravity+FLD step",
           "scenes": len(worlds), "cells": cells, "threads": threads, "steps": steps,
           "ms_per_step": round(elapsed / steps * 1e3, 3),
           "mcell_steps_per_s": round(cells * steps / elapsed / 1e6, 3),
           "one_thread_mcell_steps_per_s": round(cells / serial / 1e6, 3), "worlds_non_finite": bad}
    log(f"ensemble: {res['ms_per_step']} ms per step of all {len(worlds)}, {res['mcell_steps_per_s']} Mcell-steps/s "
        f"({threads} threads; one thread {res['one_thread_mcell_steps_per_s']})")
    del worlds
    return res


def cpu_baseline(scenes, budget_s=12.0):
    """The oracle twin (f64 C, one thread) on the same workload's statistics at 128^2 (the bounded sample): steps
    until the budget is spent."""
    import oracle
    n = 128
    kw = scenes.config_physics_kwargs(n)
    ph = oracle.reference_physics(**{k: v for k, v in kw.items() if k != "max_padded"}, max_padded=4096.0)
    planes = scenes.turbulent_scene(n).astype(np.float32).astype(np.float64)
    twin = oracle.Twin(ph, planes, gravity=True, hydro=True, heat=True)
    steps = 0
    t0 = time.perf_counter()
    while steps < 1 or time.perf_counter() - t0 < budget_s:
        twin.step()
        steps += 1
    el = time.perf_counter() - t0
    return dict(value=n * n * steps / el / 1e6, unit="Mcell-steps/s", cores=1, kind="port",
                sample=f"oracle twin (f64 C, 1 thread), full step with Jacobi-PCG FLD at the reference's 1e-12 "
                       f"tolerances, {n}^2 sample of the same fields, {steps} steps in {el:.1f} s",
                cpu=host_cpu(), host_threads=len(os.sched_getaffinity(0)))


def host_cpu():
    """The host CPU's model name, as /proc/cpuinfo gives it."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU; --gpus {args.gpus} is not supported (the path shards by design only)",
              file=sys.stderr)
        sys.exit(2)
    import torch
    import sandbox_reactions_amd as sr
    from sandbox_reactions_amd import scenes
    sr.require_device()
    name, cus, mem = sr.device_info()
    log(f"device {name}, {cus} CUs, {mem / 2**30:.0f} GiB")

    fs = full_step(sr, scenes, args.step_n, args.steps, args.warmup)
    n = args.step_n
    value = n * n * args.steps / fs["elapsed"] / 1e6
    def roof(name, counter_name, grid=None):
        launches, mean_ms, bytes_per = fs["kernels"][name]
        achieved = bytes_per / (mean_ms * 1e-3) / 1e9 if mean_ms > 0 else 0.0
        return {"bound": "hbm", "kernel": name, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": measured_traffic(counter_name, n, grid),
                "launches": launches, "mean_ms": round(mean_ms, 5), "algorithmic_bytes_per_launch": bytes_per}
    fld = None if args.no_fld else fld_solve(sr, scenes, args.fld_n)
    ens = None if args.no_ensemble else ensemble(sr, scenes, args.ensemble_threads)
    copy_gbs = copy_bandwidth()
    bracket = event_bracket_ms()
    log(f"device copies (GB/s): {copy_gbs}; timing-event bracket {bracket} ms")
    cpu = None if args.no_cpu else cpu_baseline(scenes)
    cost = fs["cost"]
    out = {
        "metric": "Mcell-steps/s full step + FLD solve ms to 1e-8: learned vs plain MG-PCG",
        "value": round(value, 3),
        "unit": "Mcell-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(fs["elapsed"] / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 state and fluxes, f64 mass / residuals / dots",
        "data": "synthetic (seeded lognormal fields, the reference's LCG)",
        "config": {
            "workload": f"full hydro+gravity+FLD step, {n}^2, lognormal rho (sigma_ln 1.0) and T (0.5), Mach 0.5, "
                        f"gamma 5/3, Kramers opacity, FLD by MG-PCG (coarse scale {COARSE_SCALE}) in inexact Newton "
                        f"(forcing 0.1), last pass to 1e-8",
            "grid": n, "fft_padded": 2 * n,
            "pass_ms": {"P1": round(cost.pass_ms[1], 4), "P2": round(cost.pass_ms[2], 4),
                        "P4": round(cost.pass_ms[4], 4), "P5": round(cost.pass_ms[5], 4),
                        "P8+P9": round(cost.pass_ms[8], 4)},
            "heat_last_step": {"passes": fs["heat"].passes, "rounds": fs["heat"].rounds,
                               "converged": fs["heat"].converged},
            "fld_solve_2048": None if fld is None else {
                "workload": f"stiff FLD solve alone, {args.fld_n}^2, diffusion number 1e3, MG-PCG V(2,2), "
                            f"coarse scale {COARSE_SCALE}, to 1e-8",
                "plain_mg_pcg_ms": round(fld["ms"], 3), "iterations": fld["iterations"], "final_rel": fld["rel"],
                "levels": fld["levels"], "tiles_pre_post": fld["tiles"], "learned_ms": fld.get("learned_ms"),
                "learned_vs_plain": round(fld["ms"] / fld["learned_ms"], 4) if fld.get("learned_ms") else None,
                "learned_iterations": fld.get("learned_iterations"),
                "learned_net_iterations": fld.get("learned_net_iterations"),
                "learned_fell_back": fld.get("learned_fell_back"),
                "learned_without_fallback": fld.get("learned_without_fallback"),
                "net_apply_ms": fld.get("net_apply_ms")},
            "ensemble_64": ens,
        },
        # the step's dominant HBM-bound kernel: the FLD's level-0 pre-smoother (P5 is most of the step); the hydro
        # update beside it. achieved uses the event-timed mean as read; event_bracket_ms is what the same events
        # read around an empty kernel, the part of that mean a kernel trace does not count
        "roofline": dict(roof("mg_pre0", "sr::k_pre<32, 32, 2>", ((n + 31) // 32) ** 2 * 256),
                         device_copy_gbs=copy_gbs, event_bracket_ms=bracket,
                         also=[roof("hydro_update", "sr::k_hydro<1, false>", ((n + 15) // 16) ** 2 * 320)]),
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
