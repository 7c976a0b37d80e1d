"""Benchmark of the grid-physics path on one MI355X.

The headline workload is the full hydro + gravity + FLD step at 1024^2 (a lognormal density with sigma_ln 1.0, a
lognormal temperature with sigma_ln 0.5, random velocities at rms Mach 0.5, box L = 1, gamma = 5/3, Kramers opacity,
FLD solved by MG-PCG to a relative residual of 1e-8 per Newton pass). `value` is Mcell-steps/s of that step with the
state resident in HBM. Beside it: the stiff FLD solve alone at 2048^2 (time to 1e-8 relative residual, plain MG-PCG),
the dominant kernel's roofline from HIP events over the timed region, and the oracle (the f64 CPU twin restated in C)
timed on a bounded sample of the same workload on one host core.

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
    return ap.parse_args()


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def full_step(sr, scenes, n, steps, warmup):
    import torch
    kw = scenes.config_physics_kwargs(n)
    ph = sr.physics(**kw)
    hs = sr.heat_settings(stop=sr.STOP_RELATIVE, solve_tol=1e-8, newton_heat=1e-6, newton_cell=1e-3)
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
    launches, mean_ms, bytes_per = world.kernel_time("hydro_update")
    heat = world.read_heat()
    log(f"full step: {elapsed / steps * 1e3:.3f} ms/step; hydro_update {mean_ms:.4f} ms x {launches}; "
        f"last heat passes {heat.passes} rounds {heat.rounds}")
    return dict(elapsed=elapsed, cost=cost, kernel=(launches, mean_ms, bytes_per), heat=heat, world=world)


def fld_solve(sr, scenes, n):
    import torch
    log(f"FLD system {n}^2: building")
    c, kx, ky, b, dt = scenes.fld_system(n)
    t = lambda a: torch.tensor(a.astype(np.float32), device="cuda").contiguous()
    f = sr.Fld(n, n)
    f.set_system(t(c), t(kx), t(ky))
    bt = torch.tensor(b, device="cuda")
    times, its, rels = [], [], []
    for rep in range(4):
        x = torch.zeros((n, n), dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        it, rel = f.solve(bt, x, tol=1e-8)
        torch.cuda.synchronize()
        if rep > 0:
            times.append(time.perf_counter() - t0)
            its.append(it)
            rels.append(rel)
    log(f"FLD solve {n}^2: {np.median(times) * 1e3:.2f} ms, {its[0]} iterations, rel {rels[0]:.2e}, "
        f"{f.levels} levels")
    return dict(ms=float(np.median(times) * 1e3), iterations=int(its[0]), rel=float(max(rels)), levels=f.levels)


def copy_bandwidth(mib=512, reps=10):
    """A plain device-to-device copy of a known size: the achievable HBM rate on this box (read + write bytes / s),
    and the calibration kernel for the FETCH_SIZE / WRITE_SIZE counters."""
    import torch
    src = torch.ones(mib * 2**20 // 4, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    dst.copy_(src)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        dst.copy_(src)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    return 2 * src.numel() * 4 / (ms * 1e-3) / 1e9


def measured_traffic(kernel, n):
    """HBM bytes per launch of `kernel` at n^2 from the committed counter summary (rocprofv3 FETCH_SIZE and
    WRITE_SIZE runs, each scaled by what the counter reads for a copy of known size), or None."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", f"traffic_full_step_{n}.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        t = json.load(f)
    k = t.get("kernels", {}).get(kernel)
    return None if k is None else k.get("traffic_bytes")


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
    while steps < 1 or (time.perf_counter() - t0 < budget_s and steps < 50):
        twin.step()
        steps += 1
    el = time.perf_counter() - t0
    return dict(value=n * n * steps / el / 1e6, unit="Mcell-steps/s", cores=1, kind="port",
                sample=f"oracle twin (f64 C, 1 thread), full step with Jacobi-PCG FLD at the reference's 1e-12 "
                       f"tolerances, {n}^2 sample of the same fields, {steps} steps in {el:.1f} s")


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
    launches, mean_ms, bytes_per = fs["kernel"]
    achieved = bytes_per / (mean_ms * 1e-3) / 1e9 if mean_ms > 0 else 0.0
    fld = None if args.no_fld else fld_solve(sr, scenes, args.fld_n)
    copy_gbs = copy_bandwidth()
    log(f"device copy {copy_gbs:.0f} GB/s")
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
                        f"gamma 5/3, Kramers opacity, FLD MG-PCG to 1e-8 per Newton pass",
            "grid": n, "fft_padded": 2 * n,
            "pass_ms": {"P1": round(cost.pass_ms[1], 4), "P2": round(cost.pass_ms[2], 4),
                        "P4": round(cost.pass_ms[4], 4), "P5": round(cost.pass_ms[5], 4),
                        "P8+P9": round(cost.pass_ms[8], 4)},
            "heat_last_step": {"passes": fs["heat"].passes, "rounds": fs["heat"].rounds,
                               "converged": fs["heat"].converged},
            "fld_solve_2048": None if fld is None else {
                "workload": f"stiff FLD solve alone, {args.fld_n}^2, diffusion number 1e3, MG-PCG V(2,2) to 1e-8",
                "plain_mg_pcg_ms": round(fld["ms"], 3), "iterations": fld["iterations"], "final_rel": fld["rel"],
                "levels": fld["levels"], "learned_ms": None},
        },
        "roofline": {"bound": "hbm", "kernel": "hydro_update", "achieved": round(achieved, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": measured_traffic("sr::k_hydro<1>", n), "launches": launches,
                     "mean_ms": round(mean_ms, 5), "algorithmic_bytes_per_launch": bytes_per,
                     "device_copy_gbs": round(copy_gbs, 1)},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
