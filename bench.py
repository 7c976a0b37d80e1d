"""Benchmark of the grid-physics path on one MI355X.

The headline workload is the full hydro + gravity + FLD step at 1024^2 (a lognormal density with sigma_ln 1.0, a
lognormal temperature with sigma_ln 0.5, random velocities at rms Mach 0.5, box L = 1, gamma = 5/3, Kramers opacity,
FLD Newton passes solved by MG-PCG, inexactly before the last one and to a relative residual of 1e-8 on it). `value`
is Mcell-steps/s of that step with the state resident in HBM, timed with the kernel clocks off. Beside it: the same
steps one at a time (median), an instrumented window (per-pass times and per-kernel clocks, from which the dominant
kernel per step is named), the step's algorithmic HBM bytes from the library's work counters (the full step's
roofline), the stiff FLD solve alone at 2048^2 (plain MG-PCG on the tuned and on the stated 7-level shape, and the
learned preconditioner; medians of 5), the 64-scene ensemble, P6 burning alone over a 1024^2 burning field, P3 alone
with 16 sinks on the turbulent 1024^2 field, and the oracle (the f64 CPU twin restated in C) on the same 1024^2
workload on one host core and on all of them.

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
    ap.add_argument("--no-burn", action="store_true", help="skip the P6 burning field and the P3 pass")
    return ap.parse_args()


def log(msg):
    print(msg, file=sys.stderr, flush=True)


# the coarse-grid correction's scale for the multigrid V-cycle (1 is the reference's plain prolongation): 1.2 cuts the
# iterations of these stiff lognormal systems by about a tenth (the f64 model: 119 -> 107 at 1024^2, 152 -> 135 at
# 2048^2)
COARSE_SCALE = 1.2


# the kernels the world clocks when timing is on: the P4 passes and the PCG iteration's kernels
KERNELS = ("hydro_update", "hydro_mid", "mg_pre0", "mg_post0", "mg_coarse", "fld_matvec", "fld_update",
           "fld_update_pre0")


def full_step(sr, scenes, n, steps, warmup):
    """K steps timed with the clocks off (the headline), the same K one at a time (their median), then an instrumented
    window (pass times, kernel clocks)."""
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
    world.read_work(reset=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    world.run(steps)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    work = world.read_work(reset=True)
    per = []
    for _ in range(steps):
        t1 = time.perf_counter()
        world.run(1)
        per.append(time.perf_counter() - t1)
    heat = world.read_heat()
    # the same steps with the solves decided on the device, their iterations replayed as captured graphs
    world.set_solve_mode(sr.SOLVE_GRAPH)
    world.upload(planes)
    world.run(warmup)
    world.read_work(reset=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    world.run(steps)
    torch.cuda.synchronize()
    dev_elapsed = time.perf_counter() - t0
    dev_work = world.read_work(reset=True)
    world.set_solve_mode(sr.SOLVE_HOST)
    # the instrumented window, host-decided (every launch does work, so the clocks' counts are the step's work)
    inst = max(5, min(steps, 10))
    world.set_timing(True)
    world.run(inst)
    cost = world.read_cost()
    kernels = {k: world.kernel_time(k) for k in KERNELS}
    world.set_timing(False)
    log(f"full step: {elapsed / steps * 1e3:.3f} ms/step (median of single steps "
        f"{np.median(per) * 1e3:.3f}; device-decided {dev_elapsed / steps * 1e3:.3f}); " +
        "; ".join(f"{k} {v[1] * 1e3:.2f} us x {v[0] / inst:.1f}/step" for k, v in kernels.items()) +
        f"; last heat passes {heat.passes} rounds {heat.rounds}")
    return dict(elapsed=elapsed, work=work, per=per, cost=cost, kernels=kernels, inst=inst, heat=heat, world=world,
                dev_elapsed=dev_elapsed, dev_work=dev_work)


def fld_solve(sr, scenes, n, reps=5):
    """The stiff FLD system alone: plain MG-PCG on the tuned shape (all levels, coarse scale 1.2) and on the
    workload's stated shape (7 levels, V(2,2), plain prolongation), then the learned preconditioner (the committed
    trained network) with its fallback to plain multigrid, then without fallback (one solve, for the record), and one
    network application timed alone. Each time is the median of `reps` solves after one untimed solve."""
    import torch
    from sandbox_reactions_amd import train
    log(f"FLD system {n}^2: building")
    c, kx, ky, b, dt = scenes.fld_system(n)
    t = lambda a: torch.tensor(a.astype(np.float32), device="cuda").contiguous()
    own, tkx, tky = t(c), t(kx), t(ky)
    bt = torch.tensor(b, device="cuda")

    def timed(f, precond, reps):
        out = []
        for rep in range(reps + 1):
            x = torch.zeros((n, n), dtype=torch.float64, device="cuda")
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            it, rel, li, fb = f.solve_stats(bt, x, tol=1e-8, precond=precond)
            torch.cuda.synchronize()
            out.append((time.perf_counter() - t0, it, rel, li, fb))
        return out[1:]

    def summary(runs):
        return dict(ms=round(float(np.median([r[0] for r in runs]) * 1e3), 3),
                    ms_min=round(float(min(r[0] for r in runs) * 1e3), 3), iterations=int(runs[0][1]),
                    final_rel=float(max(r[2] for r in runs)), reps=len(runs))

    f7 = sr.Fld(n, n, levels=7, coarse_scale=1.0)
    f7.set_system(own, tkx, tky)
    stated = summary(timed(f7, sr.PRECOND_MULTIGRID, reps))
    stated.update(levels=f7.levels)
    log(f"FLD solve {n}^2, 7 levels, scale 1: {stated['ms']:.2f} ms, {stated['iterations']} iterations")
    del f7
    f = sr.Fld(n, n, coarse_scale=COARSE_SCALE)
    f.set_system(own, tkx, tky)
    plain = summary(timed(f, sr.PRECOND_MULTIGRID, reps))
    plain.update(levels=f.levels, tiles=f.tiles())
    log(f"FLD solve {n}^2: {plain['ms']:.2f} ms, {plain['iterations']} iterations, rel {plain['final_rel']:.2e}, "
        f"{f.levels} levels")
    res = dict(plain=plain, stated_shape=stated)
    if not os.path.exists(train.DEFAULT_WEIGHTS):
        log("no trained network: the learned preconditioner is not measured")
        return res
    net = sr.Net(n, n, train.load())
    f.set_network(net, fallback_iterations=plain["iterations"])
    learned = timed(f, sr.PRECOND_LEARNED, reps)
    f.set_network(net, fallback_iterations=0)
    pure = timed(f, sr.PRECOND_LEARNED, 1)[0]
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
    ls = summary(learned)
    ls.update(net_iterations=int(learned[0][3]), fell_back=bool(learned[0][4]))
    res.update(learned=ls, learned_vs_plain=round(plain["ms"] / ls["ms"], 4), net_apply_ms=round(apply_ms, 4),
               learned_without_fallback={"ms": round(pure[0] * 1e3, 3), "iterations": pure[1], "rel": pure[2]})
    log(f"FLD learned {n}^2: {ls['ms']:.2f} ms, {ls['iterations']} iterations ({ls['net_iterations']} with the "
        f"network, fell back {ls['fell_back']}); without fallback {pure[1]} iterations in {pure[0] * 1e3:.1f} ms; "
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


def measured_traffic(kernel, grid=None, n=N_STEP):
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


# P6's algorithmic bytes per cell: Sigma (f64), E and the 10 fractions (f32) in and out, the bookings per group, the
# neutrino loss and S_nu (f64) out
BURN_BYTES_PER_CELL = 8 + 2 * (4 + 40) + 8 * 7 + 8 + 8


def burn(sr, scenes, n=N_STEP, passes=5, cpu_n=128):
    """P6 alone over the burning field (the reference's ref_burn.json made uneven, scenes.burn_scene) at n^2: each pass
    from the same upload, at the field's step dt (5e-3: one sub-step per cell) and at ten times it (sub-cycled, with
    retries); medians of `passes` after one untimed. Beside them the oracle's P6 on one host thread over a cpu_n^2
    sample of the same field."""
    import torch
    import oracle
    from sandbox_reactions_amd import react
    planes, kw = scenes.burn_scene(n)
    params = react.params(overrides={"burn_k": 10.0})
    w = sr.World(n, n, physics=sr.physics(**kw), gravity=False, hydro=False, heat=False)
    w.set_reactions(params)
    out = {"workload": f"P6 burning alone, {n}^2: Sigma 30 x lognormal (0.5), T 100 x lognormal (0.5), fresh (H 0.75, "
                       f"He 0.25) and burnt fractions mixed by a smooth field, the cold pressure on, gamma 2, every "
                       f"shipped record, burn_k 10",
           "algorithmic_bytes_per_cell": BURN_BYTES_PER_CELL}
    sample = np.ascontiguousarray(planes[:, :cpu_n, :cpu_n]).astype(np.float64)
    ph = oracle.reference_physics(**kw)
    for dt in (5e-3, 5e-2):
        times = []
        for k in range(passes + 1):
            w.upload(planes)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            w.react_apply(dt)
            torch.cuda.synchronize()
            if k:
                times.append(time.perf_counter() - t0)
        rep = w.read_react()
        ms = float(np.median(times)) * 1e3
        q = sample.copy()
        t0 = time.perf_counter()
        orep, *_ = oracle.react_step(ph, params, q, dt)
        cpu_s = time.perf_counter() - t0
        gbs = n * n * BURN_BYTES_PER_CELL / (ms * 1e-3) / 1e9
        out[f"dt_{dt:g}"] = {"ms_per_pass": round(ms, 4), "mcells_per_s": round(n * n / ms / 1e3, 1),
                             "substeps_per_cell": round(rep.substeps / (n * n), 3), "retried": rep.retried,
                             "max_substeps": rep.max_substeps, "unfinished": rep.unfinished,
                             "achieved_gbs": round(gbs, 1), "hbm_frac": round(gbs / HBM_PEAK_GBS, 4),
                             "cpu_oracle_mcells_per_s": round(cpu_n * cpu_n / cpu_s / 1e6, 4),
                             "cpu_sample": f"oracle P6 (f64 C) on a {cpu_n}^2 corner of the field, one thread, "
                                           f"{orep.substeps} sub-steps in {cpu_s:.3f} s"}
        log(f"P6 {n}^2 dt {dt:g}: {ms:.3f} ms per pass ({n * n / ms / 1e3:.1f} Mcells/s), "
            f"{rep.substeps / (n * n):.2f} sub-steps per cell; CPU oracle {cpu_n * cpu_n / cpu_s / 1e6:.3f} Mcells/s")
    w.close()
    return out


def sinks_leg(sr, scenes, n=N_STEP, passes=5):
    """P3 alone on the turbulent n^2 field with 16 sinks spread over it (each pass from the same upload: candidates,
    accretion, the gas's pull on every sink twice over all cells, the kicks and merging), median of `passes` after one
    untimed."""
    import torch
    kw = dict(scenes.config_physics_kwargs(n), dx=1.0)
    planes = scenes.turbulent_scene(n).astype(np.float32)
    w = sr.World(n, n, physics=sr.physics(**kw), gravity=False, hydro=False, heat=False)
    w.set_sinks(6000.0)
    g = np.linspace(0.1, 0.9, 4) * n
    sinks = [(float(x) + 0.5, float(y) + 0.5, 0.0, 0.0, 1.0) for y in g for x in g]
    times = []
    for k in range(passes + 1):
        w.upload(planes)
        w.put_sinks(sinks)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        w.sinks_apply(1e-3)
        torch.cuda.synchronize()
        if k:
            times.append(time.perf_counter() - t0)
    ms = float(np.median(times)) * 1e3
    w.close()
    log(f"P3 {n}^2 with 16 sinks: {ms:.3f} ms per pass")
    return {"workload": f"P3 alone on the turbulent {n}^2 field with 16 sinks on a 4 x 4 lattice (dx 1, Sigma_BH "
                        f"6000)", "ms_per_pass": round(ms, 4), "sinks": 16,
            "gas_pull_terms_per_pass": 2 * 16 * n * n}


def cpu_baseline(scenes, n=N_STEP, steps=1, omp_steps=2, ref_budget_s=4.0):
    """The oracle twin (f64 C) on the headline workload at its full size: its heat pass solved by multigrid PCG to
    1e-11 (the device's method; the reference twin's Jacobi rule would take thousands of rounds per pass here), one
    step on one host thread (liboracle.so, the reference twin is single-threaded) and a few on all of the host's
    threads (liboracle_omp.so). Beside them, the reference's own rule (Jacobi PCG at 1e-12) on a 128^2 sample of the
    same fields for a few seconds."""
    import oracle
    kw = scenes.config_physics_kwargs(n)
    ph = oracle.reference_physics(**{k: v for k, v in kw.items() if k != "max_padded"}, max_padded=4096.0)
    planes = scenes.turbulent_scene(n).astype(np.float32).astype(np.float64)
    out = {}
    threads = len(os.sched_getaffinity(0))
    for name, lib, k in (("serial", oracle.lib, steps), ("omp", oracle.omp_lib(), omp_steps)):
        lib.sro_heat_set_mg(1e-11)
        try:
            twin = oracle.Twin(ph, planes, gravity=True, hydro=True, heat=True, library=lib)
            t0 = time.perf_counter()
            twin.step(k)
            el = time.perf_counter() - t0
        finally:
            lib.sro_heat_set_mg(0.0)
        out[name] = dict(value=n * n * k / el / 1e6, steps=k, seconds=round(el, 2))
        log(f"CPU twin {n}^2 ({name}): {k} steps in {el:.2f} s")
    m = 128
    kw = scenes.config_physics_kwargs(m)
    ph = oracle.reference_physics(**{k: v for k, v in kw.items() if k != "max_padded"}, max_padded=4096.0)
    twin = oracle.Twin(ph, scenes.turbulent_scene(m).astype(np.float32).astype(np.float64), gravity=True, hydro=True,
                       heat=True)
    k = 0
    t0 = time.perf_counter()
    while k < 1 or time.perf_counter() - t0 < ref_budget_s:
        twin.step()
        k += 1
    el = time.perf_counter() - t0
    return dict(value=out["serial"]["value"], unit="Mcell-steps/s", cores=1, kind="port",
                sample=f"oracle twin (f64 C), the full step on the same {n}^2 fields with multigrid-PCG heat solves to "
                       f"1e-11: {out['serial']['steps']} step in {out['serial']['seconds']} s on one thread",
                omp=dict(value=out["omp"]["value"], cores=threads, sample=f"liboracle_omp.so, {out['omp']['steps']} "
                         f"steps in {out['omp']['seconds']} s on {threads} threads"),
                reference_rule_128=dict(value=m * m * k / el / 1e6, cores=1,
                                        sample=f"Jacobi PCG at the reference's 1e-12 rule, {m}^2 sample, {k} steps in "
                                               f"{el:.1f} s"),
                cpu=host_cpu(), host_threads=threads)


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
    # the clocked kernels' names and launch grids (threads) in a kernel trace, for the committed traffic summary
    gxm = (n + 255) // 256
    mv_rows = max(1, min(16, (n * gxm + 2047) // 2048))
    tiles32 = ((n + 31) // 32) ** 2 * 256
    trace = {"mg_pre0": ("sr::k_pre<32, 32, 2>", tiles32), "mg_post0": ("sr::k_post<32, 32, 2>", tiles32),
             "fld_matvec": ("sr::k_matvec_p", gxm * 256 * ((n + mv_rows - 1) // mv_rows)),
             "fld_update": ("sr::k_update", min(1024, (n * n + 1023) // 1024) * 256),
             "fld_update_pre0": ("sr::k_pre_upd<32, 32, 2>", tiles32),
             "mg_coarse": ("sr::k_coarse<true>", 1024),
             "hydro_update": ("sr::k_hydro<1, false, false>", ((n + 15) // 16) ** 2 * 320),
             "hydro_mid": ("sr::k_hydro<0, false, false>", ((n + 15) // 16) ** 2 * 320)}
    per_step = {}
    for name, (launches, mean_ms, bytes_per) in fs["kernels"].items():
        if launches == 0:  # e.g. the update on its own, when every update is folded into the next pre-smoother
            continue
        per_step[name] = {"launches_per_step": round(launches / fs["inst"], 2), "mean_us": round(mean_ms * 1e3, 3),
                          "ms_per_step": round(launches * mean_ms / fs["inst"], 4),
                          "algorithmic_bytes_per_launch": bytes_per}

    def roof(name):
        launches, mean_ms, bytes_per = fs["kernels"][name]
        achieved = bytes_per / (mean_ms * 1e-3) / 1e9 if mean_ms > 0 else 0.0
        return {"bound": "hbm", "kernel": name, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": measured_traffic(*trace[name], n=n),
                "launches": launches, "mean_ms": round(mean_ms, 5), "algorithmic_bytes_per_launch": bytes_per,
                "ms_per_step": per_step[name]["ms_per_step"]}
    dominant = max(per_step, key=lambda k: per_step[k]["ms_per_step"])
    fld = None if args.no_fld else fld_solve(sr, scenes, args.fld_n)
    ens = None if args.no_ensemble else ensemble(sr, scenes, args.ensemble_threads)
    burning = None if args.no_burn else burn(sr, scenes)
    sinks = None if args.no_burn else sinks_leg(sr, scenes)
    copy_gbs = copy_bandwidth()
    bracket = event_bracket_ms()
    log(f"device copies (GB/s): {copy_gbs}; timing-event bracket {bracket} ms")
    cpu = None if args.no_cpu else cpu_baseline(scenes)
    cost, work = fs["cost"], fs["work"]
    k = args.steps
    full = {"bound": "hbm", "achieved": round(work.algorithmic_bytes / fs["elapsed"] / 1e9, 1), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(work.algorithmic_bytes / fs["elapsed"] / 1e9 / HBM_PEAK_GBS, 4),
            "bytes_per_step": round(work.algorithmic_bytes / k),
            "bytes_per_step_by_pass": {"P2": round(work.bytes_p2 / k), "P4": round(work.bytes_p4 / k),
                                       "P5": round(work.bytes_p5 / k), "P8+P9": round(work.bytes_p89 / k)},
            "per_step": {"newton_passes": work.newton_passes / k, "pcg_iterations": work.pcg_iterations / k,
                         "precond_applications": work.precond_applications / k,
                         "true_residuals": work.true_residuals / k, "hierarchy_builds": work.hierarchy_builds / k,
                         "convolutions": work.convolutions / k}}
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
            "step_ms_median": round(float(np.median(fs["per"])) * 1e3, 4),
            "device_decided": {"ms_per_step": round(fs["dev_elapsed"] / k * 1e3, 4),
                               "pcg_iterations_per_step": fs["dev_work"].pcg_iterations / k,
                               "newton_passes_per_step": fs["dev_work"].newton_passes / k,
                               "what": "the same steps from the same upload with every PCG iteration's continuation "
                                       "decided on the device and the iterations replayed as captured graphs "
                                       "(sr_set_solve_mode SOLVE_GRAPH); the headline decides on the host"},
            "pcg_iterations_per_step": work.pcg_iterations / k,
            "step_ms_single": [round(t * 1e3, 3) for t in fs["per"]],
            "instrumented_steps": fs["inst"],
            "pass_ms": {"P1": round(cost.pass_ms[1], 4), "P2": round(cost.pass_ms[2], 4),
                        "P4": round(cost.pass_ms[4], 4), "P5": round(cost.pass_ms[5], 4),
                        "P8+P9": round(cost.pass_ms[8], 4)},
            "kernels_per_step": per_step,
            "heat_last_step": {"passes": fs["heat"].passes, "rounds": fs["heat"].rounds,
                               "converged": fs["heat"].converged},
            "fld_solve_2048": None if fld is None else dict(
                workload=f"stiff FLD solve alone, {args.fld_n}^2, diffusion number 1e3, to 1e-8: `plain` on the "
                         f"tuned solver (all levels, V(2,2), coarse scale {COARSE_SCALE}), `stated_shape` on 7 "
                         f"levels, V(2,2), plain prolongation; medians of 5", **fld),
            "ensemble_64": ens,
            "burning_1024": burning,
            "sinks_1024": sinks,
        },
        # the kernel with the most device time per step (from the clocks of the instrumented window), with every other
        # clocked kernel beside it; `full_step` is the whole step's algorithmic bytes over the headline window's time.
        # achieved uses the event-timed means as read; event_bracket_ms is what the same events read around an empty
        # kernel, the part of a mean that a kernel trace does not count
        "roofline": dict(roof(dominant), full_step=full, device_copy_gbs=copy_gbs, event_bracket_ms=bracket,
                         also=[roof(k) for k in sorted(per_step, key=lambda k: -per_step[k]["ms_per_step"])
                               if k != dominant]),
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
