"""A headless run of a scene on this library (the reference's `headless` driver, crates/sr-engine/src/headless.rs,
for the passes this path carries: gravity, hydro, heat).

    python -m sandbox_reactions_amd.headless --scene FILE --steps N [--out DIR] [--timing] [--dump-every K]
                                             [--load-dump FILE] [--gamma G]

stdout: the first line `SR-ADAPTER name=... backend=hip driver=...`; the last, on success, `SR-HEADLESS DONE
steps=<n> sim_time=<t> until=none`; errors start `SR-ERROR`. Exit codes as the reference's: 0 done, 3 no device or a
device error, 4 bad arguments or scene (also a scene this path does not carry), 5 a non-finite value in the state
(the reference's line, naming the step, cell and channel, headless.rs:620-625). The guard scans the state at every
step's P9 here (fused into the floors pass, so it costs nothing), where the reference's scans steps 0, 64, 128, ... and
the state after a run's last step: the step named is the first at whose end the value was in the state, which the
reference names only when that step is one it scans.

With --out, DIR/summary.json holds the reference's run-summary fields this path has -- version, tool, adapter, scene,
world {width, height, edge}, steps, sim_time, wall_s and, with --timing, `timing` {per_pass_ms by the reference's pass
labels, step_ms_mean, source "gpu-timestamp"} (headless.rs:497-531, :1356-1381) -- and DIR/dumps/s_<step>.bin the
SRDUMP01 dumps (every K steps with --dump-every, and the last). The floors and the step's reductions run fused here,
so their time is reported under "P8 floors". --load-dump continues from a dump's cells, step index and sim time, and
--steps N is N more steps (the DONE line counts them, the dumps' indices count from the scene); the next dt restarts
from the cells (P9 on the loaded state), and no F plane is read (this library writes none).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

PASS_LABELS = {1: "P1 dt", 2: "P2 gravity", 4: "P4 hydro", 5: "P5 heat", 8: "P8 floors"}
# the state's channel names, as the guard's line names them (crates/sr-engine/src/state.rs:29-31)
STATE_CHANNELS = ["sigma", "mom_x", "mom_y", "energy", "x_H", "x_He", "x_C", "x_O", "x_Ne", "x_Mg", "x_Si", "x_S",
                  "x_Fe", "x_n"]


def run(argv=None) -> int:
    ap = argparse.ArgumentParser(prog="sandbox_reactions_amd.headless")
    ap.add_argument("--scene", required=True)
    ap.add_argument("--steps", type=int, required=True)
    ap.add_argument("--out")
    ap.add_argument("--timing", action="store_true")
    ap.add_argument("--dump-every", type=int, default=0)
    ap.add_argument("--load-dump")
    ap.add_argument("--gamma", type=float, default=2.0)
    a = ap.parse_args(argv)
    from . import dump, scene_json
    try:
        scene = scene_json.load(a.scene)
        scene.check_path()
        scene.planes(a.gamma)  # every cell laid down before a device is touched (a bad scene is exit 4)
        if a.steps < 0:
            raise scene_json.SceneError("--steps must be >= 0")
    except (OSError, ValueError) as e:
        print(f"SR-ERROR {e}", flush=True)
        return 4
    try:
        import sandbox_reactions_amd as sr
        sr.require_device()
        name, cus, mem = sr.device_info()
    except Exception as e:  # no device or library
        print(f"SR-ERROR {e}", flush=True)
        return 3
    print(f"SR-ADAPTER name={name} backend=hip driver=libsrhip cus={cus}", flush=True)
    try:
        world = scene_json.world(scene, gamma=a.gamma)
        step0, clock0 = 0, 0.0
        if a.load_dump:
            d = dump.read(a.load_dump)
            if (d["width"], d["height"]) != (scene.width, scene.height):
                print(f"SR-ERROR the dump's world {d['width']} x {d['height']} is not the scene's", flush=True)
                return 4
            world.upload(d["planes"])
            step0, clock0 = int(d["step"]), float(d["sim_time"])
        if a.timing:
            world.set_timing(True)
        dumps = os.path.join(a.out, "dumps") if a.out else None
        if dumps:
            os.makedirs(dumps, exist_ok=True)
        t0 = time.perf_counter()
        done = 0
        while done < a.steps:
            k = a.steps - done if a.dump_every <= 0 else min(a.dump_every, a.steps - done)
            world.run(k)
            done += k
            r = world.read_dt()
            if r.non_finite:
                print(f"SR-ERROR non-finite value in the state at step {step0 + r.nf_step}, cell ({r.nf_x}, "
                      f"{r.nf_y}), channel {STATE_CHANNELS[r.nf_channel]}", flush=True)
                return 5
            if dumps and (a.dump_every > 0 or done == a.steps):
                dump.write(os.path.join(dumps, f"s_{step0 + done}.bin"), world.readback(), step=step0 + done,
                           sim_time=clock0 + r.sim_time, dt=r.dt, dt_next=r.dt_next)
        wall = time.perf_counter() - t0
        r = world.read_dt()
    except sr.SrError as e:
        print(f"SR-ERROR {e}", flush=True)
        return 5 if e.code == sr._lib.SR_ERR_NON_FINITE else 3
    if a.out:
        summary = {"version": 1, "tool": "sandbox-reactions-amd", "adapter": {"name": name, "backend": "hip",
                                                                            "driver": "libsrhip"},
                   "scene": a.scene, "world": {"width": scene.width, "height": scene.height,
                                               "edge": "leave" if scene.edge == 1 else "bounce"},
                   "steps": step0 + done, "sim_time": clock0 + r.sim_time, "wall_s": wall}
        if a.timing:
            c = world.read_cost()
            summary["timing"] = {"per_pass_ms": {PASS_LABELS[k]: c.pass_ms[k] for k in sorted(PASS_LABELS)},
                                 "step_ms_mean": c.step_ms, "source": "gpu-timestamp"}
        with open(os.path.join(a.out, "summary.json"), "w") as f:
            f.write(json.dumps(summary, indent=2) + "\n")
    # the steps of this run and the clock to 6 decimals (headless.rs:1396)
    print(f"SR-HEADLESS DONE steps={done} sim_time={clock0 + r.sim_time:.6f} until=none", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(run())
