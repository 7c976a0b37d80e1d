"""A headless run of a scene on this library (the reference's `headless` driver, crates/sr-engine/src/headless.rs,
for the passes this path carries: gravity, hydro, heat).

    python -m sandbox_reactions_amd.headless --scene FILE (--steps N | --until sim_time:T) [--max-steps M]
                                             [--world WxH] [--edge leave|bounce] [--out DIR] [--timing]
                                             [--dump-every K] [--load-dump FILE] [--gamma G]

stdout: the first line `SR-ADAPTER name=... backend=hip driver=...`; the last, when the run ends without an error,
`SR-HEADLESS DONE steps=<n> sim_time=<t> until=<none|met|unmet>`; errors start `SR-ERROR`. `--until sim_time:T` stops
at the first step whose clock reaches T (its last steps taken one at a time); the observers' conditions (event:,
stage:, ending) are not carried. Exit codes as the reference's: 0 done, 2 `--steps` or `--until` not reached within
`--max-steps` (default 2,000,000), 3 no device or a device error, 4 bad arguments or scene (also a scene or condition
this path does not carry), 5 a non-finite value in the state
(the reference's line, naming the step, cell and channel, headless.rs:620-625). The guard scans the state at every
step's P9 here (fused into the floors pass, so it costs nothing), where the reference's scans steps 0, 64, 128, ... and
the state after a run's last step: the step named is the first at whose end the value was in the state, which the
reference names only when that step is one it scans.

With --out, DIR/summary.json holds the reference's run-summary fields this path has -- version, tool, adapter, scene,
world {width, height, edge}, steps, sim_time, wall_s, with --until `until` {condition, met}, and with --timing,
`timing` {per_pass_ms by the reference's pass labels, step_ms_mean, source "gpu-timestamp"} (headless.rs:497-531,
:1356-1381) -- and DIR/dumps/s_<step>.bin the SRDUMP01 dumps: at every multiple K of the step index (counted from the
scene) with --dump-every K, and after the last step (headless.rs:1211-1213, :1298-1345). A dump of a run with heat
carries F (resume.flux); beside every dump, s_<step>.bin.sigma64 holds Sigma's FP64 plane. The floors and the step's
reductions run fused here, so their time is reported under "P8 floors".

--load-dump (read and checked before anything else: a bad dump is exit 4, `SR-ERROR bad dump: ...`) takes the world,
the cells, the step index, the clock (sim time, dt, dt_next) and F from the dump, and Sigma's FP64 plane from its
.sigma64 file when there is one: such a resume goes on bit for bit, one without that file from the dump's f32 Sigma.
--world must then agree with the dump's world. --steps N is N more steps (the DONE line counts them, the dumps'
indices count from the scene). --world resizes the scene's world, its objects shifted to keep their place relative to
the centre (scene.rs:428-435). Refusals of the library (a world size, an FFT axis past the cap, an argument) are
exit 4.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

PASS_LABELS = {1: "P1 dt", 2: "P2 gravity", 3: "P3 sinks", 4: "P4 hydro", 5: "P5 heat", 6: "P6 reactions",
               8: "P8 floors"}
# the switch that records each optional pass (the reference names only the passes a step records, headless.rs:521-528)
PASS_SWITCH = {2: "gravity", 3: "sinks", 4: "hydro", 5: "heat", 6: "reactions"}
MAX_STEPS = 2_000_000  # --max-steps's default (headless.rs:77)
# the state's channel names, as the guard's line names them (crates/sr-engine/src/state.rs:29-31)
STATE_CHANNELS = ["sigma", "mom_x", "mom_y", "energy", "x_H", "x_He", "x_C", "x_O", "x_Ne", "x_Mg", "x_Si", "x_S",
                  "x_Fe", "x_n"]


class _Args(argparse.ArgumentParser):
    """Bad arguments are exit 4, as the reference's (exit 2 is a run that did not reach its target)."""

    def error(self, message):
        raise ValueError(f"bad arguments: {message}")


def parse_until(text):
    """`--until sim_time:<t>` (t finite, >= 0), the one condition this path reads (watch.rs:115-135); the
    observers' conditions (event:, stage:, ending) are refused as not carried."""
    kind, _, value = text.partition(":")
    if kind == "sim_time" and value:
        try:
            t = float(value)
        except ValueError:
            t = float("nan")
        if not (t >= 0.0 and t != float("inf")):
            raise ValueError(f"--until {text!r}: sim_time takes a time of 0 or more, got {value!r}")
        return t
    if kind in ("event", "stage") or text == "ending":
        raise ValueError(f"--until {text!r}: the observers' conditions are not carried by this path (sim_time:<t> is)")
    raise ValueError(f"--until {text!r}: expected sim_time:<t>, event:<kind>, stage:<id> or ending")


def _world_arg(text):
    wh = text.split("x")
    if len(wh) != 2 or not all(v.isdigit() for v in wh):
        raise ValueError(f"bad arguments: --world takes WxH, got {text!r}")
    ww, hh = int(wh[0]), int(wh[1])
    if not (8 <= ww <= 4096 and 8 <= hh <= 4096 and ww % 8 == 0 and hh % 8 == 0):
        raise ValueError(f"bad arguments: --world: {ww} x {hh} is not a world size (multiples of 8 in [8, 4096])")
    return ww, hh


def run(argv=None) -> int:
    ap = _Args(prog="sandbox_reactions_amd.headless")
    ap.add_argument("--scene", required=True)
    ap.add_argument("--steps", type=int)
    ap.add_argument("--until")
    ap.add_argument("--max-steps", type=int, default=MAX_STEPS)
    ap.add_argument("--out")
    ap.add_argument("--timing", action="store_true")
    ap.add_argument("--dump-every", type=int)
    ap.add_argument("--load-dump")
    ap.add_argument("--gamma", type=float, default=2.0)
    ap.add_argument("--world")
    ap.add_argument("--edge")
    from . import dump, scene_json
    try:
        a = ap.parse_args(argv)
        # headless.rs:190-202: one of --steps and --until, at least one step, a positive bound, a dump cadence of at
        # least one step and somewhere to write it
        if (a.steps is None) == (a.until is None):
            raise ValueError("bad arguments: give --steps N or --until <condition>, one of the two")
        if a.steps is not None and a.steps < 1:
            raise ValueError("bad arguments: --steps must be at least 1")
        if a.max_steps < 1:
            raise ValueError("bad arguments: --max-steps must be at least 1")
        if a.dump_every is not None and a.dump_every < 1:
            raise ValueError("bad arguments: --dump-every must be at least 1")
        if a.dump_every is not None and not a.out:
            raise ValueError("bad arguments: --dump-every needs --out (dumps go to DIR/dumps)")
        until = parse_until(a.until) if a.until is not None else None
        world_over = _world_arg(a.world) if a.world is not None else None
        if a.edge is not None and a.edge not in ("leave", "bounce"):
            raise ValueError(f"bad arguments: --edge takes leave or bounce, got {a.edge!r}")
    except ValueError as e:
        msg = str(e)
        print(f"SR-ERROR {msg if msg.startswith('bad arguments') else 'bad arguments: ' + msg}", flush=True)
        return 4
    # a dump to start from, read and checked before anything else (headless.rs:943-957); it gives the world
    d = None
    if a.load_dump:
        try:
            d = dump.read(a.load_dump)
        except (OSError, ValueError) as e:
            print(f"SR-ERROR bad dump: {e}", flush=True)
            return 4
        if world_over is not None and world_over != (d["width"], d["height"]):
            print(f"SR-ERROR bad arguments: --world {world_over[0]}x{world_over[1]} against the dump's "
                  f"{d['width']}x{d['height']}", flush=True)
            return 4
    try:
        scene = scene_json.load(a.scene)
        if d is not None:
            scene.resize(d["width"], d["height"])
        elif world_over is not None:
            scene.resize(*world_over)
        # --edge leave|bounce stands over the scene's (headless.rs:182-187)
        if a.edge is not None:
            scene.edge = 1 if a.edge == "leave" else 0
        scene.check_path()
        if d is None:
            scene.planes(a.gamma)  # every cell laid down before a device is touched (a bad scene is exit 4)
    except (OSError, ValueError) as e:
        print(f"SR-ERROR bad scene: {e}", flush=True)
        return 4
    try:
        import sandbox_reactions_amd as sr
        sr.require_device()
        name, cus, mem = sr.device_info()
    except Exception as e:  # no device or library
        print(f"SR-ERROR {e}", flush=True)
        return 3
    print(f"SR-ADAPTER name={name} backend=hip driver=libsrhip cus={cus}", flush=True)
    refused = (sr._lib.SR_ERR_WORLD_SIZE, sr._lib.SR_ERR_BOX_AXIS, sr._lib.SR_ERR_ARGUMENT, sr._lib.SR_ERR_LENGTH)
    try:
        try:
            # the passes over the active box, as the reference's headless runs them (headless.rs:1081)
            world = scene_json.world(scene, gamma=a.gamma, planes=d["planes"] if d is not None else None,
                                     box_mode=sr.BOX_FIT)
        except sr.SrError as e:
            if e.code in refused:
                print(f"SR-ERROR bad arguments: {e}", flush=True)
                return 4
            raise
        step0 = 0
        if d is not None:
            # the dump's records: its clock and F (and Sigma's FP64 plane beside it, for a bit-for-bit resume); the
            # world's clock goes on from the dump's sim time, the step index from its step
            step0 = int(d["step"])
            flux = d["flux"] if scene.switches["heat"] else None
            world.resume(float(d["dt"]), float(d["dt_next"]), float(d["sim_time"]), flux=flux, sigma=d["sigma64"])
            # and its active box with the step index its refits count from (Step::restore, step/latch.rs:496-505)
            world.resume_box(d["box"], step0)
            # and P3's sinks from its header (dump.rs:87-99)
            if scene.switches["sinks"]:
                world.put_sinks(d["sinks"])
        if a.timing:
            world.set_timing(True)
        dumps = os.path.join(a.out, "dumps") if a.out else None
        if dumps:
            os.makedirs(dumps, exist_ok=True)

        def write_dump(r):
            # F follows the channels whenever the run runs P5 (dump.rs:16-19); every dump here follows a step
            flux = np.stack([world.read_field(sr.FIELD_FLUX_X), world.read_field(sr.FIELD_FLUX_Y)]) \
                if scene.switches["heat"] else None
            b = world.read_box()
            dump.write(os.path.join(dumps, f"s_{step0 + done}.bin"), world.readback(), step=step0 + done,
                       sim_time=r.sim_time, dt=r.dt, dt_next=r.dt_next, flux=flux, sigma64=world.readback_sigma(),
                       box=[b.x0, b.y0, b.w, b.h, b.fft[0], b.fft[1], b.from_step, b.fits],
                       box_args=list(b.tiles) + list(b.cells),
                       sinks=world.read_sinks()[0] if scene.switches["sinks"] else ())

        t0 = time.perf_counter()
        target = min(a.steps if a.steps is not None else a.max_steps, a.max_steps)
        done, met = 0, False
        r = world.read_dt()
        while done < target and not met:
            k = target - done
            if until is not None and r.dt_next > 0.0:
                # headless.rs:1206-1210: a chunk covers at most half the time left at the next step's dt, so the
                # last steps before the condition are single ones and the run stops at the first step past it
                k = min(k, max(1, int((until - (r.sim_time)) / (2.0 * r.dt_next))))
            if a.dump_every is not None:
                # headless.rs:1211-1213: chunks end on the multiples of K of the step index counted from the scene
                k = min(k, a.dump_every - (step0 + done) % a.dump_every)
            world.run(k)
            done += k
            r = world.read_dt()
            met = until is not None and r.sim_time >= until
            if r.non_finite:
                print(f"SR-ERROR non-finite value in the state at step {step0 + r.nf_step}, cell ({r.nf_x}, "
                      f"{r.nf_y}), channel {STATE_CHANNELS[r.nf_channel]}", flush=True)
                return 5
            # headless.rs:1298-1305: a dump at every multiple of K before the last step, the last always
            if dumps and a.dump_every is not None and (step0 + done) % a.dump_every == 0 and done < target \
                    and not met:
                write_dump(r)
        wall = time.perf_counter() - t0
        r = world.read_dt()
        if dumps:
            write_dump(r)
    except sr.SrError as e:
        print(f"SR-ERROR {e}", flush=True)
        return 5 if e.code == sr._lib.SR_ERR_NON_FINITE else (4 if e.code in refused else 3)
    if a.out:
        summary = {"version": 1, "tool": "sandbox-reactions-amd", "adapter": {"name": name, "backend": "hip",
                                                                            "driver": "libsrhip"},
                   "scene": a.scene, "world": {"width": scene.width, "height": scene.height,
                                               "edge": "leave" if scene.edge == 1 else "bounce"},
                   "steps": step0 + done, "sim_time": r.sim_time, "wall_s": wall}
        if until is not None:
            summary["until"] = {"condition": f"sim_time:{a.until.partition(':')[2]}", "met": met}
        if a.timing:
            c = world.read_cost()
            ran = [k for k in sorted(PASS_LABELS) if k not in PASS_SWITCH or scene.switches[PASS_SWITCH[k]]]
            summary["timing"] = {"per_pass_ms": {PASS_LABELS[k]: c.pass_ms[k] for k in ran},
                                 "step_ms_mean": c.step_ms, "source": "gpu-timestamp"}
        with open(os.path.join(a.out, "summary.json"), "w") as f:
            f.write(json.dumps(summary, indent=2) + "\n")
    # the steps of this run, the clock to 6 decimals and the condition's verdict: none for a run on --steps that
    # took them all, met, or unmet when --max-steps cut the run (exit 2) (headless.rs:1389-1400)
    reached = met if until is not None else done >= (a.steps or 0)
    verdict = ("met" if until is not None else "none") if reached else "unmet"
    print(f"SR-HEADLESS DONE steps={done} sim_time={r.sim_time:.6f} until={verdict}", flush=True)
    return 0 if reached else 2


if __name__ == "__main__":
    sys.exit(run())
