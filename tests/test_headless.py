"""The headless driver (sandbox_reactions_amd/headless.py): refusals and exit codes without a device (CPU), and a run
of one of the reference's scenes with its summary.json, timing and dumps (GPU)."""
import json
import os
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SCENES = os.path.join(ROOT, "tests", "golden", "scenes")


def headless(*args):
    r = subprocess.run([sys.executable, "-m", "sandbox_reactions_amd.headless", *args], cwd=ROOT, capture_output=True,
                       text=True, timeout=300)
    return r.returncode, r.stdout


def test_a_missing_scene_is_exit_4():
    code, out = headless("--scene", "no/such/scene.json", "--steps", "1")
    assert code == 4 and out.startswith("SR-ERROR")


def test_a_scene_the_path_does_not_carry_is_exit_4(tmp_path):
    code, out = headless("--scene", os.path.join(SCENES, "asym_blob.json"), "--steps", "1")
    assert code == 4 and "neutrinos" in out  # its default switches run P7 and P3
    bad = {"version": 1, "world": {"width": 32, "height": 32},
           "switches": {"reactions": False, "neutrinos": False, "sinks": False},
           "objects": [{"kind": "rect", "x0": 0, "y0": 0, "w": 8, "h": 8, "sigma": 1.0, "pressure": 1e-3,
                        "composition": {"H": 1.0}}]}
    path = tmp_path / "below_cold.json"
    path.write_text(json.dumps(bad))
    code, out = headless("--scene", str(path), "--steps", "1")
    assert code == 4 and "below the cold pressure" in out


@pytest.mark.parametrize("args, text", [
    (["--steps", "3", "--until", "sim_time:1"], "one of the two"),
    ([], "one of the two"),
    (["--steps", "0"], "at least 1"),
    (["--until", "sim_time:-1"], "0 or more"),
    (["--until", "sim_time:abc"], "0 or more"),
    (["--until", "event:supernova"], "not carried"),
    (["--until", "ending"], "not carried"),
    (["--until", "sometime"], "expected sim_time"),
    (["--steps", "3", "--max-steps", "0"], "at least 1"),
    (["--steps", "x"], "bad arguments"),
    (["--steps", "1", "--world", "7x8"], "not a world size"),
    (["--steps", "1", "--world", "64by64"], "WxH"),
    (["--steps", "1", "--edge", "sideways"], "leave or bounce"),
])
def test_bad_arguments_are_exit_4(args, text):
    """headless.rs:190-202 and watch.rs:115-135: one of --steps and --until, at least one step, sim_time of 0 or more;
    the observers' conditions are refused here as not carried."""
    code, out = headless("--scene", os.path.join(SCENES, "sod.json"), *args)
    assert code == 4 and out.startswith("SR-ERROR") and text in out, out


def test_no_device_is_exit_3():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a device is visible")
    code, out = headless("--scene", os.path.join(SCENES, "sod.json"), "--steps", "1")
    assert code == 3 and out.startswith("SR-ERROR")


@pytest.mark.gpu
def test_a_run_writes_the_summary_and_dumps(gpu_lib, tmp_path):
    from sandbox_reactions_amd import dump
    print("headless heat_rim 3 steps", flush=True)
    code, out = headless("--scene", os.path.join(SCENES, "heat_rim.json"), "--steps", "3", "--out", str(tmp_path),
                         "--timing", "--dump-every", "1")
    assert code == 0, out
    lines = out.strip().split("\n")
    assert lines[0].startswith("SR-ADAPTER") and lines[-1].startswith("SR-HEADLESS DONE steps=3")
    s = json.load(open(tmp_path / "summary.json"))
    assert s["version"] == 1 and s["steps"] == 3 and s["world"] == {"width": 64, "height": 64, "edge": "leave"}
    assert s["timing"]["source"] == "gpu-timestamp" and s["timing"]["per_pass_ms"]["P5 heat"] > 0
    for k in (1, 2, 3):
        d = dump.read(str(tmp_path / "dumps" / f"s_{k}.bin"))
        assert d["step"] == k and np.all(np.isfinite(d["planes"]))
    assert d["flux"] is not None  # a run with heat writes F (dump.rs:16-19)
    # resuming from the second dump for one step lands on the third's state, bit for bit (dump.rs:8-20)
    code, out = headless("--scene", os.path.join(SCENES, "heat_rim.json"), "--steps", "1", "--out",
                         str(tmp_path / "resumed"), "--load-dump", str(tmp_path / "dumps" / "s_2.bin"))
    assert code == 0, out
    a = dump.read(str(tmp_path / "resumed" / "dumps" / "s_3.bin"))
    b = dump.read(str(tmp_path / "dumps" / "s_3.bin"))
    assert np.array_equal(a["planes"].view(np.uint32), b["planes"].view(np.uint32))
    assert np.array_equal(a["flux"].view(np.uint32), b["flux"].view(np.uint32))
    assert np.array_equal(a["sigma64"].view(np.uint64), b["sigma64"].view(np.uint64))
    assert (a["sim_time"], a["dt"], a["dt_next"], a["step"]) == (b["sim_time"], b["dt"], b["dt_next"], b["step"])
    assert a["box"] == b["box"] and a["box_args"] == b["box_args"]


@pytest.mark.gpu
def test_a_burning_scene_runs_p6(gpu_lib, tmp_path):
    """ref_burn.json (heat and reactions on, gravity and hydro off) for 4 steps with timing: summary.json names the
    passes the steps ran, P6 among them with its own time, and none that they did not; resuming from the dump at 2
    lands on the dump at 4 bit for bit (P6 keeps no state between steps)."""
    from sandbox_reactions_amd import dump
    print("headless ref_burn 4 steps", flush=True)
    scene = os.path.join(SCENES, "ref_burn.json")
    code, out = headless("--scene", scene, "--steps", "4", "--out", str(tmp_path), "--timing", "--dump-every", "2")
    assert code == 0, out
    s = json.load(open(tmp_path / "summary.json"))
    passes = s["timing"]["per_pass_ms"]
    assert set(passes) == {"P1 dt", "P5 heat", "P6 reactions", "P8 floors"} and passes["P6 reactions"] > 0
    code, out = headless("--scene", scene, "--steps", "2", "--out", str(tmp_path / "resumed"), "--load-dump",
                         str(tmp_path / "dumps" / "s_2.bin"))
    assert code == 0, out
    a = dump.read(str(tmp_path / "resumed" / "dumps" / "s_4.bin"))
    b = dump.read(str(tmp_path / "dumps" / "s_4.bin"))
    assert np.array_equal(a["planes"].view(np.uint32), b["planes"].view(np.uint32))
    start = dump.read(str(tmp_path / "dumps" / "s_2.bin"))["planes"]
    assert not np.array_equal(b["planes"][4:], start[4:])


@pytest.mark.gpu
def test_a_resume_across_a_box_refit_goes_on_bit_for_bit(gpu_lib, tmp_path):
    """box_blob.json (the blob moving through its active box): 20 steps with a dump every 10, and 10 steps resumed
    from the dump at 10 (the box refit after step 16's P9 falls in the resumed part): the two dumps at 20 hold the same
    cells, box record and clock; the box is the blob's, not the world's."""
    from sandbox_reactions_amd import dump
    print("headless box_blob resume across a refit", flush=True)
    scene = os.path.join(SCENES, "box_blob.json")
    code, out = headless("--scene", scene, "--steps", "20", "--out", str(tmp_path / "a"), "--dump-every", "10")
    assert code == 0, out
    code, out = headless("--scene", scene, "--steps", "10", "--out", str(tmp_path / "b"), "--load-dump",
                         str(tmp_path / "a" / "dumps" / "s_10.bin"))
    assert code == 0, out
    a = dump.read(str(tmp_path / "a" / "dumps" / "s_20.bin"))
    b = dump.read(str(tmp_path / "b" / "dumps" / "s_20.bin"))
    assert np.array_equal(a["planes"].view(np.uint32), b["planes"].view(np.uint32))
    assert np.array_equal(a["sigma64"].view(np.uint64), b["sigma64"].view(np.uint64))
    assert a["box"] == b["box"] and a["box_args"] == b["box_args"]
    assert (a["sim_time"], a["dt"], a["dt_next"]) == (b["sim_time"], b["dt"], b["dt_next"])
    x0, y0, w, h, fw, fh, frm, fits = a["box"]
    fft = lambda n: max(32, 1 << (2 * n - 1).bit_length())
    assert (frm, fits) == (17, 3) and 0 < w < 600 and 0 < h < 400 and (fw, fh) == (fft(w), fft(h))
    assert a["box_args"] == [w // 8, h // 8, 1, -(-(w * h) // 256), 1, 1]


@pytest.mark.gpu
def test_until_sim_time_and_max_steps(gpu_lib, tmp_path):
    """--until sim_time:T stops at the first step whose clock reaches T (exit 0, until=met, the step before short of
    T); --world and --edge stand over the scene's world; --max-steps cutting a run on --steps is exit 2 with
    until=unmet."""
    from sandbox_reactions_amd import dump
    print("headless --until", flush=True)
    code, out = headless("--scene", os.path.join(SCENES, "sod.json"), "--until", "sim_time:0.05", "--out",
                         str(tmp_path))
    lines = out.strip().split("\n")
    assert code == 0 and lines[-1].startswith("SR-HEADLESS DONE") and lines[-1].endswith("until=met"), out
    s = json.load(open(tmp_path / "summary.json"))
    assert s["until"] == {"condition": "sim_time:0.05", "met": True}
    d = dump.read(str(tmp_path / "dumps" / f"s_{s['steps']}.bin"))
    assert d["sim_time"] >= 0.05 > d["sim_time"] - d["dt"] * (1 + 1e-6)
    code, out = headless("--scene", os.path.join(SCENES, "sod.json"), "--steps", "2", "--world", "400x16", "--edge",
                         "leave", "--out", str(tmp_path / "over"))
    assert code == 0, out
    assert json.load(open(tmp_path / "over" / "summary.json"))["world"] == {"width": 400, "height": 16, "edge": "leave"}
    code, out = headless("--scene", os.path.join(SCENES, "sod.json"), "--steps", "50", "--max-steps", "10")
    last = out.strip().split("\n")[-1]
    assert code == 2 and last.startswith("SR-HEADLESS DONE steps=10 ") and last.endswith("until=unmet"), out


@pytest.mark.parametrize("args, text", [
    (["--steps", "1", "--dump-every", "0", "--out", "x"], "--dump-every must be at least 1"),
    (["--steps", "1", "--dump-every", "2"], "needs --out"),
])
def test_dump_cadence_arguments_are_exit_4(args, text):
    """headless.rs:197-202: --dump-every of at least one step, and only with --out."""
    code, out = headless("--scene", os.path.join(SCENES, "sod.json"), *args)
    assert code == 4 and out.startswith("SR-ERROR bad arguments") and text in out, out


def test_bad_dumps_are_exit_4(tmp_path):
    """headless.rs:943-957: the dump is read and checked before anything else (no device touched): a missing or cut
    dump is exit 4 with `SR-ERROR bad dump`, and --world against the dump's world is a bad argument."""
    scene = os.path.join(SCENES, "heat_rim.json")
    code, out = headless("--scene", scene, "--steps", "1", "--load-dump", str(tmp_path / "none.bin"))
    assert code == 4 and out.startswith("SR-ERROR bad dump"), out
    src = os.path.join(ROOT, "tests", "golden", "dumps", "heat_overflow.bin")
    cut = tmp_path / "cut.bin"
    cut.write_bytes(open(src, "rb").read()[:-100])
    code, out = headless("--scene", scene, "--steps", "1", "--load-dump", str(cut))
    assert code == 4 and out.startswith("SR-ERROR bad dump") and "planes hold" in out, out
    code, out = headless("--scene", scene, "--steps", "1", "--load-dump", src, "--world", "128x64")
    assert code == 4 and out.startswith("SR-ERROR bad arguments") and "against the dump's 64x64" in out, out


@pytest.mark.gpu
def test_a_sink_scene_runs_p3_and_resumes_its_sinks(gpu_lib, tmp_path):
    """sink_cloud.json with a sink placed at its centre by a dump (the header's sinks, dump.rs:87-99): steps 2 to 4
    with timing name P3 among the passes; the sink accretes; resuming from the dump at 2 lands on the dump at 4 bit for
    bit, the sinks included."""
    from sandbox_reactions_amd import dump
    print("headless sink_cloud 4 steps", flush=True)
    scene = os.path.join(SCENES, "sink_cloud.json")
    code, out = headless("--scene", scene, "--steps", "1", "--out", str(tmp_path / "a0"))
    assert code == 0, out
    first = str(tmp_path / "a0" / "dumps" / "s_1.bin")
    d1 = dump.read(first)
    start = tmp_path / "start.bin"  # the first step's dump with a sink added to its header
    dump.write(str(start), d1["planes"], sigma64=dump.read_sigma64(first + ".sigma64", 64, 64), step=1,
               sim_time=d1["sim_time"], dt=d1["dt"], dt_next=d1["dt_next"], box=d1["box"], box_args=d1["box_args"],
               sinks=[(32.5, 32.5, 0.0, 0.0, 2000.0)])
    code, out = headless("--scene", scene, "--steps", "3", "--out", str(tmp_path / "a"), "--timing",
                         "--dump-every", "2", "--load-dump", str(start))
    assert code == 0, out
    passes = json.load(open(tmp_path / "a" / "summary.json"))["timing"]["per_pass_ms"]
    assert {"P1 dt", "P2 gravity", "P3 sinks", "P4 hydro", "P8 floors"} == set(passes) and passes["P3 sinks"] > 0
    d2 = dump.read(str(tmp_path / "a" / "dumps" / "s_2.bin"))
    d4 = dump.read(str(tmp_path / "a" / "dumps" / "s_4.bin"))
    assert len(d4["sinks"]) == 1 and d4["sinks"][0][4] > d2["sinks"][0][4] > 2000.0
    code, out = headless("--scene", scene, "--steps", "2", "--out", str(tmp_path / "b"), "--load-dump",
                         str(tmp_path / "a" / "dumps" / "s_2.bin"))
    assert code == 0, out
    b4 = dump.read(str(tmp_path / "b" / "dumps" / "s_4.bin"))
    assert np.array_equal(b4["planes"].view(np.uint32), d4["planes"].view(np.uint32)) and b4["sinks"] == d4["sinks"]
