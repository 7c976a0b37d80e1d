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


def test_a_scene_with_cold_pressure_is_exit_4():
    code, out = headless("--scene", os.path.join(SCENES, "asym_blob.json"), "--steps", "1")
    assert code == 4 and "cold pressure" in out


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
    # resuming from the second dump for one step lands on the third's state
    code, out = headless("--scene", os.path.join(SCENES, "heat_rim.json"), "--steps", "1", "--out",
                         str(tmp_path / "resumed"), "--load-dump", str(tmp_path / "dumps" / "s_2.bin"))
    assert code == 0, out
    a = dump.read(str(tmp_path / "resumed" / "dumps" / "s_3.bin"))["planes"]
    b = dump.read(str(tmp_path / "dumps" / "s_3.bin"))["planes"]
    assert np.max(np.abs(a[0] - b[0])) <= 1e-6 * np.abs(b[0]).max()
