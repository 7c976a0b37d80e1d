"""The traffic summary of two counter runs (sandbox_reactions_amd.pmc) on synthetic counter files: the calibration
copy's scale, per-kernel and per-grid bytes per launch, and the launches of one kernel at several grids kept apart."""
import csv
import json
import os

from sandbox_reactions_amd import pmc


def write_run(d, rows):
    os.makedirs(d, exist_ok=True)
    with open(os.path.join(d, "run_counter_collection.csv"), "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=["Kernel_Name", "Grid_Size", "Counter_Value"])
        w.writeheader()
        for name, grid, kib in rows:
            w.writerow({"Kernel_Name": name, "Grid_Size": grid, "Counter_Value": kib})


def test_summary_scales_by_the_copy_and_keeps_grids_apart(tmp_path):
    mib = 1024.0  # KiB per MiB
    # FETCH reads half the copy (a scale of 2), WRITE reads it exactly
    write_run(tmp_path / "fetch", [("k_calib_copy1(float const*, float*, unsigned long)", 2097152, 256 * mib),
                                   ("sr::k_pre<32, 32, 2>(sr::Lv, float const*)", 262144, 10 * mib),
                                   ("sr::k_pre<32, 32, 2>(sr::Lv, float const*)", 262144, 12 * mib),
                                   ("sr::k_pre<32, 32, 2>(sr::Lv, float const*)", 65536, 3 * mib)])
    write_run(tmp_path / "write", [("k_calib_copy1(float const*, float*, unsigned long)", 2097152, 512 * mib),
                                   ("sr::k_pre<32, 32, 2>(sr::Lv, float const*)", 262144, 5 * mib),
                                   ("sr::k_pre<32, 32, 2>(sr::Lv, float const*)", 262144, 5 * mib),
                                   ("sr::k_pre<32, 32, 2>(sr::Lv, float const*)", 65536, 1 * mib)])
    s = pmc.summarize(str(tmp_path / "fetch"), str(tmp_path / "write"))
    assert abs(s["copy_scale"]["fetch"] - 2.0) < 1e-12 and abs(s["copy_scale"]["write"] - 1.0) < 1e-12
    big = s["by_grid"]["sr::k_pre<32, 32, 2>|262144"]
    assert big["launches"] == 2
    assert abs(big["fetch_bytes"] - 2 * 11 * 2**20) < 1 and abs(big["write_bytes"] - 5 * 2**20) < 1
    small = s["by_grid"]["sr::k_pre<32, 32, 2>|65536"]
    assert abs(small["traffic_bytes"] - (2 * 3 + 1) * 2**20) < 1
    assert s["kernels"]["sr::k_pre<32, 32, 2>"]["grid"] == 262144  # the grid with the most traffic
    s["workload"], s["command"] = "synthetic", "none"
    out = tmp_path / "traffic.json"
    pmc.write_outputs(s, str(out))
    assert json.loads(out.read_text())["by_grid"].keys() == s["by_grid"].keys()
    md = (tmp_path / "traffic.md").read_text()
    assert "| sr::k_pre<32, 32, 2> | 65536 | 1 |" in md
