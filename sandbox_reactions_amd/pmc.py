"""HBM traffic per kernel launch from two rocprofv3 counter runs (FETCH_SIZE, WRITE_SIZE; KiB per dispatch), and the
per-kernel durations of a rocprofv3 kernel trace.

Each counter is scaled by what it reads for a device copy of known size: the 512 MiB copy `bench.py` makes with
one float per lane (k_calib_copy1, the access pattern of the stencil kernels), which reads and writes 512 MiB. The
float4 and runtime copies are reported beside it (wide requests can count differently). Usage (after the two runs, on any machine):

    python -m sandbox_reactions_amd.pmc FETCH_DIR WRITE_DIR OUT_JSON [--workload TEXT] [--command TEXT]
    python -m sandbox_reactions_amd.pmc --trace TRACE_DIR OUT_MD [--workload TEXT] [--command TEXT]
    python -m sandbox_reactions_amd.pmc --counters OUT_MD PASS_DIR... [--kernel REGEX] [--workload TEXT] [--command TEXT]

OUT_JSON gets {"copy_scale": {...}, "kernels": {name: {"launches", "fetch_bytes", "write_bytes", "traffic_bytes",
"grid"}}} for the dispatch grid with the most total traffic per kernel, and "by_grid": {"name|grid": ...} for every
grid a kernel was launched with; a Markdown table goes beside it.
"""
from __future__ import annotations

import argparse
import collections
import csv
import glob
import json
import os

import numpy as np

COPY_BYTES = 512 * 2**20


def read_counters(d):
    """The counter rows of a rocprofv3 --pmc run: its CSV output (-f csv) or its rocpd database (the default)."""
    paths = glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True)
    rows = []
    for p in paths:
        with open(p) as f:
            rows += list(csv.DictReader(f))
    if not paths:
        import sqlite3
        dbs = glob.glob(os.path.join(d, "**", "*_results.db"), recursive=True)
        if not dbs:
            raise FileNotFoundError(f"no counter_collection.csv or rocpd database under {d}")
        for p in dbs:
            con = sqlite3.connect(p)
            q = "select kernel_name, grid_size, counter_name, value from counters_collection"
            rows += [{"Kernel_Name": k, "Grid_Size": g, "Counter_Name": c, "Counter_Value": v}
                     for k, g, c, v in con.execute(q)]
            con.close()
    return rows


def short(name):
    return name.split("(")[0].replace("void ", "").strip()


def per_kernel(rows):
    agg = collections.defaultdict(list)
    for r in rows:
        agg[(short(r["Kernel_Name"]), int(r["Grid_Size"]))].append(float(r["Counter_Value"]) * 1024.0)
    return agg


def copy_reading(agg, kernel="k_calib_copy1"):
    """The counter's bytes for the 512 MiB calibration copy (median over its dispatches)."""
    vals = [x for (k, g), v in agg.items() if k.endswith(kernel) for x in v]
    if not vals:
        raise ValueError(f"no {kernel} dispatch in the counter run: the calibration copy is missing")
    return float(np.median(vals))


def summarize(fetch_dir, write_dir):
    fa, wa = per_kernel(read_counters(fetch_dir)), per_kernel(read_counters(write_dir))
    sf = COPY_BYTES / copy_reading(fa)
    sw = COPY_BYTES / copy_reading(wa)
    others = {}
    for k in ("k_calib_copy4", "copyBuffer"):
        try:
            others[k] = {"fetch": COPY_BYTES / copy_reading(fa, k), "write": COPY_BYTES / copy_reading(wa, k)}
        except ValueError:
            pass
    kernels, by_grid = {}, {}
    for key in fa.keys() & wa.keys():
        name, grid = key
        if "copyBuffer" in name or "fillBuffer" in name or "k_calib" in name:
            continue
        f = sum(fa[key]) / len(fa[key]) * sf
        w = sum(wa[key]) / len(wa[key]) * sw
        e = {"grid": grid, "launches": len(fa[key]), "fetch_bytes": f, "write_bytes": w, "traffic_bytes": f + w}
        by_grid[f"{name}|{grid}"] = e
        cur = kernels.get(name)
        if cur is None or (f + w) * len(fa[key]) > cur["traffic_bytes"] * cur["launches"]:
            kernels[name] = e
    # `kernels`: per kernel name, its grid with the most traffic over the run; `by_grid`: every (name, grid)
    return {"copy_scale": {"fetch": sf, "write": sw}, "other_copy_scales": others, "kernels": kernels,
            "by_grid": by_grid}


def trace_table(trace_dir):
    """Per (kernel, grid): launches, mean and total device time (us, ms) from a kernel trace's CSV
    (rocprofv3 --kernel-trace -f csv), sorted by total time."""
    paths = glob.glob(os.path.join(trace_dir, "**", "*kernel_trace.csv"), recursive=True)
    if not paths:
        raise FileNotFoundError(f"no kernel_trace.csv under {trace_dir}")
    agg = collections.defaultdict(list)
    for p in paths:
        with open(p) as f:
            for r in csv.DictReader(f):
                grid = "x".join(r.get(f"Grid_Size_{a}", "1") for a in "XY")
                dur = (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) * 1e-3
                agg[(short(r["Kernel_Name"]), grid, r.get("Workgroup_Size_X", ""))].append(dur)
    rows = [(k, len(v), float(np.mean(v)), float(np.sum(v)) * 1e-3) for k, v in agg.items()]
    return sorted(rows, key=lambda r: -r[3])


def write_trace(trace_dir, out_md, workload, command):
    rows = trace_table(trace_dir)
    with open(out_md, "w") as f:
        f.write(f"# Kernel trace: {workload}\n\nCommand: `{command}`\n\n")
        f.write("Per kernel and grid, from the trace's CSV, sorted by total time.\n\n")
        f.write("| kernel | grid (threads) | wg | launches | avg us | total ms |\n|---|---|---|---|---|---|\n")
        for (name, grid, wg), n, mean, total in rows:
            f.write(f"| {name} | {grid} | {wg} | {n} | {mean:.2f} | {total:.2f} |\n")


def counters_table(dirs, kernel_re):
    """Per (kernel, grid) matching kernel_re: each counter's median per dispatch over the runs in dirs (one counter
    pass each), and the dispatches it was read on."""
    import re
    pat = re.compile(kernel_re)
    agg = collections.defaultdict(lambda: collections.defaultdict(list))
    for d in dirs:
        for r in read_counters(d):
            name = short(r["Kernel_Name"])
            if not pat.search(name):
                continue
            agg[(name, int(r["Grid_Size"]))][r["Counter_Name"]].append(float(r["Counter_Value"]))
    return {k: {c: (float(np.median(v)), len(v)) for c, v in cs.items()} for k, cs in agg.items()}


def write_counters(dirs, out_md, kernel_re, workload, command):
    t = counters_table(dirs, kernel_re)
    with open(out_md, "w") as f:
        f.write(f"# Counters: {workload}\n\nCommands: `{command}`\n\n")
        f.write("Median per dispatch of each counter (one rocprofv3 --pmc pass per group of counters), per kernel and "
                "grid (threads).\n\n")
        for (name, grid), cs in sorted(t.items()):
            f.write(f"## {name}, grid {grid}\n\n| counter | median per dispatch | dispatches |\n|---|---|---|\n")
            for c, (v, n) in sorted(cs.items()):
                f.write(f"| {c} | {v:.6g} | {n} |\n")
            f.write("\n")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--trace", action="store_true", help="summarise a kernel trace: TRACE_DIR OUT_MD")
    ap.add_argument("--counters", action="store_true", help="tabulate counter passes: OUT_MD PASS_DIR...")
    ap.add_argument("--kernel", default=".", help="with --counters: the kernels to keep (a regular expression)")
    ap.add_argument("paths", nargs="+")
    ap.add_argument("--workload", default="")
    ap.add_argument("--command", default="")
    a = ap.parse_args()
    if a.trace:
        write_trace(a.paths[0], a.paths[1], a.workload, a.command)
        return
    if a.counters:
        write_counters(a.paths[1:], a.paths[0], a.kernel, a.workload, a.command)
        return
    if len(a.paths) != 3:
        ap.error("FETCH_DIR WRITE_DIR OUT_JSON")
    a.fetch_dir, a.write_dir, a.out_json = a.paths
    s = summarize(a.fetch_dir, a.write_dir)
    s["workload"] = a.workload
    s["command"] = a.command
    write_outputs(s, a.out_json)


def write_outputs(s, out_json):
    """The summary as JSON and, beside it, a Markdown table."""
    with open(out_json, "w") as f:
        json.dump(s, f, indent=1, sort_keys=True)
    md = os.path.splitext(out_json)[0] + ".md"
    with open(md, "w") as f:
        f.write(f"# HBM traffic per launch: {s['workload']}\n\nCommands: `{s['command']}`\n\n")
        f.write(f"Counters scaled by the 512 MiB one-float-per-lane calibration copy: FETCH "
                f"x{s['copy_scale']['fetch']:.4f}, WRITE x{s['copy_scale']['write']:.4f} (the same counters over other "
                f"copies of 512 MiB would need: {json.dumps(s['other_copy_scales'])}).\n\n")
        f.write("Every kernel and grid (threads) the run launched, by total traffic.\n\n")
        f.write("| kernel | grid | launches | fetch MB | write MB | traffic MB |\n|---|---|---|---|---|---|\n")
        rows = s.get("by_grid") or {f"{k}|{v['grid']}": v for k, v in s["kernels"].items()}
        for k, v in sorted(rows.items(), key=lambda kv: -kv[1]["traffic_bytes"] * kv[1]["launches"]):
            f.write(f"| {k.split('|')[0]} | {v['grid']} | {v['launches']} | {v['fetch_bytes'] / 1e6:.2f} | "
                    f"{v['write_bytes'] / 1e6:.2f} | {v['traffic_bytes'] / 1e6:.2f} |\n")


if __name__ == "__main__":
    main()
