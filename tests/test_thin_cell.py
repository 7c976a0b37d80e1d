"""The HOT_FACE bracket (crates/sr-physics/src/reference/hydro.rs:305-331) on the reference's own case
(crates/sr-engine/tests/gpu/sod.rs:19-24, :671-686): tests/golden/dumps/thin_cell.bin, 64 x 64 cells of an evolved
life with a dense cell between a thin one and denser gas, the cold pressure on. One step of P4 alone, leave edges:
no free gas cell's |v| above 10 x the largest of its eight free gas neighbours'. Without the bracket the thin cell
leaves at |v| ~ 2,100, 54.6 x its neighbours' (the reference's measurement); the oracle with its bracket switched off
shows the check catches exactly that, so the check pins the bracket in the oracle, and the GPU is held to the same
bound and to the oracle's state. The dump's held cells (the hold is not carried by this path) are left out of the
check as the reference leaves them out; here they step as free gas."""
import os

import numpy as np
import pytest

import oracle
from sandbox_reactions_amd import cold, dump

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "dumps")
LEAP = 10.0


def case():
    d = dump.read(os.path.join(HERE, "thin_cell.bin"))
    kw = dict(gamma=2.0, dx=1.0, edge=oracle.EDGE_LEAVE, **cold.SHIPPED)
    held = d["held"] if d["held"] is not None else np.zeros((64, 64), np.uint32)
    return d["planes"].astype(np.float64), held.astype(bool), kw


def worst_leap(planes, held, sigma_vac=1e-6):
    s, mx, my = planes[0], planes[1], planes[2]
    gas = (~held) & (s >= sigma_vac)
    speed = np.where(s > 0, np.hypot(mx, my) / np.where(s > 0, s, 1.0), 0.0)
    h, w = s.shape
    worst = (0.0, (0, 0), 0.0)
    for y in range(1, h - 1):
        for x in range(1, w - 1):
            if not gas[y, x]:
                continue
            near = [(y + dy, x + dx) for dy in (-1, 0, 1) for dx in (-1, 0, 1) if (dy, dx) != (0, 0)]
            most = max([speed[q] for q in near if gas[q]], default=0.0)
            if most > 0.0 and speed[y, x] / most > worst[0]:
                worst = (speed[y, x] / most, (x, y), speed[y, x])
    return worst


def twin_step(hot_face=True):
    planes, held, kw = case()
    ph = oracle.reference_physics(**kw)
    oracle.lib.sro_hydro_set_hot_face(1 if hot_face else 0)
    try:
        dt = oracle.hydro_dt(ph, planes)
        out = oracle.hydro_step(ph, planes, dt)
    finally:
        oracle.lib.sro_hydro_set_hot_face(1)
    return out, held, dt


def test_the_twin_takes_no_leap_and_would_without_the_bracket():
    out, held, dt = twin_step()
    ratio, cell, v = worst_leap(out, held)
    print(f"with the bracket: worst {ratio:.3f} at {cell}, |v| {v:.3e} (dt {dt:.4e})")
    assert ratio <= LEAP
    bad, held, _ = twin_step(hot_face=False)
    ratio_off, cell_off, v_off = worst_leap(bad, held)
    print(f"without it: worst {ratio_off:.3f} at {cell_off}, |v| {v_off:.3e}")
    assert ratio_off > LEAP


@pytest.mark.gpu
def test_the_gpu_takes_no_leap_and_matches_the_twin(gpu_lib):
    sr = gpu_lib
    planes, held, kw = case()
    print("thin cell on the GPU", flush=True)
    world = sr.World(64, 64, physics=sr.physics(**kw), gravity=False, hydro=True, heat=False)
    world.upload(planes.astype(np.float32))
    dt = world.read_dt().dt_next
    ph = oracle.reference_physics(**kw)
    want = oracle.hydro_step(ph, planes.astype(np.float32).astype(np.float64), dt)
    world.set_dt_cap(dt)
    world.run(1)
    g = world.readback().astype(np.float64)
    ratio, cell, v = worst_leap(g, held)
    print(f"GPU: worst {ratio:.3f} at {cell}, |v| {v:.3e}", flush=True)
    assert ratio <= LEAP
    assert np.abs(g[0] - want[0]).sum() / np.abs(want[0]).sum() <= 1e-5
