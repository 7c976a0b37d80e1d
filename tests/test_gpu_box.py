"""The active box (crates/sr-engine/src/step/boxfit.rs:1-14; the reference's own cases in tests/gpu/box.rs): the box is
the bounding box of the cells with Sigma >= Sigma_vac, grown by 8, rounded out to multiples of 8 and clamped to the
world; its FFT size per axis the smallest power of two >= 2 x the box, at least 32; refit at the upload and after P9
at step indices = 0 (mod 16); P8 and P9 over its cells only; the whole-world switch holding it at the world on the same
schedule. The oracle is that rule in integers on the state the library holds (Sigma's FP64 plane); the hand-worked
boxes are box.rs:148-162's."""
import os

import numpy as np
import pytest

from sandbox_reactions_amd import scene_json

pytestmark = pytest.mark.gpu

SCENES = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "scenes")
MARGIN, ROUND, FFT_MIN, REFIT = 8, 8, 32, 16


def oracle_box(sigma, svac):
    """[x0, y0, w, h] of the cells with sigma >= svac, by the rule (box.rs:70-83)."""
    ys, xs = np.nonzero(sigma >= svac)
    if xs.size == 0:
        return [0, 0, 0, 0]
    h, w = sigma.shape
    lo = lambda v: max(0, ((v - MARGIN) // ROUND) * ROUND)
    hi = lambda v, n: min(n, ((v + MARGIN + ROUND - 1) // ROUND) * ROUND)
    x0, y0 = lo(int(xs.min())), lo(int(ys.min()))
    return [x0, y0, hi(int(xs.max()) + 1, w) - x0, hi(int(ys.max()) + 1, h) - y0]


def fft(n):
    f = 1
    while f < 2 * n:
        f *= 2
    return max(f, FFT_MIN)


def rect(b):
    return [b.x0, b.y0, b.w, b.h]


def svac32(sr):
    return float(np.float32(sr.physics().sigma_vac))


def vacuum(w, h):
    """A world of vacuum: Sigma 0, at rest, E 0, pure hydrogen (box.rs:55-60)."""
    p = np.zeros((14, h, w), np.float32)
    p[4] = 1.0
    return p


def with_gas(w, h, cells, sigma=1.0):
    p = vacuum(w, h)
    for x, y in cells:
        p[0, y, x] = sigma
    return p


def world(sr, w, h, mode, hydro=False, physics=None):
    wd = sr.World(w, h, physics=physics or sr.physics(), gravity=False, hydro=hydro, heat=False)
    wd.set_box_mode(mode)
    return wd


def test_the_box_at_known_places_edges_and_corners(gpu_lib):
    sr = gpu_lib
    W = H = 800
    sv = svac32(sr)
    block = [(x, y) for y in range(50, 60) for x in range(100, 200)]
    places = [
        ("empty world", [], [0, 0, 0, 0]),
        ("one cell, middle", [(300, 200)], [288, 192, 24, 24]),
        ("on the 8-grid", [(8, 8)], [0, 0, 24, 24]),
        ("on the 8-grid, inside", [(16, 16)], [8, 8, 24, 24]),
        ("one off the grid", [(7, 7)], [0, 0, 16, 16]),
        ("corner (0, 0)", [(0, 0)], [0, 0, 16, 16]),
        ("far corner", [(W - 1, H - 1)], [W - 16, H - 16, 16, 16]),
        ("left edge", [(0, 200)], [0, 192, 16, 24]),
        ("right edge", [(W - 1, 200)], [W - 16, 192, 16, 24]),
        ("top edge", [(300, 0)], [288, 0, 24, 16]),
        ("a 100 x 10 block", block, [88, 40, 120, 32]),
        ("two far corners", [(10, H - 10), (W - 10, 5)], [0, 0, W, H]),
    ]
    wd = world(sr, W, H, sr.BOX_FIT)
    for name, cells, want in places:
        print(f"box: {name}", flush=True)
        start = with_gas(W, H, cells)
        assert oracle_box(start[0].astype(np.float64), sv) == want, name
        wd.upload(start)
        b = wd.read_box()
        assert (rect(b), b.from_step, b.fits) == (want, 0, 1), name
        wd.run(1)  # P8, P9, then step 0's refit
        b = wd.read_box()
        assert rect(b) == oracle_box(wd.readback_sigma(), sv) == want, name
        assert (b.from_step, b.fits) == (1, 2), name
        assert list(b.tiles) == [b.w // 8, b.h // 8, 1] and list(b.cells) == [(b.w * b.h + 255) // 256, 1, 1]
    # Sigma_vac itself counts as matter, the next f32 below it does not
    p = with_gas(W, H, [])
    p[0, 40, 40] = np.float32(sv)
    p[0, 300, 500] = np.nextafter(np.float32(sv), np.float32(0))
    wd.upload(p)
    assert rect(wd.read_box()) == oracle_box(p[0].astype(np.float64), sv) == [32, 32, 24, 24]


@pytest.mark.parametrize("wh,cells,want", [
    ((600, 400), [], [32, 32]),
    ((600, 400), [(4, 4)], [32, 32]),
    ((600, 400), [(300, 200)], [64, 64]),
    ((600, 400), [(0, 0), (599, 399)], [2048, 1024]),
    ((2048, 64), [(0, 0), (2047, 63)], [4096, 128]),
])
def test_the_fft_size_per_axis(gpu_lib, wh, cells, want):
    sr = gpu_lib
    w, h = wh
    print(f"box fft {w}x{h} {cells}", flush=True)
    wd = world(sr, w, h, sr.BOX_FIT)
    wd.upload(with_gas(w, h, cells))
    wd.run(1)
    b = wd.read_box()
    assert [fft(b.w), fft(b.h)] == want and list(b.fft) == want


def _box_blob():
    with open(os.path.join(SCENES, "box_blob.json")) as f:
        return scene_json.parse(f.read())


def test_the_box_is_refit_after_p9_at_step_indices_0_mod_16_and_follows_the_gas(gpu_lib):
    """box_blob.json (a 16x16 blob at 0.35 cells per step, bounce walls): stepped one step at a time, the box after
    every step is the rule applied to the state after the last refit step's P9, from that step + 1, one fit more per
    refit; the gas never leaves the box; the box moves with the blob."""
    sr = gpu_lib
    s = _box_blob()
    kw = s.physics_kwargs()
    sv = float(np.float32(sr.physics(**kw).sigma_vac))
    wd = sr.World(s.width, s.height, physics=sr.physics(**kw), gravity=False, hydro=True, heat=False)
    wd.set_box_mode(sr.BOX_FIT)
    wd.upload(s.planes())
    want = oracle_box(wd.readback_sigma(), sv)
    b = wd.read_box()
    assert (rect(b), b.from_step, b.fits) == (want, 0, 1)
    fits, rects = 1, {tuple(want)}
    for k in range(40):
        wd.run(1)
        sig = wd.readback_sigma()
        if k % REFIT == 0:
            want, fits = oracle_box(sig, sv), fits + 1
            rects.add(tuple(want))
        b = wd.read_box()
        assert (rect(b), b.from_step, b.fits) == (want, k // REFIT * REFIT + 1, fits), f"after step {k}"
        ys, xs = np.nonzero(sig >= sv)
        assert xs.min() >= b.x0 and xs.max() < b.x0 + b.w and ys.min() >= b.y0 and ys.max() < b.y0 + b.h, k
    print(f"boxes over 40 steps: {sorted(rects)}", flush=True)
    assert len(rects) >= 2


def test_p8_and_p9_run_over_the_box(gpu_lib):
    """Iron gas on x 100..120, y 50..60 of a 256 x 128 world, every cell of its box iron, hydrogen vacuum outside
    (box.rs:241-283): the box [88, 40, 40, 32]; dt_0 over the box is the dt of a world made of the box's cells alone
    (bit for bit) and differs from the whole world's; P8 floors the vacuum inside the box and leaves the cells
    outside as uploaded (Sigma 0), where the whole-world switch floors them all."""
    sr = gpu_lib
    W, H = 256, 128
    cells = [(x, y) for y in range(50, 60) for x in range(100, 120)]
    start = with_gas(W, H, cells)
    x0, y0, bw, bh = oracle_box(start[0].astype(np.float64), svac32(sr))
    assert [x0, y0, bw, bh] == [88, 40, 40, 32]
    start[4, y0:y0 + bh, x0:x0 + bw] = 0.0
    start[12, y0:y0 + bh, x0:x0 + bw] = 1.0  # x_Fe (channels: x_H, x_He, x_C, x_O, x_Ne, x_Mg, x_Si, x_S, x_Fe, x_n)
    crop = sr.World(bw, bh, physics=sr.physics(), gravity=False, hydro=False, heat=False)
    crop.upload(np.ascontiguousarray(start[:, y0:y0 + bh, x0:x0 + bw]))
    crop.run(1)
    dt_crop = crop.read_dt().dt
    out = {}
    for mode in (sr.BOX_FIT, sr.BOX_WHOLE_WORLD):
        wd = world(sr, W, H, mode)
        wd.upload(start)
        wd.run(1)
        b = wd.read_box()
        want = [x0, y0, bw, bh] if mode == sr.BOX_FIT else [0, 0, W, H]
        assert rect(b) == want
        assert list(b.tiles) == [b.w // 8, b.h // 8, 1] and list(b.cells) == [(b.w * b.h + 255) // 256, 1, 1]
        out[mode] = (wd.read_dt().dt, wd.readback())
    assert out[sr.BOX_FIT][0] == dt_crop and out[sr.BOX_WHOLE_WORLD][0] != dt_crop
    floor = np.float32(sr.physics().sigma_floor)
    vac = start[0] < np.float32(sr.physics().sigma_vac)
    inside = np.zeros((H, W), bool)
    inside[y0:y0 + bh, x0:x0 + bw] = True
    fit, whole = out[sr.BOX_FIT][1], out[sr.BOX_WHOLE_WORLD][1]
    assert np.all(fit[0][vac & inside] == floor) and np.all(fit[0][vac & ~inside] == 0.0)
    assert np.all(whole[0][vac] == floor)
    assert np.array_equal(fit[:, ~inside], start[:, ~inside])


def test_the_whole_world_switch_holds_the_box_at_the_world(gpu_lib):
    sr = gpu_lib
    W = H = 800
    for cells in ([], [(0, 0)], [(300, 200)]):
        wd = world(sr, W, H, sr.BOX_WHOLE_WORLD)
        wd.upload(with_gas(W, H, cells))
        for n, frm, fits in ((1, 1, 2), (16, 17, 3)):
            wd.run(n)
            b = wd.read_box()
            assert (rect(b), list(b.fft), list(b.tiles)) == ([0, 0, W, H], [fft(W), fft(H)], [W // 8, H // 8, 1])
            assert (b.from_step, b.fits) == (frm, fits), cells


def test_the_blob_in_the_box_is_the_whole_worlds(gpu_lib):
    """box_blob.json 96 steps in the box and in the whole world: every cell of the box the same bits (the hydro pass
    reads a cell below Sigma_vac as the floor state, so the floors the whole world adds outside the box change no
    flux), and outside it vacuum (uploaded, or floored while an earlier box held it) against the floor state; the
    blob's mass the same; its speed the scene's 0.35 cells per step (0.30-0.40 over each 16-step chunk,
    box.rs:480-505)."""
    sr = gpu_lib
    s = _box_blob()
    kw = s.physics_kwargs()
    sv = float(np.float32(sr.physics(**kw).sigma_vac))

    def centroid(sig):
        gas = sig >= sv
        return (sig[gas] * np.nonzero(gas)[1]).sum() / sig[gas].sum()

    runs = {}
    for mode in (sr.BOX_FIT, sr.BOX_WHOLE_WORLD):
        wd = sr.World(s.width, s.height, physics=sr.physics(**kw), gravity=False, hydro=True, heat=False)
        wd.set_box_mode(mode)
        wd.upload(s.planes())
        x, speeds = centroid(wd.readback_sigma()), []
        for chunk in range(6):
            wd.run(16)
            c = centroid(wd.readback_sigma())
            speeds.append((x - c) / 16.0)
            x = c
        runs[mode] = (wd.readback(), wd.readback_sigma(), wd.read_box(), speeds)
    (pf, sf, b, speeds), (pw, sw, _, _) = runs[sr.BOX_FIT], runs[sr.BOX_WHOLE_WORLD]
    print(f"box {rect(b)}, speeds {speeds}", flush=True)
    inside = np.zeros(sf.shape, bool)
    inside[b.y0:b.y0 + b.h, b.x0:b.x0 + b.w] = True
    assert np.array_equal(sf[inside].view(np.uint64), sw[inside].view(np.uint64))
    assert np.array_equal(pf[:, inside].view(np.uint32), pw[:, inside].view(np.uint32))
    # outside the last box: vacuum, as the scene uploads it (the floor state) or as an earlier box's P8 left it
    floor = np.float32(sr.physics(**kw).sigma_floor)
    out = pf[0][~inside]
    assert np.all((out == 0.0) | (out == floor)) and np.all(pw[0][~inside] == floor)
    assert sf[sf >= sv].sum() == sw[sw >= sv].sum()
    assert all(0.30 <= v <= 0.40 for v in speeds), speeds
