"""State dumps in the reference's SRDUMP01 format (crates/sr-engine/src/dump.rs:1-228): the eight bytes `SRDUMP01`,
a little-endian u32 header length, a UTF-8 JSON header (version 1, width, height, step, sim_time, channels, sinks,
resume), then the fourteen canonical f32 planes, row-major; then, as the header's `resume` flags say, F's two planes
(`flux`), P5's rate plane (`rate`) and the hold's mask plane (`held`, 1 held, 0 free).

read() refuses, naming the cause, what the reference's Dump::from_bytes refuses (dump.rs:146-330): a wrong magic, a
file cut short or with bytes past its last plane, a header that is not UTF-8 JSON, an unknown or missing key, another
version, a world size outside multiples of 8 in [8, 4096], a channel list other than the canonical one (the names of
state.rs:29-31, in that order), more than 16 sinks, a flag the planes disagree with, a rate without F, a mask value
other than 0 or 1. write() makes a dump this module and the reference's reader accept.

Beside a dump, write() can leave `<dump>.sigma64`: Sigma's plane in FP64. The library carries Sigma in FP64 (mass is
conserved to FP64 round-off), and the dump's f32 plane is that value rounded; a resume from the dump plus this file
goes on bit for bit, where one from the dump alone starts from the rounded Sigma. The reference's reader never sees
the extra file. Its layout: the eight bytes `SRSIG64\\0`, width and height as little-endian u32, then w h
little-endian f64, row-major.
"""
from __future__ import annotations

import json
import os
import struct

import numpy as np

MAGIC = b"SRDUMP01"
SIGMA_MAGIC = b"SRSIG64\0"
# the state's channel names in their order (crates/sr-engine/src/state.rs:29-31)
CHANNELS = ["sigma", "mom_x", "mom_y", "energy", "x_H", "x_He", "x_C", "x_O", "x_Ne", "x_Mg", "x_Si", "x_S", "x_Fe",
            "x_n"]
MAX_SINKS = 16
HEADER_KEYS = {"version", "width", "height", "step", "sim_time", "channels", "sinks", "resume"}
RESUME_KEYS = {"dt", "dt_next", "latch_since", "box", "box_args", "flux"}
RESUME_OPTIONAL = {"rate", "held", "mass_carry"}


class DumpError(ValueError):
    pass


def _keys(obj, at, keys, optional=frozenset()):
    if not isinstance(obj, dict):
        raise DumpError(f"{at}: not an object")
    for k in obj:
        if k not in keys and k not in optional:
            raise DumpError(f"{at}.{k}: unknown key")
    for k in sorted(keys):
        if k not in obj:
            raise DumpError(f"{at}.{k}: missing")
    return obj


def _whole(v, at):
    if isinstance(v, bool) or not isinstance(v, int) or v < 0:
        raise DumpError(f"{at}: not a whole number: {v!r}")
    return v


def _number(v, at):
    if isinstance(v, bool) or not isinstance(v, (int, float)) or not np.isfinite(v):
        raise DumpError(f"{at}: not a finite number: {v!r}")
    return float(v)


def _flag(v, at):
    if not isinstance(v, bool):
        raise DumpError(f"{at}: not true or false: {v!r}")
    return v


def read(path: str) -> dict:
    with open(path, "rb") as f:
        out = from_bytes(f.read())
    out["sigma64"] = read_sigma64(path + ".sigma64", out["width"], out["height"])
    return out


def from_bytes(b: bytes) -> dict:
    if len(b) < 8:
        raise DumpError(f"truncated: {len(b)} bytes, the magic alone is 8")
    if b[:8] != MAGIC:
        raise DumpError(f"not a state dump: magic {b[:8]!r}, expected \"SRDUMP01\" (wrong magic)")
    if len(b) < 12:
        raise DumpError("truncated in the header length")
    n = struct.unpack("<I", b[8:12])[0]
    if 12 + n > len(b):
        raise DumpError(f"truncated in the header: {n} bytes announced, {len(b) - 12} present")
    try:
        h = json.loads(b[12:12 + n].decode("utf-8"))
    except (UnicodeDecodeError, json.JSONDecodeError) as e:
        raise DumpError(f"header: not UTF-8 JSON: {e}")
    _keys(h, "header", HEADER_KEYS)
    if _whole(h["version"], "header.version") != 1:
        raise DumpError(f"header.version: {h['version']}, this build reads 1")
    w, hh = _whole(h["width"], "header.width"), _whole(h["height"], "header.height")
    if not (8 <= w <= 4096 and 8 <= hh <= 4096 and w % 8 == 0 and hh % 8 == 0):
        raise DumpError(f"header: world {w} x {hh} is not a world size (multiples of 8 in [8, 4096])")
    if h["channels"] != CHANNELS:
        raise DumpError(f"header.channels: {h['channels']}, expected {CHANNELS} in that order")
    step = _whole(h["step"], "header.step")
    sim_time = _number(h["sim_time"], "header.sim_time")
    sinks = h["sinks"]
    if not isinstance(sinks, list):
        raise DumpError("header.sinks: not a list")
    if len(sinks) > MAX_SINKS:
        raise DumpError(f"header.sinks: {len(sinks)} sinks, at most {MAX_SINKS}")
    for i, s in enumerate(sinks):
        _keys(s, f"header.sinks[{i}]", {"x", "y", "vx", "vy", "m"})
        for k in ("x", "y", "vx", "vy", "m"):
            _number(s[k], f"header.sinks[{i}].{k}")
    r = _keys(h["resume"], "header.resume", RESUME_KEYS, RESUME_OPTIONAL)
    dt, dt_next = _number(r["dt"], "header.resume.dt"), _number(r["dt_next"], "header.resume.dt_next")
    _number(r["latch_since"], "header.resume.latch_since")
    for k, m in (("box", 8), ("box_args", 6)):
        v = r[k]
        if not (isinstance(v, list) and len(v) == m and all(not isinstance(x, bool) and isinstance(x, int)
                                                           and 0 <= x < 2 ** 32 for x in v)):
            raise DumpError(f"header.resume.{k}: not {m} whole numbers below 2^32: {v!r}")
    flux = _flag(r["flux"], "header.resume.flux")
    rate = _flag(r.get("rate", False), "header.resume.rate")
    held = _flag(r.get("held", False), "header.resume.held")
    if "mass_carry" in r:
        _number(r["mass_carry"], "header.resume.mass_carry")
    if rate and not flux:
        raise DumpError("header.resume.rate: true without F (flux false): the rate is P5's")
    cells = w * hh
    planes = 14 + (2 if flux else 0) + (1 if rate else 0) + (1 if held else 0)
    body = b[12 + n:]
    if len(body) != 4 * cells * planes:
        raise DumpError(f"the planes hold {len(body)} bytes, the header says {4 * cells * planes}")
    data = np.frombuffer(body, dtype="<f4").reshape(planes, hh, w)
    out = {"header": h, "width": w, "height": hh, "step": step, "sim_time": sim_time, "dt": dt, "dt_next": dt_next,
           "planes": data[:14].astype(np.float32), "box": list(r["box"]), "box_args": list(r["box_args"]),
           "sinks": [tuple(float(s[k]) for k in ("x", "y", "vx", "vy", "m")) for s in sinks]}
    k = 14
    out["flux"] = data[k:k + 2].astype(np.float32) if flux else None
    k += 2 if flux else 0
    out["rate"] = data[k].astype(np.float32) if rate else None
    k += 1 if rate else 0
    if held:
        m = data[k]
        if not np.all((m == 0.0) | (m == 1.0)):
            raise DumpError("header.resume.held: a mask value other than 0 or 1")
        out["held"] = m.astype(np.uint32)
    else:
        out["held"] = None
    return out


def whole_world_box(w: int, h: int):
    """The box record and its indirect sizes of a world without an active box, as the reference's twin writes them
    (headless.rs:635-651): the whole world, its FFT size per axis, no step, no fit."""
    fft = lambda n: max(32, 1 << (2 * n - 1).bit_length())
    return [0, 0, w, h, fft(w), fft(h), 0, 0], [-(-w // 8), -(-h // 8), 1, -(-(w * h) // 256), 1, 1]


def to_bytes(planes: np.ndarray, step: int = 0, sim_time: float = 0.0, dt: float = 0.0, dt_next: float = 0.0,
             flux: np.ndarray | None = None, box=None, box_args=None, sinks=()) -> bytes:
    """The dump's bytes. dt and dt_next are written at full f64 (this library's clock is FP64; the reference reads them
    as f32), sim_time at full f64 as the reference writes it. `box` (8 words: x0, y0, w, h, fft_w, fft_h, from_step,
    fits) and `box_args` (6: the 8 x 8 tiles, then the 256-cell groups over it) are the active box's (World.read_box);
    without them the whole world's record. `sinks`: P3's sinks, (x, y, vx, vy, m) each (dump.rs:87-99)."""
    p = np.ascontiguousarray(planes, dtype="<f4")
    if p.ndim != 3 or p.shape[0] != 14:
        raise DumpError("planes: (14, h, w) expected")
    _, hh, w = p.shape
    if flux is not None and np.shape(flux) != (2, hh, w):
        raise DumpError(f"flux: (2, {hh}, {w}) expected")
    if len(sinks) > MAX_SINKS or any(len(sk) != 5 for sk in sinks):
        raise DumpError(f"sinks: at most {MAX_SINKS}, (x, y, vx, vy, m) each")
    if box is None:
        box, box_args = whole_world_box(w, hh)
    box, box_args = [int(v) for v in box], [int(v) for v in box_args]
    if len(box) != 8 or len(box_args) != 6 or not all(0 <= v < 2 ** 32 for v in box + box_args):
        raise DumpError("box: 8 words and box_args: 6 words below 2^32 expected")
    header = {"version": 1, "width": w, "height": hh, "step": int(step), "sim_time": float(sim_time),
              "channels": CHANNELS,
              "sinks": [dict(zip(("x", "y", "vx", "vy", "m"), (float(v) for v in sk))) for sk in sinks],
              "resume": {"dt": float(dt), "dt_next": float(dt_next), "latch_since": 0.0,
                         "box": box, "box_args": box_args, "flux": flux is not None, "rate": False}}
    hb = json.dumps(header).encode("utf-8")
    parts = [MAGIC, struct.pack("<I", len(hb)), hb, p.tobytes()]
    if flux is not None:
        parts.append(np.ascontiguousarray(flux, dtype="<f4").tobytes())
    return b"".join(parts)


def write(path: str, planes: np.ndarray, sigma64: np.ndarray | None = None, **kw):
    """Writes the dump (to_bytes' keywords) and, with `sigma64`, Sigma's FP64 plane beside it."""
    with open(path, "wb") as f:
        f.write(to_bytes(planes, **kw))
    side = path + ".sigma64"
    if sigma64 is not None:
        s = np.ascontiguousarray(sigma64, dtype="<f8")
        if s.shape != planes.shape[1:]:
            raise DumpError(f"sigma64: {planes.shape[1:]} expected, got {s.shape}")
        with open(side, "wb") as f:
            f.write(SIGMA_MAGIC + struct.pack("<II", s.shape[1], s.shape[0]) + s.tobytes())
    elif os.path.exists(side):
        os.remove(side)


def read_sigma64(path: str, w: int, h: int):
    """Sigma's FP64 plane from `path` when it is there and of the world (w, h); None otherwise."""
    if not os.path.exists(path):
        return None
    with open(path, "rb") as f:
        b = f.read()
    if len(b) != 16 + 8 * w * h or b[:8] != SIGMA_MAGIC or struct.unpack("<II", b[8:16]) != (w, h):
        raise DumpError(f"{path}: not the FP64 Sigma plane of a {w} x {h} world")
    return np.frombuffer(b[16:], dtype="<f8").reshape(h, w).astype(np.float64)
