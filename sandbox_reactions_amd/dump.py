"""State dumps in the reference's SRDUMP01 format (crates/sr-engine/src/dump.rs:1-228): the eight bytes `SRDUMP01`,
a little-endian u32 header length, a UTF-8 JSON header (version 1, width, height, step, sim_time, channels, sinks,
resume), then the fourteen canonical f32 planes, row-major; then, as the header's `resume` flags say, F's two planes
(`flux`), P5's rate plane (`rate`) and the hold's mask plane (`held`, 1 held, 0 free).

read() refuses, naming the cause: a wrong magic, a file cut short or with bytes past its last plane, a header that is
not JSON, another version, a world size outside multiples of 8 in [8, 4096], a channel list other than the canonical
one, a flag the planes disagree with. write() makes a dump this module and the reference's reader accept.
"""
from __future__ import annotations

import json
import struct

import numpy as np

MAGIC = b"SRDUMP01"
CHANNELS = ["sigma", "mom_x", "mom_y", "energy", "x_h", "x_he", "x_c", "x_o", "x_ne", "x_mg", "x_si", "x_s", "x_fe",
            "x_n"]


class DumpError(ValueError):
    pass


def read(path: str) -> dict:
    with open(path, "rb") as f:
        return from_bytes(f.read())


def from_bytes(b: bytes) -> dict:
    if len(b) < 12 or b[:8] != MAGIC:
        raise DumpError("not an SRDUMP01 file (wrong magic)")
    n = struct.unpack("<I", b[8:12])[0]
    if 12 + n > len(b):
        raise DumpError("the file is cut inside its header")
    try:
        h = json.loads(b[12:12 + n].decode("utf-8"))
    except (UnicodeDecodeError, json.JSONDecodeError) as e:
        raise DumpError(f"the header is not UTF-8 JSON: {e}")
    if h.get("version") != 1:
        raise DumpError(f"version: {h.get('version')} is not 1")
    w, hh = h.get("width"), h.get("height")
    if not (isinstance(w, int) and isinstance(hh, int) and 8 <= w <= 4096 and 8 <= hh <= 4096 and w % 8 == 0
            and hh % 8 == 0):
        raise DumpError(f"world: {w} x {hh} is not a world size")
    if h.get("channels") is not None and len(h["channels"]) != 14:
        raise DumpError("channels: not the 14 canonical channels")
    r = h.get("resume", {})
    cells = w * hh
    planes = 14 + (2 if r.get("flux") else 0) + (1 if r.get("rate") else 0) + (1 if r.get("held") else 0)
    if r.get("rate") and not r.get("flux"):
        raise DumpError("resume.rate without resume.flux")
    body = b[12 + n:]
    if len(body) != 4 * cells * planes:
        raise DumpError(f"the planes hold {len(body)} bytes, the header says {4 * cells * planes}")
    data = np.frombuffer(body, dtype="<f4").reshape(planes, hh, w)
    out = {"header": h, "width": w, "height": hh, "step": h.get("step", 0), "sim_time": h.get("sim_time", 0.0),
           "dt": r.get("dt"), "dt_next": r.get("dt_next"), "planes": data[:14].astype(np.float32)}
    k = 14
    out["flux"] = data[k:k + 2].astype(np.float32) if r.get("flux") else None
    k += 2 if r.get("flux") else 0
    out["rate"] = data[k].astype(np.float32) if r.get("rate") else None
    k += 1 if r.get("rate") else 0
    if r.get("held"):
        m = data[k]
        if not np.all((m == 0.0) | (m == 1.0)):
            raise DumpError("held: a mask value other than 0 or 1")
        out["held"] = m.astype(np.uint32)
    else:
        out["held"] = None
    return out


def to_bytes(planes: np.ndarray, step: int = 0, sim_time: float = 0.0, dt: float = 0.0, dt_next: float = 0.0,
             flux: np.ndarray | None = None) -> bytes:
    p = np.ascontiguousarray(planes, dtype="<f4")
    if p.ndim != 3 or p.shape[0] != 14:
        raise DumpError("planes: (14, h, w) expected")
    _, hh, w = p.shape
    header = {"version": 1, "width": w, "height": hh, "step": int(step), "sim_time": float(sim_time),
              "channels": CHANNELS, "sinks": [],
              "resume": {"dt": float(np.float32(dt)), "dt_next": float(np.float32(dt_next)), "latch_since": 0.0,
                         "box": [0, 0, w, hh, 0, 0, int(step), 0], "box_args": [0, 0, 0, 0, 0, 0],
                         "flux": flux is not None, "rate": False}}
    hb = json.dumps(header).encode("utf-8")
    parts = [MAGIC, struct.pack("<I", len(hb)), hb, p.tobytes()]
    if flux is not None:
        parts.append(np.ascontiguousarray(flux, dtype="<f4").tobytes())
    return b"".join(parts)


def write(path: str, planes: np.ndarray, **kw):
    with open(path, "wb") as f:
        f.write(to_bytes(planes, **kw))
