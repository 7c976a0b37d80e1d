"""The C ABI without a GPU: libsrhip.so loads, and exports every function include/sr_hip.h declares.

The declarations are read from the header text (not from the Python binding's list), so a function added to the header
and forgotten in the library fails here. No compute entry point is called: this container has no GPU.
"""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "sr_hip.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", " ", text, flags=re.S)
    text = re.sub(r"//[^\n]*", " ", text)
    names = re.findall(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_]*\s*\**\s*(sr_[a-z0-9_]+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_header_declares_the_boundary():
    names = declared_functions()
    for expected in ["sr_world_create", "sr_upload", "sr_run", "sr_readback", "sr_read_dt", "sr_read_heat",
                     "sr_read_field", "sr_set_precond", "sr_fld_solve", "sr_fft2", "sr_gravity_field",
                     "sr_set_timing", "sr_read_cost", "sr_last_error", "sr_destroy"]:
        assert expected in names
    assert len(names) >= 25


def test_library_exports_every_declared_symbol():
    from sandbox_reactions_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.fail(f"{_lib.LIB_PATH} is not built (run __graft_entry__.build())")
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert missing == []


def test_python_binding_covers_the_header():
    from sandbox_reactions_amd import _lib
    bound = {name for name, *_ in _lib.SIGNATURES}
    assert set(declared_functions()) <= bound


def test_host_only_calls():
    """Calls that touch no device: defaults and the error message."""
    from sandbox_reactions_amd import _lib
    lib = _lib.load()
    ph = _lib.Physics()
    lib.sr_physics_defaults(ctypes.byref(ph))
    assert ph.gamma == 2.0 and ph.G == 1.0 and ph.dx == 1.0
    assert ph.cfl == pytest.approx(0.4) and ph.eta_g == pytest.approx(0.3)
    hs = _lib.HeatSettings()
    lib.sr_heat_settings_defaults(ctypes.byref(hs))
    assert hs.max_passes == 128 and hs.nu1 == 2 and hs.nu2 == 2
    assert isinstance(lib.sr_last_error(), (bytes, type(None)))
