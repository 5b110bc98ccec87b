"""The C-ABI library loads without a GPU and exports every entry point declared in
include/trt.h; host-side argument checks return status codes instead of aborting."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            text = open(os.path.join(ROOT, "include", fn)).read()
            names |= set(re.findall(r"^\s*(?:[\w\s\*]+?)\b(trt_\w+)\s*\(", text, re.M))
    return names


def test_header_declares_entry_points():
    names = _declared()
    assert {"trt_create", "trt_render", "trt_trace_closest", "trt_trace_any", "trt_build"} <= names


def test_library_exports_every_declared_symbol():
    from trayracer_amd import _lib
    lib = _lib.load()
    missing = [n for n in sorted(_declared()) if not hasattr(lib, n)]
    assert not missing, missing
    declared_in_binding = {p[0] for p in _lib.PROTOTYPES}
    assert _declared() <= declared_in_binding | {"trt_version"}


def test_null_and_invalid_arguments_return_status():
    from trayracer_amd import _lib
    lib = _lib.load()
    assert lib.trt_create(0, None) == 1  # TRT_INVALID_ARG
    assert lib.trt_render(None, None, None) == 1
    assert lib.trt_set_phase_timing(None, 0) == 1
    assert lib.trt_build(None, None) == 1
    assert lib.trt_trace_closest(None, None, 0, None, None) == 1
    assert lib.trt_last_error(None) == b"null context"
    v = ctypes.c_double()
    assert lib.trt_measure_gather(0, 1 << 14, 1, None) == 1
    assert lib.trt_measure_gather_lanes(0, 1 << 14, 1, 0, 0, ctypes.byref(v)) == 1   # no active lane
    assert lib.trt_measure_gather_lanes(0, 1 << 14, 1, 65, 0, ctypes.byref(v)) == 1  # more lanes than a wave
    assert lib.trt_measure_gather_lanes(0, 1 << 14, 1, 64, 2, ctypes.byref(v)) == 1  # unknown pattern
    assert lib.trt_measure_gather_lanes(0, 16, 1, 64, 0, ctypes.byref(v)) == 1       # not one record
    assert lib.trt_version().startswith(b"trt")
    # no device in this container: creation reports a device error instead of aborting
    ctx = ctypes.c_void_p()
    st = lib.trt_create(0, ctypes.byref(ctx))
    assert st in (0, 3)
    if st == 0:
        lib.trt_destroy(ctx)
