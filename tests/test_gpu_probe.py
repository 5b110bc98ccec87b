"""Measurement hooks of the load-path ceilings (trt_measure_gather, trt_measure_gather_lanes):
they run, and their figures are positive and ordered as the access patterns require (a
pseudo-random record per lane is never cheaper than one record per wave)."""
import ctypes
import math

import pytest

pytestmark = pytest.mark.gpu


def test_gather_probes_report_positive_rates():
    from trayracer_amd import _lib
    lib = _lib.load()
    g = ctypes.c_double()
    assert lib.trt_measure_gather(0, 16 << 10, 64, ctypes.byref(g)) == 0
    assert math.isfinite(g.value) and g.value > 0
    ns = {}
    for active, same in ((64, 0), (8, 0), (64, 1), (8, 1)):
        v = ctypes.c_double()
        assert lib.trt_measure_gather_lanes(0, 16 << 10, 64, active, same, ctypes.byref(v)) == 0
        assert math.isfinite(v.value) and v.value > 0
        ns[(active, same)] = v.value
    print("ns per wave-level 16-byte load per CU:", ns)
    assert ns[(64, 0)] >= 0.9 * ns[(64, 1)]
