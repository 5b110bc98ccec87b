"""trt_trace_closest_noalloc (the allocation-free closest-hit entry point) against
trt_trace_closest through the C ABI: bit-identical hits on two calls in a row, the second
smaller, on a single-level and an instanced scene; sizes beyond its limit are refused."""
import ctypes

import numpy as np
import pytest

from tests.gpu_util import random_rays
from trayracer_amd import Renderer, _lib, scenes

pytestmark = pytest.mark.gpu


def _noalloc(rend, rays):
    rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
    n = rays.shape[0]
    hits = np.full((n, 4), np.nan, np.float32)
    inst = np.zeros(n, np.uint32)
    st = rend.lib.trt_trace_closest_noalloc(rend.ctx, rays.ctypes.data_as(_lib.c_float_p), n,
                                            hits.ctypes.data_as(_lib.c_float_p), inst.ctypes.data_as(_lib.c_u32_p))
    return st, hits, inst


@pytest.mark.parametrize("scene", ["soup", "instanced"])
def test_noalloc_equals_trace_closest(scene):
    r = Renderer(0)
    try:
        if scene == "soup":
            sc = scenes.soup(n_tris=50_000, seed=3, width=32, height=32)
            mk = lambda n, seed: random_rays(n, seed=seed, lo=-0.3, hi=1.3, tmin=1e-5)  # noqa: E731
        else:
            sc = scenes.instanced(n_instances=40, subdiv=2, seed=4, width=32, height=32, emissive_per_instance=2)
            mk = lambda n, seed: random_rays(n, seed=seed, lo=-10, hi=110, toward=(0, 100), tmin=1e-4)  # noqa: E731
        r.UpdateScene(sc)
        limit = int(_lib.TRT_NOALLOC_MAX_RAYS)
        for n, seed in ((limit, 1), (777, 2)):  # the second call smaller than the first
            rays = mk(n, seed)
            st, h, i = _noalloc(r, rays)
            assert st == 0, r.lib.trt_last_error(r.ctx)
            hr, ir = r.trace_closest(rays)
            assert np.array_equal(h.view(np.uint32), hr.view(np.uint32))
            assert np.array_equal(i, ir)
            assert (h[:, 3].view(np.uint32) != 0xFFFFFFFF).any()
        st, _, _ = _noalloc(r, mk(limit + 1, 3))
        assert st == 1 and b"at most" in r.lib.trt_last_error(r.ctx)
    finally:
        r.close()
