"""GPU: a scene file (reference JSON format, OBJ models) loaded by trt_load_scene_file
renders the same per-path radiance as the oracle on the same file (parsed on the host by
trt_scene_file_read), and as the same scene uploaded through UpdateScene."""
import pathlib

import numpy as np
import pytest

from trayracer_amd import Renderer, scenes

pytestmark = pytest.mark.gpu
FIX = pathlib.Path(__file__).resolve().parent / "golden" / "scenes" / "box_scene.json"


def test_scene_file_matches_oracle_and_update_scene():
    from oracle import Oracle
    r = Renderer(0)
    try:
        info = r.LoadScene(str(FIX), presplit=1.0)
        assert (info["n_instances"], info["n_triangles"]) == (4, 32)
        s = r.settings
        assert (s.spp, s.max_depth) == (4, 5)  # from the file's renderer section
        s.width, s.height, s.seed, s.rr_depth = 48, 32, 3, 3
        r.RenderFrame()
        loaded = r.path_radiance()
        assert r.bvh_info()["two_level"] == 1

        sc, _, _ = scenes.from_file(str(FIX), width=48, height=32, rr_depth=3, seed=3)
        ref = Oracle(sc).render()["path_rgb"]
        rel = np.abs(loaded - ref) / np.maximum(np.abs(ref), 1e-6)
        assert (rel <= 1e-4).all(), f"{(rel > 1e-4).any(axis=1).sum()} paths differ from the oracle"
        assert np.abs(ref).sum() > 0  # the light is visible

        r.UpdateScene(sc, presplit=0.0)
        r.RenderFrame()
        assert np.array_equal(r.path_radiance().view(np.uint32), loaded.view(np.uint32))
    finally:
        r.close()
