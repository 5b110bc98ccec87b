"""Error paths of the C ABI: every refusal is a status code plus a message
(trt_last_error), before anything is launched, and the context stays usable. The
reference instead reports failures in a message box and exits
(Tracer/Utility/Errors.cpp:15-37, Tracer/CUDA/CudaError.cpp:13-25)."""
import copy

import numpy as np
import pytest

from trayracer_amd import Renderer, scenes
from trayracer_amd._lib import TrtError

pytestmark = pytest.mark.gpu

INVALID_ARG, OUT_OF_MEMORY, NOT_BUILT, UNSUPPORTED = 1, 2, 4, 5


@pytest.fixture(scope="module")
def rend():
    r = Renderer(0)
    yield r
    r.close()


def _status(exc):
    return exc.value.status


def test_render_before_build_and_camera(rend):
    fresh = Renderer(0)
    try:
        with pytest.raises(TrtError, match="not built") as e:
            fresh.RenderFrame()
        assert _status(e) == NOT_BUILT
    finally:
        fresh.close()


def test_material_id_out_of_range(rend):
    sc = copy.deepcopy(scenes.cornell(width=16, height=16, spp=1))
    sc["mat_ids"] = sc["mat_ids"].copy()
    sc["mat_ids"][7] = 99
    with pytest.raises(TrtError, match="material id out of range at triangle 7") as e:
        rend.UpdateScene(sc)
    assert _status(e) == INVALID_ARG
    rend.UpdateScene(scenes.cornell(width=16, height=16, spp=1))  # usable afterwards
    rend.RenderFrame()


def test_path_count_limit(rend):
    rend.UpdateScene(scenes.cornell(width=16, height=16, spp=1))
    s = rend.settings
    s.width, s.height, s.spp = 32768, 32768, 1  # 2^30 paths
    with pytest.raises(TrtError, match="2\\^30") as e:
        rend.RenderFrame()
    assert _status(e) == INVALID_ARG
    s.width, s.height, s.spp = 16, 16, 1
    rend.RenderFrame()


def test_invalid_render_options(rend):
    rend.UpdateScene(scenes.cornell(width=16, height=16, spp=1))
    s = rend.settings
    for field, value, msg in (("max_depth", 0, "positive"), ("max_depth", 40, "< 32"), ("spp", 0, "positive"),
                              ("ray_epsilon", -1.0, "epsilon"), ("mode", 99, "render mode")):
        old = getattr(s, field)
        setattr(s, field, value)
        with pytest.raises(TrtError, match=msg):
            rend.RenderFrame()
        setattr(s, field, old)
    rend.RenderFrame()


def test_scene_without_emitters_renders_black(rend):
    """No emitter and no sky: next-event estimation is off and every path carries zero
    radiance (nothing to refuse); AOVs are still written."""
    sc = copy.deepcopy(scenes.cornell(width=24, height=24, spp=2))
    sc["materials"] = sc["materials"].copy()
    sc["materials"][:, 4:7] = 0.0
    rend.UpdateScene(sc)
    st = rend.RenderFrame()
    assert st["shadow_rays"] == 0
    assert not rend.path_radiance().any()
    assert rend.albedo()[..., :3].max() > 0


def test_negative_tmin_is_rejected(rend):
    rend.UpdateScene(scenes.cornell(width=16, height=16, spp=1))
    rays = np.zeros((4, 8), np.float32)
    rays[:, 0:3] = 0.5
    rays[:, 4] = 1.0
    rays[:, 7] = 1e30
    rays[2, 3] = -1.0
    with pytest.raises(TrtError, match="ray 2 has a negative tmin"):
        rend.trace_closest(rays)
    with pytest.raises(TrtError, match="ray 2 has a negative tmin"):
        rend.trace_any(rays)


def test_denoiser_errors(rend):
    fresh = Renderer(0)
    try:
        with pytest.raises(TrtError, match="not initialised") as e:
            fresh.denoise()
        assert _status(e) == NOT_BUILT
    finally:
        fresh.close()
    rend.UpdateScene(scenes.cornell(width=200, height=64, spp=1))
    rend.RenderFrame()
    rend.denoiser_init(seed=1)
    with pytest.raises(TrtError, match="width % 128") as e:
        rend.denoise()
    assert _status(e) == UNSUPPORTED
    with pytest.raises(TrtError, match="num_params"):
        rend.denoiser_set_params(np.zeros(10, np.float32))


def test_out_of_memory_is_reported(rend):
    """A device allocation that cannot succeed (a denoiser training set of ~550 GB, well
    beyond the GPU's memory) maps to TRT_OUT_OF_MEMORY, and the context keeps working."""
    rend.UpdateScene(scenes.cornell(width=128, height=128, spp=1))
    rend.RenderFrame()
    rend.denoiser_init(seed=1)
    with pytest.raises(TrtError) as e:
        rend.denoiser_dataset(1000, 4096, 4096)
    assert _status(e) == OUT_OF_MEMORY
    rend.denoiser_dataset(1, 128, 128)
    rend.RenderFrame()
    rend.denoiser_capture(0, 0)
