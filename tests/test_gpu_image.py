"""Whole GPU images of the Cornell-box configuration (256x256, max depth 4, Russian
roulette from depth 3, seed 0) scored by rel-MSE = mean (x - r)^2 / (r^2 + 1e-2) against
the oracle's FP64-accumulated 4096-spp image (tests/golden/config0_ref4096.npz, written
by tests/golden/make_config0_reference.py).

* 4096 spp with the reference's own seed, accumulated over 16 progressive frames: the
  same paths as the reference, so the image differs from it only by FP32 accumulation.
* 16 spp (the configuration's own frame) and 4096 spp with another seed: independent
  estimates, whose rel-MSE must equal the variance the reference's per-pixel second
  moment predicts (an unbiased estimator with that variance), within 20 % (the 16-spp
  figure averaged over eight seeds).
"""
import os

import numpy as np
import pytest

from trayracer_amd import Renderer, scenes

pytestmark = pytest.mark.gpu

REF = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "config0_ref4096.npz")


@pytest.fixture(scope="module")
def ref():
    d = np.load(REF)
    m = d["mean"].astype(np.float64)
    var = np.maximum(d["second_moment"].astype(np.float64) - m * m, 0.0)
    return m, var, int(d["spp"])


@pytest.fixture(scope="module")
def rend():
    r = Renderer(0)
    yield r
    r.close()


def _relmse(x, r):
    return float(((x - r) ** 2 / (r * r + 1e-2)).mean())


def _render(rend, spp_total, seed, chunk=256):
    s = rend.settings
    s.seed = seed
    s.spp = min(chunk, spp_total)
    rend.Reset()
    for _ in range(spp_total // s.spp):
        rend.RenderFrame()
    assert rend.SampleCount() == spp_total
    return rend.color()[..., :3].astype(np.float64)


def test_config0_same_samples_equal_fp64_reference(rend, ref):
    m, var, n_ref = ref
    sc = scenes.cornell()
    rend.UpdateScene(sc)
    img = _render(rend, n_ref, seed=sc["render"]["seed"])
    rel = np.abs(img - m) / np.maximum(np.abs(m), 1e-3)
    print("max relative difference to the FP64 mean:", rel.max(), "relMSE:", _relmse(img, m))
    assert rel.max() < 1e-4
    assert _relmse(img, m) < 1e-9


@pytest.mark.parametrize("spp,seeds", [(16, (0,)), (16, tuple(range(1, 9))), (4096, (1,))],
                         ids=["16spp-reference-seed", "16spp-8-seeds", "4096spp-other-seed"])
def test_config0_relmse_matches_predicted_variance(rend, ref, spp, seeds):
    """rel-MSE of independent estimates equals the prediction from the reference's
    per-pixel variance. The per-frame value is heavy-tailed (a rare firefly weighs a lot
    in the expectation), so the 16-spp check averages eight seeds."""
    m, var, n_ref = ref
    sc = scenes.cornell()
    rend.UpdateScene(sc)
    den = m * m + 1e-2
    own = (var / n_ref / den).mean()
    got, expect = [], []
    for seed in seeds:
        img = _render(rend, spp, seed=seed)
        # the seed-0 frame's samples are the reference's first ones (covariance var / n_ref)
        expect.append(float((var / spp / den).mean() + (-own if seed == sc["render"]["seed"] else own)))
        got.append(_relmse(img, m))
    g, e = float(np.mean(got)), float(np.mean(expect))
    print(f"{spp} spp, seeds {seeds}: relMSE {g:.5f} (per seed {np.round(got, 5)}), predicted {e:.5f}")
    assert abs(g / e - 1.0) < 0.2
