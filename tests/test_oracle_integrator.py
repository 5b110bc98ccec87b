"""Pins the oracle's integrator (camera -> trace -> shade -> NEE/MIS -> shadow) with
closed-form scenes.

* Parallel planes: a large Lambert floor of albedo rho under a large parallel
  emitter of radiance Le. Every path from the floor ends on the emitter, so the
  radiance leaving the floor is exactly rho * Le (up to the ~1e-6 solid angle the
  finite planes miss). Exercises light selection, light pdf, BSDF sampling and both
  MIS weights.
* The same with a GGX metal floor: the answer is Le times the directional albedo
  int f(wo, wi) cos dwi, computed by quadrature of the BSDF.
* A camera looking straight at an emitter sees Le times the cos^4 vignetting weight,
  exactly.
"""
import math

import numpy as np

from oracle import oracle_py as O
from trayracer_amd.scenes import MODEL_LAMBERT, material


def _planes(floor_mat, le=2.5, half=1000.0, fov_deg=6.0, width=8, height=8, spp=512):
    P = lambda *a: np.array(a, np.float32)  # noqa: E731
    A = half
    floor = [np.concatenate([P(-A, 0, -A), P(-A, 0, A), P(A, 0, A)]), np.concatenate([P(-A, 0, -A), P(A, 0, A), P(A, 0, -A)])]
    # emitter at y = 1 with normal -y
    y = 1.0
    em = [np.concatenate([P(-A, y, -A), P(A, y, -A), P(A, y, A)]), np.concatenate([P(-A, y, -A), P(A, y, A), P(-A, y, A)])]
    light = material((0.5, 0.5, 0.5), roughness=1.0, emission=(le, le, le), model=MODEL_LAMBERT)
    return dict(
        positions=np.stack(floor + em).astype(np.float32), normals=None,
        mat_ids=np.array([0, 0, 1, 1], np.uint32), meshes=[(0, 4)],
        materials=np.stack([floor_mat, light]), instances=None,
        camera=dict(position=(0.0, 0.5, 0.0), target=(0.0, 0.0, 0.0), up=(0.0, 0.0, 1.0),
                    fov=math.radians(fov_deg)),
        render=dict(width=width, height=height, spp=spp, max_depth=3, rr_depth=10, seed=5, ray_epsilon=1e-4,
                    clamp=0.0),
    )


def _vignetting(scene):
    """cos^4 weight of every path's camera ray, from the same PCG camera stream."""
    r = scene["render"]
    W, H, spp = r["width"], r["height"], r["spp"]
    n = W * H * spp
    px = np.empty((n, 2), np.uint32)
    u = np.empty((n, 5), np.float32)
    for p in range(n):
        pix, s = p % (W * H), p // (W * H)
        px[p] = (pix % W, pix // W)
        u[p] = O.rng_floats(r["seed"], pix, s, 0xFFFF, 5)
    return O.camera_rays(scene["camera"], W, H, px, u)[:, 6].astype(np.float64)


def _check_mean(L, expect, label, extra=1e-3):
    se = L.std() / math.sqrt(L.size)
    assert abs(L.mean() - expect) < 4 * se + extra * abs(expect), f"{label}: {L.mean()} vs {expect} (se {se})"


def test_planes_lambert_floor():
    rho, le = 0.6, 2.5
    sc = _planes(material((rho, rho, rho), roughness=1.0, model=MODEL_LAMBERT), le=le)
    out = O.Oracle(sc).render()
    vig = _vignetting(sc)
    L = out["path_rgb"][:, 0].astype(np.float64) / vig
    _check_mean(L, rho * le, "lambert planes")


def test_planes_lambert_floor_with_roulette():
    rho, le = 0.45, 1.0
    sc = _planes(material((rho, rho, rho), roughness=1.0, model=MODEL_LAMBERT), le=le)
    sc["render"]["rr_depth"] = 0
    out = O.Oracle(sc).render()
    L = out["path_rgb"][:, 0].astype(np.float64) / _vignetting(sc)
    _check_mean(L, rho * le, "lambert planes + RR")


def _directional_albedo(m, wo, n=800):
    # midpoint quadrature over (mu = cos theta, phi): dw = dmu dphi
    mu = (np.arange(n) + 0.5) / n
    phi = 2 * np.pi * (np.arange(2 * n) + 0.5) / (2 * n)
    MU, PH = np.meshgrid(mu, phi, indexing="ij")
    r = np.sqrt(1 - MU ** 2)
    wi = np.stack([r * np.cos(PH), r * np.sin(PH), MU], axis=-1).reshape(-1, 3).astype(np.float32)
    f, _ = O.bsdf_eval(m, np.tile(wo, (wi.shape[0], 1)), wi)
    return (f.astype(np.float64) * wi[:, 2:3]).sum(axis=0) * (1.0 / n) * (2 * np.pi / (2 * n))


def test_planes_ggx_floor():
    m = material((0.9, 0.6, 0.3), roughness=0.5, metallic=1.0)
    le = 1.5
    sc = _planes(m, le=le, fov_deg=2.0, spp=1024)
    out = O.Oracle(sc).render()
    vig = _vignetting(sc)
    expect = le * _directional_albedo(m, np.array([0, 0, 1], np.float32))
    for c in range(3):
        L = out["path_rgb"][:, c].astype(np.float64) / vig
        _check_mean(L, expect[c], f"ggx planes channel {c}", extra=5e-3)


def test_direct_view_of_emitter_is_exact():
    le = 3.0
    sc = _planes(material((0.5, 0.5, 0.5), roughness=1.0, model=MODEL_LAMBERT), le=le, spp=4)
    sc["camera"] = dict(position=(0.0, 0.5, 0.0), target=(0.0, 1.0, 0.0), up=(0.0, 0.0, 1.0), fov=math.radians(30.0))
    sc["render"]["max_depth"] = 1
    out = O.Oracle(sc).render()
    vig = _vignetting(sc).astype(np.float32)
    np.testing.assert_array_equal(out["path_rgb"][:, 0], (vig * np.float32(le)).astype(np.float32))


def test_render_crop_equals_window_of_full_frame():
    """A cropped oracle render traces the crop's paths with the whole film's keys and
    camera: its per-path radiance, images and AOVs equal the same window of a full render
    (the full-size GPU parity tests compare a window of a full-size GPU frame this way)."""
    from oracle import Oracle
    from trayracer_amd import scenes
    sc = scenes.cornell(width=40, height=30, spp=3, max_depth=4, seed=4)
    orc = Oracle(sc)
    full = orc.render()
    x0, y0, w, h = 7, 11, 13, 9
    cut = orc.render(crop=(x0, y0, w, h))
    paths = full["path_rgb"].reshape(3, 30, 40, 3)[:, y0:y0 + h, x0:x0 + w].reshape(-1, 3)
    assert np.array_equal(cut["path_rgb"], paths)
    for k in ("image_f32", "image_f64", "aov"):
        assert np.array_equal(cut[k], full[k][y0:y0 + h, x0:x0 + w]), k
    assert cut["counts"].sum() < full["counts"].sum()


def test_config0_reference_fixture_is_the_oracle_image():
    """The committed FP64 4096-spp Cornell image (tests/golden/config0_ref4096.npz) is the
    oracle's: windows at a corner, at the light and at the metal box re-rendered here."""
    import os

    from oracle import Oracle
    from trayracer_amd import scenes
    d = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "config0_ref4096.npz"))
    sc = scenes.cornell()
    assert (int(d["width"]), int(d["height"]), int(d["max_depth"]), int(d["rr_depth"]), int(d["seed"])) == \
        (256, 256, 4, 3, 0)
    orc = Oracle(sc)
    for x0, y0 in ((0, 0), (126, 230), (150, 60)):
        out = orc.render(render=dict(spp=int(d["spp"])), crop=(x0, y0, 3, 3))
        want = d["mean"][y0:y0 + 3, x0:x0 + 3]
        assert np.array_equal(out["image_f64"].astype(np.float32), want), (x0, y0)


def test_last_vertex_light_sample_has_full_weight():
    """With the balance heuristic, a light sample's weight pl / (pl + pb) leaves the rest
    to the BSDF sample that the extension ray traces; the last vertex traces none, so its
    light sample must carry full weight: then MIS equals light sampling alone in
    expectation at any max depth (checked at depth 2 on the Cornell box, where the last
    bounce's light is a large part of the image)."""
    from oracle import Oracle
    from trayracer_amd import scenes
    sc = scenes.cornell(width=8, height=8, spp=4096, max_depth=2, rr_depth=9, seed=2)
    orc = Oracle(sc)
    res = []
    for strategy in (0, 1):
        L = orc.render(render=dict(strategy=strategy))["path_rgb"].astype(np.float64).sum(axis=1)
        res.append((L.mean(), L.std() / math.sqrt(L.size)))
    (m0, s0), (m1, s1) = res
    assert abs(m0 - m1) < 4 * math.hypot(s0, s1), res
