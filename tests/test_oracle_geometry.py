"""Pins the oracle's ray/triangle test, BVH traversal and camera model.

* Watertight ray/triangle (Woop et al. 2013) against an independent float64
  Moller-Trumbore in numpy, away from edges; watertightness on shared edges.
* The oracle's BVH closest-hit and any-hit equal brute force over every triangle
  (same triangle test), which is what makes prim ids comparable across structures.
* The camera against the closed forms of the reference model
  (Tracer/Optix/GPU/optix.cu:142-174): centre ray = forward, horizontal field of view.
"""
import math

import numpy as np
import pytest

from oracle import oracle_py as O
from tests.gpu_util import prim_of, random_rays
from trayracer_amd import scenes


def _mt64(rays, tris):
    o, d = rays[:, 0:3].astype(np.float64), rays[:, 4:7].astype(np.float64)
    v0, v1, v2 = (tris[:, 3 * k:3 * k + 3].astype(np.float64) for k in range(3))
    e1, e2 = v1 - v0, v2 - v0
    p = np.cross(d, e2)
    det = np.einsum("ij,ij->i", e1, p)
    inv = 1.0 / det
    s = o - v0
    u = np.einsum("ij,ij->i", s, p) * inv
    q = np.cross(s, e1)
    v = np.einsum("ij,ij->i", d, q) * inv
    t = np.einsum("ij,ij->i", e2, q) * inv
    hit = (u >= 0) & (v >= 0) & (u + v <= 1) & (t >= rays[:, 3]) & (t <= rays[:, 7])
    return hit, t, u, v


def test_triangle_test_against_fp64():
    rng = np.random.default_rng(0)
    n = 200000
    tris = rng.uniform(-1, 1, (n, 9)).astype(np.float32)
    rays = random_rays(n, seed=1, lo=-2, hi=2, toward=(-0.7, 0.7), tmin=0.0)
    out = O.intersect(rays, tris)
    hit, t, u, v = _mt64(rays, tris)
    margin = np.minimum(np.minimum(u, v), 1 - u - v)
    clear = np.abs(margin) > 1e-4
    assert np.array_equal(out[clear, 0] > 0, hit[clear])
    h = clear & hit
    assert h.sum() > 10000
    np.testing.assert_allclose(out[h, 1], t[h], rtol=2e-5, atol=1e-6)
    np.testing.assert_allclose(out[h, 2], u[h], atol=2e-4)
    np.testing.assert_allclose(out[h, 3], v[h], atol=2e-4)


def test_shared_edge_is_watertight():
    # planar quad (parallelogram) split along its diagonal; rays through points of the
    # shared edge must hit one of the two triangles (no crack between them)
    rng = np.random.default_rng(2)
    a0, a1, a2 = np.array([0.1, 0.2, 0.3]), np.array([1.3, 0.25, 0.4]), np.array([1.1, 1.4, 0.2])
    a = a0, a1, a2, a0 + a2 - a1
    t1 = np.concatenate([a[0], a[1], a[2]]).astype(np.float32)
    t2 = np.concatenate([a[0], a[2], a[3]]).astype(np.float32)
    n = 50000
    s = rng.random(n)
    p = a[0][None] + s[:, None] * (a[2] - a[0])[None]
    o = p + rng.normal(size=(n, 3)) * 2.0
    d = p - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.zeros((n, 8), np.float32)
    rays[:, 0:3], rays[:, 4:7], rays[:, 7] = o, d, 1e30
    h1 = O.intersect(rays, np.tile(t1, (n, 1)))[:, 0] > 0
    h2 = O.intersect(rays, np.tile(t2, (n, 1)))[:, 0] > 0
    assert (h1 | h2).all(), f"{(~(h1 | h2)).sum()} rays leak through the shared edge"


def test_bvh_equals_bruteforce():
    sc = scenes.soup(n_tris=20000, seed=7, width=8, height=8)
    orc = O.Oracle(sc)
    rays = random_rays(20000, seed=3, lo=-0.2, hi=1.2, tmin=1e-5)
    hb, ib = orc.trace_closest(rays)
    hf, i_f = orc.trace_closest(rays, brute=True)
    assert np.array_equal(prim_of(hb), prim_of(hf))
    assert np.array_equal(hb.view(np.uint32), hf.view(np.uint32))
    rays[:, 7] = np.random.default_rng(5).uniform(0.001, 0.2, rays.shape[0])
    assert np.array_equal(orc.trace_any(rays), orc.trace_any(rays, brute=True))


def test_bvh_equals_bruteforce_instanced():
    sc = scenes.instanced(n_instances=12, subdiv=1, seed=3, width=8, height=8, emissive_per_instance=2)
    orc = O.Oracle(sc)
    rays = random_rays(10000, seed=4, lo=-5, hi=105, toward=(0, 100), tmin=1e-4)
    hb, ib = orc.trace_closest(rays)
    hf, i_f = orc.trace_closest(rays, brute=True)
    assert np.array_equal(prim_of(hb), prim_of(hf)) and np.array_equal(ib, i_f)
    assert (prim_of(hb) != 0xFFFFFFFF).sum() > 10


def test_closest_hit_matches_fp64_nearest():
    # closest hit of the oracle = nearest FP64 intersection (up to near-ties)
    sc = scenes.soup(n_tris=3000, seed=9, width=8, height=8)
    orc = O.Oracle(sc)
    rays = random_rays(2000, seed=6, lo=-0.2, hi=1.2, tmin=1e-5)
    h, _ = orc.trace_closest(rays, brute=True)
    tris = sc["positions"]
    best_t = np.full(rays.shape[0], np.inf)
    best_p = np.full(rays.shape[0], -1)
    for k in range(0, tris.shape[0]):
        hit, t, u, v = _mt64(rays, np.repeat(tris[k:k + 1], rays.shape[0], axis=0))
        upd = hit & (t < best_t)
        best_t[upd], best_p[upd] = t[upd], k
    got = prim_of(h).astype(np.int64)
    got[got == 0xFFFFFFFF] = -1
    agree = got == best_p
    # disagreements only where FP64 itself sees a near-tie or an edge graze
    assert agree.mean() > 0.995


def test_camera_closed_forms():
    cam = dict(position=(1.0, 2.0, 3.0), target=(1.0, 2.0, 10.0), up=(0, 1, 0), fov=math.radians(60.0))
    # 1x1 film, jitter 0.5: screen (0,0) -> forward
    r = O.camera_rays(cam, 1, 1, [[0, 0]], [[0.5, 0.5, 0.5, 0.5, 0.0]])[0]
    np.testing.assert_allclose(r[3:6], [0, 0, 1], atol=1e-7)
    np.testing.assert_allclose(r[0:3], cam["position"], atol=1e-7)
    assert r[6] == 1.0
    # right edge of the film (x + jitter = W): angle to forward = fov/2 (horizontal)
    W, H = 64, 32
    r = O.camera_rays(cam, W, H, [[W - 1, H // 2]], [[1.0, 0.0, 0.5, 0.5, 0.0]])[0]
    ang = math.degrees(math.acos(r[5]))
    assert abs(ang - 30.0) < 1e-4
    # side = cross(forward, up) = -x for forward +z, up +y (reference CameraNode::Transform)
    assert r[3] < 0
    # vignetting weight = cos^4 of the angle to forward
    assert abs(r[6] - math.cos(math.radians(30.0)) ** 4) < 1e-6
    # top edge: y + jitter = H -> screen.y = 1/aspect
    r = O.camera_rays(cam, W, H, [[W // 2, H - 1]], [[0.0, 1.0, 0.5, 0.5, 0.0]])[0]
    assert abs(r[4] / r[5] - math.tan(math.radians(30.0)) / (W / H)) < 1e-6


def test_camera_thin_lens_focus():
    # all lens samples of one pixel converge on the focal plane
    cam = dict(position=(0.0, 0.0, 0.0), target=(0.0, 0.0, 1.0), up=(0, 1, 0), fov=math.radians(40.0),
               aperture=0.1, focal_dist=3.0, bokeh_sides=6, bokeh_rotation=0.3)
    rng = np.random.default_rng(1)
    n = 200
    u = np.concatenate([np.full((n, 2), 0.25), rng.random((n, 3))], axis=1)
    r = O.camera_rays(cam, 16, 16, np.tile([[5, 9]], (n, 1)), u)
    o, d = r[:, 0:3].astype(np.float64), r[:, 3:6].astype(np.float64)
    p = o + d * ((3.0 - o[:, 2]) / d[:, 2])[:, None]
    assert np.ptp(p[:, 0]) < 1e-5 and np.ptp(p[:, 1]) < 1e-5
    # lens points stay inside the hexagon of circumradius `aperture`
    assert np.max(np.linalg.norm(o[:, 0:2], axis=1)) <= 0.1 + 1e-6


@pytest.mark.parametrize("distortion", [0.0, 0.5, 1.2])
def test_camera_barrel_distortion_closed_form(distortion):
    """Barrel distortion (Tracer/Optix/GPU/optix.cu:117-138): a film point at lens radius
    r = tan(fov/2) x |screen| leaves at the angle arctan(tan(r d k) / d) to the forward axis
    (k = atan(t d / 2) / (t d / 2), t = tan(fov/2)). At the horizontal film edge (r = t)
    the double-angle identity gives tan(angle) = t / (1 - (t d / 2)^2), independent of the
    formula's evaluation order; without distortion the angle is arctan(r)."""
    fov = math.radians(70.0)
    t = math.tan(fov / 2)
    cam = dict(position=(0.0, 0.0, 0.0), target=(0.0, 0.0, 1.0), up=(0, 1, 0), fov=fov, distortion=distortion)
    W, H = 64, 64
    # right edge of the middle row: screen = (1, 0)
    r = O.camera_rays(cam, W, H, [[W - 1, H // 2]], [[1.0, 0.0, 0.5, 0.5, 0.0]])[0]
    tan_angle = math.hypot(r[3], r[4]) / r[5]
    want = t if distortion == 0.0 else t / (1.0 - (t * distortion / 2) ** 2)
    assert abs(tan_angle - want) < 2e-6 * want
    # interior points: radial scale tan(r d k) / (r d), direction kept
    rng = np.random.default_rng(2)
    px = np.stack([rng.integers(0, W, 200), rng.integers(0, H, 200)], axis=1).astype(np.uint32)
    jit = rng.random((200, 2))
    out = O.camera_rays(cam, W, H, px, np.concatenate([jit, np.full((200, 3), 0.5)], axis=1))
    sx = ((px[:, 0] + jit[:, 0]) / W) * 2 - 1
    sy = ((px[:, 1] + jit[:, 1]) / H) * 2 - 1
    lx, ly = t * sx, t * sy
    rad = np.hypot(lx, ly)
    if distortion > 0:
        d = t * distortion / 2
        k = np.tan(rad * distortion * math.atan(d) / d) / (rad * distortion)
    else:
        k = np.ones_like(rad)
    tx, ty = -out[:, 3] / out[:, 5], out[:, 4] / out[:, 5]  # side = -x for this camera
    np.testing.assert_allclose(np.hypot(tx, ty), rad * k, rtol=2e-5)
    np.testing.assert_allclose(np.arctan2(ty, tx), np.arctan2(ly, lx), atol=2e-5)
