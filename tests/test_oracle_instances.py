"""Pins the oracle's two-level scene handling with checks that do not run its code.

The GPU and the oracle share the instance transforms (trt/scene.h: invert_affine,
xform_point, xform_vector, xform_normal) and the light table (trt/lights_build.h), so a
GPU-vs-oracle comparison cannot see an error in them. Here numpy restates the geometry
in float64 from the instance matrices alone (3x4, rows applied to column vectors, the
convention the reference's traversal reads, Tracer/Resources/Instance.cpp:39-49; the
reference's host light gathering reads them transposed, SURVEY Q18,
Tracer/Renderer/Scene.cpp:291-294, which a rotated, non-uniformly scaled instance
exposes):

* rays aimed at numpy-computed points of chosen (instance, triangle) pairs, along the
  numpy-computed world normal, must hit exactly those pairs at the expected distance;
* the surface attributes at those hits (position, geometric normal, shading normal from
  vertex normals under the inverse transpose, world-space area) equal numpy's;
* every instanced light record (world vertices, edges, normal, area, power) equals
  numpy's, in (instance, triangle) order;
* on an instanced scene lit by an instanced emissive dome, light sampling alone, BSDF
  sampling alone and both under MIS (the renderer's estimator) converge to the same
  per-pixel means.
"""
import math

import numpy as np

from oracle import Oracle
from trayracer_amd import scenes


def _affine(rng):
    """A rotated, non-uniformly scaled and sheared linear part (R1 S R2)."""
    def rot():
        q = rng.normal(size=4)
        q /= np.linalg.norm(q)
        w, x, y, z = q
        return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                         [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                         [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])
    return rot() @ np.diag(rng.uniform(0.5, 2.0, 3)) @ rot()


def _scene(seed=3, n_inst=12, emissive=(5, 17)):
    rng = np.random.default_rng(seed)
    pos, nrm = scenes.icosphere(2)  # 320 triangles, smooth normals
    T = pos.shape[0]
    mat_ids = np.zeros(T, np.uint32)
    for k, t in enumerate(emissive):
        mat_ids[t] = 1 + k
    pal = 1 + len(emissive)
    mats = np.zeros((n_inst * pal, 17), np.float32)
    inst = np.zeros((n_inst, 14), np.float32)
    for i in range(n_inst):
        M = np.zeros((3, 4))
        M[:, :3] = _affine(rng)
        M[:, 3] = (10.0 * (i % 4), 10.0 * ((i // 4) % 3), 5.0 * (i % 2)) + rng.uniform(-1, 1, 3)
        inst[i, :12] = M.reshape(-1)
        inst[i, 12:14] = np.array([0, i * pal], np.uint32).view(np.float32)
        mats[i * pal] = scenes.material(tuple(rng.uniform(0.2, 0.9, 3)), roughness=0.7)
        for k in range(len(emissive)):
            mats[i * pal + 1 + k] = scenes.material((0.5, 0.5, 0.5), emission=tuple(rng.uniform(1.0, 5.0, 3)))
    cam = dict(position=(15.0, 10.0, -30.0), target=(15.0, 10.0, 2.0), up=(0, 1, 0), fov=math.radians(60.0))
    return dict(positions=pos, normals=nrm, mat_ids=mat_ids, meshes=[(0, T)], materials=mats, instances=inst,
                camera=cam, render=dict(width=32, height=24, spp=4, max_depth=3, rr_depth=9, seed=1,
                                        ray_epsilon=1e-4, clamp=0.0))


def _world(M, p):
    return p @ M[:, :3].T + M[:, 3]


def _rays_at(sc, pairs, dist=0.3):
    """Rays from `dist` outside the world centroid of each (instance, triangle), along
    minus the outward world normal; returns rays and the numpy centroids and normals."""
    pos = sc["positions"].astype(np.float64).reshape(-1, 3, 3)
    rays = np.zeros((len(pairs), 8), np.float32)
    cents, nws = [], []
    for j, (i, k) in enumerate(pairs):
        M = sc["instances"][i, :12].astype(np.float64).reshape(3, 4)
        v = _world(M, pos[k])
        c = v.mean(axis=0)
        n = np.cross(v[1] - v[0], v[2] - v[0])
        n /= np.linalg.norm(n)
        if np.dot(n, c - M[:, 3]) < 0:  # outward: away from the instance's centre
            n = -n
        o = c + dist * n
        rays[j, 0:3], rays[j, 3], rays[j, 4:7], rays[j, 7] = o, 1e-4, -n, 1e30
        cents.append(c)
        nws.append(n)
    return rays, np.array(cents), np.array(nws)


def test_rays_hit_numpy_placed_instance_triangles():
    sc = _scene()
    orc = Oracle(sc)
    rng = np.random.default_rng(5)
    pairs = [(i, int(k)) for i in range(len(sc["instances"])) for k in rng.choice(320, 25, replace=False)]
    rays, cents, _ = _rays_at(sc, pairs)
    hits, inst = orc.trace_closest(rays)
    prim = hits[:, 3].view(np.uint32)
    want_i = np.array([p[0] for p in pairs])
    want_k = np.array([p[1] for p in pairs])
    assert np.array_equal(inst, want_i) and np.array_equal(prim, want_k)
    np.testing.assert_allclose(hits[:, 0], 0.3, rtol=2e-4)
    # the same through the brute-force loop over all instances
    hb, ib = orc.trace_closest(rays, brute=True)
    assert np.array_equal(ib, want_i) and np.array_equal(hb[:, 3].view(np.uint32), want_k)


def test_surface_attributes_at_instanced_hits():
    sc = _scene()
    orc = Oracle(sc)
    rng = np.random.default_rng(6)
    pairs = [(i, int(k)) for i in range(len(sc["instances"])) for k in rng.choice(320, 10, replace=False)]
    rays, cents, nws = _rays_at(sc, pairs)
    hits, _ = orc.trace_closest(rays)
    srf = orc.surface(rays)
    pos = sc["positions"].astype(np.float64).reshape(-1, 3, 3)
    nrm = sc["normals"].astype(np.float64).reshape(-1, 3, 3)
    for j, (i, k) in enumerate(pairs):
        M = sc["instances"][i, :12].astype(np.float64).reshape(3, 4)
        A = M[:, :3]
        b1, b2 = float(hits[j, 1]), float(hits[j, 2])
        v = _world(M, pos[k])
        p = (1 - b1 - b2) * v[0] + b1 * v[1] + b2 * v[2]
        np.testing.assert_allclose(srf["p"][j], p, rtol=0, atol=2e-5 * (1 + np.abs(p).max()))
        np.testing.assert_allclose(p, cents[j], atol=1e-3)  # the aimed-at point
        cw = np.cross(v[1] - v[0], v[2] - v[0])
        np.testing.assert_allclose(srf["ng"][j], cw / np.linalg.norm(cw), atol=1e-5)  # winding side
        assert abs(srf["area"][j] - 0.5 * np.linalg.norm(cw)) <= 1e-5 * srf["area"][j]  # world-space area
        n_obj = (1 - b1 - b2) * nrm[k, 0] + b1 * nrm[k, 1] + b2 * nrm[k, 2]
        ns = np.linalg.inv(A).T @ n_obj  # normals transform by the inverse transpose
        ns /= np.linalg.norm(ns)
        if np.dot(ns, -rays[j, 4:7]) < 0:  # turned to face the incoming ray
            ns = -ns
        np.testing.assert_allclose(srf["ns"][j], ns, atol=2e-5)
        assert abs(np.dot(ns, nws[j])) > 0.5  # smooth normal close to the face normal


def test_instanced_light_records():
    sc = _scene()
    orc = Oracle(sc)
    L, prob, alias, total = orc.lights()
    pos = sc["positions"].astype(np.float64).reshape(-1, 3, 3)
    mats = sc["materials"]
    want = []
    for i, row in enumerate(sc["instances"]):
        M = row[:12].astype(np.float64).reshape(3, 4)
        moff = int(row[13:14].view(np.uint32)[0])
        for k in np.nonzero(sc["mat_ids"])[0]:  # triangle order within the mesh
            le = mats[moff + sc["mat_ids"][k], 4:7].astype(np.float64)
            v = _world(M, pos[k])
            c = np.cross(v[1] - v[0], v[2] - v[0])
            area = 0.5 * np.linalg.norm(c)
            want.append(np.concatenate([v[0], [area], v[1] - v[0], [le.sum() * area], v[2] - v[0], [0],
                                        c / np.linalg.norm(c), [0], le, [0]]))
    want = np.array(want)
    assert L.shape == want.shape == (24, 20)
    cols = [0, 1, 2, 4, 5, 6, 8, 9, 10, 12, 13, 14, 16, 17, 18]
    np.testing.assert_allclose(L[:, cols], want[:, cols], rtol=2e-5, atol=2e-5)
    np.testing.assert_allclose(L[:, [3, 7]], want[:, [3, 7]], rtol=2e-5)  # area, power
    assert abs(total - want[:, 7].sum()) <= 1e-5 * total
    # alias table reproduces the power distribution
    n = len(prob)
    p = prob.astype(np.float64) / n
    for j in range(n):
        p[alias[j]] += (1.0 - prob[j]) / n
    np.testing.assert_allclose(p, want[:, 7] / want[:, 7].sum(), rtol=1e-5, atol=1e-7)


def _dome_scene(seed=9):
    """Twelve diffuse instanced icospheres inside an instanced, non-uniformly scaled
    emissive dome (an icosphere mesh with reversed winding, so it emits inwards): light
    reaches the spheres from 320 emitting triangles, the case where the estimators differ."""
    sc = _scene(seed=seed, emissive=())
    pos, nrm = scenes.icosphere(2)
    dome = pos.reshape(-1, 3, 3)[:, ::-1].reshape(-1, 9).copy()  # reversed winding: normals point inwards
    T = pos.shape[0]
    sc["positions"] = np.concatenate([sc["positions"], dome])
    sc["normals"] = np.concatenate([sc["normals"], -nrm.reshape(-1, 3, 3)[:, ::-1].reshape(-1, 9)])
    sc["mat_ids"] = np.concatenate([sc["mat_ids"], np.zeros(T, np.uint32)])
    sc["meshes"] = [(0, T), (T, T)]
    n = len(sc["materials"])
    sc["materials"] = np.concatenate([sc["materials"], scenes.material((0.5, 0.5, 0.5), emission=(0.8, 0.7, 0.5))[None]])
    M = np.zeros((3, 4))
    M[:, :3] = np.diag([40.0, 30.0, 45.0])
    M[:, 3] = (15.0, 10.0, 2.0)
    row = np.zeros(14, np.float32)
    row[:12] = M.reshape(-1)
    row[12:14] = np.array([1, n], np.uint32).view(np.float32)
    sc["instances"] = np.concatenate([sc["instances"], row[None]])
    sc["camera"] = dict(position=(15.0, 10.0, -20.0), target=(15.0, 10.0, 2.0), up=(0, 1, 0), fov=math.radians(70.0))
    return sc


def test_two_level_estimators_agree():
    """Light sampling alone, BSDF sampling alone and both under MIS are each unbiased:
    per pixel their means agree within the standard errors, on an instanced scene lit by
    an instanced emissive dome (long paths: depth 12 with Russian roulette from depth 3,
    since BSDF sampling alone cannot reach light through the last vertex)."""
    sc = _dome_scene()
    sc["render"] = dict(sc["render"], width=24, height=16, spp=512, max_depth=12, rr_depth=3)
    orc = Oracle(sc)
    assert orc.lights()[0].shape[0] == 320
    r = sc["render"]
    stats = []
    for strategy in (0, 1, 2):
        L = orc.render(render=dict(strategy=strategy))["path_rgb"].astype(np.float64).sum(axis=1)
        L = L.reshape(r["spp"], -1)
        stats.append((L.mean(axis=0), L.std(axis=0) / math.sqrt(r["spp"]), L.mean()))
    print("image means (MIS, light, BSDF):", [s_[2] for s_ in stats])
    m0, s0, t0 = stats[0]
    lit = s0 > 0
    assert lit.mean() > 0.5
    for k in (1, 2):
        mk, sk, tk = stats[k]
        z = (mk[lit] - m0[lit]) / np.hypot(sk[lit], s0[lit])
        assert abs(z.mean()) < 4.0 / math.sqrt(z.size) + 0.1 and (np.abs(z) > 5).mean() < 0.01, (k, z.mean())
        assert abs(tk - t0) < 0.02 * t0
