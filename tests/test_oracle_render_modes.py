"""Debug / AOV render modes (trt/render_modes.h) on the host build, against closed forms."""
import math

import numpy as np
import pytest

from oracle import Oracle
from trayracer_amd import _lib, scenes

M = _lib.MODES


def render(sc, **r):
    return Oracle(sc).render(render=r)


def quad_scene(quads, mats, camera, width=16, height=16, spp=1):
    tris, mids = [], []
    for (p0, p1, p2, p3), m in zip(quads, range(len(quads))):
        p0, p1, p2, p3 = (np.asarray(p, np.float32) for p in (p0, p1, p2, p3))
        tris += [np.concatenate([p0, p1, p2]), np.concatenate([p0, p2, p3])]
        mids += [m, m]
    pos = np.array(tris, np.float32)
    return dict(positions=pos, normals=None, mat_ids=np.array(mids, np.uint32), meshes=[(0, len(pos))],
                materials=np.array(mats, np.float32), camera=camera,
                render=dict(width=width, height=height, spp=spp, max_depth=4, rr_depth=3, seed=0, ray_epsilon=1e-4,
                            clamp=0.0))


def test_id_colors_follow_the_bit_layout():
    sc = scenes.cornell(width=32, height=32, spp=1)
    out = render(sc, mode=M["material_id"])["image_f32"].reshape(-1, 3)
    # channel c, bit 7-k <- id bit 3k+c: ids 1..8 give these 8-bit colours
    table = {1: (128, 0, 0), 2: (0, 128, 0), 3: (128, 128, 0), 4: (0, 0, 128), 5: (128, 0, 128), 6: (0, 128, 128),
             7: (128, 128, 128), 8: (64, 0, 0)}
    allowed = {tuple(np.float32(np.array(v, np.float32) * np.float32(1.0 / 255.0))) for v in table.values()}
    allowed.add((0.0, 0.0, 0.0))  # misses (the box is open towards the camera)
    got = {tuple(px) for px in out}
    assert got <= allowed and len(got) >= 4


def test_geometric_normal_and_material_property():
    sc = scenes.cornell(width=32, height=32, spp=1)
    gn = render(sc, mode=M["geometric_normal"])["image_f32"].reshape(-1, 3)
    axes = {tuple(((np.eye(3)[i] * s + 1) * 0.5).astype(np.float32)) for i in range(3) for s in (-1, 1)}
    wall = (np.abs(gn - 0.5).max(axis=1) == 0.5) & gn.any(axis=1)  # axis-aligned walls: exactly (n + 1) / 2
    assert wall.mean() > 0.5
    assert {tuple(p) for p in gn[wall]} <= axes
    rough = render(sc, mode=M["material_property"], mode_flags=_lib.MATERIAL_PROPERTIES.index("roughness"))
    vals = set(np.unique(rough["image_f32"]))
    assert vals <= set(np.float32(sc["materials"][:, 3])) | {np.float32(0.0)}


def test_zdepth_of_a_facing_plane():
    cam = dict(position=(0.0, 0.0, -1.0), target=(0.0, 0.0, 1.0), up=(0, 1, 0), fov=math.radians(40.0))
    q = [((-5, -5, 1), (-5, 5, 1), (5, 5, 1), (5, -5, 1))]
    sc = quad_scene(q, [scenes.material()], cam)
    img = render(sc, mode=M["zdepth"], zdepth_max=4.0)["image_f32"]
    np.testing.assert_allclose(img, 0.5, atol=2e-7)  # view depth 2 everywhere
    img = render(sc, mode=M["zdepth"], zdepth_max=1.5)["image_f32"]
    assert (img == 1.0).all()  # clamped


def test_ambient_occlusion_between_two_planes():
    """Floor at y=0, ceiling at y=h: a cosine-distributed ray from the floor reaches the
    ceiling at t = h / cos(theta); E[min(t / a, 1)] = 2x - x^2 with x = h / a (x <= 1)."""
    h, a = 0.5, 1.0
    big = 200.0
    floor = ((-big, 0, -big), (-big, 0, big), (big, 0, big), (big, 0, -big))
    ceil = ((-big, h, -big), (big, h, -big), (big, h, big), (-big, h, big))
    cam = dict(position=(0.0, 0.25, 0.0), target=(0.0, 0.0, 1.0), up=(0, 1, 0), fov=math.radians(20.0))
    sc = quad_scene([floor, ceil], [scenes.material(base=(0.5, 0.25, 1.0)), scenes.material()], cam,
                    width=8, height=8, spp=512)
    r = render(sc, mode=M["ambient_occlusion"], ao_dist=a)
    img = r["image_f64"][..., 0]
    x = h / a
    # per-pixel mean of 512 samples; Var[min(t/a,1)] = E[v^2] - E[v]^2 with E[v^2] = x^2 (1 - 2 ln x)
    ev = 2 * x - x * x
    var = x * x * (1 - 2 * math.log(x)) - ev * ev
    se = math.sqrt(var / 512)
    assert abs(img.mean() - ev) < 4 * se / math.sqrt(img.size)
    assert np.abs(img - ev).max() < 5 * se
    shading = render(sc, mode=M["ambient_occlusion_shading"], ao_dist=a)["path_rgb"]
    ao = r["path_rgb"]
    # shading mode = occlusion x base colour of the floor hit (ceiling pixels: x base 0.8)
    ratio = shading / np.maximum(ao, 1e-30)
    floor_px = np.isclose(ratio[:, 1] / np.maximum(ratio[:, 0], 1e-30), 0.5)
    assert floor_px.any()
    np.testing.assert_allclose(ratio[floor_px], np.broadcast_to([0.5, 0.25, 1.0], ratio[floor_px].shape), rtol=1e-6)


def test_frames_coordinates_and_wireframe():
    sc = scenes.cornell(width=24, height=24)
    r = dict(width=24, height=24, spp=1)
    n_img = render(sc, mode=M["shading_normal"], **r)["image_f32"].reshape(-1, 3)
    n = n_img * 2 - 1
    t = render(sc, mode=M["tangent"], **r)["image_f32"].reshape(-1, 3) * 2 - 1
    b = render(sc, mode=M["bitangent"], **r)["image_f32"].reshape(-1, 3) * 2 - 1
    hit = n_img.any(axis=1)  # misses are black
    assert hit.mean() > 0.5
    for v in (n, t, b):
        np.testing.assert_allclose(np.linalg.norm(v[hit], axis=1), 1.0, atol=2e-6)
    np.testing.assert_allclose(np.einsum("ij,ij->i", n[hit], t[hit]), 0.0, atol=2e-6)
    np.testing.assert_allclose(np.einsum("ij,ij->i", n[hit], b[hit]), 0.0, atol=2e-6)
    np.testing.assert_allclose(np.cross(t[hit], b[hit]), n[hit], atol=4e-6)  # right-handed (t, b, n)
    uv = render(sc, mode=M["texture_coordinate"], **r)["image_f32"].reshape(-1, 3)
    assert (uv[:, :2] >= 0).all() and (uv[:, :2].sum(axis=1) <= 1 + 1e-6).all() and (uv[:, 2] == 0).all()
    wf = render(sc, mode=M["wireframe"], **r)["image_f32"]
    assert set(np.unique(wf)) <= {0.0, 1.0} and 0 < wf.mean() < 0.5
    oid = render(sc, mode=M["object_id"], **r)["image_f32"].reshape(-1, 3)
    assert {tuple(p) for p in oid[hit]} == {(np.float32(128 / 255.0), 0.0, 0.0)}  # one object: id 1


def test_misses_are_black_and_path_tracing_unchanged():
    sc = scenes.cornell(width=16, height=16, spp=2)
    base = render(sc)["path_rgb"]
    same = render(sc, mode=M["path_tracing"], mode_flags=5, ao_dist=3.0)["path_rgb"]
    assert np.array_equal(base.view(np.uint32), same.view(np.uint32))
