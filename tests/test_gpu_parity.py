"""GPU parity: libtrt.so (gfx950 kernels) against the CPU oracle on identical inputs.

Bar: closest-hit primitive/instance ids, t and barycentrics bit-exact (the tie rule is
order independent, so traversal order and tree shape cannot change them); any-hit
flags exact; per-path radiance bit-exact (the shared arithmetic avoids contraction and
libm calls on both sides, and both add a path's terms in vertex order: the emission of
vertex d, then its next-event estimate), with the relative error reported first.
"""
import numpy as np
import pytest

from tests.gpu_util import prim_of, random_rays
from trayracer_amd import Renderer, scenes

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def rend():
    r = Renderer(0)
    yield r
    r.close()


def _oracle(scene):
    from oracle import Oracle
    return Oracle(scene)


def _check_hits(h_gpu, i_gpu, h_ref, i_ref, label):
    pg, pr = prim_of(h_gpu), prim_of(h_ref)
    bad = np.nonzero((pg != pr) | (i_gpu != i_ref))[0]
    assert bad.size == 0, f"{label}: {bad.size} prim/instance mismatches, first {bad[:5]}: gpu {pg[bad[:5]]} ref {pr[bad[:5]]}"
    hit = pr != 0xFFFFFFFF
    assert np.array_equal(h_gpu[hit].view(np.uint32), h_ref[hit].view(np.uint32)), f"{label}: t/bary bits differ"


@pytest.mark.parametrize("presplit", [0.0, 2.0])
def test_trace_cornell_exact(rend, presplit):
    sc = scenes.cornell()
    rend.UpdateScene(sc, presplit=presplit)
    orc = _oracle(sc)
    rays = random_rays(20000, seed=11, lo=0.02, hi=0.98)
    hg, ig = rend.trace_closest(rays)
    hr, ir = orc.trace_closest(rays)
    hb, ib = orc.trace_closest(rays, brute=True)
    _check_hits(hr, ir, hb, ib, "oracle bvh vs brute")
    _check_hits(hg, ig, hr, ir, "gpu vs oracle")
    assert (prim_of(hg) != 0xFFFFFFFF).mean() > 0.7  # the box is open at z = 0
    occ_g = rend.trace_any(rays[:5000].copy())
    sr = rays[:5000].copy()
    sr[:, 7] = np.random.default_rng(3).uniform(0.05, 1.5, 5000)
    occ_g = rend.trace_any(sr)
    occ_r = orc.trace_any(sr)
    assert np.array_equal(occ_g, occ_r)


@pytest.mark.parametrize("presplit", [0.0, 1.0, 3.0, 6.0])
def test_trace_soup_exact(rend, presplit):
    sc = scenes.soup(n_tris=200_000, seed=5, width=64, height=64)
    rend.UpdateScene(sc, presplit=presplit)
    info = rend.bvh_info()
    assert info["depth"] >= 10
    assert (info["references"] > 200_000) == (presplit > 0)
    orc = _oracle(sc)
    rays = random_rays(50000, seed=12, lo=-0.3, hi=1.3, tmin=1e-5)
    hg, ig = rend.trace_closest(rays)
    hr, ir = orc.trace_closest(rays)
    _check_hits(hg, ig, hr, ir, "soup gpu vs oracle")
    hb, ib = orc.trace_closest(rays[:2000], brute=True)
    _check_hits(hr[:2000], ir[:2000], hb, ib, "soup oracle bvh vs brute")
    sr = rays.copy()
    sr[:, 7] = np.random.default_rng(4).uniform(0.001, 0.3, sr.shape[0])
    assert np.array_equal(rend.trace_any(sr), orc.trace_any(sr))


def _axis_rays(n, seed, scale=1.0):
    """Degenerate rays for the slab test: directions along +-x, +-y, +-z (two zero
    components, +0.0 or -0.0), along face diagonals (one zero component), and
    origins snapped to the Cornell box's wall planes and to a coarse grid, so that rays
    start on or graze box faces."""
    rng = np.random.default_rng(seed)
    dirs = []
    for k in range(3):
        for sgn in (1.0, -1.0):
            d = np.zeros(3)
            d[k] = sgn
            dirs.append(d)
    for k in range(3):
        for s1 in (1.0, -1.0):
            for s2 in (1.0, -1.0):
                d = np.zeros(3)
                d[(k + 1) % 3] = s1
                d[(k + 2) % 3] = s2
                dirs.append(d / np.linalg.norm(d))
    dirs = np.array(dirs)
    o = rng.uniform(0.0, 1.0, (n, 3))
    snap = rng.integers(0, 4, (n, 3))
    o = np.where(snap == 0, np.round(o * 8.0) / 8.0, o)    # coarse grid
    o = np.where(snap == 1, np.round(o), o)                # on the 0 / 1 wall planes
    r = np.zeros((n, 8), np.float32)
    r[:, 0:3] = o * scale
    r[:, 3] = np.where(rng.random(n) < 0.5, 0.0, 1e-4)  # tmin 0: hits at an origin on a wall
    d = dirs[rng.integers(0, len(dirs), n)]
    neg0 = (d == 0.0) & (rng.random(d.shape) < 0.5)  # -0.0 components: reciprocal -inf
    r[:, 4:7] = np.where(neg0, -0.0, d)
    r[:, 7] = 1e30
    return r


@pytest.mark.parametrize("scene", ["cornell", "soup"])
def test_trace_axis_aligned_rays(rend, scene):
    """Closest and any hits of axis-aligned and face-diagonal rays (zero direction
    components) starting on wall planes and grid points equal the oracle's brute force."""
    sc = scenes.cornell() if scene == "cornell" else scenes.soup(n_tris=50_000, seed=3, width=64, height=64)
    rend.UpdateScene(sc)
    orc = _oracle(sc)
    rays = _axis_rays(20000, seed=21)
    hg, ig = rend.trace_closest(rays)
    hb, ib = orc.trace_closest(rays, brute=True)
    _check_hits(hg, ig, hb, ib, f"{scene} axis rays gpu vs brute force")
    sr = rays.copy()
    sr[:, 7] = np.random.default_rng(6).uniform(0.01, 0.6, sr.shape[0])
    assert np.array_equal(rend.trace_any(sr), orc.trace_any(sr))


@pytest.mark.parametrize("n_tris", [1, 2, 3])
def test_trace_tiny_scenes(rend, n_tris):
    """Scenes of 1-3 triangles: the tree's root is a leaf, stored as a node with an empty
    second child (a zero box at the origin, which rays through the origin visit)."""
    import copy
    sc = copy.deepcopy(scenes.cornell())
    sc["positions"] = np.asarray(sc["positions"], np.float32)[:n_tris]
    sc["mat_ids"] = np.asarray(sc["mat_ids"])[:n_tris]
    sc["meshes"] = [(0, n_tris)]
    rend.UpdateScene(sc)
    orc = _oracle(sc)
    rays = np.concatenate([random_rays(4000, seed=30 + n_tris, lo=-0.2, hi=1.2, tmin=1e-4, toward=(0.0, 1.0)),
                           _axis_rays(4000, seed=40 + n_tris)])
    rays[:500, 0:3] = 0.0  # from the origin, where the empty child's box sits
    hg, ig = rend.trace_closest(rays)
    hb, ib = orc.trace_closest(rays, brute=True)
    _check_hits(hg, ig, hb, ib, f"{n_tris}-triangle scene gpu vs brute force")
    assert np.array_equal(rend.trace_any(rays), orc.trace_any(rays, brute=True))


def test_empty_scene_is_rejected():
    """A scene without triangles is refused by the build with an error (host check, no
    kernel runs); the context stays usable for a valid scene afterwards, and a trace of
    no rays returns empty results."""
    import copy

    from trayracer_amd import Renderer
    from trayracer_amd._lib import TrtError
    r = Renderer(0)
    sc = copy.deepcopy(scenes.cornell())
    sc["positions"] = np.zeros((0, 3, 3), np.float32)
    sc["mat_ids"] = np.zeros(0, np.uint32)
    sc["meshes"] = []
    with pytest.raises(TrtError, match="no triangles"):
        r.UpdateScene(sc)
    good = scenes.cornell()
    r.UpdateScene(good)
    rays = random_rays(1000, seed=5, lo=-0.2, hi=1.2, tmin=1e-4, toward=(0.0, 1.0))
    hg, ig = r.trace_closest(rays)
    hb, ib = _oracle(good).trace_closest(rays, brute=True)
    _check_hits(hg, ig, hb, ib, "cornell after a refused empty scene")
    h0, i0 = r.trace_closest(rays[:0])  # no rays: nothing launched, empty results
    assert len(h0) == 0 and len(i0) == 0 and len(r.trace_any(rays[:0])) == 0
    r.close()


def test_non_finite_rays_are_rejected(rend):
    """Rays with a NaN or infinite origin or direction component, or a NaN bound, are
    refused before anything is launched; an infinite tmax is a valid unbounded ray."""
    from trayracer_amd._lib import TrtError
    sc = scenes.cornell()
    rend.UpdateScene(sc)
    rays = random_rays(256, seed=6, lo=-0.2, hi=1.2, tmin=1e-4, toward=(0.0, 1.0))
    for col, val in ((0, np.nan), (2, np.inf), (5, -np.inf), (3, np.nan), (7, np.nan)):
        bad = rays.copy()
        bad[100, col] = val
        with pytest.raises(TrtError, match="ray 100 "):
            rend.trace_closest(bad)
        with pytest.raises(TrtError, match="ray 100 "):
            rend.trace_any(bad)
    unbounded = rays.copy()
    unbounded[:, 7] = np.inf
    hg, ig = rend.trace_closest(unbounded)
    hb, ib = _oracle(sc).trace_closest(unbounded, brute=True)
    _check_hits(hg, ig, hb, ib, "cornell with tmax = +inf")


@pytest.mark.parametrize("max_leaf", [1, 7])
def test_trace_soup_leaf_sizes(rend, max_leaf):
    """Leaves of 1 and of up to 7 triangles: with 7, a wave's leaf phase holds up to 896
    triangle tasks, more than one 128-task window, so lanes that do not fit keep their
    leaves for the next round (trt_trace_closest / _any run the render's own kernels)."""
    sc = scenes.soup(n_tris=100_000, seed=9, width=64, height=64)
    rend.UpdateScene(sc, max_leaf=max_leaf, presplit=1.0)
    orc = _oracle(sc)
    rays = random_rays(40000, seed=13, lo=-0.3, hi=1.3, tmin=1e-5)
    hg, ig = rend.trace_closest(rays)
    hr, ir = orc.trace_closest(rays)
    _check_hits(hg, ig, hr, ir, f"soup max_leaf={max_leaf} gpu vs oracle")
    sr = rays.copy()
    sr[:, 7] = np.random.default_rng(5).uniform(0.001, 0.3, sr.shape[0])
    assert np.array_equal(rend.trace_any(sr), orc.trace_any(sr))


def _compare_paths(pg, pr, label, tol=1e-4):
    denom = np.maximum(np.abs(pr), 1e-6)
    rel = np.abs(pg - pr) / denom
    bad = np.nonzero((rel > tol).any(axis=1))[0]
    assert bad.size == 0, f"{label}: {bad.size}/{pg.shape[0]} paths differ > {tol}, first {bad[:5]} gpu {pg[bad[:3]]} ref {pr[bad[:3]]}"
    diff = np.nonzero((pg.view(np.uint32) != pr.view(np.uint32)).any(axis=1))[0]
    assert diff.size == 0, f"{label}: {diff.size}/{pg.shape[0]} paths not bit-exact (within {tol}), first {diff[:5]}"


def test_path_radiance_cornell(rend):
    sc = scenes.cornell(width=64, height=48, spp=4, max_depth=4, rr_depth=3, seed=7)
    rend.UpdateScene(sc)
    rend.RenderFrame()
    pg = rend.path_radiance()
    ref = _oracle(sc).render()
    _compare_paths(pg, ref["path_rgb"], "cornell")
    # image = mean of paths in fixed sample order (same float ops as the oracle)
    img = rend.color()[..., :3]
    np.testing.assert_allclose(img, ref["image_f32"], rtol=1e-5, atol=1e-7)
    aov = ref["aov"]
    np.testing.assert_array_equal(rend.albedo()[..., :3], aov[..., 0:3])
    np.testing.assert_array_equal(rend.normal(), aov[..., 4:8])


def test_phase_timing_off_changes_no_result(rend):
    """A frame rendered without the per-phase timing events (only those that order the two
    streams) has the same paths, counts and image as one with them; its phase times are 0."""
    sc = scenes.cornell(width=48, height=40, spp=2, max_depth=5, rr_depth=3, seed=11)
    rend.UpdateScene(sc)
    on = rend.RenderFrame()
    p_on = rend.path_radiance().copy()
    rend.Reset()
    rend.SetPhaseTiming(False)
    try:
        off = rend.RenderFrame()
    finally:
        rend.SetPhaseTiming(True)
    assert np.array_equal(rend.path_radiance().view(np.uint32), p_on.view(np.uint32))
    for k in ("paths", "extend_rays", "shadow_rays", "primary_rays", "extend_launches", "overflow"):
        assert off[k] == on[k], k
    assert off["frame_ms"] > 0 and on["shade_ms"] > 0
    for k in ("raygen_ms", "extend_ms", "shade_ms", "shadow_ms", "finalize_ms", "trace_ms"):
        assert off[k] == 0.0, k


def test_aov_means_accumulate_over_frames(rend):
    """Albedo / normal / view depth are per-pixel means over every accumulated sample:
    two frames of 3 samples equal the oracle's 6-sample frame (up to the sums'
    association: frame sums are added, the oracle sums the 6 in order); Reset restarts them."""
    sc = scenes.cornell(width=40, height=32, spp=3, max_depth=2, seed=13)
    sc["camera"] = dict(sc["camera"], aperture=0.05, focal_dist=1.6)  # samples differ in their first hits
    rend.UpdateScene(sc)
    rend.RenderFrame()
    one = _oracle(sc).render()
    np.testing.assert_array_equal(rend.normal(), one["aov"][..., 4:8])
    rend.RenderFrame()
    two = _oracle(sc).render(render=dict(spp=6))
    np.testing.assert_allclose(rend.albedo()[..., :3], two["aov"][..., 0:3], rtol=1e-6, atol=1e-7)
    np.testing.assert_allclose(rend.normal(), two["aov"][..., 4:8], rtol=1e-6, atol=1e-6)
    assert not np.allclose(one["aov"][..., 4:8], two["aov"][..., 4:8])  # the second frame changed them
    rend.Reset()
    rend.RenderFrame()
    np.testing.assert_array_equal(rend.normal(), one["aov"][..., 4:8])


def test_path_radiance_cornell_rr_deep(rend):
    sc = scenes.cornell(width=32, height=32, spp=8, max_depth=10, rr_depth=1, seed=9)
    rend.UpdateScene(sc)
    rend.RenderFrame()
    ref = _oracle(sc).render()
    _compare_paths(rend.path_radiance(), ref["path_rgb"], "cornell rr")


@pytest.mark.parametrize("width,height,spp", [(1, 1, 1), (3, 5, 2), (67, 1, 3), (1, 130, 1), (67, 13, 2), (13, 67, 1)])
def test_path_radiance_ragged_frames(rend, width, height, spp):
    """Frames whose path counts are far from a whole wave or queue segment (1 path up to
    1742), including single-row and single-column films, and films whose sides are not
    multiples of RayGen's 8x8 tiles (tiles, then the leftover bottom rows and right
    columns: every path traced once, each with its own pixel's ray)."""
    sc = scenes.cornell(width=width, height=height, spp=spp, max_depth=4, rr_depth=3, seed=13)
    rend.UpdateScene(sc)
    rend.RenderFrame()
    ref = _oracle(sc).render()
    assert rend.path_radiance().shape[0] == width * height * spp
    _compare_paths(rend.path_radiance(), ref["path_rgb"], f"cornell {width}x{height}x{spp}")
    np.testing.assert_allclose(rend.color()[..., :3], ref["image_f32"], rtol=1e-5, atol=1e-7)


@pytest.mark.parametrize("presplit", [0.0, 1.0, 6.0])
def test_path_radiance_soup(rend, presplit):
    sc = scenes.soup(n_tris=50_000, seed=3, width=48, height=32, spp=2, max_depth=8)
    rend.UpdateScene(sc, presplit=presplit)
    rend.RenderFrame()
    ref = _oracle(sc).render()
    _compare_paths(rend.path_radiance(), ref["path_rgb"], "soup")


def test_instanced_exact(rend):
    sc = scenes.instanced(n_instances=40, subdiv=2, seed=4, width=48, height=32, spp=2, max_depth=6,
                          emissive_per_instance=4)
    rend.UpdateScene(sc)
    assert rend.bvh_info()["two_level"] == 1
    orc = _oracle(sc)
    rays = random_rays(20000, seed=13, lo=-10, hi=110, toward=(0, 100), tmin=1e-4)
    hg, ig = rend.trace_closest(rays)
    hr, ir = orc.trace_closest(rays)
    _check_hits(hg, ig, hr, ir, "instanced gpu vs oracle")
    assert (prim_of(hg) != 0xFFFFFFFF).any()
    ar = _axis_rays(20000, seed=22, scale=100.0)  # axis-aligned rays through the instance field
    hg, ig = rend.trace_closest(ar)
    hr, ir = orc.trace_closest(ar)
    _check_hits(hg, ig, hr, ir, "instanced axis rays gpu vs oracle")
    rend.RenderFrame()
    ref = orc.render()
    _compare_paths(rend.path_radiance(), ref["path_rgb"], "instanced")


def test_instanced_duplicates_tie_to_lower_instance(rend):
    """Instances with identical transforms give identical hits; the (t, instance, prim)
    order picks the lower instance id, whatever the traversal visits first."""
    sc = scenes.instanced(n_instances=8, subdiv=2, seed=5, width=16, height=16, spp=1, max_depth=2,
                          emissive_per_instance=1)
    inst = np.asarray(sc["instances"])
    sc["instances"] = np.concatenate([inst, inst[[3, 0, 5, 0]]])  # ids 8-11 repeat 3, 0, 5, 0
    rend.UpdateScene(sc)
    orc = _oracle(sc)
    rays = random_rays(20000, seed=17, lo=-10, hi=110, toward=(0, 100), tmin=1e-4)
    # aim half of them at the repeated instances' translations
    tgt = inst[[3, 0, 5], :][:, [3, 7, 11]]
    aim = tgt[np.arange(10000) % 3] + np.random.default_rng(3).normal(0.0, 0.5, (10000, 3))
    d = aim - rays[:10000, 0:3]
    rays[:10000, 4:7] = d / np.linalg.norm(d, axis=1, keepdims=True)
    hg, ig = rend.trace_closest(rays)
    hb, ib = orc.trace_closest(rays, brute=True)
    _check_hits(hg, ig, hb, ib, "duplicated instances gpu vs brute force")
    hit = prim_of(hb) != 0xFFFFFFFF
    assert hit.sum() > 1000 and not np.isin(ib[hit], [8, 9, 10, 11]).any()


def test_progressive_accumulation(rend):
    sc = scenes.cornell(width=32, height=32, spp=2, seed=1)
    rend.UpdateScene(sc)
    rend.RenderFrame()
    a1 = rend.accum().copy()
    rend.RenderFrame()
    a2 = rend.accum()
    assert rend.SampleCount() == 4
    assert np.all(a2[..., 3] == 4.0)
    orc = _oracle(sc)
    r0 = orc.render(sample_offset=0)["image_f32"]
    r1 = orc.render(sample_offset=2)["image_f32"]
    np.testing.assert_allclose(a1[..., :3] / 2.0, r0, rtol=1e-5, atol=1e-7)
    np.testing.assert_allclose((a2[..., :3] - a1[..., :3]) / 2.0, r1, rtol=1e-4, atol=1e-6)
    rend.Reset()
    rend.RenderFrame()
    assert rend.SampleCount() == 2


def test_pick_ray(rend):
    sc = scenes.cornell(width=64, height=64)
    rend.UpdateScene(sc)
    p = rend.PickRay(32, 32)
    assert p["primitive"] != 0xFFFFFFFF
    assert abs(p["direction"][2] - 1.0) < 0.01


@pytest.mark.parametrize("sun_nee", [False, True], ids=["sky", "sky+sun-nee"])
def test_path_radiance_with_sky(rend, sun_nee):
    """Open Cornell box under the daylight sky: rays leaving through the open side and
    camera rays that miss pick up sky radiance (sun disc for camera rays); with sun_nee
    the sun disc is also sampled by next-event estimation (shadow rays to infinity) and
    weighted by MIS where BSDF-sampled rays reach it."""
    sc = scenes.cornell(width=48, height=40, spp=2, max_depth=5, seed=11)
    sc["camera"] = dict(sc["camera"], position=(0.5, 0.6, -2.2), fov=1.2)  # box plus surrounding sky
    sc["sky"] = dict(enabled=True, draw_sun=True, dir=(0.2, 0.3, -1.0), angular_diameter=32.0 if not sun_nee else 300.0,
                     intensity=1000.0, turbidity=4.0, sun_nee=sun_nee)
    rend.UpdateScene(sc)
    rend.RenderFrame()
    ref = _oracle(sc).render()
    _compare_paths(rend.path_radiance(), ref["path_rgb"], "sky")
    np.testing.assert_array_equal(rend.albedo()[..., :3], ref["aov"][..., 0:3])
    assert (ref["aov"][..., 0:3] > 0).all()  # every first hit or miss has an albedo (sky for misses)


@pytest.mark.parametrize("mode", [m for m in range(1, 13)])
def test_render_modes_match_oracle(rend, mode):
    """Debug / AOV render modes: per-path values equal to the oracle's (bit-exact)."""
    sc = scenes.cornell(width=40, height=32, spp=2, max_depth=4, seed=5)
    sc["render"] = dict(sc["render"], mode=mode, mode_flags=7 if mode == 6 else 0, ao_dist=0.4, zdepth_max=3.0)
    rend.UpdateScene(sc)
    s = rend.settings
    assert (s.mode, s.ao_dist) == (mode, 0.4)
    rend.RenderFrame()
    ref = _oracle(sc).render()
    assert np.array_equal(rend.path_radiance().view(np.uint32), ref["path_rgb"].view(np.uint32)), mode
    np.testing.assert_array_equal(rend.color()[..., :3], ref["image_f32"])


@pytest.mark.parametrize("lens", [dict(aperture=0.05, focal_dist=1.6, bokeh_sides=0, bokeh_rotation=0.0, distortion=0.0),
                                  dict(aperture=0.08, focal_dist=1.2, bokeh_sides=6, bokeh_rotation=0.4,
                                       distortion=0.25)],
                         ids=["thin-lens disc", "hexagonal bokeh + distortion"])
def test_camera_lens_models(rend, lens):
    """Thin lens (disc or polygonal aperture) and barrel distortion: per-path radiance and
    AOVs equal to the oracle's."""
    sc = scenes.cornell(width=40, height=32, spp=4, max_depth=4, seed=21)
    sc["camera"] = dict(sc["camera"], **lens)
    rend.UpdateScene(sc)
    rend.RenderFrame()
    ref = _oracle(sc).render()
    _compare_paths(rend.path_radiance(), ref["path_rgb"], "lens")
    np.testing.assert_array_equal(rend.normal(), ref["aov"][..., 4:8])


def test_tonemap_matches_host_transform(rend, tmp_path):
    """Device display transform (all curves, colour and denoised sources) equals the
    host transform of the same floats; PNG export of the result."""
    from oracle.oracle_py import tonemap
    sc = scenes.cornell(width=64, height=48, spp=4, max_depth=4, seed=2)
    rend.UpdateScene(sc)
    rend.RenderFrame()
    col = rend.color()
    for method in range(8):
        for exposure, gamma in ((1.0, 2.2), (3.5, 1.8)):
            got = rend.tonemap(method, exposure, gamma)
            np.testing.assert_array_equal(got, tonemap(col, method, exposure, gamma))
    rend.save_png(str(tmp_path / "c.png"), method="aces")
    assert (tmp_path / "c.png").stat().st_size > 64 * 48 * 4


def test_path_radiance_disney_lobes(rend):
    """Cornell geometry with every Disney lobe in use (sheen and sheen tint, clearcoat,
    subsurface, anisotropic and tinted specular, partial metallic): per-path radiance
    against the oracle."""
    import copy
    sc = copy.deepcopy(scenes.cornell(width=48, height=40, spp=4, max_depth=5, seed=23))
    m = sc["materials"]
    lobes = [dict(sheen=0.9, sheen_tint=0.6, roughness=0.7), dict(clearcoat=0.8, clearcoat_gloss=0.7, roughness=0.4),
             dict(subsurface=0.9, roughness=0.6), dict(anisotropic=0.85, roughness=0.35, metallic=0.6),
             dict(specular_tint=0.8, specular=0.9, roughness=0.25, sheen=0.3, clearcoat=0.3, subsurface=0.4)]
    for i, kw in enumerate(lobes):
        if i == 3:  # keep the emitter (material 3)
            continue
        base = tuple(float(x) for x in m[i, 0:3])
        m[i] = scenes.material(base, **kw)
    m[3] = scenes.material((0.78, 0.78, 0.78), emission=(17.0, 12.0, 4.0), roughness=1.0)
    extra = scenes.material((0.4, 0.6, 0.8), **lobes[3])
    sc["materials"] = np.concatenate([m, extra[None]])
    sc["mat_ids"] = np.where(sc["mat_ids"] == 1, 5, sc["mat_ids"]).astype(np.uint32)  # red wall: anisotropic
    rend.UpdateScene(sc)
    rend.RenderFrame()
    ref = _oracle(sc).render()
    assert (ref["path_rgb"] > 0).any(axis=1).mean() > 0.5
    _compare_paths(rend.path_radiance(), ref["path_rgb"], "disney lobes")


def test_path_radiance_with_firefly_clamp(rend):
    """The firefly clamp (a contribution's largest component limited to `clamp`, the
    reference's SafeColor, CudaUtility.h:91-107): per-path radiance against the oracle,
    with the clamp active on many paths."""
    sc = scenes.cornell(width=48, height=40, spp=4, max_depth=5, seed=29)
    sc["render"] = dict(sc["render"], clamp=0.5)
    rend.UpdateScene(sc)
    rend.RenderFrame()
    ref = _oracle(sc).render()
    unclamped = _oracle(scenes.cornell(width=48, height=40, spp=4, max_depth=5, seed=29)).render()
    assert (ref["path_rgb"] < unclamped["path_rgb"] - 1e-6).any(axis=1).mean() > 0.01  # the clamp acts
    _compare_paths(rend.path_radiance(), ref["path_rgb"], "clamp")


def test_pick_ray_matches_oracle(rend):
    """PickRay (DeviceRenderer::PickRay, Tracer/Renderer/DeviceRenderer.cpp:71-93): the camera
    ray through the pixel's centre and its closest hit, against the oracle's camera and
    traversal, at pixels over the film."""
    from oracle import oracle_py as O
    sc = scenes.cornell(width=64, height=48)
    rend.UpdateScene(sc)
    orc = _oracle(sc)
    for x, y in ((0, 0), (32, 24), (63, 47), (10, 40), (50, 5)):
        p = rend.PickRay(x, y)
        cr = O.camera_rays(sc["camera"], 64, 48, [[x, y]], [[0.5, 0.5, 0.5, 0.5, 0.5]])[0]
        np.testing.assert_array_equal(np.array(p["origin"], np.float32), cr[0:3])
        np.testing.assert_array_equal(np.array(p["direction"], np.float32), cr[3:6])
        ray = np.concatenate([cr[0:3], [0.0], cr[3:6], [1e30]]).astype(np.float32)
        h, i = orc.trace_closest(ray[None])
        assert p["primitive"] == h[0, 3:4].view(np.uint32)[0]
        if p["primitive"] != 0xFFFFFFFF:
            assert np.float32(p["t"]) == h[0, 0]


def test_accumulator_checkpoint_resumes_the_render(rend):
    """A progressive render read back (trt_read_accum), restored into a fresh context
    (trt_write_accum) and continued equals the uninterrupted render bit for bit: the
    resumed frames trace the following samples."""
    from trayracer_amd import Renderer
    sc = scenes.cornell(width=40, height=32, spp=2, max_depth=4, seed=17)
    rend.UpdateScene(sc)
    for _ in range(2):
        rend.RenderFrame()
    saved = rend.accum().copy()
    for _ in range(3):
        rend.RenderFrame()
    full = rend.accum().copy()
    other = Renderer(0)
    try:
        other.UpdateScene(sc)
        other.write_accum(saved)
        assert other.SampleCount() == 4
        np.testing.assert_array_equal(other.color()[..., :3], saved[..., :3] / np.float32(4))
        for _ in range(3):
            other.RenderFrame()
        assert np.array_equal(other.accum().view(np.uint32), full.view(np.uint32))
        bad = saved.copy()
        bad[0, 0, 3] = 3.0
        from trayracer_amd._lib import TrtError
        with pytest.raises(TrtError, match="same sample count"):
            other.write_accum(bad)
    finally:
        other.close()
