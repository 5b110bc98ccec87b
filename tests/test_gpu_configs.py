"""GPU parity at the benchmark configurations' own sizes.

The other GPU tests use scenes small enough for brute force; these run the shapes the
benchmark times, where the deep traversal paths (stack spill, refill, multi-segment
queues) and the large-tile convolution kernels run:

* 1M-triangle soup, pre-split budget 6 (6.7M references, depth ~27), 1920x1080, 8 spp,
  depth 8: closest/any hits of random and camera rays bit-exact against the oracle's
  independent BVH; the bench frame rendered in full, its first-hit AOVs over the whole
  film equal to the oracle's, and (under the sky, since the opaque soup renders black) a
  256x256 window of the frame's paths bit-exact against the oracle (same keys); the
  traversal stack never overflows and its spill area is used.
* 1000 instances of a 20,480-triangle icosphere with 10,000 emitters, 3840x2160, 4 spp,
  depth 6: the same ray and window checks.
* The U-Net at 1920x1080 on features of 1-spp renders of both scenes (trt_denoise, the
  product path): every layer against the FP64 reference network applied to the device's
  own input activation, on row bands at the top, middle and bottom of each level, and the
  denoised radiance on the same bands; and a 4x256x512 batch (16-row tiles) compared
  layer by layer in full.

Ray semantics follow optixTrace closest / terminate-on-first-hit (Tracer/Optix/GPU/optix.cu:264-265,
333-335); the bar is the one of test_gpu_parity.py.
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


def _check_hits(h_gpu, i_gpu, h_ref, i_ref, label):
    pg, pr = prim_of(h_gpu), prim_of(h_ref)
    bad = np.nonzero((pg != pr) | (i_gpu != i_ref))[0]
    assert bad.size == 0, f"{label}: {bad.size} prim/instance mismatches, first {bad[:5]}: gpu {pg[bad[:5]]} ref {pr[bad[:5]]}"
    hit = pr != 0xFFFFFFFF
    assert np.array_equal(h_gpu[hit].view(np.uint32), h_ref[hit].view(np.uint32)), f"{label}: t/bary bits differ"
    return hit.mean()


def _camera_rays(sc, n, seed, tmin):
    """Camera rays of the scene's own camera through random pixels of its film with
    random jitter (the oracle's camera code)."""
    from oracle import oracle_py as O
    r = sc["render"]
    W, H = r["width"], r["height"]
    rng = np.random.default_rng(seed)
    px = np.stack([rng.integers(0, W, n), rng.integers(0, H, n)], axis=1).astype(np.uint32)
    u = rng.random((n, 5)).astype(np.float32)
    cr = O.camera_rays(sc["camera"], W, H, px, u)
    rays = np.zeros((n, 8), np.float32)
    rays[:, 0:3] = cr[:, 0:3]
    rays[:, 3] = tmin
    rays[:, 4:7] = cr[:, 3:6]
    rays[:, 7] = 1e30
    return rays


def _compare_paths(pg, pr, label, tol=1e-4):
    rel = np.abs(pg - pr) / np.maximum(np.abs(pr), 1e-6)
    bad = np.nonzero((rel > tol).any(axis=1))[0]
    assert bad.size == 0, f"{label}: {bad.size}/{pg.shape[0]} paths differ > {tol}, first {bad[:5]} gpu {pg[bad[:3]]} ref {pr[bad[:3]]}"
    diff = np.nonzero((pg.view(np.uint32) != pr.view(np.uint32)).any(axis=1))[0]
    assert diff.size == 0, f"{label}: {diff.size}/{pg.shape[0]} paths not bit-exact (within {tol}), first {diff[:5]}"


def _window(paths, spp, W, H, crop):
    x0, y0, w, h = crop
    return paths.reshape(spp, H, W, 3)[:, y0:y0 + h, x0:x0 + w].reshape(-1, 3)


# ---------------------------------------------------------------------------------------
# config 1: 1M-triangle soup
# ---------------------------------------------------------------------------------------
@pytest.fixture(scope="module")
def soup_full():
    from oracle import Oracle
    sc = scenes.soup()  # 1,000,000 triangles, 1920x1080, 8 spp, depth 8
    return sc, Oracle(sc)


def test_config1_rays_exact(rend, soup_full):
    sc, orc = soup_full
    rend.UpdateScene(sc, presplit=6.0)
    info = rend.bvh_info()
    print("config-1 BVH:", info)
    assert info["references"] > 6_000_000 and info["depth"] >= 20
    rays = np.concatenate([random_rays(200_000, seed=101, lo=-0.3, hi=1.3, tmin=1e-5),
                           _camera_rays(sc, 200_000, seed=102, tmin=1e-5)])
    hg, ig = rend.trace_closest(rays)
    hr, ir = orc.trace_closest(rays)
    frac = _check_hits(hg, ig, hr, ir, "1M soup gpu vs oracle")
    assert frac > 0.5
    sr = rays.copy()
    sr[:, 7] = np.random.default_rng(103).uniform(1e-3, 0.5, sr.shape[0]).astype(np.float32)
    occ_g, occ_r = rend.trace_any(sr), orc.trace_any(sr)
    assert np.array_equal(occ_g, occ_r), f"{(occ_g != occ_r).sum()} any-hit mismatches"
    assert 0.05 < occ_g.mean() < 0.95


SKY = dict(enabled=True, draw_sun=True, dir=(0.3, 0.8, -0.5), angular_diameter=32.0, intensity=1000.0,
           turbidity=4.0)


def test_config1_bench_frame(rend, soup_full):
    """The benchmark's frame (1920x1080, 8 spp, depth 8, RR from depth 3) rendered in
    full on the GPU. The soup is opaque (1M triangles in the unit cube: a mean free path
    of ~0.01) and its 64 emitters lie inside it, so no light reaches the camera and every
    path's radiance is zero, on the GPU and in the oracle alike; the frame is checked
    through its first hits instead: the albedo, normal and view-depth AOVs of all 2M
    pixels (means over the 8 samples) against the oracle's. The counting pass over the
    same frame reports the traversal stack's spills."""
    sc, orc = soup_full
    rend.UpdateScene(sc, presplit=6.0)
    st = rend.RenderFrame()
    r = sc["render"]
    W, H, spp = r["width"], r["height"], r["spp"]
    assert st["overflow"] == 0 and st["paths"] == W * H * spp
    print("config-1 frame:", {k: st[k] for k in ("extend_rays", "shadow_rays", "frame_ms")})
    assert not rend.path_radiance().any()
    ref = orc.render(render=dict(max_depth=1))  # every sample's first hit
    np.testing.assert_array_equal(rend.albedo()[..., :3], ref["aov"][..., 0:3])
    np.testing.assert_array_equal(rend.normal(), ref["aov"][..., 4:8])
    assert (ref["aov"][..., 0:3] > 0).any(axis=2).mean() > 0.5
    rend.Reset()  # the counting frame then traces the same paths (sample offset 0)
    tc = rend.measure_traversal()
    print("config-1 traversal counts:", tc)
    assert tc["closest_rays"] == st["extend_rays"] and tc["any_rays"] == st["shadow_rays"]
    assert tc["closest_spills"] > 0, "the 1M soup never used the traversal stack's spill area"


def test_config1_sky_lit_window_paths(rend, soup_full):
    """The same frame under the daylight sky (the reference's default sky is on): paths
    that leave the soup carry sky radiance, so a 256x256 window at the film's centre (8
    spp, the GPU frame rendered in full) compares non-zero per-path radiance."""
    from oracle import Oracle
    sc, _ = soup_full
    sc = dict(sc, sky=SKY)
    rend.UpdateScene(sc, presplit=6.0)
    st = rend.RenderFrame()
    r = sc["render"]
    W, H, spp = r["width"], r["height"], r["spp"]
    assert st["overflow"] == 0
    pg = rend.path_radiance()
    crop = (832, 412, 256, 256)
    ref = Oracle(sc).render(crop=crop)
    assert (ref["path_rgb"] > 0).any(axis=1).mean() > 0.3
    _compare_paths(_window(pg, spp, W, H, crop), ref["path_rgb"], "1M soup 1080p window, sky")


# ---------------------------------------------------------------------------------------
# config 2: 1000 instanced icospheres, 10,000 emitters
# ---------------------------------------------------------------------------------------
@pytest.fixture(scope="module")
def instanced_full():
    from oracle import Oracle
    sc = scenes.instanced()  # 1000 x 20,480 triangles, 3840x2160, 4 spp, depth 6
    return sc, Oracle(sc)


def test_config2_rays_exact(rend, instanced_full):
    sc, orc = instanced_full
    assert len(sc["instances"]) == 1000 and sc["positions"].shape[0] == 20480
    rend.UpdateScene(sc)
    info = rend.bvh_info()
    assert info["two_level"] == 1
    L, _, _, _ = orc.lights()
    assert L.shape[0] == 10_000
    rays = np.concatenate([random_rays(200_000, seed=201, lo=-10, hi=110, toward=(0, 100), tmin=1e-4),
                           _camera_rays(sc, 200_000, seed=202, tmin=1e-4)])
    hg, ig = rend.trace_closest(rays)
    hr, ir = orc.trace_closest(rays)
    frac = _check_hits(hg, ig, hr, ir, "instanced gpu vs oracle")
    assert frac > 0.1
    assert np.unique(ig[prim_of(hg) != 0xFFFFFFFF]).size > 500  # most instances are hit
    sr = rays.copy()
    sr[:, 7] = np.random.default_rng(203).uniform(1.0, 80.0, sr.shape[0]).astype(np.float32)
    occ_g, occ_r = rend.trace_any(sr), orc.trace_any(sr)
    assert np.array_equal(occ_g, occ_r), f"{(occ_g != occ_r).sum()} any-hit mismatches"


def test_config2_frame_paths(rend, instanced_full):
    sc, orc = instanced_full
    rend.UpdateScene(sc)
    st = rend.RenderFrame()
    r = sc["render"]
    W, H, spp = r["width"], r["height"], r["spp"]
    assert st["overflow"] == 0 and st["paths"] == W * H * spp
    print("config-2 frame:", {k: st[k] for k in ("extend_rays", "shadow_rays", "frame_ms")})
    pg = rend.path_radiance()
    crop = (1792, 952, 256, 256)
    ref = orc.render(crop=crop)
    win = _window(pg, spp, W, H, crop)
    assert (ref["path_rgb"] > 0).any(axis=1).mean() > 0.01  # the window sees light
    _compare_paths(win, ref["path_rgb"], "instanced 4K window")


# ---------------------------------------------------------------------------------------
# config 4: the U-Net at 1920x1080 on renders of configs 1 and 2
# ---------------------------------------------------------------------------------------
# (pooling / upsampling level of each convolution's output, input of each convolution)
_LEVEL = {0: 1, 1: 1, 2: 2, 3: 2, 4: 4, 5: 4, 6: 8, 7: 8, 8: 4, 9: 4, 10: 2, 11: 2, 12: 1, 13: 1, 14: 1}
_UPCAT = {8: (7, 5), 10: (9, 3), 12: (11, 1)}
_POOLED = {2: 1, 4: 3, 6: 5}


def _host_features(color, albedo, normal):
    """The denoiser's input features from the frame buffers (log(1 + radiance), albedo,
    normal, log(1 + view depth), zero padding to 16), as float32."""
    f = np.zeros(color.shape[:2] + (16,), np.float32)
    f[..., 0:3] = np.log1p(np.maximum(color[..., :3], 0.0))
    f[..., 3:6] = np.clip(albedo[..., :3], 0.0, 1.0)
    f[..., 6:9] = normal[..., :3]
    f[..., 9] = np.log1p(np.maximum(normal[..., 3], 0.0))
    return f


def _input_rows(i, lo, hi, feats, acts):
    """Rows [lo, hi) of convolution i's input, built from the device's own activations
    (pooling and upsampling restated as the reference network does them)."""
    from oracle import unet_ref as U
    if i == 0:
        return U.bf16(feats[:, lo:hi])
    if i in _POOLED:
        return U.pool(acts[_POOLED[i]][:, 2 * lo:2 * hi].astype(np.float64), U.bf16)
    if i in _UPCAT:
        c, s = _UPCAT[i]
        up = np.repeat(np.repeat(acts[c][:, lo // 2:(hi + 1) // 2], 2, axis=1), 2, axis=2)
        up = up[:, lo % 2:lo % 2 + (hi - lo)]
        return np.concatenate([up, acts[s][:, lo:hi]], axis=3).astype(np.float64)
    return acts[i - 1][:, lo:hi].astype(np.float64)


def _conv_rows(i, r0, r1, H, feats, acts, layers):
    """FP64 output rows [r0, r1) of convolution i (before its activation function)."""
    from oracle import unet_ref as U
    lo, hi = max(r0 - 1, 0), min(r1 + 1, H)
    x = _input_rows(i, lo, hi, feats, acts)
    pad_top, pad_bot = (1 if r0 == 0 else 0), (1 if r1 == H else 0)
    x = np.pad(x, ((0, 0), (pad_top, pad_bot), (0, 0), (0, 0)))
    w, b = layers[i]
    y = U.conv3x3(x, w, b)  # 'same' in both axes; rows 1..-1 are the valid ones
    return y[:, 1:-1]


def _bands(h, rows):
    return [(0, rows), (h // 2 - rows // 2, h // 2 + rows // 2), (h - rows, h)]


def _check_network(rend, feats, params, N, H, W, rad=None, full=False):
    """Every layer of the last forward against the reference network applied to the
    device's own input activation (one bf16 ulp of the layer's largest value), on row
    bands (or in full); with `rad`, the denoised radiance on the same bands."""
    from oracle import unet_ref as U
    ref = U.UnetRef(params, round_bf16=True)
    acts = {}
    for i, (ci, co, _) in enumerate(U.SPEC[:14]):
        lv = _LEVEL[i]
        acts[i] = rend.denoiser_activation(i, (N, H // lv, W // lv, co))
    worst = 0.0
    for i in range(14):
        h = H // _LEVEL[i]
        for r0, r1 in ([(0, h)] if full else _bands(h, max(2, 16 // _LEVEL[i]))):
            y = U.bf16(np.maximum(_conv_rows(i, r0, r1, h, feats, acts, ref.layers), 0.0))
            got = acts[i][:, r0:r1]
            err = float(np.abs(y - got).max() / max(np.abs(acts[i]).max(), 1e-30))
            worst = max(worst, err)
            assert err < 1e-2, f"conv {i} rows {r0}:{r1}: max error {err:.3e} of the layer's largest activation"
    print(f"per-layer max error {worst:.2e} (bound 1e-2)")
    if rad is not None:
        for r0, r1 in ([(0, H)] if full else _bands(H, 16)):
            z = U.bf16(feats[:, r0:r1, :, :3]) + _conv_rows(14, r0, r1, H, feats, acts, ref.layers)[..., :3]
            want = np.expm1(np.maximum(z, 0.0))
            got = rad[:, r0:r1, :, :3].astype(np.float64)
            d = np.abs(got - want) / (1.0 + want)  # error of log(1 + radiance)
            # the device rounds its own log1p of the radiance to bf16: that feature may be
            # one bf16 ulp from the host's (same rounding of a value within an ulp of a tie)
            f = np.abs(U.bf16(feats[:, r0:r1, :, :3]))
            ulp = np.where(f > 0, 2.0 ** (np.floor(np.log2(np.maximum(f, 1e-30))) - 7), 0.0)
            assert (d < 2e-4 * max(1.0, float(np.abs(z).max())) + ulp).all(), float(d.max())
            assert (rad[:, r0:r1, :, 3] == 1.0).all()
    return acts


@pytest.mark.parametrize("which", ["soup", "instanced"])
def test_config4_denoise_1080p_on_renders(rend, which):
    """trt_denoise on 1920x1080 1-spp frames of the config-1 and config-2 scenes (the
    persistent 16-row-tile convolutions run at the full and half resolution levels)."""
    from oracle import unet_ref as U
    W, H = 1920, 1080
    if which == "soup":  # sky-lit: the soup alone renders black (see test_config1_bench_frame)
        sc = dict(scenes.soup(width=W, height=H, spp=1), sky=SKY)
    else:
        sc = scenes.instanced(width=W, height=H, spp=1)
    rend.UpdateScene(sc)
    rend.RenderFrame()
    params = U.init_params(21 if which == "soup" else 22)
    rend.denoiser_set_params(params)
    rend.denoise()
    feats = _host_features(rend.color(), rend.albedo(), rend.normal())[None]
    assert np.isfinite(feats).all() and feats[..., 0:3].max() > 0
    rad = rend.denoised()[None]
    _check_network(rend, feats, params, 1, H, W, rad=rad)


def test_unet_16_row_tiles_in_full(rend):
    """A 4 x 256 x 512 batch: the full-resolution layers (2048 16-row tiles) run the
    16-row-tile kernels, whose every output is compared with the reference here, along
    with the inference path's radiance."""
    from oracle import unet_ref as U
    rng = np.random.default_rng(31)
    N, H, W = 4, 256, 512
    f = np.zeros((N, H, W, 16), np.float32)
    f[..., :10] = rng.normal(0.5, 0.5, (N, H, W, 10))
    params = U.init_params(17)
    rend.denoiser_set_params(params)
    rad = rend.denoiser_apply(f)
    _check_network(rend, f, params, N, H, W, rad=rad, full=True)
    out = rend.denoiser_forward(f)  # training path: fp32 output of the last convolution
    want = U.UnetRef(params, round_bf16=True).forward(f)
    assert np.sqrt(((out - want) ** 2).mean()) / np.sqrt((want ** 2).mean()) < 5e-3
