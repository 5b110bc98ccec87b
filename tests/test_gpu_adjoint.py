"""GPU adjoint (path-replay backward kernels + deterministic reduction) against the
oracle's FP64-accumulated gradient and against central finite differences of the GPU
forward render, on the differentiable Cornell box (36 triangles x 5 parameters)."""
import numpy as np
import pytest

from trayracer_amd import Renderer, scenes

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def rend():
    r = Renderer(0)
    yield r
    r.close()


SKY = dict(enabled=True, draw_sun=True, dir=(0.3, 0.8, -0.5), angular_diameter=32.0, intensity=1000.0,
           turbidity=4.0)


def _setup(rend, width, height, spp, seed=3, sky=None, rr_depth=None):
    sc, target = scenes.cornell_diff(width=width, height=height, spp=spp, max_depth=4, seed=seed, rr_depth=rr_depth)
    if sky is not None:
        sc["sky"] = sky
    rend.UpdateScene(sc)
    theta = scenes.material_params(sc["materials"])
    rend.set_material_params(target, which=3)
    rend.RenderFrame()
    target_img = rend.color()[..., :3].astype(np.float64)
    rend.set_material_params(theta, which=3)
    rend.RenderFrame()
    img = rend.color()[..., :3]
    dldi = (2.0 * (img.astype(np.float64) - target_img)).astype(np.float32)
    return sc, theta, target_img, img, dldi


def _loss(img, target_img):
    d = img.astype(np.float64) - target_img
    return float((d * d).sum())


def _significant(g, frac=1e-3):
    return np.abs(g) > frac * np.abs(g).max()


@pytest.mark.parametrize("sky,rr_depth", [(None, None), (SKY, None), (None, 1)],
                         ids=["emitters", "emitters+sky", "russian-roulette-from-depth-1"])
def test_backward_matches_oracle_fp64(rend, sky, rr_depth):
    from oracle import Oracle
    sc, theta, target_img, img, dldi = _setup(rend, 64, 64, 16, sky=sky, rr_depth=rr_depth)
    g = rend.backward(dldi)
    orc = Oracle(sc)
    orc.set_material_params(theta, which=3)
    ref = orc.render()
    np.testing.assert_array_equal(img, ref["image_f32"])  # forward is bit-exact
    gr = orc.backward(dldi)
    sig = _significant(gr)
    rel = np.abs(g[sig] - gr[sig]) / np.abs(gr[sig])
    assert rel.max() < 1e-3, rel.max()
    assert np.abs(g[~sig] - gr[~sig]).max() <= 1e-4 * np.abs(gr).max()


def test_backward_is_deterministic(rend):
    _, _, _, _, dldi = _setup(rend, 48, 48, 8)
    g1 = rend.backward(dldi)
    g2 = rend.backward(dldi)
    assert np.array_equal(g1.view(np.uint32), g2.view(np.uint32))


@pytest.mark.parametrize("size,rr_depth", [((64, 64, 16), None), ((512, 512, 64), None), ((512, 512, 64), 1)],
                         ids=["64x64x16", "512x512x64", "512x512x64-rr-from-depth-1"])
def test_backward_matches_central_differences(rend, size, rr_depth):
    """Central differences of the GPU forward render with the sampling parameters held
    fixed; with Russian roulette from depth 1 its survival probability is one of the
    sampling decisions held fixed (PathTracing.h:148-154)."""
    W, H, spp = size
    sc, theta, target_img, img, dldi = _setup(rend, W, H, spp, rr_depth=rr_depth)
    g = rend.backward(dldi)
    eps = 1e-3
    idx = np.argsort(-np.abs(g).reshape(-1))[:10]
    rend.set_material_params(theta, which=rend.SAMPLING)
    worst = 0.0
    for i in idx:
        m, k = divmod(int(i), 5)
        tp, tm = theta.copy(), theta.copy()
        tp[m, k] += eps
        tm[m, k] -= eps
        rend.set_material_params(tp, which=rend.EVAL)
        rend.RenderFrame()
        lp = _loss(rend.color()[..., :3], target_img)
        rend.set_material_params(tm, which=rend.EVAL)
        rend.RenderFrame()
        lm = _loss(rend.color()[..., :3], target_img)
        fd = (lp - lm) / (2 * eps)
        worst = max(worst, abs(fd - g[m, k]) / abs(g[m, k]))
    print(f"{W}x{H}x{spp}: worst relative FD error over the 10 largest gradients {worst:.2e}")
    assert worst < 1e-3


def test_torch_autograd_wrapper(rend):
    import torch

    from trayracer_amd.autograd import render
    sc, target = scenes.cornell_diff(width=32, height=32, spp=8, max_depth=4, seed=5)
    rend.UpdateScene(sc)
    theta = torch.tensor(scenes.material_params(sc["materials"]), requires_grad=True)
    with torch.no_grad():
        target_img = render(torch.tensor(target), rend)
    img = render(theta, rend)
    loss = ((img.double() - target_img.double()) ** 2).sum()
    loss.backward()
    dldi = (2 * (img.detach().double() - target_img.double())).float().numpy()
    rend.set_material_params(theta.detach().numpy(), which=3)
    rend.RenderFrame()
    g = rend.backward(dldi)
    np.testing.assert_array_equal(theta.grad.numpy(), g)
    # a few gradient steps reduce the loss
    opt = torch.optim.Adam([theta], lr=0.02)
    losses = []
    for _ in range(6):
        opt.zero_grad()
        img = render(theta, rend)
        l2 = ((img.double() - target_img.double()) ** 2).sum()
        l2.backward()
        opt.step()
        with torch.no_grad():
            theta.clamp_(torch.tensor(scenes.PARAM_LO), torch.tensor(scenes.PARAM_HI))
        losses.append(float(l2))
    assert losses[-1] < losses[0]


def test_backward_refuses_a_stale_frame(rend):
    """trt_backward replays the last frame: after a camera, material-parameter or scene
    change it refuses (the frame's per-path radiance no longer matches the state) until
    the frame is rendered again, and it refuses options other than the frame's."""
    from trayracer_amd._lib import TrtError
    sc, theta, target_img, img, dldi = _setup(rend, 32, 32, 4)
    g = rend.backward(dldi)
    rend.SetCamera(dict(sc["camera"], position=(0.5, 0.5, -1.5)))
    with pytest.raises(TrtError, match="changed since"):
        rend.backward(dldi)
    rend.SetCamera(sc["camera"])
    rend.RenderFrame()  # the same frame again (accumulation restarted at sample 0)
    assert np.array_equal(rend.backward(dldi), g)
    rend.set_material_params(theta, which=rend.EVAL)
    with pytest.raises(TrtError, match="changed since"):
        rend.backward(dldi)
    rend.RenderFrame()
    rend.settings.seed += 1
    with pytest.raises(TrtError, match="same options"):
        rend.backward(dldi)
    rend.settings.seed -= 1
    assert np.array_equal(rend.backward(dldi), g)
    rend.UpdateScene(sc)
    with pytest.raises(TrtError, match="changed since"):
        rend.backward(dldi)


def test_autograd_nodes_rendered_before_backward(rend):
    """Two renders before one backward: each node's gradient is the adjoint of its own
    frame (the wrapper renders a node's frame again when the renderer has rendered
    another one since)."""
    import torch

    from trayracer_amd.autograd import render
    sc, target = scenes.cornell_diff(width=32, height=32, spp=8, max_depth=4, seed=6)
    rend.UpdateScene(sc)
    a = torch.tensor(scenes.material_params(sc["materials"]), requires_grad=True)
    b = torch.tensor(target, requires_grad=True)
    rng = np.random.default_rng(4)
    wa = torch.tensor(rng.normal(size=(32, 32, 3)).astype(np.float32))
    wb = torch.tensor(rng.normal(size=(32, 32, 3)).astype(np.float32))
    ia, ib = render(a, rend), render(b, rend)
    ((ia * wa).sum() + (ib * wb).sum()).backward()
    for leaf, w in ((a, wa), (b, wb)):
        rend.set_material_params(leaf.detach().numpy(), which=rend.BOTH)
        rend.RenderFrame()
        np.testing.assert_array_equal(leaf.grad.numpy(), rend.backward(w.numpy()))
