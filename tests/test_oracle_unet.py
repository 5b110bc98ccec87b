"""Pins the FP64 reference of the U-Net denoiser: its convolution against
torch.nn.functional.conv2d (an independent library), and its backward pass against
central finite differences of its own forward."""
import numpy as np
import torch

from oracle import unet_ref as U


def test_conv3x3_matches_torch():
    rng = np.random.default_rng(0)
    x = rng.normal(size=(2, 9, 12, 5))
    w = rng.normal(size=(7, 9, 5))
    b = rng.normal(size=7)
    y = U.conv3x3(x, w, b)
    tw = torch.tensor(w.reshape(7, 3, 3, 5).transpose(0, 3, 1, 2))  # [Co, Ci, ky, kx]
    ty = torch.nn.functional.conv2d(torch.tensor(x.transpose(0, 3, 1, 2)), tw, torch.tensor(b), padding=1)
    np.testing.assert_allclose(y, ty.numpy().transpose(0, 2, 3, 1), rtol=1e-10, atol=1e-10)


def test_conv_gradients_match_torch_autograd():
    rng = np.random.default_rng(1)
    x = rng.normal(size=(1, 6, 8, 4))
    w = rng.normal(size=(5, 9, 4))
    dz = rng.normal(size=(1, 6, 8, 5))
    tx = torch.tensor(x.transpose(0, 3, 1, 2), requires_grad=True)
    tw = torch.tensor(w.reshape(5, 3, 3, 4).transpose(0, 3, 1, 2), requires_grad=True)
    ty = torch.nn.functional.conv2d(tx, tw, padding=1)
    (ty * torch.tensor(dz.transpose(0, 3, 1, 2))).sum().backward()
    np.testing.assert_allclose(U.conv3x3_dgrad(dz, w), tx.grad.numpy().transpose(0, 2, 3, 1), rtol=1e-10, atol=1e-10)
    np.testing.assert_allclose(U.conv3x3_wgrad(x, dz), tw.grad.numpy().transpose(0, 2, 3, 1).reshape(5, 9, 4),
                               rtol=1e-10, atol=1e-10)


def test_unet_backward_matches_finite_differences():
    rng = np.random.default_rng(2)
    params = U.init_params(3).astype(np.float64)
    feats = rng.normal(size=(1, 8, 8, U.FEATURES))
    feats[..., 10:] = 0
    dout = rng.normal(size=(1, 8, 8, 32))
    net = U.UnetRef(params, round_bf16=False)
    net.forward(feats)
    g = net.backward(dout)
    assert g.size == U.num_params()
    # probe parameters of several layers (weights and biases) with large gradients
    eps = 1e-5
    offs = np.cumsum([0] + [co * 9 * ci + co for ci, co, _ in U.SPEC])
    picks = []
    for li in (0, 3, 7, 8, 12, 14):
        lo, hi = offs[li], offs[li + 1]
        picks.append(lo + int(np.argmax(np.abs(g[lo:hi]))))
    for i in picks:
        pp, pm = params.copy(), params.copy()
        pp[i] += eps
        pm[i] -= eps
        lp = (U.UnetRef(pp, round_bf16=False).forward(feats) * dout).sum()
        lm = (U.UnetRef(pm, round_bf16=False).forward(feats) * dout).sum()
        fd = (lp - lm) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-5 * max(1.0, abs(g[i])), (i, fd, g[i])


def test_bf16_rounding_is_nearest_even():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, -2.5, 3.0e38], np.float32)
    r = U.bf16(x)
    assert r[0] == 1.0 and r[1] == 1.0  # tie -> even
    assert r[2] == 1.0 + 2 ** -7
    assert r[3] == -2.5
