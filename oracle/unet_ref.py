"""FP64 reference of the U-Net denoiser (TEST INFRASTRUCTURE ONLY).

The reference repository's denoiser is the closed OptiX AI denoiser
(Tracer/Optix/Denoiser.cpp:36-42, 101-112), which cannot run here; this module
restates the project's own network (4 levels, 32/64/96/128 channels, 3x3 'same'
convolutions with ReLU, 2x2 average pooling, nearest upsampling + skip concatenation,
final 3x3 convolution to 3 channels padded to 32) in float64 numpy. With
`round_bf16=True` it rounds to bfloat16 exactly where the device stores bfloat16
(input features, weights, every activation, and the ReLU-masked gradients fed to the
gradient convolutions), so the device result differs only by FP32-vs-FP64 accumulation.
Its convolution is pinned against torch.nn.functional.conv2d and its backward against
central finite differences (tests/test_oracle_unet.py).

Parameter layout (matches the device): per convolution, weights [cout][9][cin]
(tap = 3*ky + kx) then bias [cout], convolutions in the order of SPEC.
"""
import numpy as np

FEATURES = 16
# (cin, cout, relu) in device order: e0a e0b e1a e1b e2a e2b e3a e3b d2a d2b d1a d1b d0a d0b out
SPEC = [(16, 32, True), (32, 32, True), (32, 64, True), (64, 64, True), (64, 96, True), (96, 96, True),
        (96, 128, True), (128, 128, True), (224, 96, True), (96, 96, True), (160, 64, True), (64, 64, True),
        (96, 32, True), (32, 32, True), (32, 32, False)]


def num_params():
    return sum(co * 9 * ci + co for ci, co, _ in SPEC)


def unpack(params):
    out, off = [], 0
    for ci, co, _ in SPEC:
        w = params[off:off + co * 9 * ci].reshape(co, 9, ci)
        off += co * 9 * ci
        b = params[off:off + co]
        off += co
        out.append((w, b))
    return out


def bf16(x):
    """Round-to-nearest-even to bfloat16 (finite inputs), returned as float64."""
    a = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    r = (((a + 0x7FFF + ((a >> 16) & 1)) >> 16) << 16).astype(np.uint32)
    return r.view(np.float32).astype(np.float64)


def conv3x3(x, w, b=None):
    """'same' 3x3 convolution, x [N,H,W,Ci], w [Co,9,Ci] -> [N,H,W,Co] (float64)."""
    N, H, W, _ = x.shape
    xp = np.zeros((N, H + 2, W + 2, x.shape[3]))
    xp[:, 1:-1, 1:-1] = x
    y = np.zeros((N, H, W, w.shape[0]))
    for t in range(9):
        ky, kx = divmod(t, 3)
        y += np.einsum("nhwc,oc->nhwo", xp[:, ky:ky + H, kx:kx + W], w[:, t, :], optimize=True)
    if b is not None:
        y += b
    return y


def conv3x3_wgrad(x, dz):
    """dW[o,t,c] = sum_p dz[p,o] x[p + s(t), c]."""
    N, H, W, C = x.shape
    xp = np.zeros((N, H + 2, W + 2, C))
    xp[:, 1:-1, 1:-1] = x
    dw = np.zeros((dz.shape[3], 9, C))
    for t in range(9):
        ky, kx = divmod(t, 3)
        dw[:, t, :] = np.einsum("nhwo,nhwc->oc", dz, xp[:, ky:ky + H, kx:kx + W], optimize=True)
    return dw


def conv3x3_dgrad(dz, w):
    """dX = 3x3 convolution of dz with the flipped, transposed kernel."""
    wf = np.transpose(w[:, ::-1, :], (2, 1, 0))  # [Ci, 9, Co]
    return conv3x3(dz, wf)


def pool(x, r):
    y = 0.25 * ((x[:, 0::2, 0::2] + x[:, 0::2, 1::2]) + (x[:, 1::2, 0::2] + x[:, 1::2, 1::2]))
    return r(y)


def pool_bwd(dy):
    return 0.25 * np.repeat(np.repeat(dy, 2, axis=1), 2, axis=2)


def upcat(coarse, skip):
    up = np.repeat(np.repeat(coarse, 2, axis=1), 2, axis=2)
    return np.concatenate([up, skip], axis=3)


def upcat_bwd(dout, c1):
    du, ds = dout[..., :c1], dout[..., c1:]
    dc = (du[:, 0::2, 0::2] + du[:, 0::2, 1::2]) + (du[:, 1::2, 0::2] + du[:, 1::2, 1::2])
    return dc, ds


class UnetRef:
    def __init__(self, params, round_bf16=True):
        self.r = bf16 if round_bf16 else (lambda a: np.asarray(a, np.float64))
        p = np.asarray(params, np.float32 if round_bf16 else np.float64)
        self.layers = [(self.r(w), np.asarray(b, np.float64)) for w, b in unpack(p)]

    def _conv(self, i, x):
        w, b = self.layers[i]
        y = conv3x3(x, w, b)
        if SPEC[i][2]:
            y = np.maximum(y, 0.0)
        return y if i == 14 else self.r(y)

    def forward(self, feats):
        r = self.r
        x = r(feats)
        a = {}
        a["x"] = x
        a[0] = self._conv(0, x)
        a[1] = self._conv(1, a[0])
        a["p0"] = pool(a[1], r)
        a[2] = self._conv(2, a["p0"])
        a[3] = self._conv(3, a[2])
        a["p1"] = pool(a[3], r)
        a[4] = self._conv(4, a["p1"])
        a[5] = self._conv(5, a[4])
        a["p2"] = pool(a[5], r)
        a[6] = self._conv(6, a["p2"])
        a[7] = self._conv(7, a[6])
        a["u2"] = upcat(a[7], a[5])
        a[8] = self._conv(8, a["u2"])
        a[9] = self._conv(9, a[8])
        a["u1"] = upcat(a[9], a[3])
        a[10] = self._conv(10, a["u1"])
        a[11] = self._conv(11, a[10])
        a["u0"] = upcat(a[11], a[1])
        a[12] = self._conv(12, a["u0"])
        a[13] = self._conv(13, a[12])
        a[14] = self._conv(14, a[13])
        self.a = a
        return a[14]

    def use_activations(self, feats, acts):
        """Replaces the forward activations by given ones (e.g. the device's), recomputing
        the pooled and concatenated tensors from them, so a backward comparison is not
        perturbed by ReLU masks that flip on near-zero pre-activations."""
        r = self.r
        a = {"x": r(feats)}
        for i in range(14):
            a[i] = np.asarray(acts[i], np.float64)
        a["p0"], a["p1"], a["p2"] = pool(a[1], r), pool(a[3], r), pool(a[5], r)
        a["u2"], a["u1"], a["u0"] = upcat(a[7], a[5]), upcat(a[9], a[3]), upcat(a[11], a[1])
        a[14] = self._conv(14, a[13])
        self.a = a

    def backward(self, dout):
        """Gradient of sum(dout * output) w.r.t. the parameters (flat, device layout)."""
        a, r = self.a, self.r
        grads = [None] * 15

        def layer(i, x, gin, need_dx=True):
            w, _ = self.layers[i]
            dz = gin * (a[i] > 0) if SPEC[i][2] else gin
            dz = r(dz)
            grads[i] = (conv3x3_wgrad(x, dz), dz.sum(axis=(0, 1, 2)))
            return conv3x3_dgrad(dz, w) if need_dx else None

        g = layer(14, a[13], dout)
        g = layer(13, a[12], g)
        g = layer(12, a["u0"], g)
        g_d1b, gs0 = upcat_bwd(g, 64)
        g = layer(11, a[10], g_d1b)
        g = layer(10, a["u1"], g)
        g_d2b, gs1 = upcat_bwd(g, 96)
        g = layer(9, a[8], g_d2b)
        g = layer(8, a["u2"], g)
        g_e3b, gs2 = upcat_bwd(g, 128)
        g = layer(7, a[6], g_e3b)
        g = layer(6, a["p2"], g)
        gs2 = gs2 + pool_bwd(g)
        g = layer(5, a[4], gs2)
        g = layer(4, a["p1"], g)
        gs1 = gs1 + pool_bwd(g)
        g = layer(3, a[2], gs1)
        g = layer(2, a["p0"], g)
        gs0 = gs0 + pool_bwd(g)
        g = layer(1, a[0], gs0)
        layer(0, a["x"], g, need_dx=False)
        return np.concatenate([np.concatenate([gw.reshape(-1), gb]) for gw, gb in grads])


def init_params(seed):
    """Random parameters for tests (He-normal over the real fan-in, zero padding)."""
    rng = np.random.default_rng(seed)
    parts = []
    for i, (ci, co, _) in enumerate(SPEC):
        w = np.zeros((co, 9, ci), np.float32)
        real_ci = 10 if i == 0 else ci
        real_co = 3 if i == 14 else co
        w[:real_co, :, :real_ci] = rng.normal(0, np.sqrt(2.0 / (9 * real_ci)), (real_co, 9, real_ci))
        b = np.zeros(co, np.float32)
        b[:real_co] = rng.normal(0, 0.05, real_co)
        parts += [w.reshape(-1), b]
    return np.concatenate(parts).astype(np.float32)
