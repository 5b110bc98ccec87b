"""Helpers shared by the GPU parity tests."""
import numpy as np


def random_rays(n, seed, lo=-0.2, hi=1.2, tmin=1e-4, tmax=1e30, toward=None):
    """Rays with origins uniform in [lo,hi]^3 and uniform directions (or aimed at
    uniform points of `toward` = (lo, hi) box)."""
    rng = np.random.default_rng(seed)
    o = rng.uniform(lo, hi, (n, 3))
    if toward is None:
        d = rng.normal(size=(n, 3))
    else:
        d = rng.uniform(toward[0], toward[1], (n, 3)) - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = np.zeros((n, 8), np.float32)
    r[:, 0:3] = o
    r[:, 3] = tmin
    r[:, 4:7] = d
    r[:, 7] = tmax
    return r


def prim_of(hits):
    return hits[:, 3].view(np.uint32)
