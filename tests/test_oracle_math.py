"""Pins the oracle's random-number, transcendental and sampling primitives.

The reference has no tests or golden vectors (its RNG is TEA-2 + LCG,
thirdparty/CUDA/random.h:31-67, replaced here by PCG32), so each primitive is pinned by
a published known-answer vector, an independent re-implementation in Python integers,
float64 library functions, or a chi-square test of the distribution it must produce.
"""
import math

import numpy as np
import pytest
from scipy import stats

from oracle import oracle_py as O

M64 = (1 << 64) - 1


def _pcg_ref(initstate, initseq, n):
    """pcg32 (O'Neill 2014, pcg-c-basic) in Python integers."""
    state, inc = 0, ((initseq << 1) | 1) & M64

    def step(s):
        return (s * 6364136223846793005 + inc) & M64

    out = []
    state = step(state)
    state = (state + initstate) & M64
    state = step(state)
    for _ in range(n):
        old = state
        state = step(state)
        xs = (((old >> 18) ^ old) >> 27) & 0xFFFFFFFF
        rot = old >> 59
        out.append(((xs >> rot) | (xs << ((-rot) & 31))) & 0xFFFFFFFF)
    return out


def _mix64(z):
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
    return z ^ (z >> 31)


def test_pcg32_published_vector():
    # pcg32-demo (pcg-c-basic): pcg32_srandom_r(&rng, 42u, 54u), first six outputs
    got = O.pcg32(42, 54, 6)
    assert [int(x) for x in got] == [0xA15C02B7, 0x7B47F409, 0xBA1D3330, 0x83D2F293, 0xBFA4784B, 0xCBED606E]


@pytest.mark.parametrize("seed", [(0, 0), (1, 2), (0xDEADBEEF12345678, 0x0123456789ABCDEF), (M64, M64)])
def test_pcg32_matches_integer_reimplementation(seed):
    got = O.pcg32(seed[0], seed[1], 257)
    assert [int(x) for x in got] == _pcg_ref(seed[0], seed[1], 257)


def test_rng_streams_keyed_by_pixel_sample_depth():
    seed, pixel, sample, depth = 7, 12345, 3, 2
    k0 = _mix64((pixel << 32) | sample)
    k1 = _mix64((seed << 32) | depth)
    ref = _pcg_ref(k0, k1 ^ 0x9E3779B97F4A7C15, 16)
    got = O.rng_floats(seed, pixel, sample, depth, 16)
    np.testing.assert_array_equal(got, np.array([(x >> 8) * 2.0 ** -24 for x in ref], np.float32))
    a = O.rng_floats(seed, pixel, sample, depth + 1, 16)
    b = O.rng_floats(seed, pixel + 1, sample, depth, 16)
    assert not np.array_equal(got, a) and not np.array_equal(got, b)
    big = np.concatenate([O.rng_floats(1, p, 0, 0, 8) for p in range(2000)])
    assert big.min() >= 0.0 and big.max() < 1.0
    assert stats.kstest(big, "uniform").pvalue > 1e-3


def test_sincos_turns_accuracy():
    rng = np.random.default_rng(0)
    u = np.concatenate([rng.uniform(-3, 3, 200000), np.linspace(0, 1, 4097)]).astype(np.float32)
    s, c = O.sincos_turns(u)
    ud = u.astype(np.float64)
    assert np.max(np.abs(s - np.sin(2 * np.pi * ud))) < 3e-7
    assert np.max(np.abs(c - np.cos(2 * np.pi * ud))) < 3e-7
    # exact special values
    s, c = O.sincos_turns(np.array([0.0, 0.25, 0.5, 0.75], np.float32))
    np.testing.assert_array_equal(s, [0.0, 1.0, 0.0, -1.0])
    np.testing.assert_array_equal(c, [1.0, 0.0, -1.0, 0.0])


def test_log_exp_atan_tan_accuracy():
    rng = np.random.default_rng(1)
    x = np.exp(rng.uniform(-80, 80, 100000)).astype(np.float32)
    x = np.concatenate([x, np.array([1.0, 2.0, 0.5, 1e-38, 3e38], np.float32)])
    y = O.log(x)
    ref = np.log(x.astype(np.float64))
    assert np.max(np.abs(y - ref) / np.maximum(np.abs(ref), 1.0)) < 3e-7
    assert O.log(np.array([1.0], np.float32))[0] == 0.0
    z = rng.uniform(-80, 80, 100000).astype(np.float32)
    e = O.exp(z)
    assert np.max(np.abs(e / np.exp(z.astype(np.float64)) - 1)) < 5e-7
    a = np.concatenate([rng.uniform(-50, 50, 50000), rng.uniform(-1, 1, 50000)]).astype(np.float32)
    assert np.max(np.abs(O.atan(a) - np.arctan(a.astype(np.float64)))) < 5e-7  # <= 4 ulp near pi/2
    t = rng.uniform(-1.5, 1.5, 50000).astype(np.float32)
    tr = np.tan(t.astype(np.float64))
    # the radians->turns product rounds the argument: relative error grows with tan^2 near pi/2
    assert np.max(np.abs(O.tan(t) - tr) / np.maximum(np.abs(tr), 1.0) / (1.0 + tr ** 2) * np.maximum(np.abs(tr), 1.0)) < 1e-6


def _chi2_uniform(values, bins):
    h, _ = np.histogram(values, bins=bins, range=(0.0, 1.0))
    return stats.chisquare(h).pvalue


def test_cosine_hemisphere_distribution():
    rng = np.random.default_rng(2)
    w = O.cosine_hemisphere(rng.random((200000, 2)))
    np.testing.assert_allclose(np.linalg.norm(w, axis=1), 1.0, atol=2e-7)
    assert (w[:, 2] >= 0).all()
    # pdf cos(theta)/pi  =>  cos^2(theta) ~ U[0,1], phi ~ U[0, 2pi)
    assert _chi2_uniform(w[:, 2] ** 2, 50) > 1e-3
    phi = (np.arctan2(w[:, 1], w[:, 0]) / (2 * np.pi)) % 1.0
    assert _chi2_uniform(phi, 50) > 1e-3


def test_triangle_sampling_uniform():
    rng = np.random.default_rng(3)
    b = O.sample_triangle(rng.random((300000, 2))).astype(np.float64)
    b0 = 1 - b[:, 0] - b[:, 1]
    assert (b >= 0).all() and (b0 >= -1e-7).all()
    # 16 congruent sub-triangles of a 4x4 barycentric grid have equal probability
    n = 4
    i = np.minimum((b[:, 0] * n).astype(int), n - 1)
    j = np.minimum((b[:, 1] * n).astype(int), n - 1)
    up = (b[:, 0] * n - i) + (b[:, 1] * n - j) > 1.0
    cell = np.where(i + j < n, (i * n + j) * 2 + up, -1)
    assert (cell >= 0).all()
    counts = np.bincount(cell, minlength=2 * n * n)
    valid = [(ii * n + jj) * 2 + u for ii in range(n) for jj in range(n) for u in (0, 1)
             if ii + jj < n - 1 or (ii + jj == n - 1 and u == 0)]
    assert counts.sum() == counts[valid].sum()
    assert stats.chisquare(counts[valid]).pvalue > 1e-3


def test_concentric_disk_uniform():
    rng = np.random.default_rng(4)
    d = O.concentric_disk(rng.random((200000, 2))).astype(np.float64)
    r2 = (d ** 2).sum(axis=1)
    assert r2.max() <= 1.0 + 1e-6
    assert _chi2_uniform(np.minimum(r2, 1.0), 40) > 1e-3
    assert _chi2_uniform((np.arctan2(d[:, 1], d[:, 0]) / (2 * np.pi)) % 1.0, 40) > 1e-3


def test_frame_orthonormal():
    rng = np.random.default_rng(5)
    n = rng.normal(size=(20000, 3))
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    n = np.concatenate([n, [[0, 0, 1], [0, 0, -1], [1, 0, 0], [0, 1, 0]]]).astype(np.float32)
    F = O.frame(n).astype(np.float64)
    G = np.einsum("nij,nkj->nik", F, F)
    assert np.max(np.abs(G - np.eye(3))) < 1e-6
    np.testing.assert_allclose(np.cross(F[:, 0], F[:, 1]), F[:, 2], atol=1e-6)


@pytest.mark.parametrize("n", [1, 2, 7, 64, 10000])
def test_alias_table_is_exact(n):
    rng = np.random.default_rng(n)
    w = np.exp(rng.normal(0, 2, n))
    prob, alias = O.alias_build(w)
    assert (alias < n).all()
    p = prob.astype(np.float64) / n
    np.add.at(p, alias, (1.0 - prob.astype(np.float64)) / n)
    np.testing.assert_allclose(p, w / w.sum(), rtol=2e-6, atol=1e-12)
