"""Spatial pre-splitting (host BVH builder): coverage, budget and tree invariants.

Compiles tests/native/presplit_check.cpp against trayracer_amd/csrc/bvh_build.cpp with
g++ and runs it for several budgets; the GPU parity tests then check that traversal
results do not depend on the budget.
"""
import pathlib
import subprocess

import pytest

ROOT = pathlib.Path(__file__).resolve().parents[1]


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    exe = tmp_path_factory.mktemp("presplit") / "presplit_check"
    src = ROOT / "tests" / "native" / "presplit_check.cpp"
    subprocess.run(["g++", "-O2", "-std=c++17", "-pthread", "-I", str(ROOT / "trayracer_amd" / "csrc"), str(src),
                    str(ROOT / "trayracer_amd" / "csrc" / "bvh_build.cpp"), "-o", str(exe)], check=True)
    return exe


@pytest.mark.parametrize("beta", [0.0, 0.5, 1.0, 3.0])
def test_presplit_invariants(checker, beta):
    out = subprocess.run([str(checker), str(beta), "20000"], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0 and out.stdout.startswith("ok"), out.stdout + out.stderr
    refs = int(out.stdout.split()[2])
    assert refs >= 20000
    if beta > 0:
        assert refs > 20000  # large triangles were split


@pytest.mark.parametrize("presplit", [0.0, 6.0])
def test_product_tree_on_the_host_equals_brute_force(presplit):
    """The product's binned-SAH tree (with pre-split references), traversed on the host by
    the CPU baseline, returns the brute-force closest and any hits: a host check of the
    builder's trees independent of the GPU kernels."""
    import numpy as np

    from oracle import Oracle
    from tests.gpu_util import random_rays
    from trayracer_amd import scenes
    sc = scenes.soup(n_tris=20_000, seed=8, width=32, height=32)
    orc = Oracle(sc)
    rays = random_rays(3000, seed=21, lo=-0.3, hi=1.3, tmin=1e-5)
    hb, ib = orc.trace_closest(rays, brute=True)
    sr = rays.copy()
    sr[:, 7] = np.random.default_rng(22).uniform(1e-3, 0.4, sr.shape[0]).astype(np.float32)
    ob = orc.trace_any(sr, brute=True)
    assert orc.use_product_bvh(presplit) > 10
    hp, ip = orc.trace_closest(rays)
    assert np.array_equal(hp.view(np.uint32), hb.view(np.uint32)) and np.array_equal(ip, ib)
    assert np.array_equal(orc.trace_any(sr), ob)
