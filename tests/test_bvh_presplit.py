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
