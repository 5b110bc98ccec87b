"""The C++ façade (include/trt.hpp, the reference's Renderer verbs over the C ABI) from a
compiled C++ program: tests/native/facade_example.cpp is built with g++ against the
header and linked to libtrt.so. Without a GPU the program must report that no device is
usable (the library loads and its errors surface as exceptions); on the GPU it renders
the scene file and its mean colour equals the Python façade's render of the same file."""
import os
import pathlib
import subprocess

import numpy as np
import pytest

ROOT = pathlib.Path(__file__).resolve().parents[1]
SCENE = ROOT / "tests" / "golden" / "scenes" / "box_scene.json"


@pytest.fixture(scope="module")
def program(tmp_path_factory):
    exe = tmp_path_factory.mktemp("facade") / "facade_example"
    lib = ROOT / "trayracer_amd"
    subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-Werror", "-I", str(ROOT / "include"),
                    str(ROOT / "tests" / "native" / "facade_example.cpp"), "-L", str(lib), "-ltrt",
                    f"-Wl,-rpath,{lib}", "-o", str(exe)], check=True)
    return exe


def _run(program, frames=4):
    return subprocess.run([str(program), str(SCENE), str(frames)], capture_output=True, text=True, timeout=300)


def test_program_builds_and_reports_missing_device(program):
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present: covered by the GPU test")
    out = _run(program)
    assert out.returncode == 2 and out.stdout.startswith("no device"), out.stdout + out.stderr


@pytest.mark.gpu
def test_program_renders_like_the_python_facade(program):
    out = _run(program)
    assert out.returncode == 0, out.stdout + out.stderr
    fields = dict(kv.split("=") for kv in out.stdout.split()[1:])
    assert fields["error_caught"] == "1" and int(fields["samples"]) == 8
    from trayracer_amd import Renderer
    r = Renderer(0)
    try:
        r.LoadScene(str(SCENE), presplit=1.0)
        s = r.settings
        s.width, s.height, s.spp, s.seed = 128, 96, 2, 5
        r.Reset()
        for _ in range(4):
            r.RenderFrame()
        mean = r.color()[..., :3].astype(np.float64).reshape(-1, 3).mean(axis=0)
        pick = r.PickRay(64, 48)
    finally:
        r.close()
    got = np.array([float(x) for x in fields["mean"].split(",")])
    np.testing.assert_allclose(got, mean, rtol=1e-7)
    assert int(fields["pick_prim"]) == pick["primitive"]
