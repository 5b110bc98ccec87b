"""48-byte device node layout (host packing): child references and outward bounds.

Compiles tests/native/node_pack_check.cpp against trayracer_amd/csrc/bvh_build.cpp with
g++ and runs it; the GPU parity tests then check that traversal over packed nodes gives
the oracle's hits.
"""
import pathlib
import subprocess

ROOT = pathlib.Path(__file__).resolve().parents[1]


def test_node_pack(tmp_path):
    exe = tmp_path / "node_pack_check"
    subprocess.run(["g++", "-O2", "-std=c++17", "-pthread", "-I", str(ROOT / "trayracer_amd" / "csrc"),
                    str(ROOT / "tests" / "native" / "node_pack_check.cpp"),
                    str(ROOT / "trayracer_amd" / "csrc" / "bvh_build.cpp"), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe), "30000"], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0 and out.stdout.startswith("ok"), out.stdout + out.stderr
    assert int(out.stdout.split()[1]) > 1000
