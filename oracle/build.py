"""Builds the CPU oracle (test infrastructure only) into oracle/_build/liboracle.so.

g++ compiles oracle.cpp together with the shared trt/ headers of the package with
-ffp-contract=off and no fast-math, so every floating-point operation is the same
correctly rounded IEEE operation the gfx950 build performs.

The reference itself (CUDA 11.8 + OptiX 7.5 + MSVC project files) cannot be built in
this environment, so there is no oracle/_ref build.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
OUT_DIR = os.path.join(HERE, "_build")
LIB = os.path.join(OUT_DIR, "liboracle.so")


def build(verbose=False):
    os.makedirs(OUT_DIR, exist_ok=True)
    tmp = LIB + ".tmp"
    cmd = [
        os.environ.get("CXX", "g++"), "-O2", "-std=c++17", "-fPIC", "-shared", "-pthread",
        "-ffp-contract=off", "-fno-fast-math", "-march=x86-64-v3",
        "-I", os.path.join(ROOT, "trayracer_amd", "csrc"),
        os.path.join(HERE, "oracle.cpp"),
        # the product's host BVH builder: only the CPU-baseline mode (orc_scene_use_product_bvh) uses it
        os.path.join(ROOT, "trayracer_amd", "csrc", "bvh_build.cpp"), "-o", tmp,
    ]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
