"""Builds the gfx950 shared library ``libtrt.so`` (HIP kernels + C ABI) in-tree.

Flags: every source goes through hipcc for gfx950 with ``-ffp-contract=off`` (the
device and host math must perform the same rounded operations as the host-compiled
oracle), no fast-math, and packed-FP32 instruction selection disabled.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libtrt.so")
SOURCES = ["api.cpp", "bvh_build.cpp", "kernels.hip", "denoiser.cpp", "denoiser_kernels.hip", "scene_file.cpp",
           "scene_file_api.cpp", "image_io.cpp", "probe.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("TRT_ARCH", "gfx950")


def hip_flags():
    return [
        f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-I", os.path.join(HERE, "..", "include"), "-I", CSRC,
        "-Wno-unused-command-line-argument",
    ]


def build(verbose=False, jobs=8):
    """Compiles each source to an object in parallel, then links libtrt.so."""
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    procs = []
    objs = []
    for src in SOURCES:
        obj = os.path.join(objdir, src.replace(".", "_") + ".o")
        objs.append(obj)
        lang = ["-x", "hip"] if src.endswith(".hip") else ["-x", "hip"]
        cmd = [HIPCC] + hip_flags() + lang + ["-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
    failed = []
    for src, p in procs:
        out = p.communicate()[0].decode(errors="replace")
        if p.returncode != 0:
            failed.append((src, out))
        elif verbose and out.strip():
            print(out)
    if failed:
        for src, out in failed:
            sys.stderr.write(f"--- {src} ---\n{out}\n")
        raise RuntimeError("hipcc failed for: " + ", ".join(s for s, _ in failed))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs + ["-lpthread"]
    subprocess.check_call(cmd)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
