"""Host code under the sanitizers (CPU).

tests/native/sanitize_driver.cpp drives the host BVH builder and pre-splits, the
scene-file readers (valid and malformed inputs), the image writers and the CPU oracle on
several threads. It is built twice with g++ -- AddressSanitizer + UndefinedBehaviorSanitizer
(every finding fatal), and ThreadSanitizer -- and each build must print "ok" with no
sanitizer report. The HIP side (api.cpp and the kernels) needs a GPU and is covered by the
GPU tests instead.
"""
import pathlib
import shutil
import subprocess

import pytest

ROOT = pathlib.Path(__file__).resolve().parents[1]
CSRC = ROOT / "trayracer_amd" / "csrc"
SOURCES = [ROOT / "tests" / "native" / "sanitize_driver.cpp", ROOT / "oracle" / "oracle.cpp",
           CSRC / "bvh_build.cpp", CSRC / "scene_file.cpp", CSRC / "image_io.cpp"]
FLAGS = {
    "asan_ubsan": ["-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer"],
    "tsan": ["-fsanitize=thread"],
}


@pytest.fixture(scope="module")
def builds(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    out = tmp_path_factory.mktemp("sanitize")
    procs = {}
    for name, flags in FLAGS.items():
        cmd = ["g++", "-O1", "-g", "-std=c++17", "-pthread", "-ffp-contract=off", "-fno-fast-math", *flags,
               "-I", str(CSRC), *map(str, SOURCES), "-o", str(out / name)]
        procs[name] = subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    for name, p in procs.items():
        log, _ = p.communicate(timeout=600)
        assert p.returncode == 0, f"{name} build failed:\n{log[-4000:]}"
    return out


@pytest.mark.parametrize("name", list(FLAGS))
def test_host_code_is_clean_under_sanitizers(builds, name, tmp_path):
    env = {"ASAN_OPTIONS": "detect_leaks=1:abort_on_error=0:halt_on_error=1",
           "UBSAN_OPTIONS": "print_stacktrace=1:halt_on_error=1",
           "TSAN_OPTIONS": "halt_on_error=1:second_deadlock_stack=1", "PATH": "/usr/bin:/bin"}
    r = subprocess.run([str(builds / name), str(tmp_path)], capture_output=True, text=True, timeout=600, env=env)
    report = r.stdout + r.stderr
    assert r.returncode == 0 and r.stdout.strip().endswith("ok"), report[-6000:]
    assert "Sanitizer" not in report and "runtime error" not in report, report[-6000:]
