"""The host-only C++ (dense eigensolver, Rayleigh-Ritz step, Matrix Market
reader, and the oracle's Lanczos solve) built with g++ under AddressSanitizer
and UndefinedBehaviorSanitizer and run on small cases (tests/cpp/sanitize_host.cc).
No device code is built here; the HIP kernels are covered by the GPU tests."""

import os
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
SOURCES = ["tests/cpp/sanitize_host.cc", "trl_amd/csrc/dense_eigen.cc", "trl_amd/csrc/mmio.cc", "oracle/trl_oracle.cc"]
FLAGS = ["-std=c++20", "-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fopenmp"]


def test_host_code_under_asan_and_ubsan(tmp_path):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not found")
    exe = tmp_path / "sanitize_host"
    cmd = [gxx, *FLAGS, f"-I{ROOT / 'include'}", f"-I{ROOT / 'trl_amd' / 'csrc'}", *[str(ROOT / s) for s in SOURCES], "-o", str(exe)]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
    if r.returncode != 0 and "asan" in r.stderr.lower() and "cannot find" in r.stderr.lower():
        pytest.skip("the sanitizer runtimes are not installed")
    assert r.returncode == 0, r.stderr[-4000:]
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1", OMP_NUM_THREADS="4")
    run = subprocess.run([str(exe), str(tmp_path)], capture_output=True, text=True, timeout=300, env=env)
    print(run.stdout, run.stderr[-4000:])
    assert run.returncode == 0 and run.stdout.strip() == "OK"
    assert "runtime error" not in run.stderr and "AddressSanitizer" not in run.stderr
