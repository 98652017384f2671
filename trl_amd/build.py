"""In-tree build of the native parts of trl_amd.

Products (all written next to their sources, git-ignored, shipped with the tree):

* ``trl_amd/lib/libtrl_hip.so``        gfx950 kernels + the C-ABI of ``include/trl_hip.h``
* ``trl_amd/lib/libtrl_hip_solver.so`` C++20 block-Lanczos driver over the ``trl::hip``
  backend + the C-ABI of ``include/trl_hip_solver.h``
* ``trl_amd/lib/test_*``               C++ acceptance tests that mirror the
  reference's ``tests/test_lanczos_{extend,convergence}.hh``

Usage: ``python -m trl_amd.build [--incremental] [--jobs N]``.
"""

from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent
PKG = ROOT / "trl_amd"
CSRC = PKG / "csrc"
LIB = PKG / "lib"
OBJ = PKG / "lib" / "obj"
INCLUDE = ROOT / "include"

ARCH = "gfx950"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# Packed FP32 instructions are kept out of every code object.
HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-fPIC",
    "-std=c++20",
    "-Xclang",
    "-target-feature",
    "-Xclang",
    "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
    "-Wno-unused-variable",
    "-Wno-unused-but-set-variable",
    f"-I{INCLUDE}",
    f"-I{CSRC}",
]
HOST_FLAGS = ["-O3", "-fPIC", "-std=c++20", "-Wall", "-Wextra", f"-I{INCLUDE}", f"-I{CSRC}", "-fopenmp"]

KERNEL_SOURCES = ["capi.hip", "k_panel.hip", "k_block.hip", "k_spmm.hip", "k_extra.hip", "k_batch.hip", "k_solve.hip"]
SOLVER_SOURCES = ["solver.cc", "dense_eigen.cc", "batch_solver.cc", "mmio.cc"]


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(line for line in r.stdout.splitlines() if "packed-fp32-ops" not in line)
    if r.returncode != 0:
        raise RuntimeError(f"command failed ({r.returncode}): {' '.join(cmd)}\n{out}")
    if out.strip():
        print(out, file=sys.stderr)


def _stale(target: Path, deps: list[Path]) -> bool:
    if not target.exists():
        return True
    t = target.stat().st_mtime
    return any(d.stat().st_mtime > t for d in deps if d.exists())


def _headers() -> list[Path]:
    return list(INCLUDE.rglob("*.h")) + list(INCLUDE.rglob("*.hh")) + list(CSRC.glob("*.hh"))


def build(incremental: bool = False, jobs: int | None = None, verbose: bool = True) -> dict[str, Path]:
    LIB.mkdir(parents=True, exist_ok=True)
    OBJ.mkdir(parents=True, exist_ok=True)
    jobs = jobs or min(8, os.cpu_count() or 4)
    hdrs = _headers()

    def compile_hip(src: str) -> Path:
        s = CSRC / src
        o = OBJ / (s.stem + ".o")
        if incremental and not _stale(o, [s] + hdrs):
            return o
        _run([HIPCC, *HIP_FLAGS, "-c", str(s), "-o", str(o)])
        return o

    def compile_host(src: str) -> Path:
        s = CSRC / src
        o = OBJ / (s.stem + ".host.o")
        if incremental and not _stale(o, [s] + hdrs):
            return o
        # host-only C++ goes through hipcc too (HIP runtime headers), no device code inside
        _run([HIPCC, "-O3", "-fPIC", "-std=c++20", "-Wall", "-Wextra", f"-I{INCLUDE}", f"-I{CSRC}", "-fopenmp", "-c", str(s), "-o", str(o)])
        return o

    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        kobjs = list(ex.map(compile_hip, KERNEL_SOURCES))
        sobjs = list(ex.map(compile_host, SOLVER_SOURCES))

    lib = LIB / "libtrl_hip.so"
    if not incremental or _stale(lib, kobjs):
        _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(lib), *map(str, kobjs)])
    slib = LIB / "libtrl_hip_solver.so"
    if not incremental or _stale(slib, sobjs + [lib]):
        _run([HIPCC, "-shared", "-fPIC", "-fopenmp", "-o", str(slib), *map(str, sobjs), f"-L{LIB}", "-ltrl_hip", "-Wl,-rpath,$ORIGIN"])

    products = {"libtrl_hip": lib, "libtrl_hip_solver": slib}

    # C++ acceptance tests (mirror the reference's test templates)
    tests_dir = ROOT / "tests" / "cpp"
    test_srcs = sorted(tests_dir.glob("test_*.cc")) if tests_dir.exists() else []

    def compile_test(s: Path) -> Path:
        exe = LIB / s.stem
        if incremental and not _stale(exe, [s] + hdrs + [slib, lib]):
            return exe
        _run([HIPCC, "-O2", "-std=c++20", "-Wall", "-Wextra", f"-I{INCLUDE}", f"-I{tests_dir}", "-fopenmp", str(s), "-o", str(exe), f"-L{LIB}", "-ltrl_hip_solver",
              "-ltrl_hip", "-Wl,-rpath,$ORIGIN"])
        return exe

    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        for exe in ex.map(compile_test, test_srcs):
            products[exe.name] = exe

    if verbose:
        for k, v in products.items():
            print(f"built {k}: {v.relative_to(ROOT)}")
    return products


def clean() -> None:
    if LIB.exists():
        shutil.rmtree(LIB)


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--incremental", action="store_true")
    ap.add_argument("--jobs", type=int, default=None)
    a = ap.parse_args(argv)
    build(incremental=a.incremental, jobs=a.jobs)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
