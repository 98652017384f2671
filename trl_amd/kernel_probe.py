"""Kernel driver for profiling: runs single hot kernels of the 3D-Laplacian
workload a fixed number of times, so that a profiler (rocprofv3 kernel trace
or hardware counters) sees clean, repeated dispatches of one kernel.

    python -m trl_amd.kernel_probe --side 160 --b 8 --reps 20 --what spmm_axpby64,copy

Kernels (--what, comma separated):
  spmm32, spmm64        Y = A X with the autotuned operator (FP32 / FP64)
  spmm_axpby32/64       Y = alpha A X + beta X + gamma Y (the Chebyshev step, FP32 / FP64)
  panels32, panels64    the same step on --panels panels per launch (the refinement's filter)
  floor32               that FP32 step with the matrix cut to its diagonal (ELL of one slot,
                        tuned the same way): the same X, Z and Y streams with no gathered
                        neighbour rows, the floor of the step's memory pattern
  stream32              one-panel launches of that diagonal step, one after another over the
                        --panels panels: the same streams, one pass per panel
  forced32              that FP32 step plus the forcing term F diag(g) (the residual-driven
                        filter of the FP64 sweeps)
  gram32, update32      the block Gram-Schmidt products over k = 16 blocks (FP32)
  gram64, update64      the same over k = 32 blocks in FP64 (a full basis of the plain solve)
  rotate32, rotate64    the Ritz rotation of a refinement sweep (m = kc = --panels blocks)
  restart64             the FP64 Lanczos restart rotation (m = 32 blocks, kc = 5)
  restart32             the FP32 phase-1 restart rotation (m = 24 blocks, kc = 5)
  cholqr32, cholqr64    CholQR2 of one FP32 panel, CholQR of one FP64 panel
  copy                  a plain FP64 copy kernel of 512 MiB (the counter calibration)
Prints, per kernel, the mean time per launch (one HIP event pair on the
library stream around the timed launches, after --warmup untimed ones) and
the algorithmic bytes per launch.
"""

from __future__ import annotations

import argparse
import ctypes as C
import sys

import numpy as np

from . import _lib as L
from . import problems as P
from .solver import DeviceCSR


def diag_of(A: P.CSR) -> P.CSR:
    """The diagonal of A as a CSR matrix (one entry per row)."""
    rows = np.repeat(np.arange(A.n), np.diff(A.rowptr))
    d = np.zeros(A.n)
    on = A.col == rows
    d[rows[on]] = A.val[on]
    return P.CSR(A.n, np.arange(A.n + 1, dtype=np.int32), np.arange(A.n, dtype=np.int32), d)


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--side", type=int, default=160)
    ap.add_argument("--b", type=int, default=8)
    ap.add_argument("--reps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=10, help="untimed launches before each kernel's timed ones")
    ap.add_argument("--panels", type=int, default=5, help="panels per launch of panels32/panels64 (the refinement's blocks)")
    ap.add_argument("--share", type=int, default=-1, help="several-panel step: 1 / 0 forces in-wave gathers shared between lanes on / off "
                    "(-1: as tuned)")
    ap.add_argument("--tile", type=int, default=-1, help="several-panel step: lines per workgroup tile, 0 off (-1: as tuned)")
    ap.add_argument("--what", default="spmm32,spmm64,spmm_axpby32,spmm_axpby64,panels32,panels64,forced32,gram32,update32,cholqr32,cholqr64,copy")
    a = ap.parse_args(argv)
    what = a.what.split(",")
    cfg = P.config_c3(a.side)
    A = cfg.csr
    n, b = A.n, a.b
    ctx = L.Context(0)
    H = L.hip()
    dA = DeviceCSR(ctx, A)
    rng = np.random.default_rng(0)

    def timed(name, fn, bytes_per_launch):
        for _ in range(a.warmup):
            fn()
        avg = ctx.time_region(fn, a.reps)
        print(f"{name:14s} launches {a.reps:4d}  mean {avg * 1e6:9.1f} us  algorithmic {bytes_per_launch / 1e6:9.1f} MB  "
              f"{bytes_per_launch / avg / 1e9:7.0f} GB/s", flush=True)

    for w in what:
        if w in ("spmm32", "spmm64", "spmm_axpby32", "spmm_axpby64"):
            p = "f32" if w in ("spmm32", "spmm_axpby32") else "f64"
            dt = np.float32 if p == "f32" else np.float64
            op = dA.operator(b, p, tune=True)
            s = np.dtype(dt).itemsize
            X = ctx.to_device(rng.standard_normal((n, b)).astype(dt))
            Y = ctx.empty((n, b), dt)
            def mat_bytes(d):
                return (d.width * n * (4 + s)) if d.format == 1 else (A.nnz * (4 + s) + (n + 1) * 4)

            mat = mat_bytes(op.desc)
            if w.startswith("spmm_axpby"):
                Y = ctx.to_device(rng.standard_normal((n, b)).astype(dt))

                # Z = Y, as the Chebyshev filter calls it
                def fn(op=op, X=X, Y=Y, dt=dt):
                    L.check(H.trl_hip_spmm_axpby(ctx.h, L.dtype_code(dt), n, b, C.byref(op.desc_axpby), X.ptr, Y.ptr, Y.ptr, 0.5, -0.25, 0.125),
                            "axpby")

                timed(f"{w}[{op.name_axpby}]", fn, mat_bytes(op.desc_axpby) + 3 * n * b * s)
            else:
                timed(f"{w}[{op.name}]", lambda op=op, X=X, Y=Y: op.apply(X, Y), mat + 2 * n * b * s)
        elif w in ("panels32", "panels64"):
            # the filter step on all refinement blocks per launch, in place
            dt = np.float32 if w == "panels32" else np.float64
            s = np.dtype(dt).itemsize
            op = dA.operator(b, "f32" if w == "panels32" else "f64", tune=True)
            npan = a.panels
            op.tune_panels(npan)
            d = op.desc_panels
            if a.share >= 0:
                d.share = a.share
            if a.tile >= 0:
                d.tile_lines, d.line = a.tile, (a.side if a.tile else 0)
            X = ctx.to_device(rng.standard_normal((npan, n, b)).astype(dt))
            Y = ctx.to_device(rng.standard_normal((npan, n, b)).astype(dt))
            mat = (d.width * n * (4 + s)) if d.format == 1 else (A.nnz * (4 + s) + (n + 1) * 4)

            def fn(d=d, X=X, Y=Y, dt=dt):
                L.check(H.trl_hip_spmm_axpby_panels(ctx.h, L.dtype_code(dt), n, b, C.byref(d), npan, X.ptr, Y.ptr, Y.ptr, 0.5, -0.25, 0.125), "panels")

            timed(f"{w}[{op._name(d)}]", fn, mat + 3 * npan * n * b * s)
        elif w in ("floor32", "stream32"):
            # the filter step's streams without gathers: the diagonal of the matrix as ELL
            # (one slot), the panel step on --panels panels per launch (floor32), or
            # one panel per launch (stream32)
            Dm = diag_of(A)
            dD = DeviceCSR(ctx, Dm, precisions=("f32",))
            op = dD.operator(b, "f32", tune=True)
            npan = a.panels
            X = ctx.to_device(rng.standard_normal((npan, n, b)).astype(np.float32))
            Y = ctx.to_device(rng.standard_normal((npan, n, b)).astype(np.float32))
            if w == "floor32":
                op.tune_panels(npan)
                d = op.desc_panels

                def fn(d=d, X=X, Y=Y):
                    L.check(H.trl_hip_spmm_axpby_panels(ctx.h, L.F32, n, b, C.byref(d), npan, X.ptr, Y.ptr, Y.ptr, 0.5, -0.25, 0.125), "floor")
            else:
                d = op.desc_axpby

                def fn(d=d, X=X, Y=Y):
                    for j in range(npan):
                        L.check(H.trl_hip_spmm_axpby(ctx.h, L.F32, n, b, C.byref(d), X.offset(j * n * b), Y.offset(j * n * b), Y.offset(j * n * b), 0.5,
                                                     -0.25, 0.125), "stream")

            mat = (d.width * n * 8) if d.format == 1 else (Dm.nnz * 8 + (n + 1) * 4)
            timed(f"{w}[{op._name(d)}]", fn, mat * (1 if w == "floor32" else npan) + 3 * npan * n * b * 4)
        elif w == "forced32":
            # the residual-driven filter step: + F diag(g) read from a fourth panel set
            op = dA.operator(b, "f32", tune=True)
            npan = a.panels
            op.tune_panels(npan)
            d = op.desc_panels
            if a.share >= 0:
                d.share = a.share
            if a.tile >= 0:
                d.tile_lines, d.line = a.tile, (a.side if a.tile else 0)
            X = ctx.to_device(rng.standard_normal((npan, n, b)).astype(np.float32))
            Y = ctx.to_device(rng.standard_normal((npan, n, b)).astype(np.float32))
            F = ctx.to_device(rng.standard_normal((npan, n, b)).astype(np.float32))
            g = ctx.to_device(rng.standard_normal((npan, b)))
            mat = (d.width * n * 8) if d.format == 1 else (A.nnz * 8 + (n + 1) * 4)

            def fn(d=d, X=X, Y=Y, F=F, g=g):
                L.check(H.trl_hip_spmm_axpby_forced(ctx.h, L.F32, n, b, C.byref(d), npan, X.ptr, Y.ptr, Y.ptr, 0.5, -0.25, 0.125, F.ptr, g.ptr), "forced")

            timed(f"{w}[{op._name(d)}]", fn, mat + 4 * npan * n * b * 4)
        elif w in ("gram32", "update32", "gram64", "update64"):
            # FP32: the sweeps' k = 16; FP64: the plain Lanczos reorthogonalisation
            # against a full basis of k = 32 blocks
            f32 = w.endswith("32")
            dt, code, s_ = (np.float32, L.F32, 4) if f32 else (np.float64, L.F64, 8)
            k = 16 if f32 else 32
            V = ctx.to_device(rng.standard_normal((k + 1, n, b)).astype(dt))
            Cd = ctx.empty((k * b * b,))
            W = V.offset(k * n * b)
            P_ = n * b * s_
            L.check(H.trl_hip_gram(ctx.h, code, n, b, k, V.ptr, W, None, Cd.ptr), "gram")
            if w.startswith("gram"):
                timed(w, lambda: L.check(H.trl_hip_gram(ctx.h, code, n, b, k, V.ptr, W, None, Cd.ptr), "gram"), (k + 1) * P_)
            else:
                Cs = ctx.to_device(np.zeros(k * b * b))  # W unchanged: W -= V 0
                timed(w, lambda: L.check(H.trl_hip_update(ctx.h, code, n, b, k, V.ptr, Cs.ptr, W), "update"), (k + 2) * P_)
        elif w in ("rotate32", "rotate64", "restart64", "restart32"):
            # the Ritz rotation Out = V[0:m] Y: a refinement sweep's (m = kc =
            # --panels) in FP32 / FP64, the FP64 Lanczos restart's (m = 32
            # blocks, kc = 5) and the FP32 phase-1 restart's (m = 24, kc = 5)
            f32 = w in ("rotate32", "restart32")
            dt = np.float32 if f32 else np.float64
            s_ = 4 if f32 else 8
            m, kc = {"restart64": (32, 5), "restart32": (24, 5)}.get(w, (a.panels, a.panels))
            V = ctx.to_device(rng.standard_normal((m, n, b)).astype(dt))
            Yd = ctx.to_device(rng.standard_normal((m * b, kc * b)))
            Out = ctx.empty((kc, n, b), dt)
            timed(w, lambda: L.check(H.trl_hip_rotate(ctx.h, L.F32 if f32 else L.F64, n, b, m, V.ptr, Yd.ptr, kc, kc, Out.ptr), "rotate"),
                  (m + kc) * n * b * s_)
        elif w in ("cholqr32", "cholqr64"):
            f32 = w == "cholqr32"
            Wd = ctx.to_device(rng.standard_normal((n, b)).astype(np.float32 if f32 else np.float64))
            R = ctx.empty((b * b,))
            passes = 2 if f32 else 1
            s_ = 4 if f32 else 8
            timed(w, lambda: L.check(H.trl_hip_cholqr(ctx.h, L.F32 if f32 else L.F64, n, b, Wd.ptr, None, R.ptr, passes), "cholqr"),
                  (5 if passes == 2 else 3) * n * b * s_)
        elif w == "copy":
            cnt = 64 << 20  # 64 Mi doubles = 512 MiB each way
            src = ctx.to_device(np.ones(cnt))
            dst = ctx.empty((cnt,))
            timed("copy", lambda: L.check(H.trl_hip_convert(ctx.h, L.F64, src.ptr, L.F64, dst.ptr, cnt), "copy"), 2 * cnt * 8)
        else:
            print(f"unknown kernel {w}", file=sys.stderr)
            return 2
    ctx.close()
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
