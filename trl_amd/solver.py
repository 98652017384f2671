"""Python entry to whole solves on the GPU (include/trl_hip_solver.h).

`DeviceCSR` keeps a CSR matrix (and an optional diagonal mass) resident in
HBM in both panel precisions; `solve` / `solve_mixed` run the C++ driver
(trl_amd::BlockLanczos over trl::hip) on it.  Names and parameter meaning
follow the reference's EigensolverParams / EigensolverResult
(include/trl/eigensolvers/params.hh:5-16).
"""

from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib as L
from .problems import CSR

ORDER = {"largest_magnitude": 0, "ascending": 1, "descending": 2}

# The mixed-precision schedule of the 160^3 benchmark (the 3D 7-point
# Laplacian + diagonal potential, b = 8, 32 smallest, ncv = 256, tol 1e-8):
# one loose FP32 thick-restart cycle over 192 columns, FP32 Chebyshev /
# Rayleigh-Ritz sweeps on 5 Ritz blocks to an FP32 residual of 1e-3, then FP64
# sweeps whose filter runs in FP32 driven by the FP64 residuals, per-block
# filter degrees (160 at most), adaptive block CGS, autotuned operator.
MIXED_SCHEDULE = dict(tol_fp32=0.4, cheb_degree=160, refine_blocks=5, max_refine=14, fp32_sweeps=12, tol_fp32_refine=1e-3, degree_opt=1,
                      residual_filter=True, ncv_fp32=192, reorth="dgks", spmm_tune=True)
REORTH = {"mgs": 0, "cgs2": 1, "dgks": 2}


@dataclass
class EigensolverParams:
    nev: int
    ncv: int
    max_restarts: int = 1000
    tolerance: float = 1e-8


class DeviceCSR:
    def __init__(self, ctx: L.Context, csr: CSR, mass: np.ndarray | None = None, precisions=("f64", "f32")):
        self.ctx = ctx
        self.n = csr.n
        self.nnz = csr.nnz
        self.rowptr = ctx.to_device(csr.rowptr.astype(np.int32))
        self.col = ctx.to_device(csr.col.astype(np.int32))
        self.val = {}
        self.mass = {}
        self.minv = {}
        for p in precisions:
            dt = np.float64 if p == "f64" else np.float32
            self.val[p] = ctx.to_device(csr.val.astype(dt))
            if mass is not None:
                self.mass[p] = ctx.to_device(np.asarray(mass, dtype=np.float64).astype(dt))
                self.minv[p] = ctx.to_device((1.0 / np.asarray(mass, dtype=np.float64)).astype(dt))
        self.has_mass = mass is not None

    def ptrs(self, p: str):
        return (self.val[p].ptr, self.mass[p].ptr if self.has_mass else None, self.minv[p].ptr if self.has_mass else None)

    def operator(self, b: int, p: str = "f64", tune: bool = True, scaled: bool = False) -> "DeviceOperator":
        """The operator kernel for block size b: the untuned CSR default, or
        the fastest of the CSR variants and the device-built ELL form
        (trl_hip_spmm_tune).  scaled applies diag(1/M) for a mass matrix."""
        return DeviceOperator(self, b, p, tune, scaled)


class DeviceOperator:
    """A trl_hip_spmat descriptor plus the device arrays it refers to."""

    def __init__(self, A: DeviceCSR, b: int, p: str, tune: bool, scaled: bool):
        ctx, H = A.ctx, L.hip()
        dt = np.float64 if p == "f64" else np.float32
        self.A, self.b, self.dtype = A, b, dt
        val, _, minv = A.ptrs(p)
        dscale = minv if scaled else None
        csr = L.SpMat(0, A.rowptr.ptr, 0, A.col.ptr, val, dscale, 0)
        self.ell = None
        self.us = None
        self.desc = csr
        self.desc_axpby = csr
        self.us_axpby = None
        self.desc_panels, self.us_panels, self.npanels = csr, None, 0
        self._csr, self._ell = csr, None
        if tune:
            w = C.c_int()
            L.check(H.trl_hip_csr_max_row(ctx.h, A.n, A.rowptr.ptr, C.byref(w)), "csr_max_row")
            ell = None
            if w.value * A.n <= 1.25 * A.nnz:
                self.ell = (ctx.empty((w.value, A.n), np.int32), ctx.empty((w.value, A.n), dt))
                L.check(H.trl_hip_csr_to_ell(ctx.h, L.dtype_code(dt), A.n, A.rowptr.ptr, A.col.ptr, val, w.value, self.ell[0].ptr, self.ell[1].ptr), "csr_to_ell")
                ell = L.SpMat(1, None, w.value, self.ell[0].ptr, self.ell[1].ptr, dscale, 0)
                self._ell = ell
            best, us = L.SpMat(), C.c_double()
            L.check(H.trl_hip_spmm_tune(ctx.h, L.dtype_code(dt), A.n, b, C.byref(csr), C.byref(ell) if ell else None, C.byref(best), C.byref(us)), "spmm_tune")
            self.desc, self.us = best, us.value
            best_ax, us_ax = L.SpMat(), C.c_double()
            L.check(H.trl_hip_spmm_tune_axpby(ctx.h, L.dtype_code(dt), A.n, b, C.byref(csr), C.byref(ell) if ell else None, C.byref(best_ax),
                                              C.byref(us_ax)), "spmm_tune_axpby")
            self.desc_axpby, self.us_axpby = best_ax, us_ax.value
            if best.format == 0 and best_ax.format == 0:
                self.ell = None

    def tune_panels(self, npanels: int) -> None:
        """Tunes the axpby form on npanels panels per launch (trl_hip_spmm_tune_panels)."""
        ctx, H = self.A.ctx, L.hip()
        csr, ell = self._csr, self._ell if self.ell else None
        best, us = L.SpMat(), C.c_double()
        L.check(H.trl_hip_spmm_tune_panels(ctx.h, L.dtype_code(self.dtype), self.A.n, self.b, npanels, C.byref(csr), C.byref(ell) if ell else None,
                                           C.byref(best), C.byref(us)), "spmm_tune_panels")
        self.desc_panels, self.us_panels, self.npanels = best, us.value, npanels

    def panels_pay(self, npanels: int) -> bool:
        """The rule the solver's filter applies: one launch over the panels
        when it beats npanels one-panel launches by 3%."""
        return self.npanels == npanels and self.us_panels < 0.97 * npanels * self.us_axpby

    @staticmethod
    def _name(d) -> str:
        order = f"/plane{d.plane}:{d.plane_block}" if d.plane else ""
        extra = ("/share" if d.share else "") + (f"/tile{d.tile_lines}x{d.line}" if d.tile_lines else "")
        return f"{'ell' if d.format else 'csr'}/v{d.variant}{order}{extra}"

    @property
    def name(self) -> str:
        return self._name(self.desc)

    @property
    def name_axpby(self) -> str:
        return self._name(self.desc_axpby)

    def bytes_per_apply(self) -> float:
        """Algorithmic HBM bytes of one apply: matrix arrays once, X read once, Y written."""
        s = np.dtype(self.dtype).itemsize
        panels = 2 * self.A.n * self.b * s
        if self.desc.format == 1:
            return panels + self.desc.width * self.A.n * (4 + s)
        return panels + self.A.nnz * (4 + s) + (self.A.n + 1) * 4

    def apply(self, X: L.DeviceArray, Y: L.DeviceArray) -> None:
        L.check(L.hip().trl_hip_spmm(self.A.ctx.h, L.dtype_code(self.dtype), self.A.n, self.b, C.byref(self.desc), X.ptr, Y.ptr), "trl_hip_spmm")


@dataclass
class SolveOutput:
    evals: np.ndarray
    X: L.DeviceArray | None
    resid: np.ndarray | None
    result: dict
    result32: dict | None = None


def _params(b, params: EigensolverParams, order, reorth, cholqr_passes, lanes, tune, rr_dense=False):
    return L.SolverParams(b, params.nev, params.ncv, params.max_restarts, params.tolerance, ORDER.get(order, order), REORTH.get(reorth, reorth),
                          cholqr_passes, lanes, 1 if tune else 0, 1 if rr_dense else 0)


def solve(A: DeviceCSR, b: int, params: EigensolverParams, V0: L.DeviceArray, *, dtype="f64", order="ascending", reorth="cgs2", cholqr_passes=0,
          lanes_per_row=0, spmm_tune=False, ritz_vectors=True, rr_dense=False) -> SolveOutput:
    ctx = A.ctx
    dt = np.float64 if dtype == "f64" else np.float32
    assert V0.dtype == dt and V0.nbytes == A.n * b * np.dtype(dt).itemsize
    nblk = (params.nev + b - 1) // b
    X = ctx.empty((nblk, A.n, b), dt) if ritz_vectors else None
    evals = np.zeros(params.nev)
    resid = np.zeros(params.nev)
    res = L.SolverResult()
    p = _params(b, params, order, reorth, cholqr_passes, lanes_per_row, spmm_tune, rr_dense)
    val, mass, minv = A.ptrs(dtype)
    rc = L.solver().trl_hip_solve_csr(ctx.h, L.dtype_code(dt), A.n, A.rowptr.ptr, A.col.ptr, val, mass, minv, C.byref(p), V0.ptr,
                                      evals.ctypes.data, X.ptr if X else None, resid.ctypes.data if X else None, C.byref(res))
    L.check(rc, "trl_hip_solve_csr")
    return SolveOutput(evals, X, resid if X else None, res.as_dict())


def solve_mixed(A: DeviceCSR, b: int, params: EigensolverParams, V0: L.DeviceArray, *, tol_fp32=1e-5, max_restarts_fp32=1000, seed=7, order="ascending",
                reorth="cgs2", cholqr_passes=0, lanes_per_row=0, spmm_tune=False, refine=1, cheb_degree=80, max_refine=8, refine_blocks=0,
                phase1_fp64=False, fp32_sweeps=0, tol_fp32_refine=1e-4, degree_opt=False, residual_filter=False, ncv_fp32=0,
                rr_dense=False) -> SolveOutput:
    ctx = A.ctx
    assert V0.dtype == np.float64
    nblk = (params.nev + b - 1) // b
    X = ctx.empty((nblk, A.n, b), np.float64)
    evals = np.zeros(params.nev)
    resid = np.zeros(params.nev)
    res = L.SolverResult()
    res32 = L.SolverResult()
    p = _params(b, params, order, reorth, cholqr_passes, lanes_per_row, spmm_tune, rr_dense)
    mp = L.MixedParams(tol_fp32, max_restarts_fp32, seed, refine, cheb_degree, max_refine, 1 if phase1_fp64 else 0, refine_blocks, fp32_sweeps,
                       tol_fp32_refine, int(degree_opt), 1 if residual_filter else 0, ncv_fp32)
    v64, m64, mi64 = A.ptrs("f64")
    v32, m32, mi32 = A.ptrs("f32")
    rc = L.solver().trl_hip_solve_csr_mixed(ctx.h, A.n, A.rowptr.ptr, A.col.ptr, v64, v32, m64, mi64, m32, mi32, C.byref(p), C.byref(mp), V0.ptr,
                                            evals.ctypes.data, X.ptr, resid.ctypes.data, C.byref(res), C.byref(res32))
    L.check(rc, "trl_hip_solve_csr_mixed")
    return SolveOutput(evals, X, resid, res.as_dict(), res32.as_dict())


@dataclass
class ShiftInvertOutput:
    evals: np.ndarray  # lambda = shift + 1/mu
    mu: np.ndarray  # Ritz values of (A - shift B)^{-1} B
    X: L.DeviceArray | None
    resid: np.ndarray | None  # |A x - lambda B x| / |A x|
    result: dict
    inner: dict


def solve_shift_invert(ctx: L.Context, A, B, shift: float, b: int, params: EigensolverParams, V0: L.DeviceArray, *, order="largest_magnitude",
                       reorth="cgs2", cholqr_passes=0, inner_tol=0.0, ritz_vectors=True, rr_dense=False) -> ShiftInvertOutput:
    """A x = lambda B x by shift-invert (trl_hip_solve_csr_shift_invert; the
    reference's CSRGeneralizedEVP, examples/openmp/csrevp.hh:51-104): A, B
    host CSR matrices (problems.CSR, full symmetric storage), B SPD, A -
    shift B positive definite."""
    assert V0.dtype == np.float64 and V0.nbytes == A.n * b * 8 and A.n == B.n
    nblk = (params.nev + b - 1) // b
    X = ctx.empty((nblk, A.n, b), np.float64) if ritz_vectors else None
    evals, mu, resid = np.zeros(params.nev), np.zeros(params.nev), np.zeros(params.nev)
    res, inner = L.SolverResult(), L.InnerStats()
    p = _params(b, params, order, reorth, cholqr_passes, 0, False, rr_dense)
    arp, acol, aval = (np.ascontiguousarray(A.rowptr, np.int32), np.ascontiguousarray(A.col, np.int32), np.ascontiguousarray(A.val, np.float64))
    brp, bcol, bval = (np.ascontiguousarray(B.rowptr, np.int32), np.ascontiguousarray(B.col, np.int32), np.ascontiguousarray(B.val, np.float64))
    rc = L.solver().trl_hip_solve_csr_shift_invert(ctx.h, A.n, arp.ctypes.data, acol.ctypes.data, aval.ctypes.data, brp.ctypes.data, bcol.ctypes.data,
                                                   bval.ctypes.data, float(shift), float(inner_tol), C.byref(p), V0.ptr, evals.ctypes.data,
                                                   mu.ctypes.data, X.ptr if X else None, resid.ctypes.data if X else None, C.byref(res),
                                                   C.byref(inner))
    L.check(rc, "trl_hip_solve_csr_shift_invert")
    return ShiftInvertOutput(evals, mu, X, resid if X else None, res.as_dict(), inner.as_dict())


def dense_syev(A: np.ndarray):
    """The host FP64 eigensolver of the Rayleigh-Ritz step (for tests)."""
    A = np.ascontiguousarray(A, dtype=np.float64)
    n = A.shape[0]
    w = np.empty(n)
    Z = np.empty((n, n))
    L.check(L.solver().trl_hip_dense_syev(n, A.ctypes.data, w.ctypes.data, Z.ctypes.data), "trl_hip_dense_syev")
    return w, Z


def dense_syev_select(A: np.ndarray, sel):
    """All eigenvalues and the eigenvectors of the selected ones (for tests)."""
    A = np.ascontiguousarray(A, dtype=np.float64)
    n = A.shape[0]
    sel = np.ascontiguousarray(sel, dtype=np.int32)
    w = np.empty(n)
    Z = np.empty((n, len(sel)))
    L.check(L.solver().trl_hip_dense_syev_select(n, A.ctypes.data, w.ctypes.data, len(sel), sel.ctypes.data, Z.ctypes.data),
            "trl_hip_dense_syev_select")
    return w, Z


def dense_syev_arrow_band(A: np.ndarray, b: int, sel):
    """The structured eigensolver of the projected matrix (for tests): all
    eigenvalues, the selected eigenvectors and p, the size of the diagonal
    Ritz part it found.  Raises TrlHipError when A lacks the arrow-plus-band
    structure."""
    A = np.ascontiguousarray(A, dtype=np.float64)
    n = A.shape[0]
    sel = np.ascontiguousarray(sel, dtype=np.int32)
    w = np.empty(n)
    Z = np.empty((n, len(sel)))
    p = C.c_int(-1)
    L.check(L.solver().trl_hip_dense_syev_arrow_band(n, b, A.ctypes.data, w.ctypes.data, len(sel), sel.ctypes.data, Z.ctypes.data, C.byref(p)),
            "trl_hip_dense_syev_arrow_band")
    return w, Z, p.value
