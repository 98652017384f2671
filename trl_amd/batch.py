"""Python entry to the batched solve (trl_hip_batch_solve_csr): many
independent standard FP64 problems advanced in lock-step on one GPU."""

from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib as L
from .problems import CSR
from .solver import ORDER


@dataclass
class BatchOutput:
    evals: np.ndarray  # (nprob, nev)
    resid: np.ndarray  # (nprob, nev) true relative residuals
    X: list | None  # per problem (nblk, n, b) or None
    results: list[dict]  # per problem: converged, iterations, n_op_apply, seconds (to convergence)
    stats: dict


def solve_batch(ctx: L.Context, mats: list[CSR], b: int, nev: int, ncv: int, V0: list[np.ndarray], *, tol=1e-8, max_restarts=1000, order="ascending",
                cholqr_passes=1, tune=True, use_graph=True, ritz_vectors=False, reorth="cgs2") -> BatchOutput:
    nprob = len(mats)
    n = (C.c_size_t * nprob)(*[m.n for m in mats])
    keep = []

    def arr(a, dt):
        a = np.ascontiguousarray(a, dtype=dt)
        keep.append(a)
        return a.ctypes.data

    rp = (C.c_void_p * nprob)(*[arr(m.rowptr, np.int32) for m in mats])
    ci = (C.c_void_p * nprob)(*[arr(m.col, np.int32) for m in mats])
    va = (C.c_void_p * nprob)(*[arr(m.val, np.float64) for m in mats])
    v0 = (C.c_void_p * nprob)(*[arr(v, np.float64) for v in V0])
    for m_, v in zip(mats, V0):
        assert v.shape == (m_.n, b)
    prm = L.BatchParams(b, nev, ncv, max_restarts, tol, ORDER.get(order, order), cholqr_passes, 1 if tune else 0, 1 if use_graph else 0,
                        {"cgs2": 1, "dgks": 2}[reorth])
    evals = np.zeros((nprob, nev))
    resid = np.zeros((nprob, nev))
    nblk = (nev + b - 1) // b
    Xs = [np.zeros((nblk, m_.n, b)) for m_ in mats] if ritz_vectors else None
    xp = (C.c_void_p * nprob)(*[x.ctypes.data for x in Xs]) if Xs else None
    res = (L.SolverResult * nprob)()
    st = L.BatchStats()
    rc = L.solver().trl_hip_batch_solve_csr(ctx.h, nprob, n, rp, ci, va, v0, C.byref(prm), evals.ctypes.data, resid.ctypes.data, xp, res, C.byref(st))
    L.check(rc, "trl_hip_batch_solve_csr")
    return BatchOutput(evals, resid, Xs, [r.as_dict() for r in res], st.as_dict())
