"""CPU oracle — TEST INFRASTRUCTURE ONLY.

ctypes binding of oracle/lib/liboracle.so (built from oracle/trl_oracle.cc,
the OpenMP restatement of the reference's block-Lanczos path).  Only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg may use it, and only
as the checker.  The product (trl_amd) never imports it.
"""

from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "lib" / "liboracle.so"

_lib = None

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        L = C.CDLL(str(LIB))
        L.oracle_num_threads.restype = C.c_int
        L.oracle_set_threads.argtypes = [C.c_int]
        L.oracle_spmm_csr.argtypes = [C.c_size_t, C.c_int, _ip, _ip, _dp, _dp, _dp, _dp]
        L.oracle_panel_mult.argtypes = [C.c_size_t, C.c_int, _dp, _dp, _dp, C.c_int]
        L.oracle_panel_sub.argtypes = [C.c_size_t, C.c_int, _dp, _dp]
        L.oracle_gram.argtypes = [C.c_size_t, C.c_int, C.c_int, _dp, _dp, _dp, _dp]
        L.oracle_update.argtypes = [C.c_size_t, C.c_int, C.c_int, _dp, _dp, _dp]
        L.oracle_cholqr.argtypes = [C.c_size_t, C.c_int, _dp, _dp, _dp]
        L.oracle_cholqr.restype = C.c_int
        L.oracle_cholesky.argtypes = [C.c_int, _dp, _dp]
        L.oracle_cholesky.restype = C.c_int
        L.oracle_syev.argtypes = [C.c_int, _dp, _dp, _dp]
        L.oracle_syev.restype = C.c_int
        L.oracle_rotate.argtypes = [C.c_size_t, C.c_int, C.c_int, _dp, _dp, C.c_int, C.c_int, _dp]
        L.oracle_lanczos_solve.argtypes = [C.c_size_t, _ip, _ip, _dp, _dp, _dp, C.c_int, C.c_uint, C.c_uint, C.c_uint, C.c_double, C.c_int, _dp, _dp, _dp,
                                           C.POINTER(C.c_int)]
        L.oracle_lanczos_solve.restype = C.c_int
        L.oracle_lanczos_solve_timed.argtypes = L.oracle_lanczos_solve.argtypes + [_dp]
        L.oracle_lanczos_solve_timed.restype = C.c_int
        L.oracle_lanczos_extend.argtypes = [C.c_size_t, _ip, _ip, _dp, _dp, _dp, C.c_int, C.c_uint, C.c_uint, _dp, _dp, _dp, _dp]
        L.oracle_lanczos_extend.restype = C.c_int
        L.oracle_time_extend.argtypes = [C.c_size_t, _ip, _ip, _dp, C.c_int, C.c_uint, C.c_uint, _dp, C.c_uint]
        L.oracle_time_extend.restype = C.c_double
        _lib = L
    return _lib


def _d(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


def _i(a):
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(_ip)


def f64(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float64)


def set_threads(t: int) -> None:
    lib().oracle_set_threads(int(t))


def num_threads() -> int:
    return int(lib().oracle_num_threads())


def spmm(csr, X, dscale=None):
    X = f64(X)
    n, b = X.shape
    Y = np.empty_like(X)
    lib().oracle_spmm_csr(n, b, _i(csr.rowptr), _i(csr.col), _d(f64(csr.val)), _d(X), _d(Y), _d(f64(dscale)))
    return Y


def panel_mult(X, W, Y, mode):
    """mode 0: Y = X W, 1: Y += X W, 2: Y = X W^T, 3: X -= Y W (in place)."""
    n, b = X.shape
    lib().oracle_panel_mult(n, b, _d(X), _d(f64(W)), _d(Y), int(mode))


def panel_sub(X, Y):
    n, b = X.shape
    lib().oracle_panel_sub(n, b, _d(X), _d(f64(Y)))


def gram(V, W, m=None):
    """V: (k, n, b) blocks, W: (n, b) -> C (k*b, b)."""
    V = f64(V)
    k, n, b = V.shape
    C = np.empty((k * b, b))
    lib().oracle_gram(n, b, k, _d(V), _d(f64(W)), _d(f64(m)), _d(C))
    return C


def update(V, C, W):
    V = f64(V)
    k, n, b = V.shape
    lib().oracle_update(n, b, k, _d(V), _d(f64(C)), _d(W))


def cholqr(X, m=None):
    n, b = X.shape
    R = np.zeros((b, b))
    rc = lib().oracle_cholqr(n, b, _d(X), _d(f64(m)), _d(R))
    return rc, R


def cholesky(G):
    G = f64(G)
    b = G.shape[0]
    R = np.zeros((b, b))
    rc = lib().oracle_cholesky(b, _d(G), _d(R))
    return rc, R


def syev(A):
    A = f64(A)
    n = A.shape[0]
    w = np.empty(n)
    Z = np.empty((n, n))
    rc = lib().oracle_syev(n, _d(A), _d(w), _d(Z))
    if rc != 0:
        raise RuntimeError("oracle syev did not converge")
    return w, Z


def rotate(V, Yblocks, kc):
    """V: (m, n, b); Yblocks: block matrix array (mb_rows, ybc, b, b) -> (kc, n, b)."""
    V = f64(V)
    m, n, b = V.shape
    ybc = Yblocks.shape[1]
    Out = np.empty((kc, n, b))
    lib().oracle_rotate(n, b, m, _d(V), _d(f64(Yblocks)), ybc, kc, _d(Out))
    return Out


def lanczos_solve(csr, b, nev, ncv, V0, *, tol=1e-8, order=0, max_restarts=1000, mass=None, ritz_vectors=True):
    """Thick-restart solve; returns dict(evals, X, converged, iterations, n_op_apply, error, seconds); seconds
    is the wall time of the driver's solve() alone (max_restarts = 1: one full restart cycle)."""
    n = csr.n
    dscale = None if mass is None else 1.0 / np.asarray(mass, dtype=np.float64)
    evals = np.zeros(nev)
    X = np.zeros((n, nev)) if ritz_vectors else None
    info = (C.c_int * 4)()
    V0 = f64(V0).reshape(n, b)
    sec = C.c_double(0.0)
    lib().oracle_lanczos_solve_timed(n, _i(csr.rowptr), _i(csr.col), _d(f64(csr.val)), _d(f64(dscale)), _d(f64(mass)), b, nev, ncv, max_restarts, tol,
                                     order, _d(V0), _d(evals), _d(X), info, C.byref(sec))
    return dict(evals=evals, X=X, converged=bool(info[0]), iterations=info[1], n_op_apply=info[2], error=info[3], seconds=sec.value)


def lanczos_extend(csr, b, nev, ncv, V0, mass=None):
    n = csr.n
    dscale = None if mass is None else 1.0 / np.asarray(mass, dtype=np.float64)
    V = np.zeros(((ncv + b) // b, n, b))
    T = np.zeros((ncv // b, ncv // b, b, b))
    beta = np.zeros((b, b))
    rc = lib().oracle_lanczos_extend(n, _i(csr.rowptr), _i(csr.col), _d(f64(csr.val)), _d(f64(dscale)), _d(f64(mass)), b, nev, ncv,
                                     _d(f64(V0).reshape(n, b)), _d(V), _d(T), _d(beta))
    return rc, V, T, beta


def time_extend(csr, b, nev, ncv, V0, steps) -> float:
    """Seconds of the oracle's extend(0, steps) on this problem."""
    return float(lib().oracle_time_extend(csr.n, _i(csr.rowptr), _i(csr.col), _d(f64(csr.val)), b, nev, ncv, _d(f64(V0).reshape(csr.n, b)), steps))
