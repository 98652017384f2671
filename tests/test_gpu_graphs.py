"""Captured graphs (trl_hip_graph_*) and the driver's replayed restart cycles.

A graph of queued library calls is replayed several times, with the
program's own eager work and synchronisations between the replays, and every
replay is compared bit for bit with the same calls run eagerly.  The capture
refuses work that needs the host and profiling mode; a graph whose scratch memory the context has since replaced is
refused at launch.  The whole-solve path is compared with capture on and off:
the same kernels in the same order, so the same eigenvalues bit for bit and
the same restart counts.
"""

import ctypes as C

import numpy as np
import pytest

import oracle
from trl_amd import _lib as L
from trl_amd import problems as P
from trl_amd.solver import DeviceCSR, EigensolverParams, solve

pytestmark = pytest.mark.gpu


def H():
    return L.hip()


def queue_steps(ctx, A, dA, Xd, V, Yd, Cd, Rd, n, b, k):
    """Y = A X; C = V[0:k]^T Y; Y -= V[0:k] C; CholQR(Y) -> R (all queued)."""
    M = L.SpMat(0, dA.rowptr.ptr, 0, dA.col.ptr, dA.val["f64"].ptr, None, 0)
    L.check(H().trl_hip_spmm(ctx.h, L.F64, n, b, C.byref(M), Xd.ptr, Yd.ptr), "spmm")
    L.check(H().trl_hip_gram(ctx.h, L.F64, n, b, k, V.ptr, Yd.ptr, None, Cd.ptr), "gram")
    L.check(H().trl_hip_update(ctx.h, L.F64, n, b, k, V.ptr, Cd.ptr, Yd.ptr), "update")
    L.check(H().trl_hip_cholqr(ctx.h, L.F64, n, b, Yd.ptr, None, Rd.ptr, 1), "cholqr")


def test_graph_replays_match_eager_runs(ctx, rng):
    A = P.laplacian_3d(30, 20, 10, rng.uniform(0, 1, 6000))
    n, b, k = A.n, 8, 6
    dA = DeviceCSR(ctx, A, precisions=("f64",))
    X = rng.standard_normal((n, b))
    V = np.linalg.qr(rng.standard_normal((n, k * b)))[0].reshape(n, k, b).transpose(1, 0, 2).copy()
    Xd, Vd = ctx.to_device(X), ctx.to_device(V)
    Yd, Cd, Rd = ctx.empty((n, b)), ctx.empty((k * b * b,)), ctx.empty((b, b))
    queue_steps(ctx, A, dA, Xd, Vd, Yd, Cd, Rd, n, b, k)
    ctx.sync()
    ref = (Yd.download(), Cd.download(), Rd.download())
    # the eager result is the oracle's product, projection and CholQR
    AX = oracle.spmm(A, X)
    assert np.abs(ref[1].reshape(k, b, b) - np.einsum("kib,ic->kbc", V, AX)).max() < 1e-10 * np.abs(AX).max()
    # capture: nothing runs until a launch
    Yd.upload(np.zeros((n, b)))
    g = C.c_void_p()
    L.check(H().trl_hip_graph_begin(ctx.h), "begin")
    queue_steps(ctx, A, dA, Xd, Vd, Yd, Cd, Rd, n, b, k)
    L.check(H().trl_hip_graph_end(ctx.h, C.byref(g)), "end")
    assert g.value
    ctx.sync()
    assert not Yd.download().any()
    try:
        for rep in range(3):
            Yd.upload(rng.standard_normal((n, b)))  # eager work between replays
            ctx.sync()
            L.check(H().trl_hip_graph_launch(ctx.h, g), "launch")
            ctx.sync()
            got = (Yd.download(), Cd.download(), Rd.download())
            for a, r in zip(got, ref):
                assert np.array_equal(a, r), rep
    finally:
        H().trl_hip_graph_destroy(g)


def test_graph_capture_refusals(ctx, rng):
    A = P.laplacian_2d(40, 30)
    dA = DeviceCSR(ctx, A, precisions=("f64",))
    g = C.c_void_p()
    # a call that reads back to the host cannot be captured: no graph, and the
    # context works on afterwards
    L.check(H().trl_hip_graph_begin(ctx.h), "begin")
    w = C.c_int()
    assert H().trl_hip_csr_max_row(ctx.h, A.n, dA.rowptr.ptr, C.byref(w)) != 0
    assert H().trl_hip_graph_end(ctx.h, C.byref(g)) != 0 and not g.value
    X = rng.standard_normal((A.n, 4))
    Xd, Yd = ctx.to_device(X), ctx.empty((A.n, 4))
    M = L.SpMat(0, dA.rowptr.ptr, 0, dA.col.ptr, dA.val["f64"].ptr, None, 0)
    L.check(H().trl_hip_spmm(ctx.h, L.F64, A.n, 4, C.byref(M), Xd.ptr, Yd.ptr), "spmm")
    assert np.abs(Yd.download() - oracle.spmm(A, X)).max() < 1e-12
    # begin twice, end without begin, profiling mode, the option that turns capture off
    L.check(H().trl_hip_graph_begin(ctx.h), "begin")
    assert H().trl_hip_graph_begin(ctx.h) == -1
    L.check(H().trl_hip_graph_end(ctx.h, C.byref(g)), "end of an empty capture")
    H().trl_hip_graph_destroy(g)
    assert H().trl_hip_graph_end(ctx.h, C.byref(g)) == -1
    ctx.prof_enable(True)
    try:
        assert H().trl_hip_graph_begin(ctx.h) == -1
    finally:
        ctx.prof_enable(False)
    # the option is read back; a capture does not depend on it
    v = C.c_int(-1)
    L.check(H().trl_hip_ctx_get_option(ctx.h, L.OPT_GRAPHS, C.byref(v)), "get option")
    assert v.value == 0
    assert H().trl_hip_ctx_get_option(ctx.h, 99, C.byref(v)) == -1


def test_graph_with_replaced_scratch_is_stale(rng):
    # a fresh context: its Gram partial sums start small; a later, larger
    # Gram replaces them, and a graph that names the old buffer is refused
    ctx = L.Context(0)
    try:
        n, b = 4000, 4
        V, W, Cd = ctx.to_device(rng.standard_normal((2, n, b))), ctx.to_device(rng.standard_normal((n, b))), ctx.empty((2 * b * b,))
        g = C.c_void_p()
        L.check(H().trl_hip_graph_begin(ctx.h), "begin")
        L.check(H().trl_hip_gram(ctx.h, L.F64, n, b, 2, V.ptr, W.ptr, None, Cd.ptr), "gram")
        L.check(H().trl_hip_graph_end(ctx.h, C.byref(g)), "end")
        L.check(H().trl_hip_graph_launch(ctx.h, g), "launch")
        ctx.sync()
        nb, bb, kk = 100000, 16, 64
        VB = ctx.empty((kk + 1, nb, bb))
        L.check(H().trl_hip_fill_zero(ctx.h, VB.ptr, VB.nbytes), "zero")
        CB = ctx.empty((kk * bb * bb,))
        L.check(H().trl_hip_gram(ctx.h, L.F64, nb, bb, kk, VB.ptr, VB.offset(kk * nb * bb), None, CB.ptr), "large gram")
        ctx.sync()
        assert H().trl_hip_graph_launch(ctx.h, g) == -10
        H().trl_hip_graph_destroy(g)
    finally:
        ctx.close()


@pytest.mark.parametrize("reorth", ["mgs", "cgs2", "dgks"])
def test_solve_with_captured_cycles_matches_direct(ctx, reorth):
    # the 2D Laplacian 128^2 (b = 4, 8 smallest, ncv = 256): restart cycles
    # replayed from graphs (TRL_HIP_OPT_GRAPHS = 1) against the same solve with
    # the steps queued directly (the default)
    c = P.config_c1()
    dA = DeviceCSR(ctx, c.csr, precisions=("f64",))
    V0 = ctx.to_device(P.initial_block(c.csr.n, c.b))
    prm = EigensolverParams(c.nev, c.ncv, tolerance=c.tol)
    outs = {}
    for on in (1, 0):
        L.check(H().trl_hip_ctx_set_option(ctx.h, L.OPT_GRAPHS, on), "option")
        try:
            solve(dA, c.b, prm, V0, order="ascending", reorth=reorth)  # untimed first call
            outs[on] = solve(dA, c.b, prm, V0, order="ascending", reorth=reorth)
        finally:
            L.check(H().trl_hip_ctx_set_option(ctx.h, L.OPT_GRAPHS, 0), "option")
    g, d = outs[1], outs[0]
    print(f"{reorth}: graphs {g.result['seconds']:.4f} s, direct {d.result['seconds']:.4f} s, {g.result['iterations']} restarts", flush=True)
    assert g.result["converged"] == 1 and g.result["iterations"] == d.result["iterations"] and g.result["n_op_apply"] == d.result["n_op_apply"]
    assert g.result["graph_cycles"] == g.result["iterations"] - 1 and d.result["graph_cycles"] == 0
    assert np.array_equal(g.evals, d.evals)
    assert np.array_equal(g.X.download(), d.X.download())
    assert g.resid.max() <= c.tol
