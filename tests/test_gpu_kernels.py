"""Kernel-level parity of libtrl_hip.so against the CPU oracle.

Every entry point of include/trl_hip.h is called through the C-ABI on
seeded inputs and compared with the oracle's restatement of the reference
operation (oracle/trl_oracle.cc).  FP64 within 1e-12 relative, FP32 within
the bound stated per test.
"""

import ctypes as C

import numpy as np
import pytest

import oracle
from trl_amd import _lib as L
from trl_amd import problems as P
from trl_amd.formats import csr_to_ell

pytestmark = pytest.mark.gpu

DT = {"f64": np.float64, "f32": np.float32}
TOL = {"f64": 1e-12, "f32": 2e-5}
BS = [1, 2, 4, 6, 8, 16]


def H():
    return L.hip()


def split_ok(p, b, cs):
    # a column split of cs lanes needs b % cs == 0 and at least 8 bytes per lane
    return b % cs == 0 and (b // cs) * np.dtype(DT[p]).itemsize >= 8


def spmat(fmt, rowptr=None, width=0, col=None, val=None, dscale=None, variant=0):
    ptr = lambda a: a.ptr if a is not None else None  # noqa: E731
    return L.SpMat(fmt, ptr(rowptr), width, ptr(col), ptr(val), ptr(dscale), variant)


def rel(a, b):
    return np.abs(np.asarray(a, np.float64) - b).max() / max(np.abs(b).max(), 1e-300)


# Every device array a test creates stays referenced until the test ends:
# an inline temporary (dev(...).ptr) would otherwise be freed by its
# destructor before the queued kernel reads it.
_KEEP = []


@pytest.fixture(autouse=True)
def _keep_device_arrays(ctx):
    yield
    ctx.sync()
    _KEEP.clear()


def dev(ctx, a, dt):
    d = ctx.to_device(np.ascontiguousarray(a, dtype=dt))
    _KEEP.append(d)
    return d


@pytest.mark.parametrize("p", ["f64", "f32"])
@pytest.mark.parametrize("b", BS)
@pytest.mark.parametrize("lanes", [0, 1, 2, 4, 8, 16, -2, -4])
def test_spmm_csr(ctx, rng, p, b, lanes):
    print(f"spmm_csr {p} b={b} lanes={lanes}", flush=True)
    A = P.laplacian_3d(7, 5, 6, rng.uniform(0, 1, 210))
    X = rng.standard_normal((A.n, b))
    d = rng.uniform(0.5, 2.0, A.n)
    dt = DT[p]
    rp, ci = dev(ctx, A.rowptr, np.int32), dev(ctx, A.col, np.int32)
    va, Xd, dd = dev(ctx, A.val, dt), dev(ctx, X, dt), dev(ctx, d, dt)
    Y = ctx.empty((A.n, b), dt)
    if lanes < 0 and not split_ok(p, b, -lanes):
        assert H().trl_hip_spmm_csr(ctx.h, L.dtype_code(dt), A.n, b, rp.ptr, ci.ptr, va.ptr, Xd.ptr, Y.ptr, None, lanes) == -1
        return
    for scale in (None, dd):
        L.check(H().trl_hip_spmm_csr(ctx.h, L.dtype_code(dt), A.n, b, rp.ptr, ci.ptr, va.ptr, Xd.ptr, Y.ptr, scale.ptr if scale else None, lanes), "spmm")
        ref = oracle.spmm(A, X.astype(dt).astype(np.float64), d.astype(dt).astype(np.float64) if scale else None)
        assert rel(Y.download(), ref) < TOL[p]


@pytest.mark.parametrize("p", ["f64", "f32"])
@pytest.mark.parametrize("b", [1, 4, 8])
def test_spmm_ell(ctx, rng, p, b):
    A = P.stiffness_mass(3, 3, 2, dof=2)[0]
    X = rng.standard_normal((A.n, b))
    width, ecol, evals = csr_to_ell(A)
    dt = DT[p]
    Y = ctx.empty((A.n, b), dt)
    ec, ev, Xd = dev(ctx, ecol, np.int32), dev(ctx, evals, dt), dev(ctx, X, dt)
    L.check(H().trl_hip_spmm_ell(ctx.h, L.dtype_code(dt), A.n, b, width, ec.ptr, ev.ptr, Xd.ptr, Y.ptr, None), "ell")
    assert rel(Y.download(), oracle.spmm(A, X.astype(dt).astype(np.float64))) < TOL[p]


@pytest.mark.parametrize("p", ["f64", "f32"])
@pytest.mark.parametrize("b", BS)
@pytest.mark.parametrize("variant", [0, 1, -2, -4])
def test_spmm_ell_variants(ctx, rng, p, b, variant):
    # slot-major ELL built on the device from CSR, through the generic entry
    # point, with and without row scaling and in the fused axpby form
    print(f"spmm_ell {p} b={b} variant={variant}", flush=True)
    A = P.laplacian_3d(7, 5, 6, rng.uniform(0, 1, 210))
    dt = DT[p]
    X = rng.standard_normal((A.n, b)).astype(dt).astype(np.float64)
    Z = rng.standard_normal((A.n, b)).astype(dt).astype(np.float64)
    d = rng.uniform(0.5, 2.0, A.n).astype(dt).astype(np.float64)
    rp, ci, va = dev(ctx, A.rowptr, np.int32), dev(ctx, A.col, np.int32), dev(ctx, A.val, dt)
    w = C.c_int(-1)
    L.check(H().trl_hip_csr_max_row(ctx.h, A.n, rp.ptr, C.byref(w)), "max_row")
    assert w.value == int(np.diff(A.rowptr).max()) == 7
    ec, ev = ctx.empty((w.value, A.n), np.int32), ctx.empty((w.value, A.n), dt)
    L.check(H().trl_hip_csr_to_ell(ctx.h, L.dtype_code(dt), A.n, rp.ptr, ci.ptr, va.ptr, w.value, ec.ptr, ev.ptr), "to_ell")
    width, hcol, hval = csr_to_ell(A)
    assert width == w.value
    assert np.array_equal(ec.download().ravel(), hcol.ravel())
    assert np.array_equal(ev.download().ravel(), hval.astype(dt).ravel())
    Xd, Zd, dd = dev(ctx, X, dt), dev(ctx, Z, dt), dev(ctx, d, dt)
    Y = ctx.empty((A.n, b), dt)
    for scale in (None, dd):
        M = spmat(1, None, w.value, ec, ev, scale, variant)
        rc = H().trl_hip_spmm(ctx.h, L.dtype_code(dt), A.n, b, C.byref(M), Xd.ptr, Y.ptr)
        if variant < 0 and not split_ok(p, b, -variant):
            assert rc == -1
            return
        L.check(rc, "spmm")
        AX = oracle.spmm(A, X, d if scale else None)
        assert rel(Y.download(), AX) < TOL[p]
        L.check(H().trl_hip_spmm_axpby(ctx.h, L.dtype_code(dt), A.n, b, C.byref(M), Xd.ptr, Y.ptr, Zd.ptr, 0.7, -1.3, 0.4), "axpby")
        assert rel(Y.download(), 0.7 * AX - 1.3 * X + 0.4 * Z) < TOL[p]
    M = spmat(1, None, w.value, ec, ev, None, 4)  # nonzero lane groups need CSR
    assert H().trl_hip_spmm(ctx.h, L.dtype_code(dt), A.n, b, C.byref(M), Xd.ptr, Y.ptr) == -1


@pytest.mark.parametrize("p", ["f64", "f32"])
@pytest.mark.parametrize("b", [1, 4, 8])
@pytest.mark.parametrize("fmt,variant", [(0, 0), (0, 1), (0, 4), (1, 1), (1, -2)])
def test_spmm_panels(ctx, rng, p, b, fmt, variant):
    # npanels panels in one launch == one launch per panel (bitwise), and
    # the fused form with Z == Y (in-place older iterate)
    if variant < 0 and not split_ok(p, b, -variant):
        pytest.skip("column split needs 8 bytes per lane")
    A = P.laplacian_3d(9, 7, 6, rng.uniform(0, 1, 378))
    dt = DT[p]
    npan = 3
    X = rng.standard_normal((npan, A.n, b)).astype(dt)
    Z = rng.standard_normal((npan, A.n, b)).astype(dt)
    d = rng.uniform(0.5, 2, A.n).astype(dt)
    rp, ci, va = dev(ctx, A.rowptr, np.int32), dev(ctx, A.col, np.int32), dev(ctx, A.val, dt)
    width, hcol, hval = csr_to_ell(A)
    ec, ev, dd = dev(ctx, hcol, np.int32), dev(ctx, hval, dt), dev(ctx, d, dt)
    M = spmat(0, rp, 0, ci, va, dd, variant) if fmt == 0 else spmat(1, None, width, ec, ev, dd, variant)
    Xd = dev(ctx, X, dt)
    Y = ctx.empty((npan, A.n, b), dt)
    L.check(H().trl_hip_spmm_panels(ctx.h, L.dtype_code(dt), A.n, b, C.byref(M), npan, Xd.ptr, Y.ptr), "panels")
    Y1 = ctx.empty((A.n, b), dt)
    got = Y.download()
    for j in range(npan):
        L.check(H().trl_hip_spmm(ctx.h, L.dtype_code(dt), A.n, b, C.byref(M), Xd.offset(j * A.n * b), Y1.ptr), "one")
        assert np.array_equal(got[j], Y1.download())
        assert rel(got[j], oracle.spmm(A, X[j].astype(np.float64), d.astype(np.float64))) < TOL[p]
    Zd = dev(ctx, Z, dt)
    L.check(H().trl_hip_spmm_axpby_panels(ctx.h, L.dtype_code(dt), A.n, b, C.byref(M), npan, Xd.ptr, Zd.ptr, Zd.ptr, 0.5, -1.0, 0.25), "axpby panels")
    got = Zd.download()
    for j in range(npan):
        ref = 0.5 * oracle.spmm(A, X[j].astype(np.float64), d.astype(np.float64)) - X[j] + 0.25 * Z[j]
        assert rel(got[j], ref) < TOL[p]
    # X overlapping Y is refused
    assert H().trl_hip_spmm_axpby_panels(ctx.h, L.dtype_code(dt), A.n, b, C.byref(M), npan, Xd.ptr, Xd.offset(A.n * b), None, 1.0, 0.0, 0.0) == -1


@pytest.mark.parametrize("p", ["f64", "f32"])
@pytest.mark.parametrize("b", [1, 4, 8])
@pytest.mark.parametrize("fmt,variant", [(0, 0), (0, 4), (1, 1), (1, -2), (1, -4)])
@pytest.mark.parametrize("npan", [1, 3])
def test_spmm_forced(ctx, rng, p, b, fmt, variant, npan):
    # the step with a forcing term, Y_j = alpha d A X_j + beta X_j + gamma Z_j
    # + F_j diag(g_j), against the oracle's product; Z may be absent
    if variant < 0 and not split_ok(p, b, -variant):
        pytest.skip("column split needs 8 bytes per lane")
    A = P.laplacian_3d(9, 7, 6, rng.uniform(0, 1, 378))
    dt = DT[p]
    X = rng.standard_normal((npan, A.n, b)).astype(dt)
    Z = rng.standard_normal((npan, A.n, b)).astype(dt)
    F = rng.standard_normal((npan, A.n, b)).astype(dt)
    g = rng.standard_normal((npan, b))
    d = rng.uniform(0.5, 2, A.n).astype(dt)
    rp, ci, va = dev(ctx, A.rowptr, np.int32), dev(ctx, A.col, np.int32), dev(ctx, A.val, dt)
    width, hcol, hval = csr_to_ell(A)
    ec, ev, dd = dev(ctx, hcol, np.int32), dev(ctx, hval, dt), dev(ctx, d, dt)
    M = spmat(0, rp, 0, ci, va, dd, variant) if fmt == 0 else spmat(1, None, width, ec, ev, dd, variant)
    Xd, Zd, Fd, gd = dev(ctx, X, dt), dev(ctx, Z, dt), dev(ctx, F, dt), dev(ctx, g, np.float64)
    Y = ctx.empty((npan, A.n, b), dt)
    for with_z in (True, False):
        L.check(H().trl_hip_spmm_axpby_forced(ctx.h, L.dtype_code(dt), A.n, b, C.byref(M), npan, Xd.ptr, Y.ptr, Zd.ptr if with_z else None, 0.5, -1.0,
                                              0.25, Fd.ptr, gd.ptr), "forced")
        got = Y.download()
        for j in range(npan):
            ref = 0.5 * oracle.spmm(A, X[j].astype(np.float64), d.astype(np.float64)) - X[j] + F[j] * g[j]
            if with_z:
                ref = ref + 0.25 * Z[j]
            assert rel(got[j], ref) < TOL[p], (j, with_z)
    # F overlapping Y, or a missing coefficient array, is refused
    assert H().trl_hip_spmm_axpby_forced(ctx.h, L.dtype_code(dt), A.n, b, C.byref(M), npan, Xd.ptr, Y.ptr, None, 1.0, 0.0, 0.0, Y.ptr, gd.ptr) == -1
    assert H().trl_hip_spmm_axpby_forced(ctx.h, L.dtype_code(dt), A.n, b, C.byref(M), npan, Xd.ptr, Y.ptr, None, 1.0, 0.0, 0.0, Fd.ptr, None) == -1


@pytest.mark.parametrize("b", [1, 6, 8])
def test_residual_split_and_combine(ctx, rng, b):
    # R = AX - X diag(theta) in FP64 stored in FP32 (and times s), then
    # X diag(h) + E in FP64
    n, npan = 1000, 3
    AX = rng.standard_normal((npan, n, b))
    X = rng.standard_normal((npan, n, b))
    th = rng.uniform(-2, 2, (npan, b))
    AXd, Xd, thd = dev(ctx, AX, np.float64), dev(ctx, X, np.float64), dev(ctx, th, np.float64)
    R, E = ctx.empty((npan, n, b), np.float32), ctx.empty((npan, n, b), np.float32)
    L.check(H().trl_hip_residual_split(ctx.h, n, b, npan, AXd.ptr, Xd.ptr, thd.ptr, -0.75, R.ptr, E.ptr), "split")
    ref = AX - X * th[:, None, :]
    # FP32 rounding of the FP64 value (an FMA may move the FP64 value by an ulp)
    assert np.all(np.abs(R.download() - ref) <= 1.2e-7 * np.abs(ref))
    assert np.all(np.abs(E.download() + 0.75 * ref) <= 1.2e-7 * np.abs(0.75 * ref))
    h = rng.uniform(0.5, 1.5, (npan, b))
    hd = dev(ctx, h, np.float64)
    Ef = E.download().astype(np.float64)
    L.check(H().trl_hip_panel_combine(ctx.h, n, b, npan, Xd.ptr, hd.ptr, E.ptr), "combine")
    assert np.abs(Xd.download() - (X * h[:, None, :] + Ef)).max() <= 1e-15 * np.abs(X).max() * 4
    assert H().trl_hip_panel_combine(ctx.h, n, 5, npan, Xd.ptr, hd.ptr, E.ptr) == -1


@pytest.mark.parametrize("p", ["f64", "f32"])
@pytest.mark.parametrize("b", [1, 8])
def test_spmm_tune(ctx, rng, p, b):
    # the tuner returns a descriptor that computes the same product
    A = P.laplacian_3d(20, 18, 16, rng.uniform(0, 1, 20 * 18 * 16))
    dt = DT[p]
    X = rng.standard_normal((A.n, b)).astype(dt).astype(np.float64)
    rp, ci, va = dev(ctx, A.rowptr, np.int32), dev(ctx, A.col, np.int32), dev(ctx, A.val, dt)
    width, hcol, hval = csr_to_ell(A)
    ec, ev = dev(ctx, hcol, np.int32), dev(ctx, hval, dt)
    csr, ell = spmat(0, rp, 0, ci, va), spmat(1, None, width, ec, ev)
    for e in (None, ell):
        best, us = L.SpMat(), C.c_double(0)
        L.check(H().trl_hip_spmm_tune(ctx.h, L.dtype_code(dt), A.n, b, C.byref(csr), C.byref(e) if e else None, C.byref(best), C.byref(us)), "tune")
        print("tuned", p, b, "ell" if best.format else "csr", best.variant, us.value)
        assert us.value > 0 and best.format in ((0,) if e is None else (0, 1)) and best.variant != 0
        Xd, Y = dev(ctx, X, dt), ctx.empty((A.n, b), dt)
        L.check(H().trl_hip_spmm(ctx.h, L.dtype_code(dt), A.n, b, C.byref(best), Xd.ptr, Y.ptr), "spmm")
        assert rel(Y.download(), oracle.spmm(A, X)) < TOL[p]
        # the fused axpby form has its own tuner; its pick computes the same step
        Z = rng.standard_normal((A.n, b)).astype(dt).astype(np.float64)
        bax, usax = L.SpMat(), C.c_double(0)
        L.check(H().trl_hip_spmm_tune_axpby(ctx.h, L.dtype_code(dt), A.n, b, C.byref(csr), C.byref(e) if e else None, C.byref(bax), C.byref(usax)),
                "tune_axpby")
        print("tuned axpby", p, b, "ell" if bax.format else "csr", bax.variant, usax.value)
        assert usax.value > 0 and bax.format in ((0,) if e is None else (0, 1)) and bax.variant != 0
        Zd, Y2 = dev(ctx, Z, dt), ctx.empty((A.n, b), dt)
        L.check(H().trl_hip_spmm_axpby(ctx.h, L.dtype_code(dt), A.n, b, C.byref(bax), Xd.ptr, Y2.ptr, Zd.ptr, 0.5, -0.25, 0.125), "axpby")
        assert rel(Y2.download(), 0.5 * oracle.spmm(A, X) - 0.25 * X + 0.125 * Z) < TOL[p]
        # and the axpby form on three panels at once (in place, Z = Y)
        bp, usp = L.SpMat(), C.c_double(0)
        L.check(H().trl_hip_spmm_tune_panels(ctx.h, L.dtype_code(dt), A.n, b, 3, C.byref(csr), C.byref(e) if e else None, C.byref(bp), C.byref(usp)),
                "tune_panels")
        print("tuned panels", p, b, "ell" if bp.format else "csr", bp.variant, bp.plane, bp.plane_block, usp.value)
        assert usp.value > 0 and bp.format in ((0,) if e is None else (0, 1))
        X3 = rng.standard_normal((3, A.n, b)).astype(dt).astype(np.float64)
        Z3 = rng.standard_normal((3, A.n, b)).astype(dt).astype(np.float64)
        X3d, Z3d = dev(ctx, X3, dt), dev(ctx, Z3, dt)
        L.check(H().trl_hip_spmm_axpby_panels(ctx.h, L.dtype_code(dt), A.n, b, C.byref(bp), 3, X3d.ptr, Z3d.ptr, Z3d.ptr, 0.5, -0.25, 0.125), "panels")
        got = Z3d.download()
        for j in range(3):
            assert rel(got[j], 0.5 * oracle.spmm(A, X3[j]) - 0.25 * X3[j] + 0.125 * Z3[j]) < TOL[p]


@pytest.mark.parametrize("p", ["f64", "f32"])
@pytest.mark.parametrize("fmt,variant", [("csr", 1), ("csr", -2), ("csr", 4), ("ell", 1), ("ell", -2), ("ell", -4)])
def test_spmm_blocked_row_order(ctx, rng, p, fmt, variant):
    # rows visited plane block by plane block give the same product, for
    # blocks that do not divide the plane and planes that do not divide n
    A = P.laplacian_3d(23, 19, 17, rng.uniform(0, 1, 23 * 19 * 17))
    b, dt = 8, DT[p]
    X = rng.standard_normal((A.n, b)).astype(dt).astype(np.float64)
    Z = rng.standard_normal((A.n, b)).astype(dt).astype(np.float64)
    d = rng.uniform(0.5, 2, A.n).astype(dt)
    rp, ci, va, dd = dev(ctx, A.rowptr, np.int32), dev(ctx, A.col, np.int32), dev(ctx, A.val, dt), dev(ctx, d, dt)
    if fmt == "ell":
        width, hcol, hval = csr_to_ell(A)
        M = spmat(1, None, width, dev(ctx, hcol, np.int32), dev(ctx, hval, dt), dd, variant)
    else:
        M = spmat(0, rp, 0, ci, va, dd, variant)
    ref = oracle.spmm(A, X, d.astype(np.float64))
    for plane, pb in [(23 * 19, 23 * 19), (23 * 19, 100), (23 * 19, 7), (500, 64), (A.n, 1000), (3 * A.n, 5)]:
        M.plane, M.plane_block = plane, pb
        Xd, Y = dev(ctx, X, dt), ctx.empty((A.n, b), dt)
        L.check(H().trl_hip_spmm(ctx.h, L.dtype_code(dt), A.n, b, C.byref(M), Xd.ptr, Y.ptr), "spmm")
        assert rel(Y.download(), ref) < TOL[p], (plane, pb)
        Zd = dev(ctx, Z, dt)
        L.check(H().trl_hip_spmm_axpby(ctx.h, L.dtype_code(dt), A.n, b, C.byref(M), Xd.ptr, Zd.ptr, Zd.ptr, 0.5, -0.25, 0.125), "axpby")
        assert rel(Zd.download(), 0.5 * ref - 0.25 * X + 0.125 * Z) < TOL[p], (plane, pb)
        # two panels at once
        X2 = np.stack([X, Z])
        X2d, Y2 = dev(ctx, X2, dt), ctx.empty((2, A.n, b), dt)
        L.check(H().trl_hip_spmm_panels(ctx.h, L.dtype_code(dt), A.n, b, C.byref(M), 2, X2d.ptr, Y2.ptr), "panels")
        got = Y2.download()
        assert rel(got[0], ref) < TOL[p] and rel(got[1], oracle.spmm(A, Z, d.astype(np.float64))) < TOL[p], (plane, pb)
    # invalid orders are refused
    for plane, pb in [(-1, 1), (100, 0), (100, 101)]:
        M.plane, M.plane_block = plane, pb
        assert H().trl_hip_spmm(ctx.h, L.dtype_code(dt), A.n, b, C.byref(M), Xd.ptr, Y.ptr) == -1


def test_spmm_empty_rows(ctx, rng):
    # rows with no entries give zero rows (ragged CSR)
    n, b = 50, 8
    rowptr = np.zeros(n + 1, np.int32)
    cols, vals = [], []
    for r in range(n):
        k = r % 4  # 0..3 entries, some rows empty
        cs = sorted(rng.choice(n, k, replace=False)) if k else []
        cols += cs
        vals += list(rng.standard_normal(k))
        rowptr[r + 1] = rowptr[r] + k
    A = P.CSR(n, rowptr, np.array(cols, np.int32), np.array(vals))
    X = rng.standard_normal((n, b))
    for lanes in (1, 4, 16, -2, -4):
        Y = ctx.empty((n, b))
        L.check(H().trl_hip_spmm_csr(ctx.h, L.F64, n, b, dev(ctx, A.rowptr, np.int32).ptr, dev(ctx, A.col, np.int32).ptr, dev(ctx, A.val, np.float64).ptr,
                                     dev(ctx, X, np.float64).ptr, Y.ptr, None, lanes), "spmm")
        assert rel(Y.download(), oracle.spmm(A, X)) < 1e-14


@pytest.mark.parametrize("p", ["f64", "f32"])
@pytest.mark.parametrize("b", BS)
def test_panel_ops(ctx, rng, p, b):
    n = 1037
    dt = DT[p]
    X = rng.standard_normal((n, b)).astype(dt).astype(np.float64)
    Y0 = rng.standard_normal((n, b)).astype(dt).astype(np.float64)
    W = rng.standard_normal((b, b))
    Wd = dev(ctx, W, np.float64)
    for mode in range(4):
        Xd, Yd = dev(ctx, X, dt), dev(ctx, Y0, dt)
        L.check(H().trl_hip_panel_mult(ctx.h, L.dtype_code(dt), n, b, Xd.ptr, Wd.ptr, Yd.ptr, mode), "panel_mult")
        Xr, Yr = X.copy(), Y0.copy()
        oracle.panel_mult(Xr, W, Yr, mode)
        if mode == 3:
            assert rel(Xd.download(), Xr) < TOL[p]
        else:
            assert rel(Yd.download(), Yr) < TOL[p]
    # mode 4: X -= Y W^T (mult_transpose into a scratch block, then -=), one pass
    Xd, Yd = dev(ctx, X, dt), dev(ctx, Y0, dt)
    L.check(H().trl_hip_panel_mult(ctx.h, L.dtype_code(dt), n, b, Xd.ptr, Wd.ptr, Yd.ptr, 4), "panel_mult")
    Wt, Xr = np.zeros_like(Y0), X.copy()
    oracle.panel_mult(Y0.copy(), W, Wt, 2)
    assert rel(Xd.download(), Xr - Wt) < TOL[p]
    assert H().trl_hip_panel_mult(ctx.h, L.dtype_code(dt), n, b, Xd.ptr, Wd.ptr, Yd.ptr, 5) == -1
    Xd, Yd = dev(ctx, X, dt), dev(ctx, Y0, dt)
    L.check(H().trl_hip_panel_sub(ctx.h, L.dtype_code(dt), n, b, Xd.ptr, Yd.ptr), "sub")
    assert rel(Xd.download(), X - Y0) < TOL[p]


def test_panel_sub_unaligned(ctx, rng):
    # a panel starting at an offset that is not 16-byte aligned
    n, b = 3, 1
    X = rng.standard_normal(4 * n)
    Y = rng.standard_normal(4 * n)
    Xd, Yd = dev(ctx, X, np.float64), dev(ctx, Y, np.float64)
    L.check(H().trl_hip_panel_sub(ctx.h, L.F64, n, b, Xd.offset(1), Yd.offset(1)), "sub")
    out = Xd.download()
    ref = X.copy()
    ref[1 : 1 + n] -= Y[1 : 1 + n]
    assert np.array_equal(out, ref)


@pytest.mark.parametrize("p", ["f64", "f32"])
@pytest.mark.parametrize("b", BS)
@pytest.mark.parametrize("k", [1, 3, 17, 33])
@pytest.mark.parametrize("mass", [False, True])
def test_gram(ctx, rng, p, b, k, mass):
    print(f"gram {p} b={b} k={k} mass={mass}", flush=True)
    n = 4099
    dt = DT[p]
    V = rng.standard_normal((k, n, b)).astype(dt).astype(np.float64)
    W = rng.standard_normal((n, b)).astype(dt).astype(np.float64)
    m = rng.uniform(0.5, 2, n).astype(dt).astype(np.float64) if mass else None
    Cd = ctx.empty((k * b, b))
    md = dev(ctx, m, dt) if mass else None
    L.check(H().trl_hip_gram(ctx.h, L.dtype_code(dt), n, b, k, dev(ctx, V, dt).ptr, dev(ctx, W, dt).ptr, md.ptr if md else None, Cd.ptr), "gram")
    ref = oracle.gram(V, W, m)
    scale = np.sqrt(n)  # entries ~ sqrt(n); compare against the typical magnitude
    err = np.abs(Cd.download() - ref).max() / scale
    assert err < (1e-13 if p == "f64" else 2e-6)


def test_gram_many_blocks_and_tiny_n(ctx, rng):
    # k beyond one launch (chunked on the host) and n < one row quad
    for n, b, k in [(3, 8, 1), (1, 4, 5), (257, 1, 1100), (301, 8, 140)]:
        V = rng.standard_normal((k, n, b))
        W = rng.standard_normal((n, b))
        Cd = ctx.empty((k * b, b))
        L.check(H().trl_hip_gram(ctx.h, L.F64, n, b, k, dev(ctx, V, np.float64).ptr, dev(ctx, W, np.float64).ptr, None, Cd.ptr), "gram")
        assert rel(Cd.download(), oracle.gram(V, W)) < 1e-12


def test_gram_deterministic(ctx, rng):
    n, b, k = 100003, 8, 9
    V = dev(ctx, rng.standard_normal((k, n, b)), np.float32)
    W = dev(ctx, rng.standard_normal((n, b)), np.float32)
    outs = []
    for _ in range(3):
        Cd = ctx.empty((k * b, b))
        L.check(H().trl_hip_gram(ctx.h, L.F32, n, b, k, V.ptr, W.ptr, None, Cd.ptr), "gram")
        outs.append(Cd.download())
    assert np.array_equal(outs[0], outs[1]) and np.array_equal(outs[0], outs[2])


def test_gram_fp32_long_sum_exact(ctx):
    # 2^22 rows of x = 1, y = 1 + 2^-10.  Every partial sum inside one FP32 MFMA
    # chunk (<= 32 rows) is exactly representable, and the chunk sums are folded
    # into an error-free (hi, lo) pair, then summed across waves in FP64: the
    # result must be exact.  A plain FP32 running sum would drop the 2^-10
    # residues once it passes 2^13.
    n, b = 1 << 22, 1
    x = np.ones((n, b), np.float32)
    y = np.full((n, b), 1.0 + 2.0**-10, np.float32)
    Cd = ctx.empty((1, 1))
    L.check(H().trl_hip_gram(ctx.h, L.F32, n, b, 1, dev(ctx, x, np.float32).ptr, dev(ctx, y, np.float32).ptr, None, Cd.ptr), "gram")
    assert Cd.download()[0, 0] == n * (1.0 + 2.0**-10)


@pytest.mark.parametrize("p", ["f64", "f32"])
@pytest.mark.parametrize("b", BS)
@pytest.mark.parametrize("k", [1, 5, 33])
def test_update(ctx, rng, p, b, k):
    n = 2051
    dt = DT[p]
    V = rng.standard_normal((k, n, b)).astype(dt).astype(np.float64)
    W = rng.standard_normal((n, b)).astype(dt).astype(np.float64)
    Cm = rng.standard_normal((k * b, b)) * 0.1
    Wd = dev(ctx, W, dt)
    L.check(H().trl_hip_update(ctx.h, L.dtype_code(dt), n, b, k, dev(ctx, V, dt).ptr, dev(ctx, Cm, np.float64).ptr, Wd.ptr), "update")
    ref = W.copy()
    oracle.update(V, Cm.astype(dt).astype(np.float64), ref)
    assert rel(Wd.download(), ref) < (1e-12 if p == "f64" else 1e-5)


@pytest.mark.parametrize("b", BS)
@pytest.mark.parametrize("mass", [False, True])
def test_cholqr_fp64_matches_oracle(ctx, rng, b, mass):
    n = 3001
    X = rng.standard_normal((n, b))
    m = rng.uniform(0.5, 2, n) if mass else None
    Xd = dev(ctx, X, np.float64)
    Rd = ctx.empty((b, b))
    md = dev(ctx, m, np.float64) if mass else None
    L.check(H().trl_hip_cholqr(ctx.h, L.F64, n, b, Xd.ptr, md.ptr if md else None, Rd.ptr, 1), "cholqr")
    assert ctx.check() == 0
    Xr = X.copy()
    rc, Rr = oracle.cholqr(Xr, m)
    assert rc == 0
    assert rel(Rd.download(), Rr) < 1e-12
    assert rel(Xd.download(), Xr) < 1e-11


@pytest.mark.parametrize("p", ["f64", "f32"])
@pytest.mark.parametrize("b", [2, 8])
def test_cholqr2_contract(ctx, rng, p, b):
    # ill-conditioned block: CholQR2 restores orthogonality, X_old = X_new R
    n = 20000
    dt = DT[p]
    Q, _ = np.linalg.qr(rng.standard_normal((n, b)))
    X = (Q * np.logspace(0, 2, b)) @ np.linalg.qr(rng.standard_normal((b, b)))[0]
    X = X.astype(dt).astype(np.float64)
    Xd = dev(ctx, X, dt)
    Rd = ctx.empty((b, b))
    L.check(H().trl_hip_cholqr(ctx.h, L.dtype_code(dt), n, b, Xd.ptr, None, Rd.ptr, 2), "cholqr2")
    assert ctx.check() == 0
    Xn = Xd.download().astype(np.float64)
    R = Rd.download()
    assert np.allclose(np.tril(R, -1), 0) and (np.diag(R) > 0).all()
    assert np.abs(Xn.T @ Xn - np.eye(b)).max() < (1e-14 * n**0.5 if p == "f64" else 1e-6)
    assert rel(Xn @ R, X) < (1e-13 if p == "f64" else 1e-6)


@pytest.mark.parametrize("n", [1, 100, 3001, 600001])
@pytest.mark.parametrize("b", [1, 4, 6, 8])
@pytest.mark.parametrize("mass", [False, True])
def test_cholqr2_fp64_matches_oracle_twice(ctx, rng, n, b, mass):
    # FP64 CholQR2 (two Gram-by-rows passes, each factored in its last
    # workgroup, then the two-step scale) equals the oracle's CholQR applied
    # twice: R = R2 R1 and X R^{-1}; n = 1 and 100 run on one workgroup,
    # 600001 on the capped grid.  Two calls on the same input are bitwise equal.
    print(f"cholqr2 f64 n={n} b={b} mass={mass}", flush=True)
    X = rng.standard_normal((n, b)) * np.logspace(0, 1, b)
    m = rng.uniform(0.5, 2, n) if mass else None
    md = dev(ctx, m, np.float64) if mass else None
    Rd = ctx.empty((b, b))
    Xd = dev(ctx, X, np.float64)
    rc = H().trl_hip_cholqr(ctx.h, L.F64, n, b, Xd.ptr, md.ptr if md else None, Rd.ptr, 2)
    if n < b:
        # rank-deficient: the first pass breaks down
        L.check(rc, "cholqr2")
        assert ctx.check() > 0
        return
    L.check(rc, "cholqr2")
    assert ctx.check() == 0
    Xr = X.copy()
    rc1, R1 = oracle.cholqr(Xr, m)
    rc2, R2 = oracle.cholqr(Xr, m)
    assert rc1 == 0 and rc2 == 0
    assert rel(Rd.download(), R2 @ R1) < 1e-12
    assert rel(Xd.download(), Xr) < 1e-11
    X2 = dev(ctx, X, np.float64)
    R2d = ctx.empty((b, b))
    L.check(H().trl_hip_cholqr(ctx.h, L.F64, n, b, X2.ptr, md.ptr if md else None, R2d.ptr, 2), "cholqr2 again")
    assert np.array_equal(X2.download(), Xd.download()) and np.array_equal(R2d.download(), Rd.download())


@pytest.mark.parametrize("b", [4, 8])
def test_cholqr2_fp32_mass_and_breakdown(ctx, rng, b):
    # FP32 CholQR2 with a mass diagonal: X^T M X = I to FP32 accuracy and
    # X_old = X_new R; then a zero column: the first pass's pivot is flagged
    # and X is left unchanged
    n = 50000
    X = rng.standard_normal((n, b)).astype(np.float32)
    m = rng.uniform(0.5, 2, n).astype(np.float32)
    Xd = dev(ctx, X, np.float32)
    md = dev(ctx, m, np.float32)
    Rd = ctx.empty((b, b))
    L.check(H().trl_hip_cholqr(ctx.h, L.F32, n, b, Xd.ptr, md.ptr, Rd.ptr, 2), "cholqr2")
    assert ctx.check() == 0
    Xn = Xd.download().astype(np.float64)
    R = Rd.download()
    assert np.abs((Xn * m[:, None]).T @ Xn - np.eye(b)).max() < 1e-6
    assert rel(Xn @ R, X.astype(np.float64)) < 1e-6
    Xz = X.copy()
    Xz[:, 1] = 0.0
    Xzd = dev(ctx, Xz, np.float32)
    L.check(H().trl_hip_cholqr(ctx.h, L.F32, n, b, Xzd.ptr, None, Rd.ptr, 2), "cholqr2 breakdown")
    assert ctx.check() == 2
    assert np.array_equal(Xzd.download(), Xz)
    assert ctx.check() == 0


def test_cholqr_breakdown_flag(ctx, rng):
    n, b = 500, 4
    X = rng.standard_normal((n, b))
    X[:, 2] = 0.0
    Xd = dev(ctx, X, np.float64)
    Rd = ctx.empty((b, b))
    L.check(H().trl_hip_cholqr(ctx.h, L.F64, n, b, Xd.ptr, None, Rd.ptr, 1), "cholqr")
    assert ctx.check() == 3  # 1-based failing pivot, as the oracle reports it
    assert oracle.cholqr(X.copy())[0] == 3
    assert np.array_equal(Xd.download(), X)  # left unchanged
    assert ctx.check() == 0  # flag cleared


@pytest.mark.parametrize("p", ["f64", "f32"])
@pytest.mark.parametrize("b", BS)
@pytest.mark.parametrize("m,kc", [(3, 1), (8, 3), (32, 5), (5, 5), (12, 4), (2, 7)])
def test_rotate(ctx, rng, p, b, m, kc):
    # short contractions (a tile's operand row in registers, the next tile
    # prefetched: up to 6 super-steps) and long ones, output chunks of up to
    # 4 x 16 columns (kc = 7 at b = 16 takes two launches), a ragged tail tile
    n = 1999
    dt = DT[p]
    V = rng.standard_normal((m, n, b)).astype(dt).astype(np.float64)
    ybc = kc + 1
    Y = rng.standard_normal((m, ybc, b, b))
    Out = ctx.empty((kc, n, b), dt)
    L.check(H().trl_hip_rotate(ctx.h, L.dtype_code(dt), n, b, m, dev(ctx, V, dt).ptr, dev(ctx, Y, np.float64).ptr, ybc, kc, Out.ptr), "rotate")
    ref = oracle.rotate(V, Y.astype(dt).astype(np.float64), kc)
    assert rel(Out.download(), ref) < (1e-12 if p == "f64" else 1e-5)


@pytest.mark.parametrize("p", ["f64", "f32"])
@pytest.mark.parametrize("b,m,kc", [(8, 5, 5), (8, 32, 5), (6, 12, 4), (16, 2, 7), (1, 40, 3)])
def test_rotate_address_forms(ctx, rng, p, b, m, kc):
    # the buffer-resource forms (short and long contractions) and the
    # 64-bit-address form that buffers past 2 GiB take (forced by the
    # OFFSETS64 option), each against the oracle on a ragged tail tile
    n = 1999
    dt = DT[p]
    V = rng.standard_normal((m, n, b)).astype(dt).astype(np.float64)
    Y = rng.standard_normal((m, kc + 1, b, b))
    ref = oracle.rotate(V, Y.astype(dt).astype(np.float64), kc)
    Vd, Yd = dev(ctx, V, dt), dev(ctx, Y, np.float64)
    got = []
    for wide in (0, 1):
        Out = ctx.empty((kc, n, b), dt)
        ctx.set_option(L.OPT_OFFSETS64, wide)
        try:
            L.check(H().trl_hip_rotate(ctx.h, L.dtype_code(dt), n, b, m, Vd.ptr, Yd.ptr, kc + 1, kc, Out.ptr), "rotate")
        finally:
            ctx.set_option(L.OPT_OFFSETS64, 0)
        got.append(Out.download())
    for g in got:
        assert rel(g, ref) < (1e-12 if p == "f64" else 1e-5)


def test_small_ops(ctx, rng):
    for b in (1, 3, 8, 16):
        A = rng.standard_normal((b, b))
        Bm = rng.standard_normal((b, b))
        Ad, Bd, Cd = dev(ctx, A, np.float64), dev(ctx, Bm, np.float64), ctx.empty((b, b))
        L.check(H().trl_hip_small_mult(ctx.h, b, Ad.ptr, Bd.ptr, Cd.ptr), "mult")
        assert rel(Cd.download(), A @ Bm) < 1e-14
        L.check(H().trl_hip_small_copy(ctx.h, b, Ad.ptr, Cd.ptr, 1), "copy_t")
        assert np.array_equal(Cd.download(), A.T)
        vals = np.arange(1.0, b + 1)
        L.check(H().trl_hip_small_set_diagonal(ctx.h, b, Cd.ptr, vals.ctypes.data), "diag")
        ref = A.T.copy()
        np.fill_diagonal(ref, vals)
        assert np.array_equal(Cd.download(), ref)
        L.check(H().trl_hip_small_zero(ctx.h, b, Cd.ptr), "zero")
        assert not Cd.download().any()


@pytest.mark.parametrize("p", ["f64", "f32"])
def test_residual_sq(ctx, rng, p):
    n, b = 777, 4
    dt = DT[p]
    X = rng.standard_normal((n, b)).astype(dt).astype(np.float64)
    Y = rng.standard_normal((n, b)).astype(dt).astype(np.float64)
    m = rng.uniform(0.5, 2, n).astype(dt).astype(np.float64)
    lam = rng.standard_normal(b)
    out = np.zeros(2 * b)
    L.check(H().trl_hip_residual_sq(ctx.h, L.dtype_code(dt), n, b, dev(ctx, X, dt).ptr, dev(ctx, Y, dt).ptr, dev(ctx, m, dt).ptr, lam.ctypes.data,
                                    out.ctypes.data), "res")
    r = Y - lam * m[:, None] * X
    assert np.allclose(out[:b], (r * r).sum(0), rtol=1e-12)
    assert np.allclose(out[b:], (Y * Y).sum(0), rtol=1e-12)


@pytest.mark.parametrize("b", [1, 4, 8])
def test_sddmm(ctx, rng, b):
    A = P.laplacian_2d(13, 11)
    nblk = 3
    X = rng.standard_normal((nblk, A.n, b))
    w = rng.uniform(0, 1, nblk * b)
    lam = rng.uniform(1, 2, nblk * b)
    G = ctx.empty(A.nnz)
    gm = ctx.empty(A.n)
    L.check(H().trl_hip_sddmm_adjoint(ctx.h, L.F64, A.n, dev(ctx, A.rowptr, np.int32).ptr, dev(ctx, A.col, np.int32).ptr, b, nblk,
                                      dev(ctx, X, np.float64).ptr, dev(ctx, w, np.float64).ptr, dev(ctx, lam, np.float64).ptr, G.ptr, gm.ptr), "sddmm")
    Xf = np.concatenate(list(X), axis=1)  # n x nvec
    rows = np.repeat(np.arange(A.n), np.diff(A.rowptr))
    ref = ((Xf[rows] * w) * Xf[A.col]).sum(1)
    assert rel(G.download(), ref) < 1e-13
    assert rel(gm.download(), -((Xf**2) * (w * lam)).sum(1)) < 1e-13


def test_convert_and_fill(ctx, rng):
    x = rng.standard_normal(1001)
    d = dev(ctx, x, np.float64)
    f = ctx.empty(1001, np.float32)
    L.check(H().trl_hip_convert(ctx.h, L.F64, d.ptr, L.F32, f.ptr, 1001), "convert")
    assert np.array_equal(f.download(), x.astype(np.float32))
    L.check(H().trl_hip_fill_zero(ctx.h, d.ptr, 1001 * 8), "fill")
    assert not d.download().any()


def test_region_timer(ctx):
    # one event pair around queued work: positive, and ten times the work
    # takes longer than one
    d = ctx.empty((64 << 20,), np.float64)
    t1 = ctx.time_region(lambda: L.check(H().trl_hip_fill_zero(ctx.h, d.ptr, d.nbytes), "fill"), 1)
    t10 = ctx.time_region(lambda: L.check(H().trl_hip_fill_zero(ctx.h, d.ptr, d.nbytes), "fill"), 10) * 10
    assert 0 < t1 < t10
    s = C.c_double()
    assert H().trl_hip_region_end(None, C.byref(s)) == -1
    assert H().trl_hip_region_end(ctx.h, None) == -1


def test_invalid_arguments(ctx):
    assert H().trl_hip_gram(ctx.h, L.F64, 10, 3, 1, None, None, None, None) == -1  # b = 3 unsupported
    assert H().trl_hip_spmm_csr(ctx.h, L.F64, 10, 4, None, None, None, None, None, None, 3) == -1
    assert H().trl_hip_spmm_csr(ctx.h, L.F64, 10, 4, None, None, None, None, None, None, -3) == -1
    assert H().trl_hip_spmm(ctx.h, L.F64, 10, 4, None, None, None) == -1
    bad = L.SpMat(2, None, 0, None, None, None, 0)
    assert H().trl_hip_spmm(ctx.h, L.F64, 10, 4, C.byref(bad), None, None) == -1
    assert H().trl_hip_cholqr(ctx.h, L.F64, 10, 4, None, None, None, 3) == -1


@pytest.mark.parametrize("p", ["f64", "f32"])
@pytest.mark.parametrize("b", [1, 4, 8])
@pytest.mark.parametrize("mass", [False, True])
@pytest.mark.parametrize("contiguous", [True, False])
def test_reorth_cgs_dgks(ctx, rng, p, b, mass, contiguous):
    # W against k M-orthonormal blocks: eta <= 0 is CGS2 (compared with the
    # oracle's gram/update applied twice); eta = 1/sqrt(2) takes the second
    # pass only for a W with large components along V
    dt = DT[p]
    n, k = 3000, 5
    m = rng.uniform(0.5, 2.0, n) if mass else np.ones(n)
    Q, _ = np.linalg.qr(rng.standard_normal((n, k * b)))
    Q = Q / np.sqrt(m)[:, None]  # M-orthonormal columns: Q^T M Q = I
    Vh = np.stack([Q[:, j * b:(j + 1) * b] for j in range(k)]).astype(dt)
    md = dev(ctx, m, dt) if mass else None
    tol = 1e-12 if p == "f64" else 2e-5
    for case in ("cancel", "near"):
        if case == "cancel":  # mostly inside span(V): the first pass cancels
            W0 = 1e-3 * rng.standard_normal((n, b)) + 10.0 * Q[:, :b] @ rng.standard_normal((b, b))
        else:
            W0 = rng.standard_normal((n, b))
            W0 -= Q @ (Q.T @ (m[:, None] * W0))
            W0 += 1e-9 * Q[:, b:2 * b]
        W0 = W0.astype(dt)
        for eta in (0.0, 0.70710678118654752):
            if contiguous:
                Vd = dev(ctx, np.concatenate([Vh, W0[None]]), dt)
                Wp = Vd.offset(k * n * b)
            else:
                Vd = dev(ctx, Vh, dt)
                Wd = dev(ctx, W0, dt)
                Wp = Wd.ptr
            Cd = ctx.empty(((2 * k + 1) * b * b,))
            ctx.prof_enable(True)
            L.check(H().trl_hip_reorth(ctx.h, L.dtype_code(dt), n, b, k, Vd.ptr, Wp, md.ptr if md else None, Cd.ptr, eta), "reorth")
            calls, second = C.c_longlong(), C.c_longlong()
            L.check(H().trl_hip_prof_reorth(ctx.h, C.byref(calls), C.byref(second)), "prof_reorth")
            ctx.prof_enable(False)
            out = np.empty((n, b), dt)
            L.check(H().trl_hip_memcpy_d2h(ctx.h, out.ctypes.data, Wp, out.nbytes), "d2h")
            Vf = Vh.astype(np.float64)
            ref = W0.astype(np.float64)
            for _ in range(2 if (eta <= 0 or case == "cancel") else 1):
                oracle.update(Vf, oracle.gram(Vf, ref, m if mass else None), ref)
            assert rel(out, ref) < tol * 10
            ortho = np.abs(Q.T @ (m[:, None] * out.astype(np.float64))).max() / np.abs(out).max()
            assert ortho < (1e-13 if p == "f64" else 1e-5), (case, eta, ortho)
            if eta > 0:
                assert calls.value == 1 and second.value == (1 if case == "cancel" else 0), (case, calls.value, second.value)


@pytest.mark.parametrize("p", ["f64", "f32"])
@pytest.mark.parametrize("b", [1, 4, 6, 8, 16])
@pytest.mark.parametrize("variant", ["z", "noz", "inplace", "scaled"])
def test_spmm_axpby(ctx, rng, p, b, variant):
    # Y = alpha diag(d) A X + beta X + gamma Z
    A = P.laplacian_3d(6, 5, 7, rng.uniform(0, 1, 210))
    dt = DT[p]
    X = rng.standard_normal((A.n, b)).astype(dt).astype(np.float64)
    Z = rng.standard_normal((A.n, b)).astype(dt).astype(np.float64)
    d = rng.uniform(0.5, 2, A.n).astype(dt).astype(np.float64)
    al, be, ga = 0.7, -1.3, 0.4
    rp, ci, va = dev(ctx, A.rowptr, np.int32), dev(ctx, A.col, np.int32), dev(ctx, A.val, dt)
    Xd, Zd = dev(ctx, X, dt), dev(ctx, Z, dt)
    Yd = Zd if variant == "inplace" else dev(ctx, np.zeros_like(X), dt)
    zp = None if variant == "noz" else Zd.ptr
    dp = dev(ctx, d, dt).ptr if variant == "scaled" else None
    L.check(H().trl_hip_spmm_csr_axpby(ctx.h, L.dtype_code(dt), A.n, b, rp.ptr, ci.ptr, va.ptr, dp, Xd.ptr, Yd.ptr, zp, al, be, ga), "axpby")
    AX = oracle.spmm(A, X, d if variant == "scaled" else None)
    ref = al * AX + be * X + (0.0 if variant == "noz" else ga * Z)
    assert rel(Yd.download(), ref) < TOL[p]
    assert H().trl_hip_spmm_csr_axpby(ctx.h, L.dtype_code(dt), A.n, b, rp.ptr, ci.ptr, va.ptr, None, Xd.ptr, Xd.ptr, None, 1.0, 0.0, 0.0) == -1


@pytest.mark.parametrize("p", ["f64", "f32"])
@pytest.mark.parametrize("fmt,variant", [("csr", 1), ("csr", -2), ("csr", -4), ("csr", 2), ("csr", 8), ("ell", 1), ("ell", -2), ("ell", -4)])
def test_spmm_offsets64(ctx, rng, p, fmt, variant):
    # the 64-bit-offset instantiations (taken on their own only for panels or
    # ELL arrays beyond 4 GiB), forced by the context option: the same
    # products, bit for bit, as the 32-bit-offset ones, and the oracle's, for
    # one panel, several panels, the axpby and forced steps, in index and
    # blocked row order
    b, dt = 8, DT[p]
    if variant < 0 and not split_ok(p, b, -variant):
        pytest.skip("column split needs 8 bytes per lane")
    print(f"offsets64 {p} {fmt} {variant}", flush=True)
    A = P.laplacian_3d(11, 9, 7, rng.uniform(0, 1, 693))
    npan = 3
    X = rng.standard_normal((npan, A.n, b)).astype(dt)
    Z = rng.standard_normal((npan, A.n, b)).astype(dt)
    F = rng.standard_normal((npan, A.n, b)).astype(dt)
    g = rng.standard_normal((npan, b))
    d = rng.uniform(0.5, 2, A.n).astype(dt)
    if fmt == "ell":
        width, hcol, hval = csr_to_ell(A)
        M = spmat(1, None, width, dev(ctx, hcol, np.int32), dev(ctx, hval, dt), dev(ctx, d, dt), variant)
    else:
        M = spmat(0, dev(ctx, A.rowptr, np.int32), 0, dev(ctx, A.col, np.int32), dev(ctx, A.val, dt), dev(ctx, d, dt), variant)
    Xd, Fd, gd = dev(ctx, X, dt), dev(ctx, F, dt), dev(ctx, g, np.float64)
    t = L.dtype_code(dt)

    def run(order):
        M.plane, M.plane_block = order
        out = []
        Y = ctx.empty((npan, A.n, b), dt)
        L.check(H().trl_hip_spmm(ctx.h, t, A.n, b, C.byref(M), Xd.ptr, Y.ptr), "spmm")
        out.append(Y.download()[0])
        L.check(H().trl_hip_spmm_panels(ctx.h, t, A.n, b, C.byref(M), npan, Xd.ptr, Y.ptr), "panels")
        out.append(Y.download())
        Zd = dev(ctx, Z, dt)
        L.check(H().trl_hip_spmm_axpby_panels(ctx.h, t, A.n, b, C.byref(M), npan, Xd.ptr, Zd.ptr, Zd.ptr, 0.5, -1.0, 0.25), "axpby")
        out.append(Zd.download())
        L.check(H().trl_hip_spmm_axpby_forced(ctx.h, t, A.n, b, C.byref(M), npan, Xd.ptr, Y.ptr, None, 0.5, -1.0, 0.0, Fd.ptr, gd.ptr), "forced")
        out.append(Y.download())
        return out

    AX = np.stack([oracle.spmm(A, X[j].astype(np.float64), d.astype(np.float64)) for j in range(npan)])
    for order in [(0, 0), (11 * 9, 40)]:
        ref32 = run(order)
        ctx.set_option(L.OPT_OFFSETS64, 1)
        try:
            got = run(order)
        finally:
            ctx.set_option(L.OPT_OFFSETS64, 0)
        for a, c in zip(got, ref32):
            assert np.array_equal(a, c), order
        assert rel(got[0], AX[0]) < TOL[p]
        assert rel(got[1], AX) < TOL[p]
        assert rel(got[2], 0.5 * AX - X + 0.25 * Z) < TOL[p]
        assert rel(got[3], 0.5 * AX - X + F * g[:, None, :]) < TOL[p]
    assert H().trl_hip_ctx_set_option(ctx.h, 99, 1) == -1


@pytest.mark.parametrize("p", ["f64", "f32"])
@pytest.mark.parametrize("b,variant", [(8, 1), (8, -2), (8, -4), (4, 1), (1, 1), (6, 1)])
def test_spmm_panels_shared_gathers_and_tiles(ctx, rng, p, b, variant):
    # the several-panel ELL step with in-wave gathers taken from the owning
    # lane (share) and with workgroups visiting tiles of lines (tile_lines x
    # line): the plain product bit for bit the same as without them (the same
    # products in the same order), the axpby and forced epilogues to the
    # rounding of the compiler's fused multiply-adds, and the oracle's; in
    # index order, blocked orders the tiles cover, blocked orders they do not
    # (the step then visits rows in block order) and orders whose segments
    # end inside waves; a ragged matrix (a stencil with boundary rows and an
    # empty row)
    dt = DT[p]
    if variant < 0 and not split_ok(p, b, -variant):
        pytest.skip("column split needs 8 bytes per lane")
    print(f"shared gathers and tiles {p} b={b} variant={variant}", flush=True)
    nx, ny, nz = 64, 8, 3
    A = P.laplacian_3d(nx, ny, nz, rng.uniform(0, 1, nx * ny * nz))
    lo, hi = A.rowptr[100], A.rowptr[101]
    keep = np.ones(A.nnz, bool)
    keep[lo:hi] = False
    rp = A.rowptr.copy()
    rp[101:] -= hi - lo
    A = P.CSR(A.n, rp, A.col[keep], A.val[keep])
    npan = 3
    X = rng.standard_normal((npan, A.n, b)).astype(dt)
    Z = rng.standard_normal((npan, A.n, b)).astype(dt)
    F = rng.standard_normal((npan, A.n, b)).astype(dt)
    g = rng.standard_normal((npan, b))
    d = rng.uniform(0.5, 2, A.n).astype(dt)
    width, hcol, hval = csr_to_ell(A)
    M = spmat(1, None, width, dev(ctx, hcol, np.int32), dev(ctx, hval, dt), dev(ctx, d, dt), variant)
    Xd, Fd, gd = dev(ctx, X, dt), dev(ctx, F, dt), dev(ctx, g, np.float64)
    t = L.dtype_code(dt)

    def run():
        out = []
        Y = ctx.empty((npan, A.n, b), dt)
        L.check(H().trl_hip_spmm_panels(ctx.h, t, A.n, b, C.byref(M), npan, Xd.ptr, Y.ptr), "panels")
        out.append(Y.download())
        Zd = dev(ctx, Z, dt)
        L.check(H().trl_hip_spmm_axpby_panels(ctx.h, t, A.n, b, C.byref(M), npan, Xd.ptr, Zd.ptr, Zd.ptr, 0.5, -1.0, 0.25), "axpby")
        out.append(Zd.download())
        L.check(H().trl_hip_spmm_axpby_forced(ctx.h, t, A.n, b, C.byref(M), npan, Xd.ptr, Y.ptr, None, 0.5, -1.0, 0.0, Fd.ptr, gd.ptr), "forced")
        out.append(Y.download())
        return out

    AX = np.stack([oracle.spmm(A, X[j].astype(np.float64), d.astype(np.float64)) for j in range(npan)])
    plane = nx * ny
    for order in [(0, 0), (plane, 4 * nx), (plane, plane), (plane, 2 * nx), (200, 37)]:
        M.plane, M.plane_block = order
        M.share, M.line, M.tile_lines = 0, 0, 0
        ref = run()
        for share, tile in [(1, 0), (0, 4), (1, 4), (1, 2)]:
            M.share, M.line, M.tile_lines = share, (nx if tile else 0), tile
            got = run()
            assert np.array_equal(got[0], ref[0]), (order, share, tile)
            for a, c in zip(got[1:], ref[1:]):
                assert rel(a, c) <= 4 * np.finfo(dt).eps, (order, share, tile)
            assert rel(got[0], AX) < TOL[p]
            assert rel(got[1], 0.5 * AX - X + 0.25 * Z) < TOL[p]
            assert rel(got[2], 0.5 * AX - X + F * g[:, None, :]) < TOL[p]
    # invalid option values are refused
    M.share = 2
    assert H().trl_hip_spmm_panels(ctx.h, t, A.n, b, C.byref(M), npan, Xd.ptr, Xd.offset(0)) == -1
