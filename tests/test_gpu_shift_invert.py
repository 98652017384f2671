"""Generalized problems A x = lambda B x with a sparse SPD B by shift-invert
(the reference's CSRGeneralizedEVP, examples/openmp/csrevp.hh:51-104, driven
as in examples/openmp/gevp_test.cc), through the C-ABI.

Pinned by the closed form of the bilinear finite-element Laplacian pair
(fe_laplacian_2d: every generalized eigenvalue is lambda_k + lambda_l of the
1D factors) and by LAPACK's dense generalized eigensolver (scipy.linalg.eigh)
on a random stiffness / sparse-mass pair.  The pieces the solve is built
from are checked on their own: the B-inner-product CholQR
(trl_hip_cholqr_bprod) against its contract, and the inner conjugate
gradient solve (trl_hip_pcg) against a direct solve, with its two error
exits.
"""

import ctypes as C

import numpy as np
import pytest
import scipy.linalg as sla

from trl_amd import _lib as L
from trl_amd import problems as P
from trl_amd.solver import EigensolverParams, solve_shift_invert

pytestmark = pytest.mark.gpu

_KEEP = []


@pytest.fixture(autouse=True)
def _keep_device_arrays(ctx):
    yield
    ctx.sync()
    _KEEP.clear()


def dev(ctx, a, dt=np.float64):
    d = ctx.to_device(np.ascontiguousarray(a, dtype=dt))
    _KEEP.append(d)
    return d


def csr_desc(ctx, A):
    return L.SpMat(0, dev(ctx, A.rowptr, np.int32).ptr, 0, dev(ctx, A.col, np.int32).ptr, dev(ctx, A.val).ptr, None, 0)


def shifted(A, B, sigma):
    K = (A.to_scipy() - sigma * B.to_scipy()).tocsr()
    K.sort_indices()
    return P.CSR(A.n, K.indptr.astype(np.int32), K.indices.astype(np.int32), K.data.astype(np.float64)), K


@pytest.mark.parametrize("b", [1, 4, 6, 8])
@pytest.mark.parametrize("passes", [1, 2])
def test_cholqr_bprod_contract(ctx, rng, b, passes):
    # W^T B W = I, W_old = W_new R with R upper triangular (positive diagonal),
    # and BW leaves the call as B times the new W
    print(f"cholqr_bprod b={b} passes={passes}", flush=True)
    _, Bm = P.stiffness_sparse_mass(8, 8, 6, dof=3)
    Bs = Bm.to_scipy()
    W = rng.standard_normal((Bm.n, b))
    BW = Bs @ W
    Wd, BWd, R = dev(ctx, W), dev(ctx, BW), ctx.empty((b, b))
    L.check(L.hip().trl_hip_cholqr_bprod(ctx.h, L.F64, Bm.n, b, Wd.ptr, BWd.ptr, R.ptr, passes), "cholqr_bprod")
    Wn, BWn, Rh = Wd.download(), BWd.download(), R.download()
    assert np.abs(Wn.T @ (Bs @ Wn) - np.eye(b)).max() < 1e-12
    assert np.abs(Wn @ Rh - W).max() < 1e-12 * np.abs(W).max()
    assert np.allclose(np.tril(Rh, -1), 0.0) and np.all(np.diag(Rh) > 0)
    assert np.abs(BWn - Bs @ Wn).max() < 1e-12 * np.abs(BWn).max()
    # W and BW may not be the same panel
    assert L.hip().trl_hip_cholqr_bprod(ctx.h, L.F64, Bm.n, b, Wd.ptr, Wd.ptr, R.ptr, passes) == -1


@pytest.mark.parametrize("b", [1, 4, 8, 16])
def test_pcg_against_direct_solve(ctx, rng, b):
    # K = A - sigma B (FE pair, sigma half the smallest eigenvalue: SPD): the
    # device solve against a dense LAPACK solve, and its true residual
    print(f"pcg b={b}", flush=True)
    A, B = P.fe_laplacian_2d(30, 23)
    sigma = 0.5 * P.fe_laplacian_2d_eigenvalues(30, 23)[0]
    Kc, Ks = shifted(A, B, sigma)
    rhs = rng.standard_normal((A.n, b))
    if b > 1:
        rhs[:, 0] = 0.0  # a zero right-hand side beside others: its column stays X = 0
    dinv = 1.0 / Ks.diagonal()
    K = csr_desc(ctx, Kc)
    X, work = ctx.empty((A.n, b)), ctx.empty((L.hip().trl_hip_pcg_workspace(A.n, b) // 8 + 1,))
    it, rel = C.c_int(0), C.c_double(0)
    L.check(L.hip().trl_hip_pcg(ctx.h, L.F64, A.n, b, C.byref(K), dev(ctx, dinv).ptr, dev(ctx, rhs).ptr, X.ptr, 1e-12, 5000, work.ptr, C.byref(it),
                                C.byref(rel)), "pcg")
    Xh = X.download()
    ref = sla.solve(Ks.toarray(), rhs, assume_a="pos")
    true = np.linalg.norm(Ks @ Xh - rhs, axis=0) / np.maximum(np.linalg.norm(rhs, axis=0), 1e-300)
    print("iterations", it.value, "recurrence", rel.value, "true", true.max(), flush=True)
    assert 0 < it.value < 5000 and rel.value <= 1e-12
    assert true.max() <= 1e-11
    assert np.abs(Xh - ref).max() <= 1e-9 * np.abs(ref).max()
    if b > 1:
        assert np.all(Xh[:, 0] == 0.0)


def test_pcg_error_exits(ctx, rng):
    # a negative definite K gives p^T K p < 0 at the first step (EINDEF); an
    # iteration cap below what the tolerance needs gives ENOCONV; FP32 and an
    # in-place solve are refused
    A, B = P.fe_laplacian_2d(20, 20)
    b = 4
    Kc, Ks = shifted(A, B, 0.0)
    neg = P.CSR(Kc.n, Kc.rowptr, Kc.col, -Kc.val)
    rhs = dev(ctx, rng.standard_normal((A.n, b)))
    dinv = dev(ctx, 1.0 / Ks.diagonal())
    X, work = ctx.empty((A.n, b)), ctx.empty((L.hip().trl_hip_pcg_workspace(A.n, b) // 8 + 1,))
    it, rel = C.c_int(0), C.c_double(0)
    Kn = csr_desc(ctx, neg)
    assert L.hip().trl_hip_pcg(ctx.h, L.F64, A.n, b, C.byref(Kn), dinv.ptr, rhs.ptr, X.ptr, 1e-12, 100, work.ptr, C.byref(it), C.byref(rel)) == -9
    K = csr_desc(ctx, Kc)
    assert L.hip().trl_hip_pcg(ctx.h, L.F64, A.n, b, C.byref(K), dinv.ptr, rhs.ptr, X.ptr, 1e-14, 3, work.ptr, C.byref(it), C.byref(rel)) == -8
    assert it.value == 3 and rel.value > 1e-14
    assert L.hip().trl_hip_pcg(ctx.h, L.F32, A.n, b, C.byref(K), dinv.ptr, rhs.ptr, X.ptr, 1e-12, 100, work.ptr, None, None) == -1
    assert L.hip().trl_hip_pcg(ctx.h, L.F64, A.n, b, C.byref(K), dinv.ptr, rhs.ptr, rhs.ptr, 1e-12, 100, work.ptr, None, None) == -1


@pytest.mark.parametrize("reorth", ["mgs", "cgs2"])
@pytest.mark.parametrize("shift_frac", [0.0, 0.6])
def test_fe_pair_closed_form(ctx, reorth, shift_frac):
    # bilinear FE Laplacian pair, 48 x 40 nodes (n = 1920), b = 4, the 8
    # eigenvalues nearest the shift (|mu| largest, the reference's order),
    # ncv = 64, tol 1e-10: eigenvalues against the closed form, true
    # residuals |A x - lambda B x| / |A x| and B-orthonormal Ritz vectors
    nx, ny, b, nev = 48, 40, 4, 8
    A, B = P.fe_laplacian_2d(nx, ny)
    exact = P.fe_laplacian_2d_eigenvalues(nx, ny)
    sigma = shift_frac * exact[0]
    print(f"fe pair {reorth} sigma={sigma}", flush=True)
    V0 = ctx.to_device(P.initial_block(A.n, b))
    out = solve_shift_invert(ctx, A, B, sigma, b, EigensolverParams(nev, 64, tolerance=1e-10), V0, reorth=reorth)
    print("evals", out.evals, "\nresid", out.resid, "\n", out.result, out.inner, flush=True)
    assert out.result["converged"] == 1
    got = np.sort(out.evals)
    want = exact[:nev]
    assert np.all(np.abs(got - want) <= 1e-10 * np.abs(want)), np.abs(got - want) / want
    assert np.allclose(out.mu, 1.0 / (out.evals - sigma), rtol=1e-14)
    assert out.resid.max() <= 1e-8
    X = out.X.download().transpose(1, 0, 2).reshape(A.n, -1)[:, :nev]
    assert np.abs(X.T @ (B.to_scipy() @ X) - np.eye(nev)).max() < 1e-9
    assert out.inner["solves"] > 0 and 0 < out.inner["max_iterations"] < 10000


def test_sparse_mass_pair_against_lapack(ctx):
    # random stiffness / sparse SPD mass pair (n = 864, 3 DOF per node), 12
    # smallest by sigma = 0, b = 6, ncv = 72: against LAPACK's dense
    # generalized solve, 1e-10 relative, residuals at tol
    K, B = P.stiffness_sparse_mass(8, 6, 6, dof=3)
    b, nev = 6, 12
    ref = sla.eigh(K.to_dense(), B.to_dense(), eigvals_only=True)[:nev]
    V0 = ctx.to_device(P.initial_block(K.n, b))
    out = solve_shift_invert(ctx, K, B, 0.0, b, EigensolverParams(nev, 72, tolerance=1e-10), V0)
    print("rel", np.abs(np.sort(out.evals) - ref) / ref, "resid", out.resid, out.inner, flush=True)
    assert out.result["converged"] == 1
    assert np.all(np.abs(np.sort(out.evals) - ref) <= 1e-10 * np.abs(ref))
    assert out.resid.max() <= 1e-8


def test_shift_invert_refuses_indefinite_shift(ctx):
    # a shift past the smallest eigenvalue makes A - sigma B indefinite:
    # refused at assembly when a diagonal entry is <= 0, else by the inner
    # solve (a direction with p^T K p <= 0, or no convergence)
    A, B = P.fe_laplacian_2d(20, 16)
    exact = P.fe_laplacian_2d_eigenvalues(20, 16)
    V0 = ctx.to_device(P.initial_block(A.n, 4))
    prm = EigensolverParams(4, 32, tolerance=1e-8)
    with pytest.raises(L.TrlHipError) as e:
        solve_shift_invert(ctx, A, B, 1e9, 4, prm, V0)
    assert e.value.code == -1
    with pytest.raises(L.TrlHipError) as e:
        solve_shift_invert(ctx, A, B, 0.5 * (exact[5] + exact[6]), 4, prm, V0)
    assert e.value.code in (-8, -9)
