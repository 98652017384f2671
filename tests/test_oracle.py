"""Pinning the CPU oracle (oracle/trl_oracle.cc) before it is trusted.

Each oracle function is checked against something that fixes it
independently: numpy/LAPACK and scipy library routines, closed-form
spectra, invariants of the operation, and the reference's own known-answer
tests (tests/openmp/test_lanczos_convergence.cc:32,54;
tests/test_lanczos_extend.hh:77-111).
"""

import json
from pathlib import Path

import numpy as np
import pytest

import oracle
from trl_amd import problems as P

GOLDEN = Path(__file__).resolve().parent / "golden"


def _sym(rng, n):
    A = rng.standard_normal((n, n))
    return (A + A.T) / 2


# ---------------------------------------------------------------- dense LA


@pytest.mark.parametrize("n", [1, 2, 3, 7, 31, 64, 129])
def test_syev_matches_lapack(rng, n):
    A = _sym(rng, n)
    w, Z = oracle.syev(A)
    w_ref = np.linalg.eigvalsh(A)
    assert np.allclose(w, w_ref, rtol=0, atol=1e-12 * max(1.0, np.abs(w_ref).max()))
    assert np.abs(A @ Z - Z * w).max() < 1e-12 * max(1.0, np.abs(A).max()) * n
    assert np.abs(Z.T @ Z - np.eye(n)).max() < 1e-12 * n


def test_syev_degenerate_and_block_tridiagonal(rng):
    # exact multiplicities (as in the 2D Laplacian) and a block-tridiagonal arrow shape
    Q, _ = np.linalg.qr(rng.standard_normal((40, 40)))
    d = np.repeat(np.arange(10.0), 4)
    A = (Q * d) @ Q.T
    w, Z = oracle.syev(A)
    assert np.allclose(w, np.sort(d), atol=1e-12)
    assert np.abs(Z.T @ Z - np.eye(40)).max() < 1e-12
    T = np.zeros((24, 24))
    for i in range(6):
        blk = _sym(rng, 4)
        T[4 * i : 4 * i + 4, 4 * i : 4 * i + 4] = blk
        if i < 5:
            B = np.triu(rng.standard_normal((4, 4)))
            T[4 * i + 4 : 4 * i + 8, 4 * i : 4 * i + 4] = B
            T[4 * i : 4 * i + 4, 4 * i + 4 : 4 * i + 8] = B.T
    w, _ = oracle.syev(T)
    assert np.allclose(w, np.linalg.eigvalsh(T), atol=1e-12)


def test_syev_closed_form_tridiagonal():
    n = 100
    T = 2 * np.eye(n) - np.eye(n, k=1) - np.eye(n, k=-1)
    w, _ = oracle.syev(T)
    exact = 2 - 2 * np.cos(np.arange(1, n + 1) * np.pi / (n + 1))
    assert np.abs(w - np.sort(exact)).max() < 1e-13


def test_cholesky(rng):
    X = rng.standard_normal((50, 6))
    G = X.T @ X
    rc, R = oracle.cholesky(G)
    assert rc == 0
    assert np.allclose(R, np.linalg.cholesky(G).T, atol=1e-12)
    # breakdown: a zero column gives a zero pivot (evp_base.hh:55 throws)
    X[:, 3] = 0
    rc, _ = oracle.cholesky(X.T @ X)
    assert rc == 4
    rc, _ = oracle.cholesky(-np.eye(3))
    assert rc == 1


# ---------------------------------------------------------------- panel ops


@pytest.mark.parametrize("b", [1, 2, 4, 6, 8])
def test_panel_ops_match_numpy(rng, b):
    n = 37
    X = rng.standard_normal((n, b))
    W = rng.standard_normal((b, b))
    for mode, ref in [(0, X @ W), (2, X @ W.T)]:
        Y = np.zeros((n, b))
        oracle.panel_mult(X.copy(), W, Y, mode)
        assert np.allclose(Y, ref, atol=1e-13)
    Y0 = rng.standard_normal((n, b))
    Y = Y0.copy()
    oracle.panel_mult(X.copy(), W, Y, 1)
    assert np.allclose(Y, Y0 + X @ W, atol=1e-13)
    Xs = X.copy()
    oracle.panel_mult(Xs, W, Y0.copy(), 3)
    assert np.allclose(Xs, X - Y0 @ W, atol=1e-13)
    Xs = X.copy()
    oracle.panel_sub(Xs, Y0)
    assert np.array_equal(Xs, X - Y0)


@pytest.mark.parametrize("b,k", [(1, 5), (4, 3), (6, 2), (8, 4)])
def test_gram_update_rotate_match_numpy(rng, b, k):
    n = 53
    V = rng.standard_normal((k, n, b))
    W = rng.standard_normal((n, b))
    m = rng.uniform(0.5, 2.0, n)
    C = oracle.gram(V, W, m)
    ref = np.concatenate([V[j].T @ (m[:, None] * W) for j in range(k)])
    assert np.allclose(C, ref, atol=1e-12)
    W2 = W.copy()
    oracle.update(V, C, W2)
    assert np.allclose(W2, W - sum(V[j] @ C[j * b : (j + 1) * b] for j in range(k)), atol=1e-12)
    Y = rng.standard_normal((k, 3, b, b))
    Out = oracle.rotate(V, Y, 2)
    for j in range(2):
        assert np.allclose(Out[j], sum(V[i] @ Y[i, j] for i in range(k)), atol=1e-12)


def test_spmm_matches_scipy(rng):
    A = P.laplacian_2d(9, 7, rng.uniform(0, 1, 63))
    X = rng.standard_normal((63, 4))
    d = rng.uniform(0.5, 2, 63)
    assert np.allclose(oracle.spmm(A, X), A.to_scipy() @ X, atol=1e-13)
    assert np.allclose(oracle.spmm(A, X, d), d[:, None] * (A.to_scipy() @ X), atol=1e-13)


def test_cholqr_contract(rng):
    # X_old = X_new R, R upper triangular with positive diagonal, X_new^T M X_new = I
    X = rng.standard_normal((200, 6))
    m = rng.uniform(0.5, 2, 200)
    X0 = X.copy()
    rc, R = oracle.cholqr(X, m)
    assert rc == 0
    assert np.allclose(np.tril(R, -1), 0) and (np.diag(R) > 0).all()
    assert np.allclose(X @ R, X0, atol=1e-12)
    assert np.allclose(X.T @ (m[:, None] * X), np.eye(6), atol=1e-12)


# ---------------------------------------------------------------- driver


def _v0(n, b, seed=42):
    return P.initial_block(n, b, seed)


@pytest.mark.parametrize("b", [1, 2, 4, 8])
def test_reference_convergence_diagonal(b):
    # tests/openmp/test_lanczos_convergence.cc:169-187: N=128, nev=16, ncv=64,
    # largest |lambda| first, exact (N-i)^2, relative 1e-8
    N, nev = 128, 16
    r = oracle.lanczos_solve(P.diagonal(N), b, nev, 4 * nev, _v0(N, b), order=0)
    assert r["converged"] and r["error"] == 0
    exact = np.array([(N - i) ** 2 for i in range(nev)], dtype=float)
    assert np.all(np.abs(r["evals"] - exact) <= 1e-8 * np.abs(exact))


@pytest.mark.parametrize("b", [1, 2, 4, 8])
def test_reference_convergence_laplace1d(b):
    # tests/openmp/test_lanczos_convergence.cc:189-209
    N, nev = 128, 16
    r = oracle.lanczos_solve(P.laplacian_1d(N), b, nev, 4 * nev, _v0(N, b), order=0)
    assert r["converged"]
    exact = 2 - 2 * np.cos((N - np.arange(nev)) * np.pi / (N + 1))
    assert np.all(np.abs(r["evals"] - exact) <= 1e-8 * np.abs(exact))


@pytest.mark.parametrize("b", [1, 2, 4, 8])
def test_reference_extend_relation(b):
    # tests/test_lanczos_extend.hh: orthogonality and A V - V T = [0 ... V_m beta]
    N, nev, ncv = 128, 8, 32
    A = P.diagonal(N)
    rc, V, T, beta = oracle.lanczos_extend(A, b, nev, ncv, _v0(N, b))
    assert rc == 0
    m = ncv // b
    Vm = np.concatenate([V[i] for i in range(m + 1)], axis=1)
    assert np.abs(Vm.T @ Vm - np.eye(ncv + b)).max() < 1e-8
    Td = T.transpose(0, 2, 1, 3).reshape(ncv, ncv)
    Vk = Vm[:, :ncv]
    R = A.to_scipy() @ Vk - Vk @ Td
    for i in range(m - 1):
        assert np.linalg.norm(R[:, i * b : (i + 1) * b]) < 1e-8
    last = R[:, (m - 1) * b :] - V[m] @ beta
    assert np.linalg.norm(last) < 1e-8


def test_laplace2d_smallest_closed_form():
    # 2D Laplacian 32 x 32 (a small stand-in for the 128 x 128 case), smallest 8
    A = P.laplacian_2d(32, 32)
    r = oracle.lanczos_solve(A, 4, 8, 96, _v0(A.n, 4), order=1)
    assert r["converged"]
    exact = P.laplacian_2d_eigenvalues(32, 32)[:8]
    assert np.all(np.abs(r["evals"] - exact) <= 1e-8 * exact)
    # Ritz vectors: true residuals
    X = r["X"]
    Ax = A.to_scipy() @ X
    res = np.linalg.norm(Ax - X * r["evals"], axis=0) / np.linalg.norm(Ax, axis=0)
    assert res.max() < 1e-7


def test_generalized_against_scipy_golden():
    g = json.loads((GOLDEN / "stiffness_mass_eigsh.json").read_text())
    K, M = P.stiffness_mass()
    assert K.digest() == g["csr_sha256"]
    r = oracle.lanczos_solve(K, 6, 20, 120, _v0(K.n, 6), order=1, mass=M)
    assert r["converged"]
    ref = np.array(g["eigenvalues"])
    assert np.all(np.abs(r["evals"] - ref) <= 1e-10 * np.abs(ref))


def test_errors():
    A = P.diagonal(16)
    # lanczos.hh:45-58 parameter checks
    assert oracle.lanczos_solve(A, 4, 8, 16, _v0(16, 4))["error"] == 1  # ncv + b > n
    assert oracle.lanczos_solve(A, 2, 8, 8, _v0(16, 2))["error"] == 1  # nev >= ncv
    # Cholesky breakdown on a rank-deficient start block
    V0 = _v0(16, 2)
    V0[:, 1] = 0
    assert oracle.lanczos_solve(A, 2, 2, 8, V0)["error"] == 2


def test_mutation_is_caught():
    """A planted mistake in the operator (a dropped off-diagonal term) must
    move the spectrum away from the closed form."""
    A = P.laplacian_1d(64)
    bad = P.CSR(A.n, A.rowptr.copy(), A.col.copy(), A.val.copy())
    bad.val[1] = 0.0  # drop A[0,1]
    exact = 2 - 2 * np.cos((64 - np.arange(4)) * np.pi / 65)
    r_ok = oracle.lanczos_solve(A, 2, 4, 32, _v0(64, 2), order=0)
    r_bad = oracle.lanczos_solve(bad, 2, 4, 32, _v0(64, 2), order=0)
    assert np.abs(r_ok["evals"] - exact).max() < 1e-8
    assert np.abs(r_bad["evals"] - exact).max() > 1e-6
