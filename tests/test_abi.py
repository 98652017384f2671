"""The C-ABI libraries load and export every symbol include/*.h declares
(no compute calls: this runs without a GPU)."""

import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

from trl_amd import _lib as L

ROOT = Path(__file__).resolve().parent.parent


def declared(header: str) -> list[str]:
    text = (ROOT / "include" / header).read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:int|void\s*\*|void|const char\s*\*|size_t)\s*(trl_hip_\w+)\s*\(", text, flags=re.M)))


@pytest.mark.parametrize("header,loader", [("trl_hip.h", L.hip), ("trl_hip_solver.h", L.solver)])
def test_exports_every_declared_symbol(header, loader):
    names = declared(header)
    assert len(names) >= 3
    lib = loader()
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_ctypes_table_matches_header():
    assert sorted(n for n, _, _ in L.HIP_API) == declared("trl_hip.h")
    assert sorted(n for n, _, _ in L.SOLVER_API) == declared("trl_hip_solver.h")


def test_strerror_and_host_only_entry_points():
    assert L.strerror(-2) == "Cholesky factorization failed in orthonormalize"
    assert L.strerror(-1) == "invalid argument"
    # argument validation happens before any device call
    assert L.hip().trl_hip_gram(None, 1, 10, 4, 1, None, None, None, None) == -1
    assert L.hip().trl_hip_ctx_create(0, None) == -1


@pytest.mark.parametrize("n", [1, 2, 5, 40, 256])
def test_dense_syev_matches_lapack(n):
    # the host FP64 eigensolver of the Rayleigh-Ritz step (GVL Householder + implicit QR)
    from trl_amd.solver import dense_syev

    rng = np.random.default_rng(n)
    A = rng.standard_normal((n, n))
    A = (A + A.T) / 2
    w, Z = dense_syev(A)
    assert np.allclose(w, np.linalg.eigvalsh(A), atol=1e-12 * max(1, np.abs(w).max()))
    assert np.abs(A @ Z - Z * w).max() < 1e-12 * n
    assert np.abs(Z.T @ Z - np.eye(n)).max() < 1e-12 * n


def test_dense_syev_degenerate():
    from trl_amd.solver import dense_syev

    rng = np.random.default_rng(3)
    Q, _ = np.linalg.qr(rng.standard_normal((48, 48)))
    d = np.repeat(np.arange(12.0), 4)
    w, Z = dense_syev((Q * d) @ Q.T)
    assert np.allclose(w, np.sort(d), atol=1e-12)
    assert np.abs(Z.T @ Z - np.eye(48)).max() < 1e-12


@pytest.mark.parametrize("case", ["random", "degenerate", "laplace2d", "lanczos_t"])
def test_dense_syev_select(case):
    # eigenvalues bitwise those of the full solver; selected eigenvectors with
    # eigen-residuals and orthogonality at rounding level, also inside
    # clusters of equal eigenvalues; an eigenvector of a simple eigenvalue
    # equals the full solver's up to sign
    from trl_amd.solver import dense_syev, dense_syev_select

    rng = np.random.default_rng(7)
    if case == "random":
        A = rng.standard_normal((200, 200))
        A = (A + A.T) / 2
    elif case == "degenerate":
        Q, _ = np.linalg.qr(rng.standard_normal((96, 96)))
        A = (Q * np.repeat(np.arange(24.0), 4)) @ Q.T
    elif case == "laplace2d":  # multiplicity 2 (lambda_ij = lambda_ji)
        s = 12
        T1 = 2 * np.eye(s) - np.eye(s, k=1) - np.eye(s, k=-1)
        A = np.kron(T1, np.eye(s)) + np.kron(np.eye(s), T1)
    else:  # block tridiagonal b = 4 plus an arrow, the shape of T after a restart
        b, m, kr = 4, 16, 3
        A = np.zeros((b * m, b * m))
        A[: kr * b, : kr * b] = np.diag(np.sort(rng.uniform(0, 1, kr * b)))
        for i in range(kr, m):
            Bi = rng.standard_normal((b, b))
            A[i * b:(i + 1) * b, i * b:(i + 1) * b] = Bi + Bi.T
            if i + 1 < m:
                C = np.triu(rng.standard_normal((b, b)))
                A[(i + 1) * b:(i + 2) * b, i * b:(i + 1) * b] = C
                A[i * b:(i + 1) * b, (i + 1) * b:(i + 2) * b] = C.T
        arrow = rng.standard_normal((b, kr * b))
        A[kr * b:(kr + 1) * b, : kr * b] = arrow
        A[: kr * b, kr * b:(kr + 1) * b] = arrow.T
    n = A.shape[0]
    wf, Zf = dense_syev(A)
    sel = np.r_[np.arange(0, 20), np.arange(n - 6, n)]
    w, Z = dense_syev_select(A, sel)
    scale = np.abs(A).max()
    # the eigenvalue-only QR (root-free) and the QR that accumulates vectors
    # agree to rounding, and both with LAPACK
    assert np.abs(w - wf).max() < 1e-13 * scale * np.sqrt(n)
    assert np.abs(w - np.linalg.eigvalsh(A)).max() < 1e-13 * scale * np.sqrt(n)
    assert np.abs(A @ Z - Z * w[sel]).max() < 1e-12 * scale * n
    assert np.abs(Z.T @ Z - np.eye(len(sel))).max() < 1e-12 * n
    gaps = np.minimum(np.abs(np.diff(wf, prepend=-np.inf)), np.abs(np.diff(wf, append=np.inf)))
    for c, j in enumerate(sel):
        if gaps[j] > 1e-6 * scale:
            assert abs(abs(Z[:, c] @ Zf[:, j]) - 1) < 1e-10


def test_dense_syev_select_arguments():
    from trl_amd import _lib as L

    A = np.eye(4)
    w = np.empty(4)
    Z = np.empty((4, 2))
    bad = np.array([2, 1], np.int32)  # not increasing
    assert L.solver().trl_hip_dense_syev_select(4, A.ctypes.data, w.ctypes.data, 2, bad.ctypes.data, Z.ctypes.data) == -1
    assert L.solver().trl_hip_dense_syev_select(4, A.ctypes.data, w.ctypes.data, 5, bad.ctypes.data, Z.ctypes.data) == -1


def test_dense_results_do_not_depend_on_the_thread_count(tmp_path):
    # the threaded Householder reduction (n >= 128) sums each entry on one
    # thread in a fixed order: one and several OpenMP threads agree bitwise
    import os
    import subprocess
    import sys

    rng = np.random.default_rng(3)
    A = rng.standard_normal((160, 160))
    np.save(tmp_path / "A.npy", (A + A.T) / 2)
    code = (
        "import sys, numpy as np\n"
        "from trl_amd.solver import dense_syev, dense_syev_select\n"
        "A = np.load(sys.argv[1])\n"
        "w, Z = dense_syev(A)\n"
        "ws, Zs = dense_syev_select(A, np.arange(12))\n"
        "T = np.load(sys.argv[3])\n"
        "wb, Zb, p = dense_syev_arrow_band(T, 8, np.arange(40))\n"
        "np.savez(sys.argv[2], w=w, Z=Z, ws=ws, Zs=Zs, wb=wb, Zb=Zb)\n"
    )
    np.save(tmp_path / "T.npy", restart_projection(8, 32, 5, np.random.default_rng(5)))
    code = code.replace("from trl_amd.solver import dense_syev, dense_syev_select", "from trl_amd.solver import dense_syev, dense_syev_select, dense_syev_arrow_band")
    out = {}
    for t in (1, 5):
        env = dict(os.environ, OMP_NUM_THREADS=str(t))
        f = tmp_path / f"r{t}.npz"
        subprocess.run([sys.executable, "-c", code, str(tmp_path / "A.npy"), str(f), str(tmp_path / "T.npy")], env=env, cwd=ROOT, check=True,
                       timeout=120)
        out[t] = np.load(f)
    for k in ("w", "Z", "ws", "Zs", "wb", "Zb"):
        assert np.array_equal(out[1][k], out[5][k]), k


def restart_projection(b, m, kr, rng, ritz=None):
    """The projected matrix of the block Lanczos after a thick restart
    (include/trl/eigensolvers/lanczos.hh:105-178): diag(theta) on the first
    kr blocks, the arrow between them and block kr, and block tridiagonal
    from block kr on with upper-triangular sub-diagonal blocks (the R
    factors of the block orthonormalisation).  kr = 0: before the first
    restart."""
    n = b * m
    A = np.zeros((n, n))
    p = kr * b
    if p:
        A[:p, :p] = np.diag(rng.uniform(-1, 3, p) if ritz is None else ritz)
        arrow = rng.standard_normal((b, p)) * 1e-2
        A[p:p + b, :p] = arrow
        A[:p, p:p + b] = arrow.T
    for i in range(kr, m):
        D = rng.standard_normal((b, b))
        A[i * b:(i + 1) * b, i * b:(i + 1) * b] = D + D.T
        if i + 1 < m:
            R = np.triu(rng.standard_normal((b, b)))
            A[(i + 1) * b:(i + 2) * b, i * b:(i + 1) * b] = R
            A[i * b:(i + 1) * b, (i + 1) * b:(i + 2) * b] = R.T
    return A


@pytest.mark.parametrize("b,m,kr", [(1, 40, 3), (2, 30, 4), (4, 16, 3), (8, 32, 5), (8, 32, 0), (8, 15, 5), (3, 7, 1), (16, 16, 2)])
def test_dense_syev_arrow_band_matches_lapack(b, m, kr):
    # the structured Rayleigh-Ritz eigensolver (QL reduction of the Ritz
    # part, Householder bulge chasing on the band): eigenvalues against LAPACK
    # (numpy.linalg.eigh) and the dense path, eigenvectors by their
    # eigen-residuals, orthogonality and, for simple eigenvalues, against
    # LAPACK's up to sign
    from trl_amd.solver import dense_syev_arrow_band, dense_syev_select

    rng = np.random.default_rng(100 * b + m + kr)
    A = restart_projection(b, m, kr, rng)
    n = A.shape[0]
    wl, Zl = np.linalg.eigh(A)
    sel = np.unique(np.r_[np.arange(0, min(n, kr * b + b)), np.arange(max(0, n - 3), n)])
    w, Z, p = dense_syev_arrow_band(A, b, sel)
    # p: the Ritz part found (with kr = 0 the leading diagonal block of a
    # random band matrix is not diagonal)
    assert p == kr * b or (kr == 0 and p == 0) or (b == 1 and p >= kr)
    scale = np.abs(A).max()
    assert np.abs(w - wl).max() < 1e-13 * scale * np.sqrt(n)
    ws, Zs = dense_syev_select(A, sel)
    assert np.abs(w - ws).max() < 1e-13 * scale * np.sqrt(n)
    assert np.abs(A @ Z - Z * w[sel]).max() < 1e-12 * scale * np.sqrt(n)
    assert np.abs(Z.T @ Z - np.eye(len(sel))).max() < 1e-12 * np.sqrt(n)
    gaps = np.minimum(np.abs(np.diff(wl, prepend=-np.inf)), np.abs(np.diff(wl, append=np.inf)))
    for c, j in enumerate(sel):
        if gaps[j] > 1e-6 * scale:
            assert abs(abs(Z[:, c] @ Zl[:, j]) - 1) < 1e-9


def test_dense_syev_arrow_band_clusters_and_zero_arrow():
    # converged Ritz values repeat (a cluster of equal theta) and part of the
    # arrow vanishes (converged pairs decouple): the QL reflectors meet zero
    # columns, inverse iteration meets a cluster; then a whole zero arrow
    from trl_amd.solver import dense_syev_arrow_band

    b, m, kr = 4, 20, 4
    rng = np.random.default_rng(11)
    ritz = np.repeat([0.5, 1.5], 8)
    for zero_all in (False, True):
        A = restart_projection(b, m, kr, rng, ritz=ritz)
        p = kr * b
        A[p:p + b, : p // 2] = 0.0
        A[: p // 2, p:p + b] = 0.0
        if zero_all:
            A[p:p + b, :p] = 0.0
            A[:p, p:p + b] = 0.0
        n = A.shape[0]
        sel = np.arange(0, 24)
        w, Z, pf = dense_syev_arrow_band(A, b, sel)
        assert pf == p
        scale = np.abs(A).max()
        assert np.abs(w - np.linalg.eigvalsh(A)).max() < 1e-13 * scale * np.sqrt(n)
        assert np.abs(A @ Z - Z * w[sel]).max() < 1e-12 * scale * np.sqrt(n)
        assert np.abs(Z.T @ Z - np.eye(len(sel))).max() < 1e-12 * np.sqrt(n)


def test_dense_syev_arrow_band_rejects_other_structure():
    # the structure check: an entry beyond the band, a Ritz part coupled
    # beyond block kr, a non-diagonal Ritz part -> invalid argument, p = -1
    from trl_amd import _lib as L
    from trl_amd.solver import dense_syev_arrow_band

    rng = np.random.default_rng(2)
    b, m, kr = 4, 10, 2
    base = restart_projection(b, m, kr, rng)
    p = kr * b
    for (i, j) in [(p + 3, p + 3 + b + 1), (1, p + b + 2), (p + 10, p + 20)]:
        A = base.copy()
        A[i, j] = A[j, i] = 0.3
        with pytest.raises(L.TrlHipError):
            dense_syev_arrow_band(A, b, [0, 1])
        w = np.empty(A.shape[0])
        Z = np.empty((A.shape[0], 1))
        sel = np.array([0], np.int32)
        pout = C.c_int(7)
        rc = L.solver().trl_hip_dense_syev_arrow_band(A.shape[0], b, A.ctypes.data, w.ctypes.data, 1, sel.ctypes.data, Z.ctypes.data, C.byref(pout))
        assert rc == -1 and pout.value == -1
    # a dense symmetric matrix has no such structure either
    A = rng.standard_normal((40, 40))
    with pytest.raises(L.TrlHipError):
        dense_syev_arrow_band(A + A.T, 4, [0])
