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
    return sorted(set(re.findall(r"^\s*(?:int|void\s*\*|void|const char\s*\*)\s*(trl_hip_\w+)\s*\(", text, flags=re.M)))


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
    assert np.array_equal(w, wf)
    scale = np.abs(A).max()
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
        "np.savez(sys.argv[2], w=w, Z=Z, ws=ws, Zs=Zs)\n"
    )
    out = {}
    for t in (1, 5):
        env = dict(os.environ, OMP_NUM_THREADS=str(t))
        f = tmp_path / f"r{t}.npz"
        subprocess.run([sys.executable, "-c", code, str(tmp_path / "A.npy"), str(f)], env=env, cwd=ROOT, check=True, timeout=120)
        out[t] = np.load(f)
    for k in ("w", "Z", "ws", "Zs"):
        assert np.array_equal(out[1][k], out[5][k]), k
