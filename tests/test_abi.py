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
    return sorted(set(re.findall(r"^\s*(?:int|void\s*\*|const char\s*\*)\s*(trl_hip_\w+)\s*\(", text, flags=re.M)))


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
