"""Sparse storage formats for the operator kernels.

CSR is the primary format (int32 rowptr/col, sorted columns), as in the
reference's CSR examples (examples/sycl/csrevp.hh:16-23).  ELL is the
slot-major padded form used by trl_hip_spmm_ell: slot s of row r at index
s*n + r; padding slots point at the row itself with value 0, so they gather
a row that is already being read.
"""

from __future__ import annotations

import numpy as np

from .problems import CSR


def csr_to_ell(A: CSR) -> tuple[int, np.ndarray, np.ndarray]:
    counts = np.diff(A.rowptr).astype(np.int64)
    width = int(counts.max()) if A.n else 0
    col = np.repeat(np.arange(A.n, dtype=np.int32)[None, :], width, axis=0)
    val = np.zeros((width, A.n))
    rows = np.repeat(np.arange(A.n), counts)
    slot = np.arange(A.nnz) - np.repeat(A.rowptr[:-1].astype(np.int64), counts)
    col[slot, rows] = A.col
    val[slot, rows] = A.val
    return width, np.ascontiguousarray(col.ravel()), np.ascontiguousarray(val.ravel())


def ell_padding(A: CSR) -> float:
    """Fraction of ELL slots that are padding."""
    counts = np.diff(A.rowptr)
    width = counts.max() if A.n else 0
    return 1.0 - A.nnz / max(width * A.n, 1)
