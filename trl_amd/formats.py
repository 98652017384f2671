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


def read_matrix_market(path) -> CSR:
    """A Matrix Market coordinate file as canonical CSR (columns ascending
    within a row, duplicate entries summed), read by the library's host
    loader (trl_hip_mm_read; the file handling of the reference's
    load_matrix_market, examples/sycl/csrevp.hh:26-124: '%' comments, a
    "symmetric" banner mirrors off-diagonal entries, 1-based indices).
    Raises OSError when the file cannot be opened or parsed, with the
    loader's message ("Error: cannot open file <path>" for a missing one),
    and ValueError for a matrix that is not square."""
    import ctypes as C

    from . import _lib as L

    lib = L.solver()
    h = C.c_void_p()
    rc = lib.trl_hip_mm_read(str(path).encode(), C.byref(h))
    if rc != 0:
        msg = lib.trl_hip_mm_last_error().decode()
        raise OSError(msg or L.strerror(rc))
    try:
        rows, cols, nnz, sym = C.c_size_t(), C.c_size_t(), C.c_size_t(), C.c_int()
        L.check(lib.trl_hip_mm_info(h, C.byref(rows), C.byref(cols), C.byref(nnz), C.byref(sym)), "trl_hip_mm_info")
        if rows.value != cols.value:
            raise ValueError(f"{path}: a {rows.value} x {cols.value} matrix is not square")
        rowptr = np.empty(rows.value + 1, np.int32)
        col = np.empty(nnz.value, np.int32)
        val = np.empty(nnz.value, np.float64)
        L.check(lib.trl_hip_mm_copy(h, rowptr.ctypes.data, col.ctypes.data, val.ctypes.data), "trl_hip_mm_copy")
    finally:
        lib.trl_hip_mm_free(h)
    return CSR(int(rows.value), rowptr, col, val)

