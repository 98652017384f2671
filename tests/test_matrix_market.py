"""The Matrix Market loader (trl_hip_mm_read; the reference's
load_matrix_market, examples/sycl/csrevp.hh:26-124) against scipy.io.mmread
and hand-checked cases.  Host only: no GPU needed."""

import numpy as np
import pytest
import scipy.io
import scipy.sparse as sp

from trl_amd import problems as P
from trl_amd.formats import read_matrix_market


def canonical(M) -> sp.csr_matrix:
    M = sp.csr_matrix(M)
    M.sum_duplicates()
    M.sort_indices()
    return M


def same(A, M):
    M = canonical(M)
    assert A.n == M.shape[0] == M.shape[1]
    assert np.array_equal(A.rowptr, M.indptr) and np.array_equal(A.col, M.indices)
    assert np.array_equal(A.val, M.data)


SYMMETRIC = """%%MatrixMarket matrix coordinate real symmetric
% a comment line
%

4 4 6
1 1 4.0
2 1 -1.5
2 2 4.0
4 2 2.25e-1
3 3 5
4 4 6.5
"""


def test_symmetric_file_mirrors_the_lower_triangle(tmp_path):
    f = tmp_path / "s.mtx"
    f.write_text(SYMMETRIC)
    A = read_matrix_market(f)
    same(A, scipy.io.mmread(str(f)))
    D = A.to_dense()
    assert np.array_equal(D, D.T)
    assert D[0, 1] == D[1, 0] == -1.5 and D[1, 3] == D[3, 1] == 0.225 and D[2, 2] == 5.0
    assert A.nnz == 8  # 6 stored entries, 2 of them off the diagonal


def test_general_file_unsorted_and_duplicates(tmp_path):
    # entries out of order and a duplicate: columns come out ascending and the
    # duplicate summed, as scipy's reader with sum_duplicates gives
    f = tmp_path / "g.mtx"
    f.write_text("%%MatrixMarket matrix coordinate real general\n3 3 7\n3 1 1.0\n1 3 2.0\n1 1 3.0\n2 2 -1.0\n1 3 0.5\n3 3 7.0\n2 1 4.0\n")
    A = read_matrix_market(f)
    same(A, scipy.io.mmread(str(f)))
    assert A.to_dense()[0, 2] == 2.5


@pytest.mark.parametrize("field", ["integer", "pattern"])
def test_integer_and_pattern_fields(tmp_path, field):
    entries = ["1 1", "2 1", "2 2", "3 2", "3 3"]
    lines = [e + ("" if field == "pattern" else f" {k + 2}") for k, e in enumerate(entries)]
    f = tmp_path / f"{field}.mtx"
    f.write_text(f"%%MatrixMarket matrix coordinate {field} symmetric\n3 3 5\n" + "\n".join(lines) + "\n")
    A = read_matrix_market(f)
    same(A, scipy.io.mmread(str(f)).astype(np.float64))


def test_no_banner_reads_real_general(tmp_path):
    # the reference treats a file without a banner as non-symmetric real
    f = tmp_path / "nb.mtx"
    f.write_text("% no banner\n2 2 2\n1 2 3.0\n2 1 3.0\n")
    A = read_matrix_market(f)
    assert np.array_equal(A.to_dense(), [[0.0, 3.0], [3.0, 0.0]])


def test_round_trip_of_a_generated_laplacian(tmp_path, rng):
    # a 2D Laplacian with a potential written by scipy (symmetric, lower
    # triangle) reads back as the same canonical CSR
    A0 = P.laplacian_2d(13, 11, rng.uniform(0, 1, 143))
    f = tmp_path / "lap.mtx"
    scipy.io.mmwrite(str(f), A0.to_scipy(), symmetry="symmetric", precision=17)
    A = read_matrix_market(f)
    assert np.array_equal(A.rowptr, A0.rowptr) and np.array_equal(A.col, A0.col)
    assert np.array_equal(A.val, A0.val)


def test_missing_file_message(tmp_path):
    f = tmp_path / "absent.mtx"
    with pytest.raises(OSError, match=f"Error: cannot open file {f}"):
        read_matrix_market(f)


@pytest.mark.parametrize(
    "text,what",
    [
        ("%%MatrixMarket matrix coordinate complex general\n1 1 1\n1 1 1.0 0.0\n", "complex"),
        ("%%MatrixMarket matrix array real general\n1 1\n1.0\n", "coordinate"),
        ("%%MatrixMarket matrix coordinate real skew-symmetric\n2 2 1\n2 1 1.0\n", "skew-symmetric"),
        ("%%MatrixMarket matrix coordinate real general\n2 2 2\n1 1 1.0\n3 1 1.0\n", "outside"),
        ("%%MatrixMarket matrix coordinate real general\n2 2 3\n1 1 1.0\n2 2 1.0\n", "ends after 2 of 3"),
        ("%%MatrixMarket matrix coordinate real general\n", "size line"),
        ("%%MatrixMarket matrix coordinate real symmetric\n2 3 1\n1 1 1.0\n", "square"),
    ],
)
def test_refused_files(tmp_path, text, what):
    f = tmp_path / "bad.mtx"
    f.write_text(text)
    with pytest.raises(OSError, match=what):
        read_matrix_market(f)


def test_rectangular_general_file_is_not_an_operator(tmp_path):
    f = tmp_path / "rect.mtx"
    f.write_text("%%MatrixMarket matrix coordinate real general\n2 3 1\n1 3 1.0\n")
    with pytest.raises(ValueError, match="not square"):
        read_matrix_market(f)


def test_cpp_loader(tmp_path):
    # the C++ header form, trl::hip::load_matrix_market (host only)
    import subprocess
    from pathlib import Path

    exe = Path(__file__).resolve().parent.parent / "trl_amd" / "lib" / "test_matrix_market"
    if not exe.exists():
        pytest.skip("C++ tests not built (python -m trl_amd.build)")
    r = subprocess.run([str(exe), str(tmp_path)], capture_output=True, text=True, timeout=60)
    print(r.stdout, r.stderr)
    assert r.returncode == 0 and r.stdout.strip() == "OK"
