#pragma once
// load_matrix_market of the reference's CSR example
// (examples/sycl/csrevp.hh:26-124) over the library's host loader
// trl_hip_mm_read: the CSR arrays on the host, canonical (columns ascending
// within a row, duplicate entries summed), for trl::hip::CSREVP or
// trl_hip_solve_csr after an upload.  Throws std::runtime_error with the
// loader's message ("Error: cannot open file <path>" for a missing file, as
// the reference).

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "trl_hip_solver.h"

namespace trl::hip {

struct HostCSR {
  std::size_t rows = 0, cols = 0;
  bool symmetric = false; // the file's banner said "symmetric"
  std::vector<int32_t> rowptr, col;
  std::vector<double> val;
};

inline HostCSR load_matrix_market(const std::string& path)
{
  trl_hip_mm* m = nullptr;
  if (trl_hip_mm_read(path.c_str(), &m) != 0) throw std::runtime_error(trl_hip_mm_last_error());
  HostCSR A;
  std::size_t nnz = 0;
  int sym = 0;
  trl_hip_mm_info(m, &A.rows, &A.cols, &nnz, &sym);
  A.symmetric = sym != 0;
  A.rowptr.resize(A.rows + 1);
  A.col.resize(nnz);
  A.val.resize(nnz);
  const int rc = trl_hip_mm_copy(m, A.rowptr.data(), A.col.data(), A.val.data());
  trl_hip_mm_free(m);
  if (rc != 0) throw std::runtime_error("Matrix Market: copy failed");
  return A;
}

} // namespace trl::hip
