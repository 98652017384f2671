// Matrix Market coordinate files -> canonical CSR on the host.
//
// The file handling follows the reference's loader
// (examples/sycl/csrevp.hh:26-124): comment lines start with '%', the
// banner's "symmetric" mirrors every off-diagonal entry (j, i) of an entry
// (i, j), indices are 1-based, and a file that cannot be opened is reported
// as "Error: cannot open file <path>".  Where the reference keeps a row's
// entries in file order (duplicates included), the CSR built here is
// canonical: columns ascending within a row, duplicate entries summed (the
// same operator).  The banner's field may also be integer or pattern (value
// 1); complex fields, array (dense) files and skew-symmetric or Hermitian
// symmetry are refused, as are indices out of range and files with fewer
// entries than their size line declares.
#include <algorithm>
#include <cctype>
#include <cerrno>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iterator>
#include <numeric>
#include <sstream>
#include <string>
#include <vector>

#include "trl_hip_solver.h"

struct trl_hip_mm {
  size_t rows = 0, cols = 0;
  int symmetric = 0;
  std::vector<int32_t> rowptr, col;
  std::vector<double> val;
};

namespace {

thread_local std::string mm_error;

int fail(const std::string& msg)
{
  mm_error = msg;
  return TRL_HIP_EIO;
}

std::string lower(std::string s)
{
  std::transform(s.begin(), s.end(), s.begin(), [](unsigned char c) { return char(std::tolower(c)); });
  return s;
}

// Parses the file into m; returns 0 or TRL_HIP_EIO with mm_error set.
int parse(const std::string& path, trl_hip_mm& m)
{
  std::ifstream f(path, std::ios::binary);
  if (!f) return fail("Error: cannot open file " + path);
  const std::string text((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
  const char* p = text.c_str();
  const char* const end = p + text.size();

  // header: blank and '%' lines; the banner among them sets format, field
  // and symmetry
  bool pattern = false;
  std::string size_line;
  while (p < end) {
    const char* eol = static_cast<const char*>(std::memchr(p, '\n', size_t(end - p)));
    if (!eol) eol = end;
    std::string line(p, eol);
    p = eol < end ? eol + 1 : end;
    const size_t first = line.find_first_not_of(" \t\r");
    if (first == std::string::npos) continue;
    if (line[first] == '%') {
      const std::string l = lower(line);
      if (l.rfind("%%matrixmarket", first) == first) {
        std::istringstream bs(l.substr(first + 14));
        std::string object, format, field, symmetry;
        bs >> object >> format >> field >> symmetry;
        if (!object.empty() && object != "matrix") return fail("Matrix Market: unsupported object '" + object + "' in " + path);
        if (!format.empty() && format != "coordinate") return fail("Matrix Market: only coordinate files are supported (" + path + " is '" + format + "')");
        if (field == "complex" || field == "hermitian") return fail("Matrix Market: complex matrices are not supported (" + path + ")");
        if (!field.empty() && field != "real" && field != "double" && field != "integer" && field != "pattern")
          return fail("Matrix Market: unknown field '" + field + "' in " + path);
        pattern = field == "pattern";
        if (symmetry == "symmetric")
          m.symmetric = 1;
        else if (!symmetry.empty() && symmetry != "general")
          return fail("Matrix Market: symmetry '" + symmetry + "' is not supported (" + path + ")");
      }
      continue;
    }
    size_line = line;
    break;
  }
  long long rows = -1, cols = -1, entries = -1;
  {
    std::istringstream ss(size_line);
    if (!(ss >> rows >> cols >> entries) || rows < 0 || cols < 0 || entries < 0)
      return fail("Matrix Market: missing or malformed size line in " + path);
  }
  if (rows > INT32_MAX || cols > INT32_MAX) return fail("Matrix Market: dimensions beyond 32-bit indices in " + path);
  if (m.symmetric && rows != cols) return fail("Matrix Market: a symmetric matrix must be square (" + path + ")");
  const long long stored = m.symmetric ? 2 * entries : entries;
  if (stored > INT32_MAX) return fail("Matrix Market: more entries than 32-bit offsets hold in " + path);

  std::vector<int32_t> ri, ci;
  std::vector<double> vi;
  ri.reserve(size_t(stored));
  ci.reserve(size_t(stored));
  vi.reserve(size_t(stored));
  for (long long k = 0; k < entries; ++k) {
    char* q = nullptr;
    errno = 0;
    const long long i = std::strtoll(p, &q, 10);
    if (q == p) return fail("Matrix Market: " + path + " ends after " + std::to_string(k) + " of " + std::to_string(entries) + " entries");
    p = q;
    const long long j = std::strtoll(p, &q, 10);
    if (q == p) return fail("Matrix Market: malformed entry " + std::to_string(k + 1) + " in " + path);
    p = q;
    double v = 1.0;
    if (!pattern) {
      v = std::strtod(p, &q);
      if (q == p) return fail("Matrix Market: malformed value of entry " + std::to_string(k + 1) + " in " + path);
      p = q;
    }
    if (i < 1 || i > rows || j < 1 || j > cols)
      return fail("Matrix Market: entry " + std::to_string(k + 1) + " (" + std::to_string(i) + ", " + std::to_string(j) + ") outside the " +
                  std::to_string(rows) + " x " + std::to_string(cols) + " matrix in " + path);
    ri.push_back(int32_t(i - 1));
    ci.push_back(int32_t(j - 1));
    vi.push_back(v);
    if (m.symmetric && i != j) {
      ri.push_back(int32_t(j - 1));
      ci.push_back(int32_t(i - 1));
      vi.push_back(v);
    }
  }

  // counting sort by row, then columns ascending within each row, and
  // duplicates summed
  m.rows = size_t(rows);
  m.cols = size_t(cols);
  const size_t nz = ri.size();
  std::vector<int64_t> start(m.rows + 1, 0);
  for (size_t k = 0; k < nz; ++k) ++start[size_t(ri[k]) + 1];
  std::partial_sum(start.begin(), start.end(), start.begin());
  std::vector<int64_t> pos(start.begin(), start.end() - 1);
  std::vector<int32_t> order(nz);
  for (size_t k = 0; k < nz; ++k) order[size_t(pos[size_t(ri[k])]++)] = int32_t(k);
  m.rowptr.assign(m.rows + 1, 0);
  m.col.clear();
  m.val.clear();
  m.col.reserve(nz);
  m.val.reserve(nz);
  for (size_t r = 0; r < m.rows; ++r) {
    auto b = order.begin() + start[r], e = order.begin() + start[r + 1];
    // stable: equal columns keep file order, so duplicates add in file order
    std::stable_sort(b, e, [&](int32_t x, int32_t y) { return ci[size_t(x)] < ci[size_t(y)]; });
    for (auto it = b; it != e; ++it) {
      const int32_t c = ci[size_t(*it)];
      if (!m.col.empty() && m.col.size() > size_t(m.rowptr[r]) && m.col.back() == c)
        m.val.back() += vi[size_t(*it)];
      else {
        m.col.push_back(c);
        m.val.push_back(vi[size_t(*it)]);
      }
    }
    m.rowptr[r + 1] = int32_t(m.col.size());
  }
  return 0;
}

} // namespace

extern "C" {

int trl_hip_mm_read(const char* path, trl_hip_mm** out)
{
  if (!path || !out) return TRL_HIP_EINVAL;
  *out = nullptr;
  try {
    auto m = new trl_hip_mm;
    const int rc = parse(path, *m);
    if (rc) {
      delete m;
      return rc;
    }
    mm_error.clear();
    *out = m;
    return 0;
  }
  catch (const std::bad_alloc&) {
    return fail(std::string("Matrix Market: out of host memory reading ") + path);
  }
}

int trl_hip_mm_info(const trl_hip_mm* m, size_t* rows, size_t* cols, size_t* nnz, int* symmetric)
{
  if (!m) return TRL_HIP_EINVAL;
  if (rows) *rows = m->rows;
  if (cols) *cols = m->cols;
  if (nnz) *nnz = m->col.size();
  if (symmetric) *symmetric = m->symmetric;
  return 0;
}

int trl_hip_mm_copy(const trl_hip_mm* m, int32_t* rowptr, int32_t* col, double* val)
{
  if (!m || !rowptr || (!m->col.empty() && (!col || !val))) return TRL_HIP_EINVAL;
  std::copy(m->rowptr.begin(), m->rowptr.end(), rowptr);
  std::copy(m->col.begin(), m->col.end(), col);
  std::copy(m->val.begin(), m->val.end(), val);
  return 0;
}

void trl_hip_mm_free(trl_hip_mm* m) { delete m; }

const char* trl_hip_mm_last_error(void) { return mm_error.c_str(); }

} // extern "C"
