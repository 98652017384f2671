#pragma once
// Host FP64 Rayleigh-Ritz step of the thick-restart block Lanczos driver,
// following impl/openmp/evp_base.hh:79-154: eigen-decompose the projected
// block matrix T, order the pairs, keep the eigenvectors in block layout and
// count the first nev pairs whose residual estimate
// ||beta y_last|| / max(|theta|, eps) is below tol.  Shared by the
// single-problem eigenproblem (trl::hip::EVPBase) and the batched driver.

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <limits>
#include <numeric>
#include <stdexcept>
#include <vector>

#include "trl_amd/dense.hh"

namespace trl_amd {

// 0 = |lambda| descending (evp_base.hh:103), 1 = ascending, 2 = descending
struct RitzResult {
  std::vector<double> theta;     // nt Ritz values in the chosen order
  std::vector<double> Y;         // nb x bc blocks of b x b: block (i, j) at ((i*bc + j)*b*b), row-major
  std::vector<double> residual;  // estimates of the first min(nev, nt) pairs
  std::size_t nconv = 0;
  int structured = -1;           // the Ritz part p of the structured eigensolver; -1: the dense one
};

// T: nb x bc blocks of b x b in BlockMatrix layout (block (i, j) at
// ((i*bc + j)*b*b), row-major inside); only the leading nb x nb blocks enter
// the eigen-decomposition (bc == nb in the driver).  beta: b x b, the last
// off-diagonal block.  dense: caller-provided workspace (resized here).
// ncols: the Ritz vectors needed, in the chosen order (the driver uses the
// first k_r*b = (nev/b + 1)*b: the restart rotation and the arrow,
// lanczos.hh:105-143); Y's other columns are left zero.  Fewer columns than
// nb*b take the eigensolver that computes only those eigenvectors: the
// structured one (dense::ArrowBandEigensolver) when T has the thick-restart
// form (unless allow_structured is false), the dense Householder one
// otherwise; out.structured tells which.
inline void rayleigh_ritz(unsigned b, std::size_t nb, std::size_t bc, const double* T, const double* beta, std::size_t nev, int order, double tol,
                          RitzResult& out, std::vector<double>& dense, std::size_t ncols = std::size_t(-1), bool allow_structured = true)
{
  const std::size_t nt = nb * b;
  ncols = std::min(ncols, nt);
  dense.resize(nt * nt);
  for (std::size_t i = 0; i < nb; ++i)
    for (std::size_t j = 0; j < nb; ++j)
      for (unsigned bi = 0; bi < b; ++bi)
        for (unsigned bj = 0; bj < b; ++bj) dense[(i * b + bi) * nt + (j * b + bj)] = T[((i * bc + j) * b + bi) * b + bj];
  std::vector<double> w(nt);
  std::vector<unsigned> idx(nt);
  std::vector<double> Z;
  // zcol[j]: the column of Z holding the eigenvector of ascending eigenvalue j
  std::vector<std::size_t> zcol(nt, 0);
  std::size_t zld = nt;
  out.structured = -1;
  if (ncols == nt) {
    Z.resize(nt * nt);
    if (dense::syev(int(nt), dense.data(), w.data(), Z.data()) != 0) throw std::runtime_error("Eigendecomposition failed");
    std::iota(zcol.begin(), zcol.end(), std::size_t(0));
  }
  else {
    // T of the thick restart is a diagonal Ritz part, its arrow to block
    // k_r, and a band of half-bandwidth b (block tridiagonal with the
    // upper-triangular R factors below the diagonal): the structured solver
    // reduces it in O(nt^2 b); any other T takes the dense one
    const int p = allow_structured ? dense::arrow_band_structure(int(nt), int(b), dense.data()) : -1;
    dense::ArrowBandEigensolver ab;
    const bool structured = p >= 0 && ab.reduce(int(nt), int(b), p, dense.data(), w.data()) == 0;
    // the order depends only on the eigenvalues: find them first (the
    // selection below then asks for the eigenvectors of the leading ncols)
    std::vector<int> sel;
    if (order == 1)
      for (std::size_t j = 0; j < ncols; ++j) sel.push_back(int(j));
    else if (order == 2)
      for (std::size_t j = nt - ncols; j < nt; ++j) sel.push_back(int(j));
    else {
      if (!structured && dense::syev_select(int(nt), dense.data(), w.data(), 0, nullptr, nullptr) != 0)
        throw std::runtime_error("Eigendecomposition failed");
      std::iota(idx.begin(), idx.end(), 0u);
      std::stable_sort(idx.begin(), idx.end(), [&](unsigned x, unsigned y) { return std::abs(w[x]) > std::abs(w[y]); });
      for (std::size_t j = 0; j < ncols; ++j) sel.push_back(int(idx[j]));
      std::sort(sel.begin(), sel.end());
    }
    Z.resize(nt * ncols);
    const int rc = structured ? ab.vectors(int(ncols), sel.data(), Z.data())
                              : dense::syev_select(int(nt), dense.data(), w.data(), int(ncols), sel.data(), Z.data());
    if (rc == 0) {
      for (std::size_t c = 0; c < ncols; ++c) zcol[std::size_t(sel[c])] = c;
      zld = ncols;
      if (structured) out.structured = p;
    }
    else {
      // inverse iteration gave up (a non-finite iterate): the full solver
      Z.resize(nt * nt);
      if (dense::syev(int(nt), dense.data(), w.data(), Z.data()) != 0) throw std::runtime_error("Eigendecomposition failed");
      std::iota(zcol.begin(), zcol.end(), std::size_t(0));
    }
  }

  std::iota(idx.begin(), idx.end(), 0u);
  if (order == 0)
    std::stable_sort(idx.begin(), idx.end(), [&](unsigned x, unsigned y) { return std::abs(w[x]) > std::abs(w[y]); });
  else if (order == 2)
    std::reverse(idx.begin(), idx.end());

  out.theta.resize(nt);
  for (std::size_t i = 0; i < nt; ++i) out.theta[i] = w[idx[i]];
  out.Y.assign(nb * bc * b * b, 0.0);
  for (std::size_t i = 0; i < nb; ++i)
    for (std::size_t j = 0; j < nb; ++j)
      for (unsigned bi = 0; bi < b; ++bi)
        for (unsigned bj = 0; bj < b; ++bj) {
          const std::size_t col = j * b + bj;
          if (col < ncols) out.Y[((i * bc + j) * b + bi) * b + bj] = Z[(i * b + bi) * zld + zcol[idx[col]]];
        }

  out.nconv = 0;
  const std::size_t last = nb - 1;
  const double eps = std::numeric_limits<double>::epsilon();
  const std::size_t n_check = std::min(std::min(nev, nt), ncols);
  out.residual.assign(n_check, 0.0);
  for (std::size_t col = 0; col < n_check; ++col) {
    const std::size_t blk = col / b, cb = col % b;
    const double* vl = out.Y.data() + (last * bc + blk) * b * b;
    double nsq = 0.0;
    for (unsigned i = 0; i < b; ++i) {
      double s = 0.0;
      for (unsigned k = 0; k < b; ++k) s += beta[i * b + k] * vl[k * b + cb];
      nsq += s * s;
    }
    const double rel = std::sqrt(nsq) / std::max(std::abs(out.theta[col]), eps);
    out.residual[col] = rel;
    if (rel < tol) ++out.nconv;
  }
}

} // namespace trl_amd
