// Host code under AddressSanitizer and UndefinedBehaviorSanitizer: the dense
// eigensolver (all three paths, serial and threaded reductions), the Rayleigh-Ritz
// step in its three orders, the Matrix Market reader on good and refused
// files, and the oracle's Lanczos solve on a small 1D Laplacian.  Built with
// g++ only (no device code) by tests/test_host_sanitizers.py; prints OK and
// exits 0 when every check holds and no sanitizer reported.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <fstream>
#include <random>
#include <string>
#include <vector>

#include "trl_amd/dense.hh"
#include "trl_amd/rayleigh_ritz.hh"
#include "trl_hip_solver.h"

// the oracle's solve (oracle/trl_oracle.cc); info: converged, iterations,
// operator applications, error
extern "C" int oracle_lanczos_solve(size_t n, const int32_t* rowptr, const int32_t* col, const double* val, const double* dscale, const double* mdiag, int b,
                                    unsigned nev, unsigned ncv, unsigned max_restarts, double tol, int order, const double* V0, double* evals_out,
                                    double* X_out, int* info);

namespace {

int failures = 0;

void check(bool ok, const std::string& what)
{
  if (!ok) {
    ++failures;
    std::printf("FAIL %s\n", what.c_str());
  }
}

std::vector<double> random_symmetric(int n, unsigned seed)
{
  std::mt19937_64 g(seed);
  std::normal_distribution<double> nd;
  std::vector<double> A(std::size_t(n) * n);
  for (int i = 0; i < n; ++i)
    for (int j = 0; j <= i; ++j) A[std::size_t(i) * n + j] = A[std::size_t(j) * n + i] = nd(g);
  return A;
}

// max |A z - w z| over the columns of Z (n x k, column c the vector of w[c])
double eigen_residual(int n, const std::vector<double>& A, const double* w, const double* Z, int k, int ldz)
{
  double worst = 0.0;
  for (int c = 0; c < k; ++c)
    for (int i = 0; i < n; ++i) {
      double s = 0.0;
      for (int j = 0; j < n; ++j) s += A[std::size_t(i) * n + j] * Z[std::size_t(j) * ldz + c];
      worst = std::max(worst, std::fabs(s - w[c] * Z[std::size_t(i) * ldz + c]));
    }
  return worst;
}

void dense_cases()
{
  for (int n : {1, 2, 3, 17, 64, 150}) {
    const auto A = random_symmetric(n, 11u + unsigned(n));
    std::vector<double> w(n), Z(std::size_t(n) * n);
    check(trl_amd::dense::syev(n, A.data(), w.data(), Z.data()) == 0, "syev n=" + std::to_string(n));
    check(eigen_residual(n, A, w.data(), Z.data(), n, n) < 1e-10 * n, "syev residual n=" + std::to_string(n));
    // a few eigenvectors from both ends
    std::vector<int> sel;
    for (int j = 0; j < n && j < 3; ++j) sel.push_back(j);
    for (int j = std::max(3, n - 2); j < n; ++j) sel.push_back(j);
    const int k = int(sel.size());
    std::vector<double> ws(n), Zs(std::size_t(n) * k), wk(k);
    check(trl_amd::dense::syev_select(n, A.data(), ws.data(), k, sel.data(), Zs.data()) == 0, "syev_select n=" + std::to_string(n));
    for (int c = 0; c < k; ++c) wk[c] = ws[sel[c]];
    check(eigen_residual(n, A, wk.data(), Zs.data(), k, k) < 1e-10 * n, "syev_select residual n=" + std::to_string(n));
    // the eigenvalue-only QR (root-free) against the one that accumulates vectors
    double wmax = 0.0;
    for (int j = 0; j < n; ++j) wmax = std::max(wmax, std::fabs(w[j]));
    for (int j = 0; j < n; ++j) check(std::fabs(ws[j] - w[j]) <= 1e-13 * std::max(wmax, 1.0) * n, "syev_select eigenvalues n=" + std::to_string(n));
  }
  // a repeated eigenvalue (the 2D Laplacian's symmetric pairs), threaded size
  const int s = 12, n = s * s;
  std::vector<double> L(std::size_t(n) * n, 0.0);
  for (int i = 0; i < s; ++i)
    for (int j = 0; j < s; ++j) {
      const int r = i * s + j;
      L[std::size_t(r) * n + r] = 4.0;
      if (i > 0) L[std::size_t(r) * n + r - s] = -1.0;
      if (i + 1 < s) L[std::size_t(r) * n + r + s] = -1.0;
      if (j > 0) L[std::size_t(r) * n + r - 1] = -1.0;
      if (j + 1 < s) L[std::size_t(r) * n + r + 1] = -1.0;
    }
  std::vector<int> sel(10);
  for (int j = 0; j < 10; ++j) sel[j] = j;
  std::vector<double> w(n), Z(std::size_t(n) * 10), wk(10);
  check(trl_amd::dense::syev_select(n, L.data(), w.data(), 10, sel.data(), Z.data()) == 0, "syev_select laplace2d");
  for (int c = 0; c < 10; ++c) wk[c] = w[c];
  check(eigen_residual(n, L, wk.data(), Z.data(), 10, 10) < 1e-11, "syev_select laplace2d residual");
  // the structured path on the thick-restart form: 5 Ritz blocks (diagonal),
  // the arrow to block 5, then block tridiagonal (b = 8, 20 blocks: threaded
  // back-transformation), against the dense path
  {
    const int b = 8, kr = 5, nb = 20, nt = nb * b, p = kr * b;
    std::mt19937_64 g(9);
    std::normal_distribution<double> nd;
    std::vector<double> T(std::size_t(nt) * nt, 0.0);
    auto at = [&](int i, int j) -> double& { return T[std::size_t(i) * nt + j]; };
    for (int i = 0; i < p; ++i) at(i, i) = nd(g);
    for (int r = 0; r < b; ++r)
      for (int c = 0; c < p; ++c) at(p + r, c) = at(c, p + r) = 0.1 * nd(g);
    for (int blk = kr; blk < nb; ++blk) {
      const int o = blk * b;
      for (int r = 0; r < b; ++r)
        for (int c = 0; c <= r; ++c) at(o + r, o + c) = at(o + c, o + r) = nd(g);
      if (blk + 1 < nb)
        for (int r = 0; r < b; ++r)
          for (int c = r; c < b; ++c) at(o + b + r, o + c) = at(o + c, o + b + r) = nd(g);
    }
    check(trl_amd::dense::arrow_band_structure(nt, b, T.data()) == p, "arrow_band_structure finds the Ritz part");
    std::vector<int> sb(48);
    for (int j = 0; j < 48; ++j) sb[j] = j;
    std::vector<double> wb(nt), Zb(std::size_t(nt) * 48), wd(nt), Zd(std::size_t(nt) * 48), wk2(48);
    check(trl_amd::dense::syev_arrow_band_select(nt, b, p, T.data(), wb.data(), 48, sb.data(), Zb.data()) == 0, "syev_arrow_band_select");
    check(trl_amd::dense::syev_select(nt, T.data(), wd.data(), 48, sb.data(), Zd.data()) == 0, "syev_select on the restart form");
    for (int j = 0; j < nt; ++j) check(std::fabs(wb[j] - wd[j]) < 1e-12, "syev_arrow_band_select eigenvalues");
    for (int c = 0; c < 48; ++c) wk2[c] = wb[c];
    check(eigen_residual(nt, T, wk2.data(), Zb.data(), 48, 48) < 1e-11, "syev_arrow_band_select residual");
    T[std::size_t(3) * nt + p + b + 1] = T[std::size_t(p + b + 1) * nt + 3] = 1.0;
    check(trl_amd::dense::arrow_band_structure(nt, b, T.data()) == -1, "arrow_band_structure rejects a coupling beyond block kr");
  }
  // argument checks
  int bad[2] = {1, 1};
  check(trl_amd::dense::syev_select(4, L.data(), w.data(), 2, bad, Z.data()) == -2, "syev_select rejects a non-increasing selection");
}

void rayleigh_ritz_cases()
{
  // block tridiagonal T with b = 4, 6 blocks; all three orders, all and some columns
  const unsigned b = 4;
  const std::size_t nb = 6, nt = nb * b;
  std::mt19937_64 g(5);
  std::normal_distribution<double> nd;
  std::vector<double> T(nb * nb * b * b, 0.0), beta(b * b);
  auto blk = [&](std::size_t i, std::size_t j, unsigned r, unsigned c) -> double& { return T[((i * nb + j) * b + r) * b + c]; };
  for (std::size_t i = 0; i < nb; ++i) {
    for (unsigned r = 0; r < b; ++r)
      for (unsigned c = 0; c <= r; ++c) blk(i, i, r, c) = blk(i, i, c, r) = nd(g);
    if (i + 1 < nb)
      for (unsigned r = 0; r < b; ++r)
        for (unsigned c = r; c < b; ++c) blk(i + 1, i, r, c) = blk(i, i + 1, c, r) = nd(g);
  }
  for (double& x : beta) x = nd(g);
  std::vector<double> dense;
  for (int order : {0, 1, 2})
    for (std::size_t ncols : {nt, std::size_t(8)}) {
      trl_amd::RitzResult out;
      trl_amd::rayleigh_ritz(b, nb, nb, T.data(), beta.data(), 8, order, 1e-8, out, dense, ncols);
      check(out.theta.size() == nt, "rayleigh_ritz theta size");
      check(out.residual.size() == 8, "rayleigh_ritz residual size");
      for (std::size_t i = 1; i < nt; ++i) {
        const double a = out.theta[i - 1], c = out.theta[i];
        const bool ordered = order == 1 ? a <= c : order == 2 ? a >= c : std::fabs(a) >= std::fabs(c);
        check(ordered, "rayleigh_ritz order " + std::to_string(order));
      }
    }
}

void matrix_market_cases(const std::string& dir)
{
  const std::string good = dir + "/good.mtx";
  {
    std::ofstream f(good);
    f << "%%MatrixMarket matrix coordinate real symmetric\n% comment\n4 4 6\n1 1 4.0\n2 1 -1.5\n2 2 4.0\n4 2 0.25\n3 3 5\n4 4 6.5\n";
  }
  trl_hip_mm* m = nullptr;
  check(trl_hip_mm_read(good.c_str(), &m) == 0 && m, "mm_read good");
  size_t rows = 0, cols = 0, nnz = 0;
  int sym = 0;
  check(trl_hip_mm_info(m, &rows, &cols, &nnz, &sym) == 0 && rows == 4 && cols == 4 && nnz == 8 && sym == 1, "mm_info");
  std::vector<int32_t> rp(rows + 1), ci(nnz);
  std::vector<double> v(nnz);
  check(trl_hip_mm_copy(m, rp.data(), ci.data(), v.data()) == 0 && rp[4] == 8 && ci[0] == 0 && ci[1] == 1, "mm_copy");
  trl_hip_mm_free(m);
  const char* bad[] = {
      "%%MatrixMarket matrix coordinate real general\n2 2 3\n1 1 1.0\n2 2 1.0\n",
      "%%MatrixMarket matrix coordinate real general\n2 2 2\n1 1 1.0\n3 1 1.0\n",
      "%%MatrixMarket matrix coordinate complex general\n1 1 1\n1 1 1.0 0.0\n",
      "%%MatrixMarket matrix coordinate real general\n",
      "%%MatrixMarket matrix coordinate real general\n2 2 1\n1\n",
      "",
  };
  int k = 0;
  for (const char* text : bad) {
    const std::string p = dir + "/bad" + std::to_string(k++) + ".mtx";
    {
      std::ofstream f(p);
      f << text;
    }
    trl_hip_mm* b = nullptr;
    check(trl_hip_mm_read(p.c_str(), &b) == TRL_HIP_EIO && !b && *trl_hip_mm_last_error(), "mm_read refuses " + p);
  }
  trl_hip_mm* none = nullptr;
  check(trl_hip_mm_read((dir + "/absent.mtx").c_str(), &none) == TRL_HIP_EIO, "mm_read missing file");
}

void oracle_case()
{
  // 1D Laplacian, N = 128, b = 4, the 8 smallest: 2 - 2 cos(k pi / 129)
  const int n = 128, b = 4;
  const unsigned nev = 8, ncv = 32;
  std::vector<int32_t> rp(n + 1), ci;
  std::vector<double> v;
  for (int i = 0; i < n; ++i) {
    rp[i] = int32_t(ci.size());
    for (int j = i - 1; j <= i + 1; ++j)
      if (j >= 0 && j < n) {
        ci.push_back(j);
        v.push_back(j == i ? 2.0 : -1.0);
      }
  }
  rp[n] = int32_t(ci.size());
  std::mt19937_64 g(1);
  std::normal_distribution<double> nd;
  std::vector<double> V0(std::size_t(n) * b), ev(nev), X(std::size_t(n) * nev);
  for (double& x : V0) x = nd(g);
  int info[4] = {0, 0, 0, 0};
  const int rc = oracle_lanczos_solve(n, rp.data(), ci.data(), v.data(), nullptr, nullptr, b, nev, ncv, 200, 1e-10, 1, V0.data(), ev.data(), X.data(), info);
  check(rc == 0 && info[0] == 1, "oracle solve converged");
  for (unsigned k = 0; k < nev; ++k) {
    const double exact = 2.0 - 2.0 * std::cos(double(k + 1) * M_PI / double(n + 1));
    check(std::fabs(ev[k] - exact) < 1e-9, "oracle eigenvalue " + std::to_string(k));
  }
}

} // namespace

int main(int argc, char** argv)
{
  const std::string dir = argc > 1 ? argv[1] : ".";
  dense_cases();
  rayleigh_ritz_cases();
  matrix_market_cases(dir);
  oracle_case();
  if (failures) {
    std::printf("%d failures\n", failures);
    return 1;
  }
  std::printf("OK\n");
  return 0;
}
