// Host FP64 dense symmetric eigensolver for the Rayleigh-Ritz step
// (the role Eigen::SelfAdjointEigenSolver plays in the reference,
// impl/openmp/evp_base.hh:95).
//
// Method (Golub & Van Loan, Matrix Computations, 4th ed.):
//   1. Householder tridiagonalisation Q^T A Q = T (Alg. 8.3.1), with Q
//      accumulated backward from the stored reflectors;
//   2. implicit symmetric QR steps with the Wilkinson shift on T
//      (Alg. 8.3.2/8.3.3), chasing the bulge with Givens rotations that are
//      accumulated into Q.
// Q is kept transposed (row k = k-th basis vector) so both the reflector and
// the rotation updates run over contiguous memory.
#include <algorithm>
#include <cmath>
#include <limits>
#include <numeric>
#include <vector>

#include "trl_amd/dense.hh"

namespace trl_amd::dense {

int syev(int n, const double* A, double* w, double* Z)
{
  if (n <= 0) return 0;
  if (n == 1) {
    w[0] = A[0];
    Z[0] = 1.0;
    return 0;
  }
  const std::size_t N = std::size_t(n);
  std::vector<double> a(A, A + N * N); // working copy, row-major
  std::vector<double> vs(N * N, 0.0);  // reflector k stored in row k (entries k+1..n-1)
  std::vector<double> betas(N, 0.0);
  std::vector<double> p(N), wv(N);

  // ---- 1. tridiagonalisation ---------------------------------------------
  for (int k = 0; k + 2 < n; ++k) {
    double* v = vs.data() + std::size_t(k) * N;
    double xnorm2 = 0.0;
    for (int i = k + 1; i < n; ++i) {
      v[i] = a[std::size_t(i) * N + k];
      xnorm2 += v[i] * v[i];
    }
    const double xnorm = std::sqrt(xnorm2);
    if (xnorm == 0.0) continue;
    const double x0 = v[k + 1];
    const double alpha = (x0 >= 0.0) ? -xnorm : xnorm;
    v[k + 1] = x0 - alpha;
    double vn2 = 0.0;
    for (int i = k + 1; i < n; ++i) vn2 += v[i] * v[i];
    if (vn2 == 0.0) continue;
    const double beta = 2.0 / vn2;
    betas[k] = beta;
    // p = beta * A22 v  (A22 symmetric: use full rows)
    for (int i = k + 1; i < n; ++i) {
      const double* ai = a.data() + std::size_t(i) * N;
      double s = 0.0;
      for (int j = k + 1; j < n; ++j) s += ai[j] * v[j];
      p[i] = beta * s;
    }
    double pv = 0.0;
    for (int i = k + 1; i < n; ++i) pv += p[i] * v[i];
    const double K = 0.5 * beta * pv;
    for (int i = k + 1; i < n; ++i) wv[i] = p[i] - K * v[i];
    for (int i = k + 1; i < n; ++i) {
      double* ai = a.data() + std::size_t(i) * N;
      const double vi = v[i], wi = wv[i];
      for (int j = k + 1; j < n; ++j) ai[j] -= vi * wv[j] + wi * v[j];
    }
    a[std::size_t(k + 1) * N + k] = alpha;
    a[std::size_t(k) * N + k + 1] = alpha;
    for (int i = k + 2; i < n; ++i) a[std::size_t(i) * N + k] = a[std::size_t(k) * N + i] = 0.0;
  }
  std::vector<double> d(N), e(N, 0.0);
  for (int i = 0; i < n; ++i) d[i] = a[std::size_t(i) * N + i];
  for (int i = 0; i + 1 < n; ++i) e[i] = a[std::size_t(i + 1) * N + i];

  // Qt = (H_0 H_1 ... H_{n-3})^T, built as Q = H_0 (H_1 (... I)) then stored
  // transposed: row j of Qt is column j of Q.
  std::vector<double> Q(N * N, 0.0); // row-major Q
  for (int i = 0; i < n; ++i) Q[std::size_t(i) * N + i] = 1.0;
  for (int k = n - 3; k >= 0; --k) {
    const double beta = betas[k];
    if (beta == 0.0) continue;
    const double* v = vs.data() + std::size_t(k) * N;
    // Q[k+1:, k+1:] -= beta v (v^T Q[k+1:, k+1:])
    std::fill(p.begin(), p.end(), 0.0);
    for (int i = k + 1; i < n; ++i) {
      const double vi = v[i];
      const double* qi = Q.data() + std::size_t(i) * N;
      for (int j = k + 1; j < n; ++j) p[j] += vi * qi[j];
    }
    for (int i = k + 1; i < n; ++i) {
      const double bvi = beta * v[i];
      double* qi = Q.data() + std::size_t(i) * N;
      for (int j = k + 1; j < n; ++j) qi[j] -= bvi * p[j];
    }
  }
  std::vector<double> Qt(N * N);
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) Qt[std::size_t(j) * N + i] = Q[std::size_t(i) * N + j];

  // ---- 2. implicit symmetric QR with Wilkinson shift ----------------------
  const double eps = std::numeric_limits<double>::epsilon();
  const long max_sweeps = 60L * n;
  long sweeps = 0;
  while (true) {
    for (int i = 0; i + 1 < n; ++i)
      if (std::fabs(e[i]) <= eps * (std::fabs(d[i]) + std::fabs(d[i + 1]))) e[i] = 0.0;
    int hi = n - 1;
    while (hi > 0 && e[hi - 1] == 0.0) --hi;
    if (hi == 0) break;
    int lo = hi - 1;
    while (lo > 0 && e[lo - 1] != 0.0) --lo;
    if (++sweeps > max_sweeps) return -1;

    const double dd = 0.5 * (d[hi - 1] - d[hi]);
    const double eh = e[hi - 1];
    const double sgn = (dd >= 0.0) ? 1.0 : -1.0;
    const double mu = d[hi] - eh * eh / (dd + sgn * std::hypot(dd, eh));
    double x = d[lo] - mu;
    double z = e[lo];
    double bulge = 0.0;
    for (int k = lo; k < hi; ++k) {
      if (k > lo) {
        x = e[k - 1];
        z = bulge;
      }
      const double r = std::hypot(x, z);
      double c = 1.0, s = 0.0;
      if (r != 0.0) {
        c = x / r;
        s = -z / r;
      }
      if (k > lo) e[k - 1] = r;
      const double ak = d[k], bk = e[k], ck = d[k + 1];
      const double cc = c * c, ss = s * s, cs = c * s;
      d[k] = ak * cc - 2.0 * bk * cs + ck * ss;
      d[k + 1] = ak * ss + 2.0 * bk * cs + ck * cc;
      e[k] = cs * (ak - ck) + bk * (cc - ss);
      if (k + 1 < hi) {
        bulge = -s * e[k + 1];
        e[k + 1] = c * e[k + 1];
      }
      // accumulate: columns k, k+1 of Q  <=>  rows k, k+1 of Qt
      double* qk = Qt.data() + std::size_t(k) * N;
      double* qk1 = Qt.data() + std::size_t(k + 1) * N;
      for (int i = 0; i < n; ++i) {
        const double u = qk[i], v = qk1[i];
        qk[i] = c * u - s * v;
        qk1[i] = s * u + c * v;
      }
    }
  }

  // ---- sort ascending; Z row-major with eigenvector j in column j ---------
  std::vector<int> idx(N);
  std::iota(idx.begin(), idx.end(), 0);
  std::stable_sort(idx.begin(), idx.end(), [&](int x, int y) { return d[x] < d[y]; });
  for (int j = 0; j < n; ++j) {
    w[j] = d[idx[j]];
    const double* q = Qt.data() + std::size_t(idx[j]) * N;
    for (int i = 0; i < n; ++i) Z[std::size_t(i) * N + j] = q[i];
  }
  return 0;
}

} // namespace trl_amd::dense
