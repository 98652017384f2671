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
//
// syev_select needs only some eigenvectors (the Rayleigh-Ritz step uses the
// k_r*b leading Ritz vectors of m*b): eigenvalues by the same QR iteration
// without accumulation (O(n^2)), eigenvectors of the tridiagonal by inverse
// iteration with Gram-Schmidt inside clusters (as LAPACK's dstein, Wilkinson
// 1965; clusters: gaps below 1e-3 ||T||_1), then the reflectors applied to
// them: O(n^2 k) instead of O(n^3).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <limits>
#include <numeric>
#include <vector>

#include "trl_amd/dense.hh"

namespace trl_amd::dense {

namespace {

// Trailing blocks at least this many rows wide are reduced by all OpenMP
// threads; smaller ones serially (a parallel region costs a few microseconds,
// a serial step over 128 rows about as much).  Each entry is summed by one
// thread in a fixed order, so the result does not depend on the thread count.
constexpr int kParallelRows = 128;

// Householder tridiagonalisation of the row-major symmetric A (Golub & Van
// Loan Alg. 8.3.1): Q^T A Q = tridiag(d, e) with Q = H_0 H_1 ... H_{n-3},
// H_k = I - betas[k] v_k v_k^T, v_k stored in row k of vs (entries k+1..).
void tridiagonalise(int n, const double* A, std::vector<double>& vs, std::vector<double>& betas, std::vector<double>& d, std::vector<double>& e)
{
  const int wide = kParallelRows;
  const std::size_t N = std::size_t(n);
  // working copy, row-major.  While the trailing block is at least
  // `wide` rows wide both triangles are kept (rows contiguous for the
  // threads); after that only the lower triangle is used.
  std::vector<double> a(A, A + N * N);
  if (n >= wide)
    for (std::size_t i = 0; i < N; ++i)
      for (std::size_t j = i + 1; j < N; ++j) a[i * N + j] = a[j * N + i];
  vs.assign(N * N, 0.0);
  betas.assign(N, 0.0);
  std::vector<double> p(N), wv(N);
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
    // p = beta * A22 v.  Wide trailing blocks: full rows, one thread per
    // block of rows.  Narrow ones: from the lower triangle (row i, j <= i
    // holds A(i, j) = A(j, i)), one pass over it for both halves.
    const bool par = n - k >= wide;
    if (par) {
#pragma omp parallel for schedule(static)
      for (int i = k + 1; i < n; ++i) {
        const double* ai = a.data() + std::size_t(i) * N;
        double s = 0.0;
        for (int j = k + 1; j < n; ++j) s += ai[j] * v[j];
        p[i] = beta * s;
      }
    }
    else {
      for (int i = k + 1; i < n; ++i) p[i] = 0.0;
      for (int i = k + 1; i < n; ++i) {
        const double* ai = a.data() + std::size_t(i) * N;
        const double vi = v[i];
        double s = ai[i] * vi;
        for (int j = k + 1; j < i; ++j) {
          s += ai[j] * v[j];
          p[j] += ai[j] * vi;
        }
        p[i] += s;
      }
      for (int i = k + 1; i < n; ++i) p[i] *= beta;
    }
    double pv = 0.0;
    for (int i = k + 1; i < n; ++i) pv += p[i] * v[i];
    const double K = 0.5 * beta * pv;
    for (int i = k + 1; i < n; ++i) wv[i] = p[i] - K * v[i];
    // rank-2 update: full rows while wide, else the lower triangle only
    if (par) {
#pragma omp parallel for schedule(static)
      for (int i = k + 1; i < n; ++i) {
        double* ai = a.data() + std::size_t(i) * N;
        const double vi = v[i], wi = wv[i];
        for (int j = k + 1; j < n; ++j) ai[j] -= vi * wv[j] + wi * v[j];
      }
    }
    else {
      for (int i = k + 1; i < n; ++i) {
        double* ai = a.data() + std::size_t(i) * N;
        const double vi = v[i], wi = wv[i];
        for (int j = k + 1; j <= i; ++j) ai[j] -= vi * wv[j] + wi * v[j];
      }
    }
    a[std::size_t(k + 1) * N + k] = alpha;
  }
  d.assign(N, 0.0);
  e.assign(N, 0.0);
  for (int i = 0; i < n; ++i) d[i] = a[std::size_t(i) * N + i];
  for (int i = 0; i + 1 < n; ++i) e[i] = a[std::size_t(i + 1) * N + i];
}

// Implicit symmetric QR with the Wilkinson shift on tridiag(d, e) (Golub &
// Van Loan Alg. 8.3.2/8.3.3).  The rotations are applied to the rows of Qt
// (n x n, row k = k-th basis vector) when Qt is given.  d receives the
// eigenvalues (unsorted).  Returns 0, or -1 without convergence.
int tridiagonal_qr(int n, std::vector<double>& d, std::vector<double>& e, double* Qt)
{
  const std::size_t N = std::size_t(n);
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
      if (Qt) {
        // accumulate: columns k, k+1 of Q  <=>  rows k, k+1 of Qt
        double* qk = Qt + std::size_t(k) * N;
        double* qk1 = Qt + std::size_t(k + 1) * N;
        for (int i = 0; i < n; ++i) {
          const double u = qk[i], v = qk1[i];
          qk[i] = c * u - s * v;
          qk1[i] = s * u + c * v;
        }
      }
    }
  }
  return 0;
}

// Solves (T - lambda I) x = b in place for the tridiagonal T = tridiag(d, e)
// by Gaussian elimination with partial pivoting (the factorisation of
// LAPACK's dlagtf: U has two superdiagonals); pivots smaller than tiny are
// replaced by +-tiny (as dlagts does), which inverse iteration tolerates.
void shifted_tridiagonal_solve(int n, const double* d, const double* e, double lambda, double tiny, double* x, std::vector<double>& work)
{
  const std::size_t N = std::size_t(n);
  work.resize(4 * N);
  double* u0 = work.data();         // diagonal of U
  double* u1 = work.data() + N;     // first superdiagonal
  double* u2 = work.data() + 2 * N; // second superdiagonal
  double* l = work.data() + 3 * N;  // multipliers: row i+1 -= l[i] * row i (after the swap, if any)
  std::vector<char> swapped(N, 0);
  double a = d[0] - lambda, bsup = n > 1 ? e[0] : 0.0;
  for (int i = 0; i < n; ++i) {
    if (i + 1 == n) {
      u0[i] = std::fabs(a) < tiny ? (a < 0.0 ? -tiny : tiny) : a;
      u1[i] = u2[i] = 0.0;
      break;
    }
    const double c = e[i];                  // subdiagonal below a
    const double dn = d[i + 1] - lambda;    // next diagonal
    const double en = i + 2 < n ? e[i + 1] : 0.0;
    if (std::fabs(a) >= std::fabs(c)) {
      // no swap: rows i (a, bsup, 0) and i+1 (c, dn, en)
      const double piv = std::fabs(a) < tiny ? (a < 0.0 ? -tiny : tiny) : a;
      const double m = c / piv;
      u0[i] = piv;
      u1[i] = bsup;
      u2[i] = 0.0;
      l[i] = m;
      a = dn - m * bsup;
      bsup = en;
    }
    else {
      // swap rows i and i+1
      const double m = a / c;
      u0[i] = c;
      u1[i] = dn;
      u2[i] = en;
      l[i] = m;
      swapped[i] = 1;
      a = bsup - m * dn;
      bsup = -m * en;
    }
  }
  // forward: apply the row operations to x
  for (int i = 0; i + 1 < n; ++i) {
    if (swapped[i]) std::swap(x[i], x[i + 1]);
    x[i + 1] -= l[i] * x[i];
  }
  // back substitution with U (diagonal u0, superdiagonals u1, u2)
  for (int i = n - 1; i >= 0; --i) {
    double s = x[i];
    if (i + 1 < n) s -= u1[i] * x[i + 1];
    if (i + 2 < n) s -= u2[i] * x[i + 2];
    x[i] = s / u0[i];
  }
}

} // namespace

int syev(int n, const double* A, double* w, double* Z)
{
  if (n <= 0) return 0;
  if (n == 1) {
    w[0] = A[0];
    Z[0] = 1.0;
    return 0;
  }
  const std::size_t N = std::size_t(n);
  std::vector<double> vs, betas, d, e, p(N);
  tridiagonalise(n, A, vs, betas, d, e);

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

  if (tridiagonal_qr(n, d, e, Qt.data()) != 0) return -1;

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

int syev_select(int n, const double* A, double* w, int k, const int* sel, double* Z)
{
  if (n <= 0) return 0;
  for (int j = 0; j < k; ++j)
    if (sel[j] < 0 || sel[j] >= n || (j > 0 && sel[j] <= sel[j - 1])) return -2;
  if (n == 1) {
    w[0] = A[0];
    if (k == 1) Z[0] = 1.0;
    return 0;
  }
  const std::size_t N = std::size_t(n);
  std::vector<double> vs, betas, d, e;
  tridiagonalise(n, A, vs, betas, d, e);
  const std::vector<double> d0(d), e0(e); // the tridiagonal, for inverse iteration
  if (tridiagonal_qr(n, d, e, nullptr) != 0) return -1;
  std::sort(d.begin(), d.end());
  for (int j = 0; j < n; ++j) w[j] = d[j];
  if (k == 0) return 0;

  // ---- inverse iteration on tridiag(d0, e0) (as LAPACK's dstein) -----------
  double onenrm = 0.0;
  for (int i = 0; i < n; ++i) onenrm = std::max(onenrm, std::fabs(d0[i]) + std::fabs(e0[i]) + (i > 0 ? std::fabs(e0[i - 1]) : 0.0));
  const double eps = std::numeric_limits<double>::epsilon();
  const double ortol = 1e-3 * onenrm;         // eigenvalues closer than this form a cluster
  const double tiny = eps * std::max(onenrm, 1e-300);
  std::vector<double> y(N * std::size_t(k)), work;
  int cluster0 = 0;
  double prev = 0.0;
  for (int j = 0; j < k; ++j) {
    double lam = w[sel[j]];
    if (j > 0 && lam - w[sel[j - 1]] > ortol) cluster0 = j;
    // separate equal eigenvalues of a cluster so that each gets its own solve
    if (j > cluster0 && lam <= prev) lam = prev + 10.0 * eps * std::max(std::fabs(prev), onenrm);
    prev = lam;
    double* x = y.data() + std::size_t(j) * N;
    // start: a fixed pseudo-random vector (deterministic across runs)
    std::uint64_t st = 0x9E3779B97F4A7C15ull ^ std::uint64_t(sel[j] + 1);
    for (int i = 0; i < n; ++i) {
      st ^= st << 13;
      st ^= st >> 7;
      st ^= st << 17;
      x[i] = double(st >> 11) * 0x1.0p-53 - 0.5;
    }
    for (int it = 0; it < 4; ++it) {
      shifted_tridiagonal_solve(n, d0.data(), e0.data(), lam, tiny, x, work);
      // orthogonalise against the earlier vectors of the cluster (twice)
      for (int pass = 0; pass < 2; ++pass)
        for (int q = cluster0; q < j; ++q) {
          const double* yq = y.data() + std::size_t(q) * N;
          double dot = 0.0;
          for (int i = 0; i < n; ++i) dot += yq[i] * x[i];
          for (int i = 0; i < n; ++i) x[i] -= dot * yq[i];
        }
      // normalise by the largest entry first (no overflow in the squares)
      double big = 0.0;
      for (int i = 0; i < n; ++i) big = std::max(big, std::fabs(x[i]));
      if (!(big > 0.0) || !std::isfinite(big)) return -1;
      double nrm = 0.0;
      for (int i = 0; i < n; ++i) {
        x[i] /= big;
        nrm += x[i] * x[i];
      }
      nrm = std::sqrt(nrm);
      for (int i = 0; i < n; ++i) x[i] /= nrm;
    }
  }
  // ---- back-transformation z = H_0 H_1 ... H_{n-3} y, one vector a thread ---
#pragma omp parallel for schedule(static) if (n >= kParallelRows)
  for (int j = 0; j < k; ++j) {
    double* x = y.data() + std::size_t(j) * N;
    for (int kk = n - 3; kk >= 0; --kk) {
      const double beta = betas[kk];
      if (beta == 0.0) continue;
      const double* v = vs.data() + std::size_t(kk) * N;
      double s = 0.0;
      for (int i = kk + 1; i < n; ++i) s += v[i] * x[i];
      s *= beta;
      for (int i = kk + 1; i < n; ++i) x[i] -= s * v[i];
    }
    for (int i = 0; i < n; ++i) Z[std::size_t(i) * std::size_t(k) + j] = x[i];
  }
  return 0;
}

} // namespace trl_amd::dense
