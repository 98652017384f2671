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
// k_r*b leading Ritz vectors of m*b): eigenvalues by bisection on Sturm
// counts (threaded, eight counts at a time), eigenvectors of the tridiagonal
// by inverse iteration with Gram-Schmidt inside clusters (as LAPACK's dstein,
// Wilkinson 1965; clusters: gaps below 1e-3 ||T||_1), then the reflectors
// applied to them: O(n^2 k) instead of O(n^3).  ArrowBandEigensolver does
// the same for the thick-restart form of T after a structured reduction to
// tridiagonal form (below).
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

// All eigenvalues of tridiag(d, e), ascending in w, by bisection on Sturm
// counts (Givens' method; Kahan's count: the number of negative pivots of
// the LDL^T factorisation of T - x I, a pivot below pivmin taken as
// -pivmin).  Eigenvalue j is bracketed by lo (count <= j) and hi (count >
// j) from the Gershgorin interval and halved until hi - lo <= 2 eps
// max(|lo|, |hi|) + eps ||T|| -- the absolute accuracy eps ||T|| of the QR
// iteration.  Eight eigenvalues are bisected together: their eight
// independent count recurrences run in the same loop (the divisions
// pipeline, and the loop vectorises), and groups of eight run on separate
// threads.  Deterministic: each eigenvalue depends on its own bracket only.
void tridiagonal_eigenvalues_bisect(int n, const double* d, const double* e, double* w)
{
  if (n <= 0) return;
  if (n == 1) {
    w[0] = d[0];
    return;
  }
  const std::size_t N = std::size_t(n);
  const double eps = std::numeric_limits<double>::epsilon();
  std::vector<double> e2(N, 0.0);
  double gl = d[0], gu = d[0], e2max = 0.0;
  for (int i = 0; i < n; ++i) {
    const double r = (i > 0 ? std::fabs(e[i - 1]) : 0.0) + (i + 1 < n ? std::fabs(e[i]) : 0.0);
    gl = std::min(gl, d[i] - r);
    gu = std::max(gu, d[i] + r);
    if (i + 1 < n) {
      e2[std::size_t(i)] = e[i] * e[i];
      e2max = std::max(e2max, e2[std::size_t(i)]);
    }
  }
  const double tnorm = std::max(std::fabs(gl), std::fabs(gu));
  const double pivmin = std::numeric_limits<double>::min() * std::max(1.0, e2max);
  const double abstol = eps * tnorm;
  // widen so that the ends strictly bracket the spectrum
  gl -= 2.0 * eps * tnorm + 2.0 * pivmin;
  gu += 2.0 * eps * tnorm + 2.0 * pivmin;
  constexpr int G = 8;
  const int groups = (n + G - 1) / G;
#pragma omp parallel for schedule(static) if (n >= kParallelRows)
  for (int gi = 0; gi < groups; ++gi) {
    const int j0 = gi * G;
    double lo[G], hi[G], x[G], q[G], cnt[G]; // cnt: a count, held in a double so that the loop vectorises
    for (int g = 0; g < G; ++g) {
      lo[g] = gl;
      hi[g] = gu;
    }
    for (int it = 0; it < 128; ++it) {
      bool done = true;
      for (int g = 0; g < G; ++g) {
        x[g] = 0.5 * (lo[g] + hi[g]);
        if (hi[g] - lo[g] > 2.0 * eps * std::max(std::fabs(lo[g]), std::fabs(hi[g])) + abstol && x[g] > lo[g] && x[g] < hi[g]) done = false;
      }
      if (done) break;
      for (int g = 0; g < G; ++g) {
        q[g] = d[0] - x[g];
        q[g] = std::fabs(q[g]) < pivmin ? -pivmin : q[g];
        cnt[g] = q[g] < 0.0 ? 1.0 : 0.0;
      }
      for (int i = 1; i < n; ++i) {
        const double di = d[i], ei = e2[std::size_t(i - 1)];
        for (int g = 0; g < G; ++g) {
          double t = (di - x[g]) - ei / q[g];
          t = std::fabs(t) < pivmin ? -pivmin : t;
          q[g] = t;
          cnt[g] += t < 0.0 ? 1.0 : 0.0;
        }
      }
      for (int g = 0; g < G; ++g) {
        // eigenvalue j0 + g: count(x) <= j  ->  x is a lower bound
        if (cnt[g] <= double(j0 + g))
          lo[g] = x[g];
        else
          hi[g] = x[g];
      }
    }
    for (int g = 0; g < G && j0 + g < n; ++g) w[j0 + g] = 0.5 * (lo[g] + hi[g]);
  }
}

// LU factorisation of T - lambda I for the tridiagonal T = tridiag(d, e) by
// Gaussian elimination with partial pivoting (the factorisation of LAPACK's
// dlagtf: U has two superdiagonals), and solves with it in place.  Pivots
// smaller than tiny are replaced by +-tiny (as dlagts does), which inverse
// iteration tolerates.  The factors live in the object, so one factorisation
// serves every iteration of a shift.
struct ShiftedTridiagonalLU {
  int n = 0;
  std::vector<double> u0, u1, u2, l; // diagonal and superdiagonals of U; multipliers
  std::vector<char> swapped;         // row i and i+1 exchanged before eliminating

  void factor(int n_, const double* d, const double* e, double lambda, double tiny)
  {
    n = n_;
    const std::size_t N = std::size_t(n);
    u0.resize(N);
    u1.resize(N);
    u2.resize(N);
    l.resize(N);
    swapped.assign(N, 0);
    double a = d[0] - lambda, bsup = n > 1 ? e[0] : 0.0;
    for (int i = 0; i < n; ++i) {
      if (i + 1 == n) {
        u0[i] = std::fabs(a) < tiny ? (a < 0.0 ? -tiny : tiny) : a;
        u1[i] = u2[i] = 0.0;
        break;
      }
      const double c = e[i];               // subdiagonal below a
      const double dn = d[i + 1] - lambda; // next diagonal
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
  }

  void solve(double* x) const
  {
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
};

// Eigenvectors of tridiag(d0, e0) for the eigenvalues w[sel[0..k)] (w
// ascending, sel strictly increasing) by inverse iteration with
// Gram-Schmidt inside clusters, as LAPACK's dstein: vector j in y + j n.
// Returns 0, or -1 on a non-finite iterate.
int tridiagonal_vectors(int n, const double* d0, const double* e0, const double* w, int k, const int* sel, double* yv)
{
  const std::size_t N = std::size_t(n);
  double onenrm = 0.0;
  for (int i = 0; i < n; ++i) onenrm = std::max(onenrm, std::fabs(d0[i]) + (i + 1 < n ? std::fabs(e0[i]) : 0.0) + (i > 0 ? std::fabs(e0[i - 1]) : 0.0));
  const double eps = std::numeric_limits<double>::epsilon();
  const double ortol = 1e-3 * onenrm;         // eigenvalues closer than this form a cluster
  const double tiny = eps * std::max(onenrm, 1e-300);
  // clusters [first[c], first[c+1]): independent of each other
  std::vector<int> first;
  for (int j = 0; j < k; ++j)
    if (j == 0 || w[sel[j]] - w[sel[j - 1]] > ortol) first.push_back(j);
  first.push_back(k);
  const int nclusters = int(first.size()) - 1;
  int failed = 0;
#pragma omp parallel if (n >= kParallelRows && nclusters > 1) reduction(| : failed)
  {
    ShiftedTridiagonalLU lu;
#pragma omp for schedule(dynamic)
    for (int cl = 0; cl < nclusters; ++cl) {
      double prev = 0.0;
      for (int j = first[cl]; j < first[cl + 1] && !failed; ++j) {
        double lam = w[sel[j]];
        // separate equal eigenvalues of a cluster so that each gets its own solve
        if (j > first[cl] && lam <= prev) lam = prev + 10.0 * eps * std::max(std::fabs(prev), onenrm);
        prev = lam;
        double* x = yv + std::size_t(j) * N;
        // start: a fixed pseudo-random vector (deterministic across runs)
        std::uint64_t st = 0x9E3779B97F4A7C15ull ^ std::uint64_t(sel[j] + 1);
        for (int i = 0; i < n; ++i) {
          st ^= st << 13;
          st ^= st >> 7;
          st ^= st << 17;
          x[i] = double(st >> 11) * 0x1.0p-53 - 0.5;
        }
        lu.factor(n, d0, e0, lam, tiny);
        for (int it = 0; it < 4; ++it) {
          lu.solve(x);
          // orthogonalise against the earlier vectors of the cluster (twice)
          for (int pass = 0; pass < 2; ++pass)
            for (int q = first[cl]; q < j; ++q) {
              const double* yq = yv + std::size_t(q) * N;
              double dot = 0.0;
              for (int i = 0; i < n; ++i) dot += yq[i] * x[i];
              for (int i = 0; i < n; ++i) x[i] -= dot * yq[i];
            }
          // normalise by the largest entry first (no overflow in the squares)
          double big = 0.0;
          for (int i = 0; i < n; ++i) big = std::max(big, std::fabs(x[i]));
          if (!(big > 0.0) || !std::isfinite(big)) {
            failed = 1;
            break;
          }
          double nrm = 0.0;
          for (int i = 0; i < n; ++i) {
            x[i] /= big;
            nrm += x[i] * x[i];
          }
          nrm = std::sqrt(nrm);
          for (int i = 0; i < n; ++i) x[i] /= nrm;
        }
      }
    }
  }
  return failed ? -1 : 0;
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
  tridiagonal_eigenvalues_bisect(n, d.data(), e.data(), w);
  if (k == 0) return 0;

  std::vector<double> y(N * std::size_t(k));
  if (tridiagonal_vectors(n, d0.data(), e0.data(), w, k, sel, y.data()) != 0) return -1;
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

// ---------------------------------------------------------------------------
// The projected matrix of the thick-restart block Lanczos in band form.
//
// After a restart the projected matrix is T = [D M^T; M X]: D = diag(theta)
// on the p = k_r b kept Ritz coordinates, M (b x p) the arrow that couples
// them to block k_r only, and X block tridiagonal from block k_r on, whose
// sub-diagonal blocks are upper triangular (the R factors of the block
// orthonormalisation), so that X has half-bandwidth b.  Before the first
// restart p = 0 and T = X.  The dense Householder reduction costs
// O((mb)^3); here:
//  1. the D part is reduced, bottom up, to block tridiagonal form whose last
//     block alone couples to X, through a lower-triangular block: QL
//     factorisations of the coupling panel (Householder, from its last
//     column), each applied two-sided to the rest of the D part -- O(p^3),
//     p = 40 at the benchmark's size.  T is then banded with half-bandwidth
//     b in its natural order;
//  2. the band is reduced to tridiagonal form by chasing each column's
//     bulge off the end of the band (the reduction of Schwarz,
//     "Tridiagonalization of a symmetric band matrix", Numer. Math. 12,
//     1968, with a Householder reflector of length b where Schwarz has b - 1
//     Givens rotations, as in Lang's band reduction): O(n^2 b);
//  3. all eigenvalues of the tridiagonal, the selected eigenvectors by
//     inverse iteration (tridiagonal_vectors), and the band reflectors and
//     the D-part reflectors applied to those k vectors: O(n^2 k) -- against
//     O(n^3) for the Householder tridiagonalisation.
namespace {

// Householder reflector (Golub & Van Loan Alg. 5.1.1) that maps x (length m)
// onto a multiple of its LAST unit vector: v (v[m-1] = 1) and beta with
// (I - beta v v^T) x = alpha e_{m-1}.
void house_last(int m, const double* x, double* v, double& beta)
{
  double sigma = 0.0;
  for (int i = 0; i + 1 < m; ++i) sigma += x[i] * x[i];
  const double x1 = x[m - 1];
  for (int i = 0; i + 1 < m; ++i) v[i] = x[i];
  v[m - 1] = 1.0;
  if (sigma == 0.0) {
    beta = 0.0;
    return;
  }
  const double mu = std::sqrt(x1 * x1 + sigma);
  const double v1 = x1 <= 0.0 ? x1 - mu : -sigma / (x1 + mu);
  beta = 2.0 * v1 * v1 / (sigma + v1 * v1);
  for (int i = 0; i + 1 < m; ++i) v[i] /= v1;
}

} // namespace

int arrow_band_structure(int n, int b, const double* A)
{
  if (n <= 0 || b < 1) return -1;
  const std::size_t N = std::size_t(n);
  // p: the largest multiple of b whose leading p x p block is diagonal
  int p = 0;
  while (p + b <= n) {
    bool diag = true;
    for (int i = p; i < p + b && diag; ++i)
      for (int j = 0; j < p + b && diag; ++j)
        if (j != i && A[std::size_t(i) * N + j] != 0.0) diag = false;
    if (!diag) break;
    p += b;
  }
  if (p >= n - b) return -1; // nothing left to form the Lanczos part
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) {
      const double a = A[std::size_t(i) * N + j];
      if (a == 0.0) continue;
      const int lo = std::min(i, j), hi = std::max(i, j);
      if (lo < p) {
        if (hi >= p + b && hi != lo) return -1; // the Ritz part couples beyond block k_r
      }
      else if (hi - lo > b)
        return -1; // outside the band of the Lanczos part
    }
  return p;
}

int ArrowBandEigensolver::reduce(int n, int b, int p, const double* A, double* w)
{
  n_ = n;
  p_ = p;
  refl_.clear();
  vstore_.clear();
  if (n <= 0) return 0;
  if (b < 1 || p < 0 || p % b != 0 || p + b > n) return -2;
  const std::size_t N = std::size_t(n);
  // step 1 works on the leading m0 x m0 window (the Ritz part and the first
  // Lanczos block), copied with both triangles; the rest is read from A
  const int m0 = p + b;
  const std::size_t M0 = std::size_t(m0);
  std::vector<double> a(M0 * M0);
  for (int i = 0; i < m0; ++i) std::copy(A + std::size_t(i) * N, A + std::size_t(i) * N + M0, a.begin() + std::ptrdiff_t(std::size_t(i) * M0));
  auto at = [&](int i, int j) -> double& { return a[std::size_t(i) * M0 + j]; };

  // ---- 1. the Ritz part: block QL reduction, bottom up ----------------------
  // Qd (p x p, row-major) accumulates the reflectors: the new Ritz-part basis
  // vectors are its columns.
  std::vector<double>& Qd = Qd_;
  Qd.assign(std::size_t(p) * p, 0.0);
  std::vector<double> v(std::size_t(std::max(p, 1))), pv(std::size_t(std::max(p, 1)));
  for (int i = 0; i < p; ++i) Qd[std::size_t(i) * p + i] = 1.0;
  for (int cur = p; cur > 0;) {
    // the panel: rows 0..cur-1 of the Ritz part against the block below
    // (columns cur..cur+b-1), reduced column by column from the last so that
    // it ends lower triangular in rows cur-b..cur-1 and zero above them
    const int h = b;
    for (int c = h - 1; c >= 0; --c) {
      const int col = cur + c;         // a column of the block below
      const int m = cur - (h - 1 - c); // rows 0..m-1 take part: the rows below are already fixed
      // x = column `col`, rows 0..m-1: reflect onto its last entry (row m-1)
      for (int i = 0; i < m; ++i) pv[std::size_t(i)] = at(i, col);
      double beta = 0.0;
      house_last(m, pv.data(), v.data(), beta);
      if (beta == 0.0) continue;
      // two-sided update of the symmetric rows/columns 0..m-1, over all
      // columns they couple to: the Ritz part (0..cur-1, plus the block
      // below, cur..cur+h-1); entries beyond that are zero
      const int cend = cur + h; // <= m0
      // rows: a[0:m, 0:cend] <- H a[0:m, 0:cend]
      for (int j = 0; j < cend; ++j) {
        double t = 0.0;
        for (int i = 0; i < m; ++i) t += v[std::size_t(i)] * at(i, j);
        t *= beta;
        for (int i = 0; i < m; ++i) at(i, j) -= t * v[std::size_t(i)];
      }
      // columns: a[0:cend, 0:m] <- a[0:cend, 0:m] H
      for (int i = 0; i < cend; ++i) {
        double t = 0.0;
        for (int j = 0; j < m; ++j) t += at(i, j) * v[std::size_t(j)];
        t *= beta;
        for (int j = 0; j < m; ++j) at(i, j) -= t * v[std::size_t(j)];
      }
      // Qd <- Qd H (columns 0..m-1)
      for (int i = 0; i < p; ++i) {
        double* qi = Qd.data() + std::size_t(i) * p;
        double t = 0.0;
        for (int j = 0; j < m; ++j) t += qi[j] * v[std::size_t(j)];
        t *= beta;
        for (int j = 0; j < m; ++j) qi[j] -= t * v[std::size_t(j)];
      }
    }
    cur -= b;
  }

  // ---- 2. band (half-bandwidth b) to tridiagonal: Householder bulge chase --
  // Column j is reduced by a reflector on rows j+1..j+b; applied from the
  // right it fills the b x b block below the band (the bulge), whose first
  // column the next reflector, b rows further down, reduces in turn, and so
  // on to the end of the matrix (one sweep per column, about (n - j) / b
  // reflectors of length b).  The part of a bulge beyond its first column
  // is left in place: it lies in the columns that the next sweep's
  // reflectors reduce, so entries stay within 2b - 1 below the diagonal.
  // The lower triangle is kept in a dense n x n array; entries more than b
  // from the diagonal are zero up to rounding after step 1 and dropped.
  std::vector<double> L(N * N, 0.0);
  auto lt = [&](int r, int c) -> double& { return L[std::size_t(r) * N + std::size_t(c)]; }; // r >= c
  for (int r = 0; r < n; ++r)
    for (int t = 0; t <= b && t <= r; ++t) lt(r, r - t) = r < m0 ? at(r, r - t) : A[std::size_t(r) * N + std::size_t(r - t)];
  refl_.clear();
  vstore_.clear();
  refl_.reserve(N * N / std::size_t(2 * b) + N);
  vstore_.reserve((N * N / std::size_t(2 * b) + N) * std::size_t(b));
  const std::size_t B = static_cast<std::size_t>(b);
  std::vector<double> S(B * B), pw(B), ww(B), hv(B);
  for (int j = 0; j + 2 < n; ++j) {
    int c0 = j, r0 = j + 1;
    while (r0 + 1 < n) {
      const int len = std::min(b, n - r0);
      // the reflector: (I - tau v v^T) x = mu e_0 for x = L[r0:r0+len, c0]
      double sigma = 0.0;
      for (int i = 1; i < len; ++i) sigma += lt(r0 + i, c0) * lt(r0 + i, c0);
      const double x0 = lt(r0, c0);
      if (sigma != 0.0) {
        const double mu = std::sqrt(x0 * x0 + sigma);
        const double v0 = x0 <= 0.0 ? x0 - mu : -sigma / (x0 + mu);
        const double tau = 2.0 * v0 * v0 / (sigma + v0 * v0);
        hv[0] = 1.0;
        for (int i = 1; i < len; ++i) hv[std::size_t(i)] = lt(r0 + i, c0) / v0;
        lt(r0, c0) = mu;
        for (int i = 1; i < len; ++i) lt(r0 + i, c0) = 0.0;
        // the other columns left of the reflector's rows (the bulge)
        for (int c = c0 + 1; c < r0; ++c) {
          double t = 0.0;
          for (int i = 0; i < len; ++i) t += hv[std::size_t(i)] * lt(r0 + i, c);
          t *= tau;
          for (int i = 0; i < len; ++i) lt(r0 + i, c) -= t * hv[std::size_t(i)];
        }
        // the diagonal block: S <- H S H = S - v w^T - w v^T, with p = tau S v
        // and w = p - (tau / 2)(p^T v) v
        for (int r = 0; r < len; ++r)
          for (int c = 0; c <= r; ++c) S[std::size_t(r) * b + c] = S[std::size_t(c) * b + r] = lt(r0 + r, r0 + c);
        double pv = 0.0;
        for (int r = 0; r < len; ++r) {
          double t = 0.0;
          for (int c = 0; c < len; ++c) t += S[std::size_t(r) * b + c] * hv[std::size_t(c)];
          pw[std::size_t(r)] = tau * t;
          pv += pw[std::size_t(r)] * hv[std::size_t(r)];
        }
        for (int r = 0; r < len; ++r) ww[std::size_t(r)] = pw[std::size_t(r)] - 0.5 * tau * pv * hv[std::size_t(r)];
        for (int r = 0; r < len; ++r)
          for (int c = 0; c <= r; ++c) lt(r0 + r, r0 + c) -= hv[std::size_t(r)] * ww[std::size_t(c)] + ww[std::size_t(r)] * hv[std::size_t(c)];
        // the rows below: L[rr, r0:r0+len] <- L[rr, r0:r0+len] H; nonzero
        // only within 2b - 1 of the reflector's rows
        const int rend = std::min(n, r0 + len + 2 * b - 1);
        for (int rr = r0 + len; rr < rend; ++rr) {
          double* row = L.data() + std::size_t(rr) * N + std::size_t(r0);
          double t = 0.0;
          for (int i = 0; i < len; ++i) t += row[i] * hv[std::size_t(i)];
          t *= tau;
          for (int i = 0; i < len; ++i) row[i] -= t * hv[std::size_t(i)];
        }
        refl_.push_back({r0, len, tau});
        vstore_.insert(vstore_.end(), hv.begin(), hv.begin() + len);
      }
      c0 = r0;
      r0 += b;
    }
  }
  std::vector<double> d(N), e(N, 0.0);
  for (int r = 0; r < n; ++r) d[std::size_t(r)] = lt(r, r);
  for (int r = 0; r + 1 < n; ++r) e[std::size_t(r)] = lt(r + 1, r);
  d0_ = d;
  e0_ = e;

  // ---- 3. the eigenvalues --------------------------------------------------
  w_.resize(N);
  tridiagonal_eigenvalues_bisect(n, d.data(), e.data(), w_.data());
  std::copy(w_.begin(), w_.end(), w);
  return 0;
}

int ArrowBandEigensolver::vectors(int k, const int* sel, double* Z) const
{
  const int n = n_, p = p_;
  const std::size_t N = std::size_t(n);
  for (int j = 0; j < k; ++j)
    if (sel[j] < 0 || sel[j] >= n || (j > 0 && sel[j] <= sel[j - 1])) return -2;
  if (k == 0 || n <= 0) return 0;
  const std::vector<double>& Qd = Qd_;
  // ---- 4. the selected eigenvectors of the tridiagonal, back-transformed ---
  std::vector<double> y(N * std::size_t(k));
  if (tridiagonal_vectors(n, d0_.data(), e0_.data(), w_.data(), k, sel, y.data()) != 0) return -1;
  // strips of 8 vectors, each strip coordinate-major and contiguous
  // (xs[(st * n + i) * 8 + jj] = coordinate i of vector 8 st + jj): a
  // reflector updates len adjacent 64-byte rows of a strip, and the strips,
  // one a thread, share no cache line
  const std::size_t K = std::size_t(k);
  const int strips = (k + 7) / 8;
  std::vector<double> xs(std::size_t(strips) * N * 8 + 8, 0.0);
  double* xb = xs.data() + ((64 - (reinterpret_cast<std::uintptr_t>(xs.data()) & 63)) & 63) / sizeof(double);
  for (int jv = 0; jv < k; ++jv)
    for (int r = 0; r < n; ++r) xb[(std::size_t(jv / 8) * N + std::size_t(r)) * 8 + std::size_t(jv % 8)] = y[std::size_t(jv) * N + r];
  // the reflectors, last first (x <- H x), a strip a thread
  std::vector<std::size_t> voff(refl_.size() + 1, 0);
  for (std::size_t t = 0; t < refl_.size(); ++t) voff[t + 1] = voff[t] + std::size_t(refl_[t].len);
#pragma omp parallel for schedule(static) if (strips > 1 && n >= kParallelRows)
  for (int st = 0; st < strips; ++st) {
    double* xst = xb + std::size_t(st) * N * 8;
    for (std::size_t t = refl_.size(); t-- > 0;) {
      const Refl h = refl_[t];
      const double* v = vstore_.data() + voff[t];
      double* xr = xst + std::size_t(h.r0) * 8;
      double dot[8] = {0, 0, 0, 0, 0, 0, 0, 0};
      for (int i = 0; i < h.len; ++i)
        for (int jj = 0; jj < 8; ++jj) dot[jj] += v[i] * xr[std::size_t(i) * 8 + std::size_t(jj)];
      for (int jj = 0; jj < 8; ++jj) dot[jj] *= h.tau;
      for (int i = 0; i < h.len; ++i)
        for (int jj = 0; jj < 8; ++jj) xr[std::size_t(i) * 8 + std::size_t(jj)] -= dot[jj] * v[i];
    }
  }
  auto xat = [&](int r, int jv) { return xb[(std::size_t(jv / 8) * N + std::size_t(r)) * 8 + std::size_t(jv % 8)]; };
  // the Ritz-part reflectors: x[0:p] <- Qd x[0:p]; the rest as it is
  for (int r = 0; r < p; ++r)
    for (int jv = 0; jv < k; ++jv) {
      double acc = 0.0;
      for (int t = 0; t < p; ++t) acc += Qd[std::size_t(r) * p + t] * xat(t, jv);
      Z[std::size_t(r) * K + jv] = acc;
    }
  for (int r = p; r < n; ++r)
    for (int jv = 0; jv < k; ++jv) Z[std::size_t(r) * K + jv] = xat(r, jv);
  return 0;
}

int syev_arrow_band_select(int n, int b, int p, const double* A, double* w, int k, const int* sel, double* Z)
{
  for (int j = 0; j < k; ++j)
    if (sel[j] < 0 || sel[j] >= n || (j > 0 && sel[j] <= sel[j - 1])) return -2;
  ArrowBandEigensolver es;
  const int rc = es.reduce(n, b, p, A, w);
  return rc != 0 ? rc : es.vectors(k, sel, Z);
}

} // namespace trl_amd::dense
