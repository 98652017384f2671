// ============================================================================
// CPU ORACLE — TEST INFRASTRUCTURE ONLY.
//
// A plain C++/OpenMP restatement of the reference's block-Lanczos path, used
// as the checker for the HIP backend (tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py).  Nothing in the product links, loads or
// calls this file.
//
// Follows, function by function (citations relative to the reference root):
//   panel ops ........ include/trl/impl/openmp/blockview.hh:31-201
//   storage .......... include/trl/impl/openmp/multivector.hh:38-89,
//                      blockmatrix.hh:24-58, matrixblockview.hh:30-74
//   EVP base ......... include/trl/impl/openmp/evp_base.hh:37-154
//                      (CholQR :39-71, Rayleigh-Ritz + convergence :79-154)
//   eigen order ...... evp_base.hh:103 (|lambda| descending, the default) and
//                      examples/sycl/evp_base.hh:24,110-117 (ascending/descending)
//   operator ......... examples/openmp/csrevp.hh:36-44 (A X),
//                      :89-97 (B-inner product, here with a diagonal B)
//   MGS reorth ....... include/trl/eigensolvers/reorthogonalization.hh:12-22
//   driver ........... include/trl/eigensolvers/lanczos.hh:30-287
//
// Eigen3 (the reference's dense LA: LLT, inverse, SelfAdjointEigenSolver) is
// absent from this image, so the dense pieces are this file's own:
//   cholesky_lower_llt ... Cholesky-Banachiewicz on the lower triangle
//   invert_upper ......... back substitution
//   syev ................. Householder tridiagonalisation + implicit
//                          symmetric QR with the Wilkinson shift (Golub &
//                          Van Loan, Matrix Computations, Alg. 8.3.1-8.3.3)
// Each is pinned by tests/test_oracle_*.py against numpy/LAPACK and closed forms.
// ============================================================================
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <limits>
#include <numeric>
#include <vector>

#include <omp.h>

namespace oracle {

using Index = std::size_t;

// ---------------------------------------------------------------------------
// Panel (BlockView): n x b row-major, element (r, c) at r*b + c.
struct Panel {
  double* d;
  Index n;
  int b;
};

static void set_zero(Panel X)
{
  const Index tot = X.n * X.b;
#pragma omp parallel for
  for (Index i = 0; i < tot; ++i) X.d[i] = 0.0;
}

static void copy_from(Panel X, Panel Y)
{
  const Index tot = X.n * X.b;
#pragma omp parallel for
  for (Index i = 0; i < tot; ++i) X.d[i] = Y.d[i];
}

static void sub_assign(Panel X, Panel Y)
{
  const Index tot = X.n * X.b;
#pragma omp parallel for
  for (Index i = 0; i < tot; ++i) X.d[i] -= Y.d[i];
}

// mode 0: Y = X W ; 1: Y += X W ; 2: Y = X W^T     (blockview.hh:81-141)
static void panel_mult(Panel X, const double* W, Panel Y, int mode)
{
  const int b = X.b;
#pragma omp parallel for
  for (Index k = 0; k < X.n; ++k) {
    double tmp[16];
    for (int i = 0; i < b; ++i) {
      double s = 0.0;
      for (int j = 0; j < b; ++j) s += X.d[k * b + j] * (mode == 2 ? W[i * b + j] : W[j * b + i]);
      tmp[i] = s;
    }
    for (int i = 0; i < b; ++i) {
      if (mode == 1) Y.d[k * b + i] += tmp[i];
      else Y.d[k * b + i] = tmp[i];
    }
  }
}

// X -= Y R     (blockview.hh:189-201)
static void subtract_product(Panel X, Panel Y, const double* R)
{
  const int b = X.b;
#pragma omp parallel for
  for (Index k = 0; k < X.n; ++k) {
    for (int i = 0; i < b; ++i) {
      double s = 0.0;
      for (int j = 0; j < b; ++j) s += Y.d[k * b + j] * R[j * b + i];
      X.d[k * b + i] -= s;
    }
  }
}

// R = X^T diag(m) Y     (blockview.hh:152-182: thread-local b x b, combined
// under a critical section; m = 1 unless a mass diagonal is given)
static void dot(Panel X, Panel Y, const double* m, double* R)
{
  const int b = X.b;
  for (int i = 0; i < b * b; ++i) R[i] = 0.0;
#pragma omp parallel
  {
    double loc[256] = {0.0};
#pragma omp for
    for (Index k = 0; k < X.n; ++k) {
      const double w = m ? m[k] : 1.0;
      for (int i = 0; i < b; ++i) {
        const double xi = X.d[k * b + i] * w;
        for (int j = 0; j < b; ++j) loc[i * b + j] += xi * Y.d[k * b + j];
      }
    }
#pragma omp critical
    for (int i = 0; i < b * b; ++i) R[i] += loc[i];
  }
}

// ---------------------------------------------------------------------------
// Operator: Y = diag(dscale) A X for CSR A.  (examples/openmp/csrevp.hh:36-44
// evaluates A X from the lower triangle with selfadjointView; the CSR here
// holds both triangles of the same symmetric A.)
struct CsrOp {
  Index n;
  const int32_t* rowptr;
  const int32_t* col;
  const double* val;
  const double* dscale; // may be null
  const double* mdiag;  // may be null: inner product weight
};

static void apply(const CsrOp& A, Panel X, Panel Y)
{
  const int b = X.b;
#pragma omp parallel for schedule(static)
  for (Index r = 0; r < A.n; ++r) {
    double acc[16] = {0.0};
    for (int32_t p = A.rowptr[r]; p < A.rowptr[r + 1]; ++p) {
      const double v = A.val[p];
      const double* x = X.d + Index(A.col[p]) * b;
      for (int c = 0; c < b; ++c) acc[c] += v * x[c];
    }
    const double s = A.dscale ? A.dscale[r] : 1.0;
    for (int c = 0; c < b; ++c) Y.d[r * b + c] = s * acc[c];
  }
}

// ---------------------------------------------------------------------------
// Dense LA (own code; Eigen is absent).

// G = L L^T from the lower triangle of G (row-major b x b); writes R = L^T
// (upper).  Returns 0 on success, else the 1-based failing pivot.
static int cholesky_lower_llt(const double* G, int b, double* R)
{
  double L[256] = {0.0};
  for (int j = 0; j < b; ++j) {
    double d = G[j * b + j];
    for (int p = 0; p < j; ++p) d -= L[j * b + p] * L[j * b + p];
    if (!(d > 0.0)) return j + 1;
    const double ljj = std::sqrt(d);
    L[j * b + j] = ljj;
    for (int i = j + 1; i < b; ++i) {
      double s = G[i * b + j];
      for (int p = 0; p < j; ++p) s -= L[i * b + p] * L[j * b + p];
      L[i * b + j] = s / ljj;
    }
  }
  for (int i = 0; i < b; ++i)
    for (int j = 0; j < b; ++j) R[i * b + j] = L[j * b + i];
  return 0;
}

static void invert_upper(const double* R, int b, double* X)
{
  for (int i = 0; i < b * b; ++i) X[i] = 0.0;
  for (int j = 0; j < b; ++j) {
    X[j * b + j] = 1.0 / R[j * b + j];
    for (int i = j - 1; i >= 0; --i) {
      double s = 0.0;
      for (int p = i + 1; p <= j; ++p) s += R[i * b + p] * X[p * b + j];
      X[i * b + j] = -s / R[i * b + i];
    }
  }
}

// Symmetric eigen-decomposition, A (n x n row-major, symmetric) -> ascending
// eigenvalues w and eigenvectors as the columns of Z (row-major n x n), by
// the textbook two-stage method of Golub & Van Loan, Matrix Computations
// (4th ed.), written here on the full row-major matrix:
//  1. Householder tridiagonalisation Q^T A Q = T (Alg. 8.3.1).  Step k takes
//     the reflector H = I - beta v v^T (v(0) = 1, the stable choice of
//     Alg. 5.1.1) that maps x = A(k+1:n, k) onto |x| e_1, and updates the
//     trailing block as A22 -= v w^T + w v^T with p = beta A22 v,
//     w = p - (beta p^T v / 2) v.  Q = H_0 H_1 ... is formed forward.
//  2. Implicit symmetric QR steps with the Wilkinson shift on the unreduced
//     trailing block of T (Alg. 8.3.2), deflating off-diagonal entries below
//     eps (|d_i| + |d_{i+1}|) (Alg. 8.3.3).  A step starts from the shifted
//     first column and chases the bulge down with Givens rotations, each of
//     which is also applied to the columns of Q.
// Returns 0, or -1 if the QR iteration needs more than 30 n steps.
static int syev(int n, const double* A, double* w, double* Z)
{
  if (n <= 0) return 0;
  const size_t N = size_t(n);
  // Qt holds Q transposed (row j = column j of Q), so that both the
  // reflector and the rotation updates run along contiguous rows
  std::vector<double> a(A, A + N * N), Qt(N * N, 0.0);
  auto at = [&](int i, int j) -> double& { return a[size_t(i) * N + j]; };
  auto qrow = [&](int j) { return Qt.data() + size_t(j) * N; };
  for (int i = 0; i < n; ++i) qrow(i)[i] = 1.0;

  // ---- 1. tridiagonalisation ------------------------------------------------
  std::vector<double> v(N), pv(N), wv(N);
  for (int k = 0; k + 2 < n; ++k) {
    const double x1 = at(k + 1, k);
    double sigma = 0.0;
    for (int i = k + 2; i < n; ++i) sigma += at(i, k) * at(i, k);
    if (sigma == 0.0) continue; // already in tridiagonal form in this column
    const double mu = std::sqrt(x1 * x1 + sigma);
    const double v1 = x1 <= 0.0 ? x1 - mu : -sigma / (x1 + mu);
    const double beta = 2.0 * v1 * v1 / (sigma + v1 * v1);
    v[k + 1] = 1.0;
    for (int i = k + 2; i < n; ++i) v[i] = at(i, k) / v1;
    double pt_v = 0.0;
    for (int i = k + 1; i < n; ++i) {
      double acc = 0.0;
      for (int j = k + 1; j < n; ++j) acc += at(i, j) * v[j];
      pv[i] = beta * acc;
      pt_v += pv[i] * v[i];
    }
    const double half = 0.5 * beta * pt_v;
    for (int i = k + 1; i < n; ++i) wv[i] = pv[i] - half * v[i];
#pragma omp parallel for if (n - k >= 128)
    for (int i = k + 1; i < n; ++i)
      for (int j = k + 1; j < n; ++j) at(i, j) -= v[i] * wv[j] + wv[i] * v[j];
    at(k + 1, k) = at(k, k + 1) = mu;
    for (int i = k + 2; i < n; ++i) at(i, k) = at(k, i) = 0.0;
    // Q <- Q H: rows k+1.. of Qt lose beta v (v^T Qt(k+1:, :))
    std::fill(pv.begin(), pv.end(), 0.0);
    for (int j = k + 1; j < n; ++j) {
      const double* qj = qrow(j);
      for (int r = 0; r < n; ++r) pv[r] += v[j] * qj[r];
    }
    for (int j = k + 1; j < n; ++j) {
      double* qj = qrow(j);
      const double f = beta * v[j];
      for (int r = 0; r < n; ++r) qj[r] -= f * pv[r];
    }
  }
  std::vector<double> d(N), e(N, 0.0); // e[i] = T(i, i+1)
  for (int i = 0; i < n; ++i) d[i] = at(i, i);
  for (int i = 0; i + 1 < n; ++i) e[i] = at(i + 1, i);

  // ---- 2. implicit QR with the Wilkinson shift -------------------------------
  const double eps = std::numeric_limits<double>::epsilon();
  auto rotate_q = [&](int k, double c, double s) { // Q <- Q G, G = [c s; -s c] on columns k, k+1
    double* qa = qrow(k);
    double* qb = qrow(k + 1);
    for (int r = 0; r < n; ++r) {
      const double x0 = qa[r], x1 = qb[r];
      qa[r] = c * x0 - s * x1;
      qb[r] = s * x0 + c * x1;
    }
  };
  long steps = 0;
  int hi = n - 1;
  while (hi > 0) {
    if (std::fabs(e[hi - 1]) <= eps * (std::fabs(d[hi - 1]) + std::fabs(d[hi]))) {
      e[hi - 1] = 0.0;
      --hi;
      continue;
    }
    int lo = hi - 1;
    while (lo > 0 && std::fabs(e[lo - 1]) > eps * (std::fabs(d[lo - 1]) + std::fabs(d[lo]))) --lo;
    if (lo > 0) e[lo - 1] = 0.0;
    if (++steps > 30L * n) return -1;
    // Wilkinson shift: the eigenvalue of the trailing 2 x 2 block nearer d[hi]
    const double delta = 0.5 * (d[hi - 1] - d[hi]);
    const double eb = e[hi - 1];
    const double den = delta + (delta >= 0.0 ? 1.0 : -1.0) * std::hypot(delta, eb);
    const double shift = d[hi] - eb * eb / den;
    double x = d[lo] - shift, z = e[lo];
    for (int k = lo; k < hi; ++k) {
      // rotation on rows/columns k, k+1 zeroing z against x: c x - s z = r, s x + c z = 0
      const double r = std::hypot(x, z);
      const double c = r == 0.0 ? 1.0 : x / r, s = r == 0.0 ? 0.0 : -z / r;
      if (k > lo) e[k - 1] = r; // the entry above the bulge
      const double dk = d[k], dk1 = d[k + 1], ek = e[k];
      d[k] = c * c * dk - 2.0 * c * s * ek + s * s * dk1;
      d[k + 1] = s * s * dk + 2.0 * c * s * ek + c * c * dk1;
      e[k] = c * s * (dk - dk1) + (c * c - s * s) * ek;
      rotate_q(k, c, s);
      if (k + 1 < hi) {
        x = e[k];
        z = -s * e[k + 1]; // the new bulge T(k, k+2)
        e[k + 1] *= c;
      }
    }
  }
  // ascending order
  std::vector<int> idx(N);
  std::iota(idx.begin(), idx.end(), 0);
  std::stable_sort(idx.begin(), idx.end(), [&](int x, int y) { return d[x] < d[y]; });
  for (int j = 0; j < n; ++j) {
    w[j] = d[idx[j]];
    for (int i = 0; i < n; ++i) Z[size_t(i) * n + j] = qrow(idx[j])[i];
  }
  return 0;
}

// ---------------------------------------------------------------------------
// Block matrix: block (i, j) of a (br x bc)-block matrix at (i*bc + j)*b*b.
struct BlockMat {
  int br = 0, bc = 0, b = 0;
  std::vector<double> a;
  BlockMat() = default;
  BlockMat(int br_, int bc_, int b_) : br(br_), bc(bc_), b(b_), a(size_t(br_) * bc_ * b_ * b_, 0.0) {}
  double* blk(int i, int j) { return a.data() + (size_t(i) * bc + j) * b * b; }
  const double* blk(int i, int j) const { return a.data() + (size_t(i) * bc + j) * b * b; }
};

static void blk_copy(double* dst, const double* src, int b) { std::memcpy(dst, src, sizeof(double) * b * b); }
static void blk_copy_t(double* dst, const double* src, int b)
{
  for (int i = 0; i < b; ++i)
    for (int j = 0; j < b; ++j) dst[i * b + j] = src[j * b + i];
}
static void blk_zero(double* dst, int b) { std::fill(dst, dst + b * b, 0.0); }
static void blk_mult(const double* A, const double* B, double* C, int b) // C = A B
{
  for (int i = 0; i < b; ++i)
    for (int j = 0; j < b; ++j) {
      double s = 0.0;
      for (int k = 0; k < b; ++k) s += A[i * b + k] * B[k * b + j];
      C[i * b + j] = s;
    }
}

// ---------------------------------------------------------------------------
// EVP base (evp_base.hh:15-182) over a CSR operator.
enum Order { LargestMagnitude = 0, Ascending = 1, Descending = 2 };

struct Evp {
  CsrOp A;
  int b;
  int order = LargestMagnitude;
  double tol = 1e-8;
  std::vector<double> vtemp; // Vtemp (evp_base.hh:176)
  std::vector<double> evals;
  BlockMat evecs;
  int last_error = 0; // 1-based Cholesky pivot of the first breakdown

  Evp(const CsrOp& a, int b_) : A(a), b(b_), vtemp(a.n * b_, 0.0) {}

  void apply_op(Panel X, Panel Y) { apply(A, X, Y); }
  void dotp(Panel X, Panel Y, double* R) { dot(X, Y, A.mdiag, R); }

  // evp_base.hh:39-71 (the b = 1 branch :41-46 is the same arithmetic)
  void orthonormalize(Panel V, double* R)
  {
    dotp(V, V, R);
    double U[256], Ui[256];
    const int piv = cholesky_lower_llt(R, b, U);
    if (piv) {
      if (!last_error) last_error = piv;
      return; // the reference throws here; the caller reports it
    }
    invert_upper(U, b, Ui);
    Panel T{vtemp.data(), V.n, b};
    panel_mult(V, Ui, T, 0);
    copy_from(V, T);
    blk_copy(R, U, b);
  }

  // evp_base.hh:79-154
  Index solve_small_dense(const BlockMat& T, const double* beta, Index nev)
  {
    const int nt = T.br * b;
    std::vector<double> dense(size_t(nt) * nt), w(nt), Z(size_t(nt) * nt);
    for (int i = 0; i < T.br; ++i)
      for (int j = 0; j < T.bc; ++j)
        for (int bi = 0; bi < b; ++bi)
          for (int bj = 0; bj < b; ++bj) dense[size_t(i * b + bi) * nt + (j * b + bj)] = T.blk(i, j)[bi * b + bj];
    if (syev(nt, dense.data(), w.data(), Z.data()) != 0) {
      last_error = -1;
      return 0;
    }
    std::vector<int> idx(nt);
    std::iota(idx.begin(), idx.end(), 0);
    if (order == LargestMagnitude) std::stable_sort(idx.begin(), idx.end(), [&](int x, int y) { return std::fabs(w[x]) > std::fabs(w[y]); });
    else if (order == Descending) std::reverse(idx.begin(), idx.end());
    evals.assign(nt, 0.0);
    for (int i = 0; i < nt; ++i) evals[i] = w[idx[i]];
    if (evecs.br != T.br) evecs = BlockMat(T.br, T.bc, b);
    for (int i = 0; i < T.br; ++i)
      for (int j = 0; j < T.bc; ++j)
        for (int bi = 0; bi < b; ++bi)
          for (int bj = 0; bj < b; ++bj) evecs.blk(i, j)[bi * b + bj] = Z[size_t(i * b + bi) * nt + idx[j * b + bj]];
    Index nconv = 0;
    const int last = evecs.br - 1;
    const double eps = std::numeric_limits<double>::epsilon();
    const Index ncheck = std::min<Index>(nev, Index(nt));
    for (Index col = 0; col < ncheck; ++col) {
      const double* vl = evecs.blk(last, int(col / b));
      const int cb = int(col % b);
      double nsq = 0.0;
      for (int i = 0; i < b; ++i) {
        double s = 0.0;
        for (int k = 0; k < b; ++k) s += beta[i * b + k] * vl[k * b + cb];
        nsq += s * s;
      }
      const double rel = std::sqrt(nsq) / std::max(std::fabs(evals[col]), eps);
      if (rel < tol) ++nconv;
    }
    return nconv;
  }
};

// ---------------------------------------------------------------------------
// Driver (lanczos.hh:22-312) with the ModifiedGS strategy
// (reorthogonalization.hh:12-22).
struct Lanczos {
  Evp& evp;
  unsigned nev, ncv, max_restarts;
  int b;
  Index n;
  std::vector<double> V, W; // (ncv + b) columns each, block-major
  BlockMat T, B;
  unsigned iterations = 0, n_op_apply = 0;
  bool converged = false;
  int error = 0; // 0 ok, 1 bad params, 2 Cholesky breakdown, 3 eigensolver failure

  Lanczos(Evp& e, unsigned nev_, unsigned ncv_, unsigned maxr)
      : evp(e), nev(nev_), ncv(ncv_), max_restarts(maxr), b(e.b), n(e.A.n)
  {
    // parameter checks of lanczos.hh:45-58
    if (ncv + unsigned(b) > n || nev >= ncv || ncv / b < nev / b + 2 || ncv % b != 0) {
      error = 1;
      return;
    }
    V.assign((ncv + b) * n, 0.0);
    W.assign((ncv + b) * n, 0.0);
    T = BlockMat(ncv / b, ncv / b, b);
    B = BlockMat(1, 2, b);
  }

  Panel vblk(std::vector<double>& M, Index i) { return Panel{M.data() + i * n * b, n, b}; }

  void reorth(Panel Vn, unsigned count, double* tmp)
  {
    for (unsigned j = 0; j < count; ++j) {
      Panel Vj = vblk(V, j);
      evp.dotp(Vj, Vn, tmp);
      subtract_product(Vn, Vj, tmp);
    }
  }

  unsigned extend(unsigned k, unsigned m)
  {
    unsigned napp = 0;
    Panel W0 = vblk(W, 0);
    double* beta = B.blk(0, 0);
    double* Z0 = B.blk(0, 1);
    if (k == 0) evp.orthonormalize(vblk(V, 0), Z0);
    for (unsigned i = k; i < m; ++i) {
      Panel Vc = vblk(V, i), Vn = vblk(V, i + 1);
      evp.apply_op(Vc, Vn);
      ++napp;
      if (i > 0) {
        panel_mult(vblk(V, i - 1), beta, W0, 2);
        sub_assign(Vn, W0);
      }
      double* Tii = T.blk(int(i), int(i));
      evp.dotp(Vc, Vn, Tii);
      panel_mult(Vc, Tii, W0, 0);
      sub_assign(Vn, W0);
      reorth(Vn, i + 1, Z0);
      evp.orthonormalize(Vn, beta);
      if (int(i) + 1 < T.br) {
        blk_copy(T.blk(int(i) + 1, int(i)), beta, b);
        blk_copy_t(T.blk(int(i), int(i) + 1), beta, b);
      }
    }
    return napp;
  }

  void solve()
  {
    if (error) return;
    unsigned k = 0;
    double* beta = B.blk(0, 0);
    const unsigned mb = ncv / b;
    while (iterations < max_restarts) {
      ++iterations;
      n_op_apply += extend(k, mb);
      if (evp.last_error > 0) {
        error = 2;
        return;
      }
      const Index conv = evp.solve_small_dense(T, beta, nev);
      if (evp.last_error < 0) {
        error = 3;
        return;
      }
      if (conv >= nev) {
        converged = true;
        return;
      }
      const unsigned kr = nev / b + 1;
      if (kr >= mb) break;
      const BlockMat& Y = evp.evecs;
      for (unsigned j = 0; j < kr; ++j) {
        Panel Wj = vblk(W, j);
        set_zero(Wj);
        for (unsigned i = 0; i < mb; ++i) panel_mult(vblk(V, i), Y.blk(int(i), int(j)), Wj, 1);
      }
      copy_from(vblk(W, kr), vblk(V, mb));
      std::swap(V, W);
      for (unsigned i = 0; i < kr; ++i)
        for (unsigned j = 0; j < kr; ++j) {
          double* Tij = T.blk(int(i), int(j));
          blk_zero(Tij, b);
          if (i == j)
            for (int q = 0; q < b; ++q) Tij[q * b + q] = evp.evals[i * b + q];
        }
      for (unsigned i = 0; i < kr; ++i) {
        double* Tki = T.blk(int(kr), int(i));
        blk_mult(beta, Y.blk(int(mb) - 1, int(i)), Tki, b);
        blk_copy_t(T.blk(int(i), int(kr)), Tki, b);
      }
      k = kr;
      Panel Vk = vblk(V, k), Vk1 = vblk(V, k + 1);
      evp.apply_op(Vk, Vk1);
      ++n_op_apply;
      double* Tkk = T.blk(int(k), int(k));
      evp.dotp(Vk, Vk1, Tkk);
      Panel W0 = vblk(W, 0);
      double* Z0 = B.blk(0, 1);
      panel_mult(Vk, Tkk, W0, 0);
      sub_assign(Vk1, W0);
      if (k > 0) {
        panel_mult(vblk(V, k - 1), T.blk(int(k), int(k) - 1), W0, 2);
        sub_assign(Vk1, W0);
      }
      reorth(Vk1, k + 1, Z0);
      evp.orthonormalize(Vk1, beta);
      if (int(k) + 1 < T.br) {
        blk_copy(T.blk(int(k) + 1, int(k)), beta, b);
        blk_copy_t(T.blk(int(k), int(k) + 1), beta, b);
        k += 1;
      }
      else
        break;
    }
  }

  // Ritz vectors X = V[0:m] Y[:, 0:ncols] (not formed by the reference driver)
  void ritz_vectors(unsigned ncols, double* X)
  {
    const unsigned mb = ncv / b;
    const BlockMat& Y = evp.evecs;
    const unsigned nblk = (ncols + b - 1) / b;
    std::vector<double> tmp(n * b);
    for (unsigned j = 0; j < nblk; ++j) {
      Panel Tj{tmp.data(), n, b};
      set_zero(Tj);
      for (unsigned i = 0; i < mb; ++i) panel_mult(vblk(V, i), Y.blk(int(i), int(j)), Tj, 1);
      for (Index r = 0; r < n; ++r)
        for (int c = 0; c < b; ++c) {
          const unsigned col = j * b + c;
          if (col < ncols) X[r * ncols + col] = tmp[r * b + c];
        }
    }
  }
};

} // namespace oracle

// ============================================================================
// C entry points (ctypes).  Panels are double, row-major n x b; b x b blocks
// row-major.
extern "C" {

int oracle_num_threads(void) { return omp_get_max_threads(); }
void oracle_set_threads(int t)
{
  if (t > 0) omp_set_num_threads(t);
}

void oracle_spmm_csr(size_t n, int b, const int32_t* rowptr, const int32_t* col, const double* val, const double* X, double* Y, const double* dscale)
{
  oracle::CsrOp A{n, rowptr, col, val, dscale, nullptr};
  oracle::apply(A, oracle::Panel{const_cast<double*>(X), n, b}, oracle::Panel{Y, n, b});
}

void oracle_panel_mult(size_t n, int b, double* X, const double* W, double* Y, int mode)
{
  oracle::Panel x{X, n, b}, y{Y, n, b};
  if (mode == 3) oracle::subtract_product(x, y, W);
  else oracle::panel_mult(x, W, y, mode);
}

void oracle_panel_sub(size_t n, int b, double* X, const double* Y) { oracle::sub_assign(oracle::Panel{X, n, b}, oracle::Panel{const_cast<double*>(Y), n, b}); }

// C (k*b x b) = V[0:k]^T diag(m) W
void oracle_gram(size_t n, int b, int k, const double* V, const double* W, const double* m, double* C)
{
  for (int j = 0; j < k; ++j)
    oracle::dot(oracle::Panel{const_cast<double*>(V) + size_t(j) * n * b, n, b}, oracle::Panel{const_cast<double*>(W), n, b}, m, C + size_t(j) * b * b);
}

// W -= sum_j V_j C_j
void oracle_update(size_t n, int b, int k, const double* V, const double* C, double* W)
{
  for (int j = 0; j < k; ++j)
    oracle::subtract_product(oracle::Panel{W, n, b}, oracle::Panel{const_cast<double*>(V) + size_t(j) * n * b, n, b}, C + size_t(j) * b * b);
}

// CholQR (evp_base.hh:39-71); returns 0 or the 1-based failing pivot
int oracle_cholqr(size_t n, int b, double* X, const double* m, double* R)
{
  std::vector<double> dummy_val;
  oracle::CsrOp A{n, nullptr, nullptr, nullptr, nullptr, m};
  oracle::Evp evp(A, b);
  evp.orthonormalize(oracle::Panel{X, n, b}, R);
  return evp.last_error;
}

int oracle_cholesky(int b, const double* G, double* R) { return oracle::cholesky_lower_llt(G, b, R); }

int oracle_syev(int n, const double* A, double* w, double* Z) { return oracle::syev(n, A, w, Z); }

// Ritz rotation as the driver issues it (lanczos.hh:105-114): Out_j = sum_i V_i Y(i, j)
void oracle_rotate(size_t n, int b, int m, const double* V, const double* Y, int ybc, int kc, double* Out)
{
  for (int j = 0; j < kc; ++j) {
    oracle::Panel Oj{Out + size_t(j) * n * b, n, b};
    oracle::set_zero(Oj);
    for (int i = 0; i < m; ++i)
      oracle::panel_mult(oracle::Panel{const_cast<double*>(V) + size_t(i) * n * b, n, b}, Y + (size_t(i) * ybc + j) * b * b, Oj, 1);
  }
}

// Full thick-restart solve.  info[0] converged, [1] iterations, [2] n_op_apply,
// [3] error (0 ok, 1 bad params, 2 Cholesky breakdown, 3 eigensolver failure).
// evals_out: nev Ritz values in the chosen order; X_out (optional): n x nev
// row-major Ritz vectors.
// seconds (may be null): wall time of solve() alone, measured around the
// call as the reference's example does (examples/openmp/csrtest.cc:100-103),
// without the construction and zero fill of the basis.
int oracle_lanczos_solve_timed(size_t n, const int32_t* rowptr, const int32_t* col, const double* val, const double* dscale, const double* mdiag, int b,
                               unsigned nev, unsigned ncv, unsigned max_restarts, double tol, int order, const double* V0, double* evals_out,
                               double* X_out, int* info, double* seconds)
{
  oracle::CsrOp A{n, rowptr, col, val, dscale, mdiag};
  oracle::Evp evp(A, b);
  evp.order = order;
  evp.tol = tol;
  oracle::Lanczos lz(evp, nev, ncv, max_restarts);
  if (lz.error) {
    info[0] = 0;
    info[1] = info[2] = 0;
    info[3] = lz.error;
    return lz.error;
  }
  std::memcpy(lz.V.data(), V0, sizeof(double) * n * b);
  const auto t0 = std::chrono::steady_clock::now();
  lz.solve();
  if (seconds) *seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  info[0] = lz.converged ? 1 : 0;
  info[1] = int(lz.iterations);
  info[2] = int(lz.n_op_apply);
  info[3] = lz.error;
  if (lz.error) return lz.error;
  for (unsigned i = 0; i < nev && i < evp.evals.size(); ++i) evals_out[i] = evp.evals[i];
  if (X_out) lz.ritz_vectors(nev, X_out);
  return 0;
}

int oracle_lanczos_solve(size_t n, const int32_t* rowptr, const int32_t* col, const double* val, const double* dscale, const double* mdiag, int b,
                         unsigned nev, unsigned ncv, unsigned max_restarts, double tol, int order, const double* V0, double* evals_out, double* X_out,
                         int* info)
{
  return oracle_lanczos_solve_timed(n, rowptr, col, val, dscale, mdiag, b, nev, ncv, max_restarts, tol, order, V0, evals_out, X_out, info, nullptr);
}

// extend(0, m) only (tests/test_lanczos_extend.hh): V_out n x (ncv + b)
// block-major, T_out (ncv x ncv) block layout, beta_out b x b.
int oracle_lanczos_extend(size_t n, const int32_t* rowptr, const int32_t* col, const double* val, const double* dscale, const double* mdiag, int b,
                          unsigned nev, unsigned ncv, const double* V0, double* V_out, double* T_out, double* beta_out)
{
  oracle::CsrOp A{n, rowptr, col, val, dscale, mdiag};
  oracle::Evp evp(A, b);
  oracle::Lanczos lz(evp, nev, ncv, 1);
  if (lz.error) return lz.error;
  std::memcpy(lz.V.data(), V0, sizeof(double) * n * b);
  lz.extend(0, ncv / b);
  if (evp.last_error) return 2;
  std::memcpy(V_out, lz.V.data(), sizeof(double) * lz.V.size());
  std::memcpy(T_out, lz.T.a.data(), sizeof(double) * lz.T.a.size());
  std::memcpy(beta_out, lz.B.blk(0, 0), sizeof(double) * b * b);
  return 0;
}

// Timed extend(0, steps) of an FP64 thick-restart cycle (the CPU baseline
// sample); returns seconds.
double oracle_time_extend(size_t n, const int32_t* rowptr, const int32_t* col, const double* val, int b, unsigned nev, unsigned ncv, const double* V0,
                          unsigned steps)
{
  oracle::CsrOp A{n, rowptr, col, val, nullptr, nullptr};
  oracle::Evp evp(A, b);
  evp.order = oracle::Ascending;
  oracle::Lanczos lz(evp, nev, ncv, 1);
  if (lz.error || steps < 1 || steps > ncv / unsigned(b)) return -1.0;
  std::memcpy(lz.V.data(), V0, sizeof(double) * n * b);
  const double t0 = omp_get_wtime();
  lz.extend(0, steps);
  return omp_get_wtime() - t0;
}

} // extern "C"
