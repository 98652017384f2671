// libtrl_hip_solver.so: whole-solve C-ABI (include/trl_hip_solver.h) over the
// trl::hip backend types and the trl_amd::BlockLanczos driver.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstdint>
#include <cstring>
#include <memory>
#include <random>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "trl/hip/csrevp.hh"
#include "trl_amd/dense.hh"
#include "trl_amd/lanczos.hh"
#include "trl_hip_solver.h"

namespace {

using trl::hip::Context;
using trl::hip::SparseOperator;
using trl::hip::throw_on;

int code_of(const std::exception& e)
{
  if (dynamic_cast<const std::invalid_argument*>(&e)) return TRL_HIP_EINVAL;
  if (dynamic_cast<const std::out_of_range*>(&e)) return TRL_HIP_ERANGE;
  const std::string w = e.what();
  if (w.find("Cholesky") != std::string::npos) return TRL_HIP_ECHOL;
  if (w.find("Eigendecomposition") != std::string::npos) return TRL_HIP_EEIG;
  if (w.find("allocation") != std::string::npos) return TRL_HIP_ENOMEM;
  if (w.find(trl_hip_strerror(TRL_HIP_ENOCONV)) != std::string::npos) return TRL_HIP_ENOCONV;
  if (w.find(trl_hip_strerror(TRL_HIP_EINDEF)) != std::string::npos) return TRL_HIP_EINDEF;
  return TRL_HIP_EINVAL;
}

double now()
{
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// True residuals ||K x - lambda M x|| / ||K x|| of nev Ritz vectors stored as
// ceil(nev/B) panels at X (device), computed with the undivided operator K.
// scratch: one n x B panel for K X (allocated here when null).
template <class S, unsigned B>
void true_residuals(const Context& c, std::size_t n, const SparseOperator<S>& op, const S* mass, const S* X, unsigned nev, const double* evals, double* out,
                    S* scratch = nullptr)
{
  constexpr trl_hip_dtype t = trl::hip::dtype_of<S>();
  trl::hip::DeviceBuffer<S> own;
  if (!scratch) {
    own = trl::hip::DeviceBuffer<S>(c, n * B);
    scratch = own.data();
  }
  const unsigned nblk = (nev + B - 1) / B;
  for (unsigned j = 0; j < nblk; ++j) {
    const S* Xj = X + std::size_t(j) * n * B;
    op.apply(int(B), Xj, scratch, false);
    double lam[16] = {0.0};
    for (unsigned q = 0; q < B && j * B + q < nev; ++q) lam[q] = evals[j * B + q];
    double sums[32];
    throw_on(trl_hip_residual_sq(c.get(), t, n, B, Xj, scratch, mass, lam, sums), "residual_sq");
    for (unsigned q = 0; q < B && j * B + q < nev; ++q) out[j * B + q] = std::sqrt(sums[q]) / std::max(std::sqrt(sums[B + q]), 1e-300);
  }
}

// The same residuals from products A x already formed (FP64 panels X and
// AX, ceil(nev/B) of them used), standard problems only.
template <unsigned B>
void residuals_from_products(const Context& c, std::size_t n, const double* X, const double* AX, unsigned nev, const double* evals, double* out)
{
  const unsigned nblk = (nev + B - 1) / B;
  for (unsigned j = 0; j < nblk; ++j) {
    double lam[16] = {0.0};
    for (unsigned q = 0; q < B && j * B + q < nev; ++q) lam[q] = evals[j * B + q];
    double sums[32];
    throw_on(trl_hip_residual_sq(c.get(), TRL_HIP_F64, n, B, X + std::size_t(j) * n * B, AX + std::size_t(j) * n * B, nullptr, lam, sums), "residual_sq");
    for (unsigned q = 0; q < B && j * B + q < nev; ++q) out[j * B + q] = std::sqrt(sums[q]) / std::max(std::sqrt(sums[B + q]), 1e-300);
  }
}

// The operator of one solve: the CSR arrays with the caller's kernel
// variant, or the fastest kernel and storage for block size B when tuning
// is asked for (the fused axpby form too unless axpby is false).
template <class S, unsigned B>
std::shared_ptr<SparseOperator<S>> make_operator(const Context& c, std::size_t n, const int32_t* rowptr, const int32_t* col, const S* val, const S* minv,
                                                 const trl_hip_solver_params& p, bool axpby = true, bool tune = true)
{
  auto op = std::make_shared<SparseOperator<S>>(c, n, rowptr, col, val, minv, p.spmm_tune ? 0 : p.lanes_per_row);
  const auto t0 = std::chrono::steady_clock::now();
  if (p.spmm_tune && tune) op->tune(int(B), true, 0.25, axpby);
  const double tune_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  if (std::getenv("TRL_HIP_VERBOSE")) {
    const auto& d = op->descriptor();
    const auto& x = op->axpby_descriptor();
    std::fprintf(stderr, "[operator] %s b=%u %s/v%d/plane%d:%d (%.1f us), axpby %s/v%d/plane%d:%d (%.1f us); tuned in %.1f ms\n",
                 sizeof(S) == 4 ? "f32" : "f64", B, d.format ? "ell" : "csr", d.variant, d.plane, d.plane_block, op->tuned_us(), x.format ? "ell" : "csr",
                 x.variant, x.plane, x.plane_block, op->axpby_us(), tune_ms);
  }
  return op;
}

// Tunes the filter step of an operator on the nref panels of a refinement.
template <class S>
void tune_panels_of(SparseOperator<S>& op, unsigned B, unsigned nref)
{
  const auto t0 = std::chrono::steady_clock::now();
  op.tune_panels(int(B), int(nref));
  if (std::getenv("TRL_HIP_VERBOSE")) {
    const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    const auto& d = op.panels_descriptor();
    std::fprintf(stderr, "[operator] %s %u panels %s/v%d/plane%d:%d (%.1f us per launch, %.1f us per panel; one panel %.1f us)%s; tuned in %.1f ms\n",
                 sizeof(S) == 4 ? "f32" : "f64", nref, d.format ? "ell" : "csr", d.variant, d.plane, d.plane_block, op.panels_us(), op.panels_us() / nref,
                 op.axpby_us(), op.panels_pay(int(nref)) ? " used" : "", ms);
  }
}

// Runs one solve with panel type S and block size B.  If X_out is given it
// receives the first nblk_out Ritz blocks.
template <class S, unsigned B>
void solve_typed(const Context& c, std::size_t n, const int32_t* rowptr, const int32_t* col, const S* val, const S* mass, const S* minv,
                 const std::shared_ptr<SparseOperator<S>>& op, const trl_hip_solver_params& p, const S* V0, double* evals, S* X_out, unsigned nblk_out,
                 double* resid, trl_hip_solver_result& res, std::vector<double>* all_evals = nullptr, unsigned* nblk_eff = nullptr)
{
  using EVP = trl::hip::CSREVP<S, B>;
  auto evp = std::make_shared<EVP>(c, n, rowptr, col, val, mass, minv);
  evp->set_order(static_cast<trl::hip::EigenvalueOrder>(p.order));
  evp->set_operator(op);
  if (p.cholqr_passes > 0) evp->set_cholqr_passes(p.cholqr_passes);
  if (X_out) evp->set_ritz_blocks(nblk_out); // the mixed solve keeps more Ritz blocks than the driver needs
  evp->set_dense_rayleigh_ritz(p.rr_dense != 0);
  trl_amd::EigensolverParams ep{p.nev, p.ncv, p.max_restarts, p.tol};

  auto run = [&](auto& lz) {
    auto V0b = lz.initial_block();
    throw_on(trl_hip_memcpy_d2d(c.get(), V0b.data_, V0, n * B * sizeof(S)), "copy V0");
    c.sync();
    const double t0 = now();
    const auto r = lz.solve();
    c.sync();
    res.seconds = now() - t0;
    res.converged = r.converged ? 1 : 0;
    res.iterations = r.iterations;
    res.n_op_apply = r.n_op_apply;
    res.host_seconds = evp->host_seconds();
    res.rr_steps = evp->rr_steps();
    res.rr_structured = evp->rr_structured();
    res.graph_cycles = lz.graph_cycles();
    const auto& ev = evp->get_current_eigenvalues();
    for (unsigned i = 0; i < p.nev && i < ev.size(); ++i) evals[i] = ev[i];
    if (all_evals) *all_evals = ev;
    if (X_out) {
      const unsigned kr = p.nev / B + 1;
      if (nblk_eff) *nblk_eff = nblk_out;
      if (lz.basis_holds_ritz_blocks()) {
        // after a final restart only the leading kr blocks are Ritz vectors
        const unsigned cnt = std::min(nblk_out, kr);
        if (nblk_eff) *nblk_eff = cnt;
        throw_on(trl_hip_memcpy_d2d(c.get(), X_out, lz.get_basis().data(), std::size_t(cnt) * n * B * sizeof(S)), "ritz vectors");
      }
      else {
        const auto& Y = evp->get_current_eigenvectors();
        throw_on(trl_hip_rotate(c.get(), EVP::dtype, n, B, int(p.ncv / B), lz.get_basis().data(), Y.data(), int(Y.block_cols()), int(nblk_out), X_out),
                 "ritz vectors");
      }
      if (resid) true_residuals<S, B>(c, n, *op, mass, X_out, p.nev, evals, resid);
      c.sync();
    }
  };
  if (p.reorth == 1) {
    trl_amd::BlockLanczos<EVP, trl::hip::BlockCGS2> lz(evp, ep);
    run(lz);
  }
  else if (p.reorth == 2) {
    trl_amd::BlockLanczos<EVP, trl::hip::BlockCGSAdaptive> lz(evp, ep);
    run(lz);
  }
  else {
    trl_amd::BlockLanczos<EVP, trl_amd::ModifiedGS> lz(evp, ep);
    run(lz);
  }
}

// Block refinement of nb Ritz panels (panel type S, FP64 coefficients)
// toward one end of the spectrum:
// block CGS2 + CholQR2 orthonormalisation, Rayleigh-Ritz over the nb*B
// vectors, and a scaled Chebyshev filter (Zhou & Saad, "Self-consistent-field
// calculations using Chebyshev-filtered subspace iteration", J. Comput. Phys.
// 226 (2007), Algorithm 3.2) that keeps the value 1 at the far wanted end a0
// and damps [lo, hi].
template <class S, unsigned B>
struct BlockRefiner {
  static constexpr trl_hip_dtype t = trl::hip::dtype_of<S>();
  const Context& c;
  std::size_t n;
  const SparseOperator<S>& op;
  const S* mass;
  unsigned nb;
  bool ascending;
  double eta; // second-pass test of the block Gram-Schmidt (<= 0: always two passes)
  // KX: the true-residual check's K X panel, kept across sweeps
  trl::hip::DeviceBuffer<S> Xa, Xb, AX, KX;
  trl::hip::DeviceBuffer<double> coef;
  S* X; // current Ritz panels (Xa or Xb)
  S* spare;

  BlockRefiner(const Context& ctx, std::size_t n_, const SparseOperator<S>& o, const S* m, unsigned nb_, bool asc, double eta_)
      : c(ctx), n(n_), op(o), mass(m), nb(nb_), ascending(asc), eta(eta_), Xa(ctx, std::size_t(nb_) * n_ * B),
        Xb(ctx, std::size_t(nb_) * n_ * B), AX(ctx, std::size_t(nb_) * n_ * B), KX(ctx, n_ * B), coef(ctx, std::size_t(nb_) * nb_ * B * B + 2 * B * B)
  {
    X = Xa.data();
    spare = Xb.data();
  }

  S* panel(S* base, unsigned j) const { return base + std::size_t(j) * n * B; }

  void orthonormalize()
  {
    double* C = coef.data();
    double* R = coef.data() + std::size_t(nb) * nb * B * B;
    for (unsigned j = 0; j < nb; ++j) {
      if (j > 0) throw_on(trl_hip_reorth(c.get(), t, n, B, int(j), X, panel(X, j), mass, C, eta), "refine reorth");
      throw_on(trl_hip_cholqr(c.get(), t, n, B, panel(X, j), mass, R, 2), "refine cholqr");
    }
    c.sync(); // raises a breakdown
  }

  // Rayleigh-Ritz; theta receives the nb*B Ritz values in the wanted order.
  void rayleigh_ritz(std::vector<double>& theta)
  {
    const unsigned nt = nb * B;
    for (unsigned j = 0; j < nb; ++j) op.apply(int(B), panel(X, j), panel(AX.data(), j));
    // H = X^T A X is symmetric: only its blocks (i, j) with i <= j are
    // formed, block column j from X_0..X_j and A X_j ((j + 2) panels read
    // instead of nb + 1)
    double* C = coef.data();
    for (unsigned j = 0; j < nb; ++j)
      throw_on(trl_hip_gram(c.get(), t, n, B, int(j + 1), X, panel(AX.data(), j), mass, C + std::size_t(j) * nt * B), "refine projection");
    std::vector<double> Cj(std::size_t(nt) * nt);
    throw_on(trl_hip_memcpy_d2h(c.get(), Cj.data(), C, Cj.size() * sizeof(double)), "refine download");
    // C + j*nt*B holds rows r = 0..(j+1)*B-1 (panel i = r/B, column r%B) of
    // H[:, j*B:(j+1)*B]; the rows below are mirrored from the later columns
    std::vector<double> H(std::size_t(nt) * nt), w(nt), Z(std::size_t(nt) * nt);
    for (unsigned j = 0; j < nb; ++j)
      for (unsigned r = 0; r < (j + 1) * B; ++r)
        for (unsigned q = 0; q < B; ++q) H[std::size_t(r) * nt + j * B + q] = Cj[std::size_t(j) * nt * B + std::size_t(r) * B + q];
    for (unsigned j = 0; j < nb; ++j)
      for (unsigned r = (j + 1) * B; r < nt; ++r)
        for (unsigned q = 0; q < B; ++q) H[std::size_t(r) * nt + j * B + q] = H[std::size_t(j * B + q) * nt + r];
    for (unsigned r = 0; r < nt; ++r)
      for (unsigned q = r + 1; q < nt; ++q) {
        const double a = 0.5 * (H[std::size_t(r) * nt + q] + H[std::size_t(q) * nt + r]);
        H[std::size_t(r) * nt + q] = H[std::size_t(q) * nt + r] = a;
      }
    if (trl_amd::dense::syev(int(nt), H.data(), w.data(), Z.data()) != 0) throw std::runtime_error("Eigendecomposition failed");
    std::vector<unsigned> idx(nt);
    for (unsigned i = 0; i < nt; ++i) idx[i] = ascending ? i : nt - 1 - i;
    theta.resize(nt);
    std::vector<double> Y(std::size_t(nt) * nt);
    for (unsigned i = 0; i < nt; ++i) theta[i] = w[idx[i]];
    for (unsigned bi = 0; bi < nb; ++bi)
      for (unsigned bj = 0; bj < nb; ++bj)
        for (unsigned ci = 0; ci < B; ++ci)
          for (unsigned cj = 0; cj < B; ++cj) Y[((std::size_t(bi) * nb + bj) * B + ci) * B + cj] = Z[std::size_t(bi * B + ci) * nt + idx[bj * B + cj]];
    throw_on(trl_hip_memcpy_h2d(c.get(), C, Y.data(), Y.size() * sizeof(double)), "refine upload");
    throw_on(trl_hip_rotate(c.get(), t, n, B, int(nb), X, C, int(nb), int(nb), spare), "refine rotate");
    std::swap(X, spare);
  }

  // Chebyshev filter of degree deg[j] on panel j (0 leaves the panel as it
  // is); damps [lo, hi], value 1 at a0.  Returns the operator applications.
  std::size_t filter(const std::vector<unsigned>& deg, double lo, double hi, double a0)
  {
    const double e = 0.5 * (hi - lo), cc = 0.5 * (hi + lo);
    const double s1 = e / (a0 - cc);
    bool uniform = true;
    for (unsigned d : deg) uniform = uniform && d == deg.front();
    std::size_t applies = 0;
    // otherwise panel by panel
    bool suffix = true; // even, non-decreasing degrees: the blocks still filtered form a suffix
    for (unsigned j = 0; j < nb; ++j) suffix = suffix && deg[j] % 2 == 0 && (j == 0 || deg[j] >= deg[j - 1]);
    if ((uniform || suffix) && deg.back() > 0 && op.panels_pay(int(nb))) {
      // all blocks still being filtered in one launch per step: the matrix
      // is read once per step for all of them
      const unsigned degree = deg.back();
      std::size_t applies = 0;
      double sg = s1;
      for (unsigned i = 1; i <= degree; ++i) {
        unsigned j0 = 0;
        while (deg[j0] < i) ++j0; // deg.back() >= i
        const int np = int(nb - j0);
        // step i reads the older iterate from the buffer written two steps
        // ago: X for odd i (X and spare alternate, X holds step 0)
        S* from = panel(i % 2 == 1 ? X : spare, j0);
        S* to = panel(i % 2 == 1 ? spare : X, j0);
        if (i == 1)
          op.axpby_panels(int(B), np, from, to, nullptr, s1 / e, -s1 * cc / e, 0.0);
        else {
          const double sn = 1.0 / (2.0 / s1 - sg);
          op.axpby_panels(int(B), np, from, to, to, 2.0 * sn / e, -2.0 * sn * cc / e, -sg * sn);
          sg = sn;
        }
        applies += std::size_t(np);
      }
      if (uniform && degree % 2 == 1) std::swap(X, spare);
      return applies;
    }
    for (unsigned j = 0; j < nb; ++j) {
      const unsigned degree = deg[j];
      if (degree == 0) continue;
      S* prev = panel(X, j);
      S* cur = panel(spare, j);
      op.axpby(int(B), prev, cur, nullptr, s1 / e, -s1 * cc / e, 0.0);
      double sg = s1;
      for (unsigned i = 2; i <= degree; ++i) {
        const double sn = 1.0 / (2.0 / s1 - sg);
        op.axpby(int(B), cur, prev, prev, 2.0 * sn / e, -2.0 * sn * cc / e, -sg * sn);
        std::swap(prev, cur);
        sg = sn;
      }
      applies += degree;
      // an odd degree leaves the result in the spare panel
      if (degree % 2 == 1 && !uniform)
        throw_on(trl_hip_memcpy_d2d(c.get(), panel(X, j), panel(spare, j), n * B * sizeof(S)), "filter copy");
    }
    if (uniform && deg.front() % 2 == 1) std::swap(X, spare);
    return applies;
  }
};

// The FP64 sweeps' Chebyshev filter run in FP32 panels.  For a Ritz pair
// (theta, x) of the FP64 refiner, r = A x - theta x, and the filter's
// three-term recurrence Y_i = alpha_i A Y_{i-1} + beta_i Y_{i-1} + gamma_i
// Y_{i-2} (Y_0 = x) gives for E_i = Y_i - p_i(theta) x
//   E_i = alpha_i A E_{i-1} + beta_i E_{i-1} + gamma_i E_{i-2} + alpha_i p_{i-1}(theta) r,
// E_0 = 0, E_1 = alpha_1 r, with p_i the filter polynomial (the same
// recurrence on scalars).  |E_i| stays of the order of the error of x (the
// components p_i(lambda) - p_i(theta) are bounded by the scaled filter and
// multiply small coefficients), so FP32 rounding of E and of the FP32
// operator contributes orders of magnitude below the error being removed;
// r and x <- p_d(theta) x + E_d are formed in FP64.
template <unsigned B>
struct ResidualFilter {
  const Context& c;
  std::size_t n;
  unsigned nb;
  const SparseOperator<float>& op32;
  trl::hip::DeviceBuffer<float> R, Ea, Eb;
  trl::hip::DeviceBuffer<double> AX, table;
  unsigned max_degree;

  // max_degree: the largest filter degree of any sweep (sizes the device
  // coefficient table once)
  ResidualFilter(const Context& ctx, std::size_t n_, unsigned nb_, const SparseOperator<float>& o, unsigned max_degree_)
      : c(ctx), n(n_), nb(nb_), op32(o), R(ctx, std::size_t(nb_) * n_ * B), Ea(ctx, std::size_t(nb_) * n_ * B), Eb(ctx, std::size_t(nb_) * n_ * B),
        AX(ctx, std::size_t(nb_) * n_ * B), table(ctx, (std::size_t(max_degree_) + 2) * nb_ * B), max_degree(max_degree_)
  {
  }

  template <class T>
  static T* panel(T* base, std::size_t n, unsigned j) { return base + std::size_t(j) * n * B; }

  // Filters the Ritz panels rf.X (Ritz values theta, in panel order) with
  // degree deg[j] on panel j; the degrees are non-decreasing.  Returns the
  // operator applications (FP64 and FP32).
  // The products A x of all panels, for the true-residual check and the
  // filter's residuals alike (standard problems: no row scaling, so the
  // filter's operator and the check's are the same); filter(..., true) then
  // takes them instead of forming its own.
  void form_ax(const BlockRefiner<double, B>& rf, const SparseOperator<double>& op64)
  {
    for (unsigned j = 0; j < nb; ++j) op64.apply(int(B), rf.panel(rf.X, j), panel(AX.data(), n, j));
  }

  std::size_t filter(BlockRefiner<double, B>& rf, const SparseOperator<double>& op64, const std::vector<unsigned>& deg, const std::vector<double>& theta,
                     double lo, double hi, double a0, bool ax_formed = false)
  {
    const double e = 0.5 * (hi - lo), cc = 0.5 * (hi + lo);
    const double s1 = e / (a0 - cc);
    const unsigned dmax = deg.back();
    unsigned j1 = 0; // first panel filtered at all
    while (j1 < nb && deg[j1] == 0) ++j1;
    if (dmax == 0 || j1 == nb) return 0;
    for (unsigned j = 1; j < nb; ++j)
      if (deg[j] < deg[j - 1]) throw std::invalid_argument("residual filter: degrees must be non-decreasing");
    // step coefficients of the scaled recurrence (alpha, beta, gamma)
    std::vector<double> al(dmax + 1), be(dmax + 1), ga(dmax + 1, 0.0);
    al[1] = s1 / e;
    be[1] = -s1 * cc / e;
    double sg = s1;
    for (unsigned i = 2; i <= dmax; ++i) {
      const double sn = 1.0 / (2.0 / s1 - sg);
      al[i] = 2.0 * sn / e;
      be[i] = -2.0 * sn * cc / e;
      ga[i] = -sg * sn;
      sg = sn;
    }
    // device table: theta (nb*B), forcing coefficients of steps 2..dmax
    // (nb*B each) and the final p_d(theta) per column (nb*B)
    const std::size_t nbB = std::size_t(nb) * B;
    std::vector<double> host((dmax + 2) * nbB, 0.0);
    std::copy(theta.begin(), theta.begin() + nbB, host.begin());
    for (std::size_t k = 0; k < nbB; ++k) {
      const unsigned dj = deg[k / B];
      double pm = 1.0, pc = al[1] * theta[k] + be[1]; // p_0, p_1
      for (unsigned i = 2; i <= dj; ++i) {
        host[std::size_t(i) * nbB + k] = al[i] * pc; // alpha_i p_{i-1}(theta)
        const double pn = (al[i] * theta[k] + be[i]) * pc + ga[i] * pm;
        pm = pc;
        pc = pn;
      }
      host[std::size_t(dmax + 1) * nbB + k] = dj == 0 ? 1.0 : pc;
    }
    if (dmax > max_degree) throw std::invalid_argument("residual filter: degree above the table size");
    throw_on(trl_hip_memcpy_h2d(c.get(), table.data(), host.data(), host.size() * sizeof(double)), "residual filter table");
    const unsigned np1 = nb - j1;
    // r = A x - theta x in FP64 for the filtered panels; R = r, E_1 = alpha_1 r in FP32
    if (!ax_formed)
      for (unsigned j = j1; j < nb; ++j) op64.apply(int(B), rf.panel(rf.X, j), panel(AX.data(), n, j));
    throw_on(trl_hip_residual_split(c.get(), n, int(B), int(np1), panel(AX.data(), n, j1), rf.panel(rf.X, j1), table.data() + std::size_t(j1) * B, al[1],
                                    panel(R.data(), n, j1), panel(Eb.data(), n, j1)),
             "residual split");
    std::size_t applies = ax_formed ? 0 : np1;
    // steps 2..dmax on the panels still being filtered (a suffix): step i
    // writes Ea for even i, Eb for odd i, and reads the other as E_{i-1};
    // the older iterate E_{i-2} is the output buffer (E_0 = 0: none at i = 2)
    for (unsigned i = 2; i <= dmax; ++i) {
      unsigned j0 = j1;
      while (deg[j0] < i) ++j0;
      const int np = int(nb - j0);
      float* from = panel(i % 2 == 1 ? Ea.data() : Eb.data(), n, j0);
      float* to = panel(i % 2 == 1 ? Eb.data() : Ea.data(), n, j0);
      op32.axpby_forced(int(B), np, from, to, i == 2 ? nullptr : to, al[i], be[i], ga[i], panel(R.data(), n, j0),
                        table.data() + std::size_t(i) * nbB + std::size_t(j0) * B);
      applies += std::size_t(np);
    }
    // x <- p_d(theta) x + E_d, E_d in Ea for even d, Eb for odd d
    const double* h = table.data() + std::size_t(dmax + 1) * nbB;
    for (unsigned j = j1; j < nb;) {
      unsigned k = j;
      while (k < nb && deg[k] % 2 == deg[j] % 2) ++k;
      throw_on(trl_hip_panel_combine(c.get(), n, int(B), int(k - j), rf.panel(rf.X, j), h + std::size_t(j) * B,
                                     panel(deg[j] % 2 == 0 ? Ea.data() : Eb.data(), n, j)),
               "residual combine");
      j = k;
    }
    return applies;
  }
};

// Per-panel filter degrees for the next sweep (degree optimisation as in
// ChASE: Winkelmann, Springer, Di Napoli, "ChASE: Chebyshev accelerated
// subspace iteration eigensolver for sequences of Hermitian eigenvalue
// problems", ACM TOMS 45(2), 2019, Sec. 4.2: each vector gets the degree
// that brings its residual to the tolerance at its convergence rate).  The
// rate of wanted vector i is measured: log(r_prev / r) over the degree its
// panel had in the previous sweep.  A vector already at tol needs 0; one
// without a measured rate needs the full degree.  Panels holding vectors
// beyond nev (the buffer that keeps the filter cutoff away from the wanted
// values) take the largest wanted degree, and the degrees are made
// non-decreasing over the blocks.  The degrees aim at tol / 4, so
// that a slightly optimistic rate still ends below tol.  Degrees are rounded
// up to even.
inline std::vector<unsigned> panel_degrees(unsigned nb, unsigned B, unsigned nev, double tol, unsigned degree, const std::vector<double>& rs,
                                           const std::vector<double>& rs_prev, const std::vector<unsigned>& deg_prev)
{
  std::vector<unsigned> deg(nb, degree);
  if (rs_prev.size() != rs.size() || deg_prev.size() != nb) return deg;
  unsigned dmax = 0;
  for (unsigned j = 0; j < nb; ++j) {
    if ((j + 1) * B > nev) continue;
    unsigned dj = 0;
    for (unsigned i = j * B; i < (j + 1) * B; ++i) {
      unsigned m = degree;
      if (rs[i] <= tol)
        m = 0;
      else if (deg_prev[j] > 0 && rs_prev[i] > 1.5 * rs[i]) {
        const double rate = std::log(rs_prev[i] / rs[i]) / double(deg_prev[j]);
        const double need = std::log(4.0 * rs[i] / tol) / rate;
        m = need < double(degree) ? unsigned(std::ceil(need)) : degree;
      }
      dj = std::max(dj, m);
    }
    deg[j] = std::min(degree, dj + (dj % 2));
    dmax = std::max(dmax, deg[j]);
  }
  for (unsigned j = 0; j < nb; ++j)
    if ((j + 1) * B > nev) deg[j] = dmax;
  // non-decreasing over the blocks (the Ritz values ascend with the block
  // index, and so, mostly, do the degrees they need): the blocks still being
  // filtered at any step are then a suffix, stepped by one launch
  for (unsigned j = 1; j < nb; ++j) deg[j] = std::max(deg[j], deg[j - 1]);
  return deg;
}

// Refinement sweeps: orthonormalise, Rayleigh-Ritz, true residuals of the
// first nev pairs; stop when all are <= tol or after max_sweeps filters.
// degree_opt: per-panel degrees (panel_degrees) instead of `degree` on every
// panel from the second filter on.  Returns true on convergence; theta and
// rs hold the last values, applies counts the filter's operator applications.
// rfilt (FP64 sweeps only, may be null): filter in FP32 through it.
// ramp: the first `ramp` filters run at degree / 2^ramp, ..., degree / 2
// (at most): the cutoff of the first sweeps, the last kept Ritz value of a
// rough start, sits far above the eigenvalue it stands for, so a full-degree
// filter would spend its steps amplifying unwanted eigenvalues below it.
template <class S, unsigned B>
bool refine_sweeps(BlockRefiner<S, B>& rf, const Context& c, std::size_t n, const SparseOperator<S>& op, const S* mass, unsigned nev, double tol,
                   unsigned max_sweeps, unsigned degree, bool degree_opt, double lo_all, double hi_all, std::vector<double>& theta,
                   std::vector<double>& rs, unsigned& sweeps, std::size_t& applies, const char* tag, ResidualFilter<B>* rfilt = nullptr,
                   unsigned ramp = 0)
{
  const double margin = 0.01 * (hi_all - lo_all);
  rs.assign(nev, 0.0);
  sweeps = 0;
  applies = 0;
  std::vector<double> rs_prev;
  std::vector<unsigned> deg_prev;
  for (;;) {
    rf.orthonormalize();
    rf.rayleigh_ritz(theta);
    bool ax_formed = false;
    if constexpr (std::is_same_v<S, double>) {
      if (rfilt && !mass && !op.descriptor().dscale) {
        // one A x per panel serves the true residuals and the filter's
        // residual split (which would otherwise form them again)
        rfilt->form_ax(rf, op);
        residuals_from_products<B>(c, n, rf.X, rfilt->AX.data(), nev, theta.data(), rs.data());
        ax_formed = true;
      }
    }
    if (!ax_formed) true_residuals<S, B>(c, n, op, mass, rf.X, nev, theta.data(), rs.data(), rf.KX.data());
    double worst = 0.0;
    for (double r : rs) worst = std::max(worst, r);
    if (std::getenv("TRL_HIP_VERBOSE")) {
      std::vector<double> srt(rs);
      std::sort(srt.begin(), srt.end());
      std::fprintf(stderr, "[refine %s] sweep %u: resid max %.3e median %.3e min %.3e theta0 %.9g theta_kept %.9g\n", tag, sweeps, worst,
                   srt[srt.size() / 2], srt.front(), theta.front(), theta.back());
    }
    if (worst <= tol) return true;
    if (sweeps >= max_sweeps) return false;
    ++sweeps;
    std::vector<unsigned> deg = degree_opt ? panel_degrees(rf.nb, B, nev, tol, degree, rs, rs_prev, deg_prev) : std::vector<unsigned>(rf.nb, degree);
    if (sweeps <= ramp) {
      // even, at least 2: the all-blocks launch steps them in pairs
      const unsigned cap = std::max(2u, ((degree >> (ramp + 1 - sweeps)) + 1) & ~1u);
      for (unsigned& d : deg) d = std::min(d, cap);
    }
    if (std::getenv("TRL_HIP_VERBOSE")) {
      std::fprintf(stderr, "[refine %s] degrees", tag);
      for (unsigned d : deg) std::fprintf(stderr, " %u", d);
      std::fprintf(stderr, "\n");
    }
    const double far = theta.back(); // last kept Ritz value
    const double flo = rf.ascending ? far : lo_all - margin, fhi = rf.ascending ? hi_all + margin : far;
    if constexpr (std::is_same_v<S, double>) {
      if (rfilt) {
        applies += rfilt->filter(rf, op, deg, theta, flo, fhi, theta.front(), ax_formed);
        rs_prev = rs;
        deg_prev = deg;
        continue;
      }
    }
    applies += rf.filter(deg, flo, fhi, theta.front());
    rs_prev = rs;
    deg_prev = deg;
  }
}


// FP64 thick-restart Lanczos warm-started from the restart state of an FP32
// one (mixed mode refine = 2).  The FP32 solve (its basis, SpMM and Gram
// products in FP32, every coefficient in FP64) runs to tol_fp32; its thick
// restart state -- the k_r = nev/b + 1 Ritz blocks, the Lanczos block that
// follows them and T = diag(theta) plus the arrow (lanczos.hh:97-143) -- is
// carried to FP64: the k_r + 1 blocks are converted and orthonormalised in
// FP64 (block CGS2 + CholQR2), their Rayleigh quotient H = V^T A V is formed
// in FP64, the Ritz part H[0:k_r, 0:k_r] is diagonalised (V[0:k_r] rotated by
// its eigenvectors Z, theta its eigenvalues) and the arrow is
// H[k_r, 0:k_r] Z.  The FP64 driver then continues with the quasi-Lanczos
// step from block k_r and its usual cycles (BlockLanczos::solve_restarted).
// What the FP64 cycles cannot see is the part of A V[0:k_r] outside the
// span of the k_r + 1 seeded blocks, the FP32 basis' rounding (about
// eps_32 ||A||): it stays in the true residuals of the final Ritz vectors.
template <unsigned B, class Reorth>
void warm_restart_solve(const Context& c, std::size_t n, const int32_t* rowptr, const int32_t* col, const float* val32, const float* mass32,
                        const float* minv32, const std::shared_ptr<SparseOperator<float>>& op32, const double* val64, const double* mass64,
                        const double* minv64, const std::shared_ptr<SparseOperator<double>>& op64, const trl_hip_solver_params& p,
                        const trl_hip_solver_params& p32, const float* V0f, double* evals, double* X, double* resid, trl_hip_solver_result& res,
                        trl_hip_solver_result& r32)
{
  using E32 = trl::hip::CSREVP<float, B>;
  using E64 = trl::hip::CSREVP<double, B>;
  const unsigned kr = p.nev / B + 1;
  const unsigned nblk = (p.nev + B - 1) / B;
  const std::size_t P = n * B;
  const trl_hip_dtype f64 = TRL_HIP_F64;
  // ---- FP32 thick-restart solve ---------------------------------------------
  auto e32 = std::make_shared<E32>(c, n, rowptr, col, val32, mass32, minv32);
  e32->set_order(static_cast<trl::hip::EigenvalueOrder>(p.order));
  e32->set_operator(op32);
  e32->set_cholqr_passes(2);
  e32->set_dense_rayleigh_ritz(p.rr_dense != 0);
  trl_amd::BlockLanczos<E32, Reorth> lz32(e32, trl_amd::EigensolverParams{p32.nev, p32.ncv, p32.max_restarts, p32.tol});
  throw_on(trl_hip_memcpy_d2d(c.get(), lz32.initial_block().data_, V0f, P * sizeof(float)), "copy V0");
  c.sync();
  const double t32 = now();
  const auto q32 = lz32.solve();
  c.sync();
  r32.seconds = now() - t32;
  r32.converged = q32.converged ? 1 : 0;
  r32.iterations = q32.iterations;
  r32.n_op_apply = q32.n_op_apply;
  r32.host_seconds = e32->host_seconds();
  r32.rr_steps = e32->rr_steps();
  r32.rr_structured = e32->rr_structured();
  const unsigned m32 = p32.ncv / B;
  // ---- the FP64 driver and its seeded state ----------------------------------
  auto e64 = std::make_shared<E64>(c, n, rowptr, col, val64, mass64, minv64);
  e64->set_order(static_cast<trl::hip::EigenvalueOrder>(p.order));
  e64->set_operator(op64);
  if (p.cholqr_passes > 0) e64->set_cholqr_passes(p.cholqr_passes);
  e64->set_dense_rayleigh_ritz(p.rr_dense != 0);
  trl_amd::BlockLanczos<E64, Reorth> lz64(e64, trl_amd::EigensolverParams{p.nev, p.ncv, p.max_restarts, p.tol});
  auto& V64 = lz64.get_basis();
  auto& V32 = lz32.get_basis();
  trl::hip::DeviceBuffer<float> S32(c, std::size_t(kr) * P);
  std::vector<double> theta32(e32->get_current_eigenvalues().begin(), e32->get_current_eigenvalues().begin() + std::size_t(kr) * B);
  if (lz32.basis_holds_ritz_blocks()) {
    // stopped right after a restart: V[0:kr] are the Ritz blocks, V[kr] the block after them
    throw_on(trl_hip_convert(c.get(), TRL_HIP_F32, V32.data(), f64, V64.data(), std::size_t(kr + 1) * P), "convert restart blocks");
  }
  else {
    // stopped after a Rayleigh-Ritz step: rotate to the Ritz blocks; the next block is V_m
    const auto& Y = e32->get_current_eigenvectors();
    throw_on(trl_hip_rotate(c.get(), TRL_HIP_F32, n, B, int(m32), V32.data(), Y.data(), int(Y.block_cols()), int(kr), S32.data()), "Ritz blocks");
    throw_on(trl_hip_convert(c.get(), TRL_HIP_F32, S32.data(), f64, V64.data(), std::size_t(kr) * P), "convert Ritz blocks");
    throw_on(trl_hip_convert(c.get(), TRL_HIP_F32, V32.data() + std::size_t(m32) * P, f64, V64.data() + std::size_t(kr) * P, P), "convert V_m");
  }
  // FP64 orthonormalisation of the kr + 1 blocks
  trl::hip::DeviceBuffer<double> coef(c, std::size_t(2 * kr + 3) * B * B), R(c, B * B);
  for (unsigned j = 0; j <= kr; ++j) {
    double* Vj = V64.data() + std::size_t(j) * P;
    if (j > 0) throw_on(trl_hip_reorth(c.get(), f64, n, B, int(j), V64.data(), Vj, mass64, coef.data(), 0.0), "seed reorth");
    throw_on(trl_hip_cholqr(c.get(), f64, n, B, Vj, mass64, R.data(), 2), "seed cholqr");
  }
  // H = V^T M (M^{-1} K) V over the kr + 1 blocks, in FP64
  const unsigned ns = kr + 1, nt = ns * B;
  trl::hip::DeviceBuffer<double> AV(c, std::size_t(ns) * P), Hd(c, std::size_t(nt) * nt);
  for (unsigned j = 0; j < ns; ++j) {
    op64->apply(int(B), V64.data() + std::size_t(j) * P, AV.data() + std::size_t(j) * P);
    throw_on(trl_hip_gram(c.get(), f64, n, B, int(ns), V64.data(), AV.data() + std::size_t(j) * P, mass64, Hd.data() + std::size_t(j) * nt * B),
             "seed projection");
  }
  res.n_op_apply = ns;
  std::vector<double> Hc(std::size_t(nt) * nt), H(std::size_t(nt) * nt);
  throw_on(trl_hip_memcpy_d2h(c.get(), Hc.data(), Hd.data(), Hc.size() * sizeof(double)), "seed H");
  // Hd + j*nt*B holds column block j: row r (block r/B, column r%B), column q of block j
  for (unsigned j = 0; j < ns; ++j)
    for (unsigned r = 0; r < nt; ++r)
      for (unsigned q = 0; q < B; ++q) H[std::size_t(r) * nt + j * B + q] = Hc[std::size_t(j) * nt * B + std::size_t(r) * B + q];
  for (unsigned r = 0; r < nt; ++r)
    for (unsigned q = r + 1; q < nt; ++q) H[std::size_t(r) * nt + q] = H[std::size_t(q) * nt + r] = 0.5 * (H[std::size_t(r) * nt + q] + H[std::size_t(q) * nt + r]);
  // Rayleigh-Ritz of the Ritz part
  const unsigned nr = kr * B;
  std::vector<double> Hr(std::size_t(nr) * nr), w(nr), Z(std::size_t(nr) * nr);
  for (unsigned r = 0; r < nr; ++r)
    for (unsigned q = 0; q < nr; ++q) Hr[std::size_t(r) * nr + q] = H[std::size_t(r) * nt + q];
  if (trl_amd::dense::syev(int(nr), Hr.data(), w.data(), Z.data()) != 0) throw std::runtime_error("Eigendecomposition failed");
  std::vector<unsigned> idx(nr);
  for (unsigned i = 0; i < nr; ++i) idx[i] = p.order == 2 ? nr - 1 - i : i;
  if (p.order == 0)
    std::stable_sort(idx.begin(), idx.end(), [&](unsigned a, unsigned b_) { return std::abs(w[a]) > std::abs(w[b_]); });
  // V[0:kr] <- V[0:kr] Z (Z in block layout, kr x kr blocks)
  std::vector<double> Zb(std::size_t(nr) * nr);
  for (unsigned bi = 0; bi < kr; ++bi)
    for (unsigned bj = 0; bj < kr; ++bj)
      for (unsigned ci = 0; ci < B; ++ci)
        for (unsigned cj = 0; cj < B; ++cj) Zb[((std::size_t(bi) * kr + bj) * B + ci) * B + cj] = Z[std::size_t(bi * B + ci) * nr + idx[bj * B + cj]];
  trl::hip::DeviceBuffer<double> Zd(c, Zb.size());
  throw_on(trl_hip_memcpy_h2d(c.get(), Zd.data(), Zb.data(), Zb.size() * sizeof(double)), "seed Z");
  throw_on(trl_hip_rotate(c.get(), f64, n, B, int(kr), V64.data(), Zd.data(), int(kr), int(kr), AV.data()), "seed rotation");
  throw_on(trl_hip_memcpy_d2d(c.get(), V64.data(), AV.data(), std::size_t(kr) * P * sizeof(double)), "seed blocks");
  // T[0:kr+1, 0:kr+1] = diag(theta) + arrow (H[kr, 0:kr] Z) and its transpose
  auto& T = lz64.get_T();
  std::vector<double> Th = T.to_host();
  const std::size_t tbc = T.block_cols();
  auto tblk = [&](unsigned i, unsigned j) { return Th.data() + (std::size_t(i) * tbc + j) * B * B; };
  for (unsigned i = 0; i <= kr; ++i)
    for (unsigned j = 0; j <= kr; ++j) std::fill(tblk(i, j), tblk(i, j) + B * B, 0.0);
  for (unsigned i = 0; i < kr; ++i)
    for (unsigned q = 0; q < B; ++q) tblk(i, i)[q * B + q] = w[idx[i * B + q]];
  for (unsigned r = 0; r < B; ++r)
    for (unsigned col2 = 0; col2 < nr; ++col2) {
      double a = 0.0;
      for (unsigned t = 0; t < nr; ++t) a += H[std::size_t(nr + r) * nt + t] * Z[std::size_t(t) * nr + idx[col2]];
      const unsigned i = col2 / B, q = col2 % B;
      tblk(kr, i)[r * B + q] = a;
      tblk(i, kr)[q * B + r] = a;
    }
  T.from_host(Th);
  // ---- FP64 cycles from the seeded restart state ----------------------------
  c.sync();
  const double t64 = now();
  const auto q64 = lz64.solve_restarted();
  c.sync();
  res.converged = q64.converged ? 1 : 0;
  res.iterations = q64.iterations;
  res.n_op_apply += q64.n_op_apply;
  res.host_seconds = e64->host_seconds();
  res.rr_steps = e64->rr_steps();
  res.rr_structured = e64->rr_structured();
  (void)t64;
  const auto& ev = e64->get_current_eigenvalues();
  for (unsigned i = 0; i < p.nev && i < ev.size(); ++i) evals[i] = ev[i];
  if (X) {
    if (lz64.basis_holds_ritz_blocks())
      throw_on(trl_hip_memcpy_d2d(c.get(), X, V64.data(), std::size_t(nblk) * P * sizeof(double)), "ritz vectors");
    else {
      const auto& Y = e64->get_current_eigenvectors();
      throw_on(trl_hip_rotate(c.get(), f64, n, B, int(p.ncv / B), V64.data(), Y.data(), int(Y.block_cols()), int(nblk), X), "ritz vectors");
    }
    if (resid) true_residuals<double, B>(c, n, *op64, mass64, X, p.nev, evals, resid, AV.data());
  }
  c.sync();
}

template <class F>
int dispatch_b(int b, F&& f)
{
  switch (b) {
  case 1:
    f(std::integral_constant<unsigned, 1>{});
    return 0;
  case 2:
    f(std::integral_constant<unsigned, 2>{});
    return 0;
  case 4:
    f(std::integral_constant<unsigned, 4>{});
    return 0;
  case 6:
    f(std::integral_constant<unsigned, 6>{});
    return 0;
  case 8:
    f(std::integral_constant<unsigned, 8>{});
    return 0;
  case 16:
    f(std::integral_constant<unsigned, 16>{});
    return 0;
  default:
    return TRL_HIP_EINVAL;
  }
}

// The shift-invert solve of A x = lambda B x (examples/openmp/gevp_test.cc:
// 62-116 with CSRGeneralizedEVP): the restated driver on
// trl::hip::CSRGeneralizedEVP, Ritz values mu mapped back to lambda = sigma +
// 1/mu, Ritz vectors X = V Y, true residuals |A x - lambda B x| / |A x| with
// A uploaded for them.
template <unsigned B>
void solve_shift_invert_typed(const Context& c, std::size_t n, const std::vector<int32_t>& arp, const std::vector<int32_t>& acol,
                              const std::vector<double>& aval, const std::vector<int32_t>& brp, const std::vector<int32_t>& bcol,
                              const std::vector<double>& bval, double shift, double inner_tol, const trl_hip_solver_params& p, const double* V0,
                              double* evals, double* mu, double* X_out, double* resid, trl_hip_solver_result& res, trl_hip_inner_stats* st)
{
  using EVP = trl::hip::CSRGeneralizedEVP<B>;
  auto evp = std::make_shared<EVP>(c, n, arp, acol, aval, brp, bcol, bval, shift);
  evp->set_order(static_cast<trl::hip::EigenvalueOrder>(p.order));
  if (p.cholqr_passes > 0) evp->set_cholqr_passes(p.cholqr_passes);
  if (inner_tol > 0.0) evp->set_inner_tolerance(inner_tol);
  evp->set_dense_rayleigh_ritz(p.rr_dense != 0);
  const unsigned nblk = (p.nev + B - 1) / B;
  evp->set_ritz_blocks(nblk);
  trl_amd::EigensolverParams ep{p.nev, p.ncv, p.max_restarts, p.tol};
  auto run = [&](auto& lz) {
    auto V0b = lz.initial_block();
    throw_on(trl_hip_memcpy_d2d(c.get(), V0b.data_, V0, n * B * sizeof(double)), "copy V0");
    c.sync();
    const double t0 = now();
    const auto r = lz.solve();
    c.sync();
    res.seconds = now() - t0;
    res.converged = r.converged ? 1 : 0;
    res.iterations = r.iterations;
    res.n_op_apply = r.n_op_apply;
    res.host_seconds = evp->host_seconds();
    res.rr_steps = evp->rr_steps();
    res.rr_structured = evp->rr_structured();
    const auto& ev = evp->get_current_eigenvalues();
    for (unsigned i = 0; i < p.nev && i < ev.size(); ++i) {
      if (mu) mu[i] = ev[i];
      evals[i] = shift + 1.0 / ev[i];
    }
    if (!X_out && !resid) return;
    trl::hip::DeviceBuffer<double> own;
    double* X = X_out;
    if (!X) {
      own = trl::hip::DeviceBuffer<double>(c, std::size_t(nblk) * n * B);
      X = own.data();
    }
    const auto& Y = evp->get_current_eigenvectors();
    if (lz.basis_holds_ritz_blocks())
      throw_on(trl_hip_memcpy_d2d(c.get(), X, lz.get_basis().data(), std::size_t(nblk) * n * B * sizeof(double)), "ritz vectors");
    else
      throw_on(trl_hip_rotate(c.get(), TRL_HIP_F64, n, B, int(p.ncv / B), lz.get_basis().data(), Y.data(), int(Y.block_cols()), int(nblk), X),
               "ritz vectors");
    if (resid) {
      trl::hip::DeviceBuffer<int32_t> drp(c, arp.size()), dcol(c, std::max<std::size_t>(acol.size(), 1));
      trl::hip::DeviceBuffer<double> dval(c, std::max<std::size_t>(aval.size(), 1)), AX(c, n * B), BX(c, n * B);
      throw_on(trl_hip_memcpy_h2d(c.get(), drp.data(), arp.data(), arp.size() * sizeof(int32_t)), "upload A");
      throw_on(trl_hip_memcpy_h2d(c.get(), dcol.data(), acol.data(), acol.size() * sizeof(int32_t)), "upload A");
      throw_on(trl_hip_memcpy_h2d(c.get(), dval.data(), aval.data(), aval.size() * sizeof(double)), "upload A");
      SparseOperator<double> aop(c, n, drp.data(), dcol.data(), dval.data());
      for (unsigned j = 0; j < nblk; ++j) {
        const double* Xj = X + std::size_t(j) * n * B;
        aop.apply(int(B), Xj, AX.data(), false);
        evp->b_operator().apply(int(B), Xj, BX.data(), false);
        double lam[16] = {0.0}, sums[32];
        for (unsigned q = 0; q < B && j * B + q < p.nev; ++q) lam[q] = evals[j * B + q];
        throw_on(trl_hip_residual_sq(c.get(), TRL_HIP_F64, n, B, BX.data(), AX.data(), nullptr, lam, sums), "residual_sq");
        for (unsigned q = 0; q < B && j * B + q < p.nev; ++q) resid[j * B + q] = std::sqrt(sums[q]) / std::max(std::sqrt(sums[B + q]), 1e-300);
      }
      c.sync();
    }
  };
  if (p.reorth == 0) {
    trl_amd::BlockLanczos<EVP, trl_amd::ModifiedGS> lz(evp, ep);
    run(lz);
  }
  else {
    trl_amd::BlockLanczos<EVP, trl::hip::BlockCGS2> lz(evp, ep);
    run(lz);
  }
  if (st) {
    st->solves = evp->inner_solves();
    st->iterations = evp->inner_iterations();
    st->max_iterations = evp->max_inner_iterations();
  }
}

} // namespace

extern "C" {

int trl_hip_dense_syev(int n, const double* A, double* w, double* Z)
{
  if (n < 0 || (n > 0 && (!A || !w || !Z))) return TRL_HIP_EINVAL;
  return trl_amd::dense::syev(n, A, w, Z) == 0 ? 0 : TRL_HIP_EEIG;
}

int trl_hip_dense_syev_select(int n, const double* A, double* w, int k, const int* sel, double* Z)
{
  if (n < 0 || k < 0 || k > n || (n > 0 && (!A || !w)) || (k > 0 && (!sel || !Z))) return TRL_HIP_EINVAL;
  const int rc = trl_amd::dense::syev_select(n, A, w, k, sel, Z);
  return rc == 0 ? 0 : (rc == -2 ? TRL_HIP_EINVAL : TRL_HIP_EEIG);
}

int trl_hip_dense_syev_arrow_band(int n, int b, const double* A, double* w, int k, const int* sel, double* Z, int* p_out)
{
  if (n < 0 || b < 1 || k < 0 || k > n || (n > 0 && (!A || !w)) || (k > 0 && (!sel || !Z))) return TRL_HIP_EINVAL;
  const int p = n > 0 ? trl_amd::dense::arrow_band_structure(n, b, A) : 0;
  if (p_out) *p_out = p;
  if (p < 0) return TRL_HIP_EINVAL;
  const int rc = trl_amd::dense::syev_arrow_band_select(n, b, p, A, w, k, sel, Z);
  return rc == 0 ? 0 : (rc == -2 ? TRL_HIP_EINVAL : TRL_HIP_EEIG);
}

int trl_hip_solve_csr(trl_hip_ctx* ctx, trl_hip_dtype t, size_t n, const int32_t* rowptr, const int32_t* col, const void* val, const void* mass,
                      const void* mass_inv, const trl_hip_solver_params* p, const void* V0, double* evals, void* X, double* resid,
                      trl_hip_solver_result* res)
{
  if (!ctx || !p || !rowptr || !col || !val || !V0 || !evals || !res) return TRL_HIP_EINVAL;
  if ((mass == nullptr) != (mass_inv == nullptr)) return TRL_HIP_EINVAL;
  if (p->order < 0 || p->order > 2 || p->reorth < 0 || p->reorth > 2) return TRL_HIP_EINVAL;
  *res = trl_hip_solver_result{};
  try {
    const Context c = Context::borrow(ctx);
    const unsigned nblk = (p->nev + unsigned(p->b) - 1) / unsigned(std::max(p->b, 1));
    int rc = 0;
    if (t == TRL_HIP_F64) {
      rc = dispatch_b(p->b, [&](auto Bc) {
        constexpr unsigned B = decltype(Bc)::value;
        const auto* v = static_cast<const double*>(val);
        const auto* mi = static_cast<const double*>(mass_inv);
        solve_typed<double, B>(c, n, rowptr, col, v, static_cast<const double*>(mass), mi, make_operator<double, B>(c, n, rowptr, col, v, mi, *p), *p,
                               static_cast<const double*>(V0), evals, static_cast<double*>(X), nblk, resid, *res);
      });
    }
    else if (t == TRL_HIP_F32) {
      rc = dispatch_b(p->b, [&](auto Bc) {
        constexpr unsigned B = decltype(Bc)::value;
        const auto* v = static_cast<const float*>(val);
        const auto* mi = static_cast<const float*>(mass_inv);
        solve_typed<float, B>(c, n, rowptr, col, v, static_cast<const float*>(mass), mi, make_operator<float, B>(c, n, rowptr, col, v, mi, *p), *p,
                              static_cast<const float*>(V0), evals, static_cast<float*>(X), nblk, resid, *res);
      });
    }
    else
      return TRL_HIP_EINVAL;
    return rc;
  }
  catch (const std::exception& e) {
    return code_of(e);
  }
}

int trl_hip_solve_csr_mixed(trl_hip_ctx* ctx, size_t n, const int32_t* rowptr, const int32_t* col, const double* val64, const float* val32,
                            const double* mass64, const double* minv64, const float* mass32, const float* minv32, const trl_hip_solver_params* p,
                            const trl_hip_mixed_params* mp, const double* V0, double* evals, double* X, double* resid, trl_hip_solver_result* res,
                            trl_hip_solver_result* res32)
{
  if (!ctx || !p || !mp || !rowptr || !col || !val64 || !val32 || !V0 || !evals || !res) return TRL_HIP_EINVAL;
  if ((mass64 == nullptr) != (minv64 == nullptr) || (mass32 == nullptr) != (minv32 == nullptr) || (mass64 == nullptr) != (mass32 == nullptr))
    return TRL_HIP_EINVAL;
  *res = trl_hip_solver_result{};
  trl_hip_solver_result r32{};
  if (mp->refine == 1 && p->order == 0) return TRL_HIP_EINVAL;
  if (p->order < 0 || p->order > 2 || p->reorth < 0 || p->reorth > 2) return TRL_HIP_EINVAL;
  try {
    const Context c = Context::borrow(ctx);
    const int rc = dispatch_b(p->b, [&](auto Bc) {
      constexpr unsigned B = decltype(Bc)::value;
      const unsigned kr = p->nev / B + 1;
      const unsigned nblk = (p->nev + B - 1) / B;
      c.sync();
      const double t0 = now();
      // With the residual-driven FP32 filter the FP64 operator only forms
      // the sweeps' Rayleigh-Ritz and residual products (a few dozen):
      // instead of a tune of its own it takes the tuned FP32 operator's
      // storage and kernel (below).  Without it, it also runs the filter.
      const bool rfil = mp->refine == 1 && mp->residual_filter && !mp->phase1_fp64;
      const auto op64 = make_operator<double, B>(c, n, rowptr, col, val64, minv64, *p, !rfil, !rfil);
      // ---- FP32 phase -------------------------------------------------------
      trl::hip::DeviceBuffer<float> V0f(c, n * B);
      throw_on(trl_hip_convert(c.get(), TRL_HIP_F64, V0, TRL_HIP_F32, V0f.data(), n * B), "convert V0");
      // phase 1's Krylov space: ncv_fp32 columns (0: the solve's ncv)
      const unsigned ncv1 = mp->ncv_fp32 ? mp->ncv_fp32 : p->ncv;
      if (mp->refine == 2) {
        // FP64 thick-restart cycles warm-started from the FP32 restart state
        auto op32w = make_operator<float, B>(c, n, rowptr, col, val32, minv32, *p);
        trl_hip_solver_params p32 = *p;
        p32.ncv = ncv1;
        p32.tol = mp->tol_fp32;
        p32.max_restarts = mp->max_restarts_fp32;
        if (p->reorth == 2)
          warm_restart_solve<B, trl::hip::BlockCGSAdaptive>(c, n, rowptr, col, val32, mass32, minv32, op32w, val64, mass64, minv64, op64, *p, p32, V0f.data(),
                                                            evals, X, resid, *res, r32);
        else if (p->reorth == 1)
          warm_restart_solve<B, trl::hip::BlockCGS2>(c, n, rowptr, col, val32, mass32, minv32, op32w, val64, mass64, minv64, op64, *p, p32, V0f.data(),
                                                     evals, X, resid, *res, r32);
        else
          warm_restart_solve<B, trl_amd::ModifiedGS>(c, n, rowptr, col, val32, mass32, minv32, op32w, val64, mass64, minv64, op64, *p, p32, V0f.data(),
                                                     evals, X, resid, *res, r32);
        c.sync();
        res->seconds = now() - t0;
        return;
      }
      const unsigned m = ncv1 / B;
      unsigned nref = mp->refine == 1 ? (mp->refine_blocks ? mp->refine_blocks : kr + 6) : kr;
      if (nref < kr) nref = kr;
      if (nref > m) nref = m;
      trl::hip::DeviceBuffer<float> X32(c, std::size_t(nref) * n * B);
      std::vector<double> ev32(std::max(p->nev, nref * B), 0.0), all32;
      trl_hip_solver_params p32 = *p;
      p32.ncv = ncv1;
      p32.tol = mp->tol_fp32;
      p32.max_restarts = mp->max_restarts_fp32;
      if (p32.cholqr_passes == 0) p32.cholqr_passes = 2;
      unsigned got = nref;
      trl::hip::DeviceBuffer<double> X1(c, mp->phase1_fp64 ? std::size_t(nref) * n * B : 1);
      if (mp->phase1_fp64) {
        p32.cholqr_passes = p->cholqr_passes;
        solve_typed<double, B>(c, n, rowptr, col, val64, mass64, minv64, op64, p32, V0, ev32.data(), X1.data(), nref, nullptr, r32, &all32, &got);
      }
      std::shared_ptr<SparseOperator<float>> op32;
      if (!mp->phase1_fp64) {
        op32 = make_operator<float, B>(c, n, rowptr, col, val32, minv32, *p);
        if (rfil && p->spmm_tune) op64->adopt(int(B), op32->descriptor());
        solve_typed<float, B>(c, n, rowptr, col, val32, mass32, minv32, op32, p32, V0f.data(), ev32.data(), X32.data(), nref, nullptr, r32, &all32, &got);
      }
      nref = got;
      // spectrum ends seen by the phase-1 Krylov space
      double lo_all = all32.front(), hi_all = all32.front();
      for (double v : all32) {
        lo_all = std::min(lo_all, v);
        hi_all = std::max(hi_all, v);
      }
      const bool asc = (p->order == 1);
      const double eta = p->reorth == 2 ? 0.70710678118654752 : 0.0;
      if (mp->refine == 1 && op32 && mp->fp32_sweeps > 0) {
        // ---- FP32 refinement sweeps: the same filter in FP32 panels (half
        // the bytes) until the FP32 residuals reach tol_fp32_refine ----------
        if (p->spmm_tune) tune_panels_of(*op32, B, nref);
        BlockRefiner<float, B> rf32(c, n, *op32, mass32, nref, asc, eta);
        throw_on(trl_hip_memcpy_d2d(c.get(), rf32.X, X32.data(), std::size_t(nref) * n * B * sizeof(float)), "copy X32");
        std::vector<double> th32, rs32;
        unsigned sw32 = 0;
        std::size_t ap32 = 0;
        refine_sweeps<float, B>(rf32, c, n, *op32, mass32, p->nev, mp->tol_fp32_refine, mp->fp32_sweeps, mp->cheb_degree, mp->degree_opt == 1, lo_all,
                                hi_all, th32, rs32, sw32, ap32, "fp32", nullptr, 2);
        throw_on(trl_hip_memcpy_d2d(c.get(), X32.data(), rf32.X, std::size_t(nref) * n * B * sizeof(float)), "copy X32");
        r32.n_op_apply += (sw32 + 1) * nref * 2 + ap32;
        r32.iterations += sw32;
      }

      bool done = false;
      if (mp->refine == 1) {
        // ---- FP64 phase: block refinement of the FP32 Ritz blocks -----------
        if (p->spmm_tune && !rfil) tune_panels_of(*op64, B, nref);
        if (p->spmm_tune && rfil && op32) op64->adopt_row_order(op32->panels_descriptor());
        std::unique_ptr<ResidualFilter<B>> rfilt;
        if (rfil) rfilt = std::make_unique<ResidualFilter<B>>(c, n, nref, *op32, mp->cheb_degree);
        BlockRefiner<double, B> rf(c, n, *op64, mass64, nref, asc, eta);
        if (mp->phase1_fp64)
          throw_on(trl_hip_memcpy_d2d(c.get(), rf.X, X1.data(), std::size_t(nref) * n * B * sizeof(double)), "copy X1");
        else
          throw_on(trl_hip_convert(c.get(), TRL_HIP_F32, X32.data(), TRL_HIP_F64, rf.X, std::size_t(nref) * n * B), "convert X32");
        std::vector<double> theta, rs;
        unsigned sweep = 0;
        std::size_t ap64 = 0;
        // the degree ramp belongs to the first sweeps after phase 1: here
        // when no FP32 sweeps ran (the all-FP64 schedule among them)
        const unsigned ramp64 = (op32 && mp->fp32_sweeps > 0) ? 0 : 2;
        done = refine_sweeps<double, B>(rf, c, n, *op64, mass64, p->nev, p->tol, mp->max_refine, mp->cheb_degree, mp->degree_opt != 0, lo_all, hi_all, theta,
                                        rs, sweep, ap64, "fp64", rfilt.get(), ramp64);
        if (done) {
          for (unsigned i = 0; i < p->nev; ++i) evals[i] = theta[i];
          if (X) throw_on(trl_hip_memcpy_d2d(c.get(), X, rf.X, std::size_t(nblk) * n * B * sizeof(double)), "copy X");
          if (resid)
            for (unsigned i = 0; i < p->nev; ++i) resid[i] = rs[i];
          res->converged = 1;
          res->iterations = sweep;
          res->n_op_apply = (sweep + 1) * nref * 2 + ap64;
        }
        else {
          // fall back to an FP64 Lanczos warm start from the refined block
          trl::hip::DeviceBuffer<double> V064(c, n * B);
          throw_on(trl_hip_memcpy_d2d(c.get(), V064.data(), rf.X, n * B * sizeof(double)), "seed");
          solve_typed<double, B>(c, n, rowptr, col, val64, mass64, minv64, op64, *p, V064.data(), evals, X, nblk, resid, *res);
          done = true;
        }
      }
      if (!done) {
        // ---- FP64 phase: thick-restart Lanczos from V0 = X32 G -------------
        trl::hip::DeviceBuffer<double> X64(c, std::size_t(kr) * n * B);
        if (mp->phase1_fp64)
          throw_on(trl_hip_memcpy_d2d(c.get(), X64.data(), X1.data(), std::size_t(kr) * n * B * sizeof(double)), "copy X1");
        else
          throw_on(trl_hip_convert(c.get(), TRL_HIP_F32, X32.data(), TRL_HIP_F64, X64.data(), std::size_t(kr) * n * B), "convert X32");
        std::mt19937_64 g(mp->seed);
        std::normal_distribution<double> nd;
        std::vector<double> G(std::size_t(kr) * B * B);
        for (auto& x : G) x = nd(g);
        trl::hip::DeviceBuffer<double> Gd(c, G.size());
        throw_on(trl_hip_memcpy_h2d(c.get(), Gd.data(), G.data(), G.size() * sizeof(double)), "upload G");
        trl::hip::DeviceBuffer<double> V064(c, n * B);
        throw_on(trl_hip_rotate(c.get(), TRL_HIP_F64, n, B, int(kr), X64.data(), Gd.data(), 1, 1, V064.data()), "mix Ritz vectors");
        solve_typed<double, B>(c, n, rowptr, col, val64, mass64, minv64, op64, *p, V064.data(), evals, X, nblk, resid, *res);
      }
      c.sync();
      res->seconds = now() - t0;
    });
    if (res32) *res32 = r32;
    return rc;
  }
  catch (const std::exception& e) {
    if (res32) *res32 = r32;
    return code_of(e);
  }
}

int trl_hip_solve_csr_shift_invert(trl_hip_ctx* ctx, size_t n, const int32_t* a_rowptr, const int32_t* a_col, const double* a_val,
                                   const int32_t* b_rowptr, const int32_t* b_col, const double* b_val, double shift, double inner_tol,
                                   const trl_hip_solver_params* p, const double* V0, double* evals, double* mu, double* X, double* resid,
                                   trl_hip_solver_result* res, trl_hip_inner_stats* inner)
{
  if (!ctx || !p || !a_rowptr || !a_col || !a_val || !b_rowptr || !b_col || !b_val || !V0 || !evals || !res || n == 0) return TRL_HIP_EINVAL;
  if (p->order < 0 || p->order > 2 || p->reorth < 0 || p->reorth > 2 || !(inner_tol >= 0.0) || !std::isfinite(shift)) return TRL_HIP_EINVAL;
  *res = trl_hip_solver_result{};
  if (inner) *inner = trl_hip_inner_stats{};
  try {
    const Context c = Context::borrow(ctx);
    const std::vector<int32_t> arp(a_rowptr, a_rowptr + n + 1), brp(b_rowptr, b_rowptr + n + 1);
    if (arp[n] < 0 || brp[n] < 0) return TRL_HIP_EINVAL;
    const std::vector<int32_t> acol(a_col, a_col + arp[n]), bcol(b_col, b_col + brp[n]);
    const std::vector<double> aval(a_val, a_val + arp[n]), bval(b_val, b_val + brp[n]);
    return dispatch_b(p->b, [&](auto Bc) {
      constexpr unsigned B = decltype(Bc)::value;
      solve_shift_invert_typed<B>(c, n, arp, acol, aval, brp, bcol, bval, shift, inner_tol, *p, V0, evals, mu, X, resid, *res, inner);
    });
  }
  catch (const std::exception& e) {
    return code_of(e);
  }
}

} // extern "C"
