#pragma once
// Test helper for the HIP backend with the duck-typed interface the
// reference's backend-generic tests use (set_random, norm, to_host_data,
// sync; tests/openmp/test_helper.hh, tests/sycl/test_helper.hh), and the two
// acceptance checks of tests/test_lanczos_extend.hh and
// tests/test_lanczos_convergence.hh written for a device backend.
//
// DRIVER selects the block-Lanczos class template: trl_amd::BlockLanczos
// (default) or, in the conformance build, the reference's trl::BlockLanczos.

#include <cmath>
#include <cstdio>
#include <memory>
#include <random>
#include <vector>

#include "trl/hip/csrevp.hh"
#include "trl_amd/lanczos.hh"

#ifndef DRIVER
#define DRIVER trl_amd::BlockLanczos
#define PARAMS trl_amd::EigensolverParams
#endif

template <class EVP>
struct HipTestHelper {
  using S = typename EVP::BlockMultivector::Scalar;
  trl::hip::Context ctx;
  std::mt19937 rng{5489u};
  std::normal_distribution<double> dist;

  explicit HipTestHelper(const trl::hip::Context& c) : ctx(c) {}

  void set_random(typename EVP::BlockMultivector::BlockView V)
  {
    std::vector<S> h(V.rows() * V.cols());
    for (auto& x : h) x = S(dist(rng));
    trl::hip::throw_on(trl_hip_memcpy_h2d(ctx.get(), V.data_, h.data(), h.size() * sizeof(S)), "set_random");
  }
  std::vector<S> to_host(typename EVP::BlockMultivector::BlockView V)
  {
    std::vector<S> h(V.rows() * V.cols());
    trl::hip::throw_on(trl_hip_memcpy_d2h(ctx.get(), h.data(), V.data_, h.size() * sizeof(S)), "to_host");
    return h;
  }
  double norm(typename EVP::BlockMultivector::BlockView V)
  {
    double s = 0.0;
    for (S x : to_host(V)) s += double(x) * double(x);
    return std::sqrt(s);
  }
  std::vector<double> to_host_data(typename EVP::BlockMultivector::BlockMatrix::BlockView B) { return B.to_host(); }
  std::vector<double> to_host_data(const std::vector<double>& v) { return v; }
  void sync() { ctx.sync(); }
};


// CSR of diag((i+1)^2) (tests/openmp/diagonal.hh:15-20) or tridiag(-1,2,-1)
// (tests/openmp/laplace.hh:19-31)
inline void make_csr(const char* kind, std::size_t N, std::vector<int32_t>& rp, std::vector<int32_t>& ci, std::vector<double>& v)
{
  rp.assign(1, 0);
  ci.clear();
  v.clear();
  for (std::size_t i = 0; i < N; ++i) {
    if (kind[0] == 'd') {
      ci.push_back(int32_t(i));
      v.push_back(double((i + 1) * (i + 1)));
    }
    else {
      if (i > 0) {
        ci.push_back(int32_t(i - 1));
        v.push_back(-1.0);
      }
      ci.push_back(int32_t(i));
      v.push_back(2.0);
      if (i + 1 < N) {
        ci.push_back(int32_t(i + 1));
        v.push_back(-1.0);
      }
    }
    rp.push_back(int32_t(ci.size()));
  }
}

template <class S, unsigned bs>
std::shared_ptr<trl::hip::CSREVP<S, bs>> make_evp(const trl::hip::Context& ctx, const char* kind, std::size_t N)
{
  std::vector<int32_t> rp, ci;
  std::vector<double> v;
  make_csr(kind, N, rp, ci, v);
  return std::shared_ptr<trl::hip::CSREVP<S, bs>>(trl::hip::CSREVP<S, bs>::from_host(ctx, N, rp, ci, v).release());
}

// Orthonormality of all basis blocks: V_i^T V_i = I, V_i^T V_j = 0 (j < i).
template <class EVP, class Helper>
bool basis_orthonormal(EVP& evp, Helper& h, typename EVP::BlockMultivector& V, double tol, double& worst)
{
  constexpr unsigned bs = EVP::blocksize;
  auto tmp = evp.create_blockmatrix(1, 1);
  auto tb = tmp.block_view(0, 0);
  worst = 0.0;
  for (std::size_t i = 0; i < V.blocks(); ++i) {
    for (std::size_t j = 0; j <= i; ++j) {
      V.block_view(i).dot(V.block_view(j), tb);
      h.sync();
      const auto g = h.to_host_data(tb);
      for (unsigned r = 0; r < bs; ++r)
        for (unsigned c = 0; c < bs; ++c) {
          const double want = (i == j && r == c) ? 1.0 : 0.0;
          worst = std::max(worst, std::abs(g[r * bs + c] - want));
        }
    }
  }
  return worst < tol;
}

// extend(0, m) on the diagonal operator, then the orthogonality and the
// Lanczos relation A V - V T = [0 ... V_m beta] (the checks of
// tests/test_lanczos_extend.hh:41-111).
template <class S, unsigned bs>
bool check_extend(const trl::hip::Context& ctx, std::size_t N, double tol)
{
  using EVP = trl::hip::CSREVP<S, bs>;
  auto evp = make_evp<S, bs>(ctx, "diagonal", N);
  HipTestHelper<EVP> h(ctx);
  PARAMS params{.nev = 8, .ncv = 32, .max_restarts = 1000};
  DRIVER<EVP> lz(evp, params);
  h.set_random(lz.initial_block());
  const unsigned m = params.ncv / bs;
  lz.extend(0, m);
  h.sync();
  auto& V = lz.get_basis();
  auto& T = lz.get_T();
  double orth = 0.0;
  bool ok = basis_orthonormal(*evp, h, V, tol, orth);
  auto AV = evp->create_multivector(N, params.ncv);
  auto VT = evp->create_multivector(N, params.ncv);
  for (unsigned j = 0; j < m; ++j) {
    evp->apply(V.block_view(j), AV.block_view(j));
    auto VTj = VT.block_view(j);
    VTj.set_zero();
    for (unsigned i = 0; i < m; ++i) V.block_view(i).mult_add(T.block_view(i, j), VTj);
    auto AVj = AV.block_view(j);
    AVj -= VTj;
  }
  double rel = 0.0;
  for (unsigned j = 0; j + 1 < m; ++j) rel = std::max(rel, h.norm(AV.block_view(j)));
  auto extra = evp->create_multivector(N, bs);
  V.block_view(m).mult(lz.get_B().block_view(0, 0), extra.block_view(0));
  auto last = AV.block_view(m - 1);
  last -= extra.block_view(0);
  rel = std::max(rel, h.norm(last));
  ok = ok && rel < tol;
  std::printf("extend bs=%u: orthogonality %.2e, relation %.2e -> %s\n", bs, orth, rel, ok ? "ok" : "FAILED");
  return ok;
}

// solve() against exact eigenvalues (tests/test_lanczos_convergence.hh:12-59):
// N = 128, nev = 16, ncv = 4 nev, largest |lambda| first, relative tol.
template <class S, unsigned bs>
bool check_convergence(const trl::hip::Context& ctx, const char* kind, double tol)
{
  using EVP = trl::hip::CSREVP<S, bs>;
  const std::size_t N = 128;
  const unsigned nev = 16;
  auto evp = make_evp<S, bs>(ctx, kind, N);
  HipTestHelper<EVP> h(ctx);
  PARAMS params{.nev = nev, .ncv = 4 * nev, .max_restarts = 1000, .tolerance = tol};
  DRIVER<EVP> lz(evp, params);
  h.set_random(lz.initial_block());
  const auto res = lz.solve();
  h.sync();
  bool ok = res.converged;
  const auto ev = h.to_host_data(evp->get_current_eigenvalues());
  const double pi = std::acos(-1.0);
  double worst = 0.0;
  for (unsigned i = 0; i < nev && ok; ++i) {
    const double exact = kind[0] == 'd' ? double((N - i) * (N - i)) : 2.0 - 2.0 * std::cos(double(N - i) * pi / double(N + 1));
    const double err = std::abs(ev[i] - exact) / std::abs(exact);
    worst = std::max(worst, err);
    if (err > tol) ok = false;
  }
  std::printf("convergence %s bs=%u: %u restarts, %u applies, max rel err %.2e -> %s\n", kind, bs, res.iterations, res.n_op_apply, worst,
              ok ? "ok" : "FAILED");
  return ok;
}
