// Conformance of the HIP backend with the reference's concepts and its
// UNCHANGED driver (built only by oracle/Makefile.ref, from the reference's
// headers where they lie; output in oracle/_ref/, never shipped as product).
//
// 1. static_asserts: trl::hip types satisfy trl::Eigenproblem,
//    trl::BlockMultiVector and trl::ReorthogonalizationStrategy
//    (include/trl/concepts.hh:26-188, eigensolvers/reorthogonalization.hh:29-32).
// 2. main(): the acceptance checks of tests/cpp/hip_test_helper.hh run with
//    DRIVER = trl::BlockLanczos (include/trl/eigensolvers/lanczos.hh), with
//    the default trl::ModifiedGS and with trl::hip::BlockCGS2, and the 2D
//    Laplacian 128^2 (8 smallest, ncv = 256) against its closed form and the
//    CPU oracle; and the generalized problem with a sparse mass matrix by
//    shift-invert (trl::hip::CSRGeneralizedEVP, the reference's
//    examples/openmp/csrevp.hh:51-104) against the closed form of the
//    bilinear finite-element Laplacian pair.
#include <algorithm>
#include <cstdint>
#include <vector>

#include <trl/concepts.hh>
#include <trl/eigensolvers/lanczos.hh>
#include <trl/eigensolvers/params.hh>
#include <trl/eigensolvers/reorthogonalization.hh>

#define DRIVER trl::BlockLanczos
#define PARAMS trl::EigensolverParams
#include "hip_test_helper.hh"

using E64 = trl::hip::CSREVP<double, 4>;
using E32 = trl::hip::CSREVP<float, 8>;
static_assert(trl::Eigenproblem<E64>);
static_assert(trl::Eigenproblem<E32>);
static_assert(trl::BlockMultiVector<E64::BlockMultivector>);
static_assert(trl::MatrixBlockViewConcept<E64::BlockMatrixBlockView>);
static_assert(trl::ReorthogonalizationStrategy<trl::ModifiedGS, E64, E64::BlockMultivector>);
static_assert(trl::ReorthogonalizationStrategy<trl::hip::BlockCGS2, E64, E64::BlockMultivector>);
using EG = trl::hip::CSRGeneralizedEVP<4>;
static_assert(trl::Eigenproblem<EG>);
static_assert(trl::ReorthogonalizationStrategy<trl::ModifiedGS, EG, EG::BlockMultivector>);

// the reference driver with the fused block-CGS2 strategy
template <unsigned bs>
bool check_cgs2(const trl::hip::Context& ctx)
{
  using EVP = trl::hip::CSREVP<double, bs>;
  auto evp = make_evp<double, bs>(ctx, "laplace1d", 128);
  HipTestHelper<EVP> h(ctx);
  trl::EigensolverParams params{.nev = 16, .ncv = 64, .max_restarts = 1000, .tolerance = 1e-8};
  trl::BlockLanczos<EVP, trl::hip::BlockCGS2> lz(evp, params);
  h.set_random(lz.initial_block());
  const auto res = lz.solve();
  h.sync();
  const auto ev = evp->get_current_eigenvalues();
  const double pi = std::acos(-1.0);
  double worst = 0.0;
  for (unsigned i = 0; i < 16; ++i) {
    const double ex = 2.0 - 2.0 * std::cos(double(128 - i) * pi / 129.0);
    worst = std::max(worst, std::abs(ev[i] - ex) / ex);
  }
  const bool ok = res.converged && worst < 1e-8;
  std::printf("reference driver + BlockCGS2 bs=%u: %u restarts, max rel err %.2e -> %s\n", bs, res.iterations, worst, ok ? "ok" : "FAILED");
  return ok;
}

// The oracle's restated driver (oracle/trl_oracle.cc, linked from
// oracle/lib/liboracle.so): the same solve on the CPU from the same start block.
extern "C" int oracle_lanczos_solve(size_t n, const int32_t* rowptr, const int32_t* col, const double* val, const double* dscale, const double* mdiag,
                                    int b, unsigned nev, unsigned ncv, unsigned max_restarts, double tol, int order, const double* V0, double* evals_out,
                                    double* X_out, int* info);

// The 2D 5-point Dirichlet Laplacian 128 x 128 (n = 16384), b = 4, the 8
// smallest eigenpairs, ncv = 256, tol 1e-8, solved by the UNCHANGED
// trl::BlockLanczos (ModifiedGS) on trl::hip in ascending order.  Checked
// against the closed form 4 - 2cos(i pi/129) - 2cos(j pi/129) and against the
// oracle's restated driver from the same start block (the convergence check
// of tests/test_lanczos_convergence.hh:12-52 on this problem), with the
// restart counts of both printed.
bool check_laplace2d_128(const trl::hip::Context& ctx)
{
  constexpr unsigned bs = 4;
  const int side = 128;
  const std::size_t N = std::size_t(side) * side;
  std::vector<int32_t> rp{0}, ci;
  std::vector<double> v;
  for (int y = 0; y < side; ++y)
    for (int x = 0; x < side; ++x) {
      const int r = y * side + x;
      const int nb[5] = {r - side, r - 1, r, r + 1, r + side};
      const bool ok[5] = {y > 0, x > 0, true, x + 1 < side, y + 1 < side};
      for (int q = 0; q < 5; ++q)
        if (ok[q]) {
          ci.push_back(nb[q]);
          v.push_back(q == 2 ? 4.0 : -1.0);
        }
      rp.push_back(int32_t(ci.size()));
    }
  using EVP = trl::hip::CSREVP<double, bs>;
  auto evp = std::shared_ptr<EVP>(EVP::from_host(ctx, N, rp, ci, v).release());
  evp->set_order(trl::hip::EigenvalueOrder::Ascending);
  HipTestHelper<EVP> h(ctx);
  const unsigned nev = 8, ncv = 256;
  const double tol = 1e-8;
  trl::EigensolverParams params{.nev = nev, .ncv = ncv, .max_restarts = 1000, .tolerance = tol};
  trl::BlockLanczos<EVP> lz(evp, params);
  h.set_random(lz.initial_block());
  const std::vector<double> V0 = h.to_host(lz.initial_block());
  const auto res = lz.solve();
  h.sync();
  const auto ev = evp->get_current_eigenvalues();

  std::vector<double> exact;
  const double pi = std::acos(-1.0);
  for (int i = 1; i <= side; ++i)
    for (int j = 1; j <= side; ++j) exact.push_back(4.0 - 2.0 * std::cos(i * pi / (side + 1)) - 2.0 * std::cos(j * pi / (side + 1)));
  std::sort(exact.begin(), exact.end());

  std::vector<double> ev_or(nev, 0.0);
  int info[4] = {0, 0, 0, 0};
  const int rc = oracle_lanczos_solve(N, rp.data(), ci.data(), v.data(), nullptr, nullptr, int(bs), nev, ncv, 1000, tol, 1, V0.data(), ev_or.data(),
                                      nullptr, info);
  double err_ex = 0.0, err_or = 0.0;
  for (unsigned i = 0; i < nev && i < ev.size(); ++i) {
    err_ex = std::max(err_ex, std::abs(ev[i] - exact[i]) / exact[i]);
    err_or = std::max(err_or, std::abs(ev[i] - ev_or[i]) / std::abs(ev_or[i]));
  }
  const bool ok = res.converged && rc == 0 && info[0] == 1 && ev.size() >= nev && err_ex < 1e-10 && err_or < 1e-10;
  std::printf("reference driver, 2D Laplacian 128^2 b=4 nev=8 ncv=256: %u restarts (oracle %d), max rel err vs closed form %.2e, vs oracle %.2e -> %s\n",
              res.iterations, info[1], err_ex, err_or, ok ? "ok" : "FAILED");
  return ok;
}

// K x = lambda M x for bilinear finite elements on the unit square (nx x ny
// interior nodes, Dirichlet): K = K1x (x) M1y + M1x (x) K1y, M = M1x (x) M1y,
// K1 = tridiag(-1, 2, -1) / h, M1 = h tridiag(1, 4, 1) / 6.  Shift 0, the 8
// eigenvalues nearest it (|mu| largest, the reference's order) by the
// UNCHANGED trl::BlockLanczos on (K - 0 M)^{-1} M in the M-inner product,
// mapped back by lambda = 1/mu (examples/openmp/gevp_test.cc:135-141),
// against lambda_k + lambda_l, lambda_k = (6/h^2)(1 - cos t_k)/(2 + cos t_k).
bool check_shift_invert(const trl::hip::Context& ctx)
{
  constexpr unsigned bs = 4;
  const int nx = 48, ny = 40;
  const std::size_t N = std::size_t(nx) * ny;
  const double hx = 1.0 / (nx + 1), hy = 1.0 / (ny + 1);
  auto k1 = [](int d, double h) { return (d == 0 ? 2.0 : -1.0) / h; };
  auto m1 = [](int d, double h) { return h * (d == 0 ? 4.0 : 1.0) / 6.0; };
  std::vector<int32_t> rp{0}, ci;
  std::vector<double> kv, mv;
  for (int y = 0; y < ny; ++y)
    for (int x = 0; x < nx; ++x) {
      for (int dy = -1; dy <= 1; ++dy)
        for (int dx = -1; dx <= 1; ++dx) {
          if (x + dx < 0 || x + dx >= nx || y + dy < 0 || y + dy >= ny) continue;
          ci.push_back((y + dy) * nx + x + dx);
          kv.push_back(k1(dx, hx) * m1(dy, hy) + m1(dx, hx) * k1(dy, hy));
          mv.push_back(m1(dx, hx) * m1(dy, hy));
        }
      rp.push_back(int32_t(ci.size()));
    }
  auto evp = std::make_shared<trl::hip::CSRGeneralizedEVP<bs>>(ctx, N, rp, ci, kv, rp, ci, mv, 0.0);
  HipTestHelper<trl::hip::CSRGeneralizedEVP<bs>> h(ctx);
  const unsigned nev = 8;
  trl::EigensolverParams params{.nev = nev, .ncv = 64, .max_restarts = 1000, .tolerance = 1e-10};
  trl::BlockLanczos<trl::hip::CSRGeneralizedEVP<bs>> lz(evp, params);
  h.set_random(lz.initial_block());
  const auto res = lz.solve();
  h.sync();
  std::vector<double> lam = evp->back_transformed();
  lam.resize(nev);
  std::sort(lam.begin(), lam.end());
  std::vector<double> exact;
  const double pi = std::acos(-1.0);
  auto one = [&](int k, int m, double hh) {
    const double t = k * pi / (m + 1);
    return 6.0 / (hh * hh) * (1.0 - std::cos(t)) / (2.0 + std::cos(t));
  };
  for (int k = 1; k <= nx; ++k)
    for (int l = 1; l <= ny; ++l) exact.push_back(one(k, nx, hx) + one(l, ny, hy));
  std::sort(exact.begin(), exact.end());
  double worst = 0.0;
  for (unsigned i = 0; i < nev; ++i) worst = std::max(worst, std::abs(lam[i] - exact[i]) / exact[i]);
  const bool ok = res.converged && worst < 1e-10;
  std::printf("reference driver, shift-invert FE pair 48x40 (sparse mass) b=4 nev=8: %u restarts, %u inner solves, max rel err vs closed form %.2e -> %s\n",
              res.iterations, evp->inner_solves(), worst, ok ? "ok" : "FAILED");
  return ok;
}

int main()
{
  trl::hip::Context ctx(0);
  int failed = 0;
  failed += !check_extend<double, 1>(ctx, 128, 1e-8);
  failed += !check_extend<double, 2>(ctx, 128, 1e-8);
  failed += !check_extend<double, 4>(ctx, 128, 1e-8);
  failed += !check_extend<double, 8>(ctx, 128, 1e-8);
  for (const char* kind : {"diagonal", "laplace1d"}) {
    failed += !check_convergence<double, 1>(ctx, kind, 1e-8);
    failed += !check_convergence<double, 2>(ctx, kind, 1e-8);
    failed += !check_convergence<double, 4>(ctx, kind, 1e-8);
    failed += !check_convergence<double, 8>(ctx, kind, 1e-8);
  }
  failed += !check_cgs2<4>(ctx);
  failed += !check_cgs2<8>(ctx);
  failed += !check_laplace2d_128(ctx);
  failed += !check_shift_invert(ctx);
  std::printf("reference driver on trl::hip: %s\n", failed ? "FAILED" : "ALL PASSED");
  return failed;
}
