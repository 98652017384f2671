// Conformance of the HIP backend with the reference's concepts and its
// UNCHANGED driver (built only by oracle/Makefile.ref, from the reference's
// headers where they lie; output in oracle/_ref/, never shipped as product).
//
// 1. static_asserts: trl::hip types satisfy trl::Eigenproblem,
//    trl::BlockMultiVector and trl::ReorthogonalizationStrategy
//    (include/trl/concepts.hh:26-188, eigensolvers/reorthogonalization.hh:29-32).
// 2. main(): the acceptance checks of tests/cpp/hip_test_helper.hh run with
//    DRIVER = trl::BlockLanczos (include/trl/eigensolvers/lanczos.hh), with
//    the default trl::ModifiedGS and with trl::hip::BlockCGS2.
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
  std::printf("reference driver on trl::hip: %s\n", failed ? "FAILED" : "ALL PASSED");
  return failed;
}
