// Acceptance tests of the HIP backend, following the reference's
// tests/openmp/test_lanczos_extend.cc and test_lanczos_convergence.cc:
// block sizes 1, 2, 4, 8 in double; returns the number of failures.
#include "hip_test_helper.hh"

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
  std::printf("%s\n", failed ? "FAILED" : "ALL PASSED");
  return failed;
}
