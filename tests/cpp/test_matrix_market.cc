// trl::hip::load_matrix_market (the reference's load_matrix_market,
// examples/sycl/csrevp.hh:26-124) on a small symmetric file: mirrored
// entries, 1-based indices, canonical rows; the missing-file message.
// Host only: runs without a GPU.
#include <cmath>
#include <cstdio>
#include <fstream>
#include <stdexcept>
#include <string>

#include "trl/hip/matrix_market.hh"

static int failures = 0;
#define CHECK(c)                                                                                                                                     \
  do {                                                                                                                                               \
    if (!(c)) {                                                                                                                                      \
      std::fprintf(stderr, "FAILED %s:%d: %s\n", __FILE__, __LINE__, #c);                                                                            \
      ++failures;                                                                                                                                    \
    }                                                                                                                                                \
  } while (0)

int main(int argc, char** argv)
{
  const std::string dir = argc > 1 ? argv[1] : ".";
  const std::string path = dir + "/mm_test.mtx";
  {
    std::ofstream f(path);
    f << "%%MatrixMarket matrix coordinate real symmetric\n% comment\n3 3 4\n3 1 -2.0\n1 1 4.0\n2 2 5.0\n3 3 6.0\n";
  }
  const auto A = trl::hip::load_matrix_market(path);
  CHECK(A.rows == 3 && A.cols == 3 && A.symmetric);
  CHECK(A.rowptr.size() == 4 && A.rowptr[0] == 0 && A.rowptr[1] == 2 && A.rowptr[2] == 3 && A.rowptr[3] == 5);
  // row 0: (0, 4) (2, -2); row 1: (1, 5); row 2: (0, -2) (2, 6)
  const int cols[] = {0, 2, 1, 0, 2};
  const double vals[] = {4.0, -2.0, 5.0, -2.0, 6.0};
  for (int k = 0; k < 5; ++k) CHECK(A.col[k] == cols[k] && A.val[k] == vals[k]);
  bool thrown = false;
  try {
    trl::hip::load_matrix_market(dir + "/absent.mtx");
  }
  catch (const std::runtime_error& e) {
    thrown = std::string(e.what()) == "Error: cannot open file " + dir + "/absent.mtx";
  }
  CHECK(thrown);
  std::printf("%s\n", failures ? "FAILED" : "OK");
  return failures ? 1 : 0;
}
