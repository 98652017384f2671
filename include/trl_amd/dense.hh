#pragma once
// Host FP64 dense linear algebra of the Rayleigh-Ritz step.

#include <vector>

namespace trl_amd::dense {

// Symmetric eigen-decomposition of the n x n row-major symmetric matrix A:
// ascending eigenvalues in w[0..n), eigenvector j in column j of the
// row-major n x n matrix Z.  Returns 0, or -1 if the QR iteration did not
// converge.
int syev(int n, const double* A, double* w, double* Z);

// The same eigenvalues (ascending, all n in w) and only the eigenvectors of
// eigenvalues w[sel[0]] < w[sel[1]] < ... (k strictly increasing indices):
// eigenvector of w[sel[j]] in column j of the row-major n x k matrix Z.
// Returns 0, -1 without convergence, -2 for a bad selection.
int syev_select(int n, const double* A, double* w, int k, const int* sel, double* Z);

// The structure of the thick-restart block Lanczos projected matrix in the
// row-major symmetric A: returns p (a multiple of b, possibly 0) if A[0:p,
// 0:p] is diagonal, the rows 0..p-1 couple beyond it only to the b
// coordinates p..p+b-1, and A[p:, p:] has half-bandwidth b; -1 otherwise.
int arrow_band_structure(int n, int b, const double* A);

// syev_select for a matrix of that structure (p from arrow_band_structure):
// the Ritz part reduced to band form, the band to tridiagonal form by
// Householder bulge chasing, O(n^2 b + n^2 k) instead of O(n^3).  The same results as
// syev_select to rounding.
int syev_arrow_band_select(int n, int b, int p, const double* A, double* w, int k, const int* sel, double* Z);

// The same in two calls, for a caller that picks the eigenvectors from the
// eigenvalues: reduce() stores the reduction and gives all eigenvalues
// (ascending, in w; 0, -1 without convergence, -2 bad arguments), vectors()
// the eigenvectors of w[sel[0..k)] as syev_select does (0, -1, -2).
class ArrowBandEigensolver {
public:
  int reduce(int n, int b, int p, const double* A, double* w);
  int vectors(int k, const int* sel, double* Z) const;

private:
  struct Refl {
    int r0, len; // I - tau v v^T acts on coordinates r0 .. r0 + len - 1
    double tau;
  };
  int n_ = 0, p_ = 0;
  std::vector<double> w_;       // eigenvalues, ascending
  std::vector<double> d0_, e0_; // the tridiagonal form
  std::vector<double> Qd_;      // p x p: the Ritz-part reflectors, accumulated
  std::vector<Refl> refl_;      // the band reduction, in order
  std::vector<double> vstore_;  // their vectors, one after the other (v[0] = 1)
};

} // namespace trl_amd::dense
