#pragma once
// Host FP64 dense linear algebra of the Rayleigh-Ritz step.

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

} // namespace trl_amd::dense
