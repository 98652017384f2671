#pragma once
// Host FP64 dense linear algebra of the Rayleigh-Ritz step.

namespace trl_amd::dense {

// Symmetric eigen-decomposition of the n x n row-major symmetric matrix A:
// ascending eigenvalues in w[0..n), eigenvector j in column j of the
// row-major n x n matrix Z.  Returns 0, or -1 if the QR iteration did not
// converge.
int syev(int n, const double* A, double* w, double* Z);

} // namespace trl_amd::dense
