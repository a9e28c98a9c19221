// Host FP64 dense linear algebra for the metric factorisation (row-major).
#pragma once

namespace p4a {

// In-place lower Cholesky A = L L^T (blocked, right-looking).  The strict upper
// triangle is zeroed.  Returns the index of the first non-positive pivot, or -1.
int cholesky_lower(double* A, int n, int nthreads);

// Inverse of a lower-triangular L (row-major) into Linv (row-major, lower).
void lower_inverse(const double* L, int n, double* Linv, int nthreads);

}  // namespace p4a
