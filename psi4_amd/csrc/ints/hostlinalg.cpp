// Host FP64 dense linear algebra for the metric factorisation (row-major).
#include "hostlinalg.hpp"
#include <omp.h>

#include <algorithm>
#include <cmath>
#include <vector>

namespace p4a {

// In-place lower Cholesky A = L L^T (blocked, right-looking).  The strict upper
// triangle is zeroed.  Returns the index of the first non-positive pivot, or -1.
int cholesky_lower(double* A, int n, int nthreads) {
    const int NB = 64;
    for (int k0 = 0; k0 < n; k0 += NB) {
        const int k1 = std::min(n, k0 + NB);
        for (int j = k0; j < k1; ++j) {
            double d = A[static_cast<size_t>(j) * n + j];
            for (int p = k0; p < j; ++p) d -= A[static_cast<size_t>(j) * n + p] * A[static_cast<size_t>(j) * n + p];
            if (!(d > 0.0)) return j;
            d = std::sqrt(d);
            A[static_cast<size_t>(j) * n + j] = d;
            for (int i = j + 1; i < k1; ++i) {
                double s = A[static_cast<size_t>(i) * n + j];
                for (int p = k0; p < j; ++p) s -= A[static_cast<size_t>(i) * n + p] * A[static_cast<size_t>(j) * n + p];
                A[static_cast<size_t>(i) * n + j] = s / d;
            }
        }
#pragma omp parallel for schedule(static) num_threads(nthreads)
        for (int i = k1; i < n; ++i) {
            double* Ai = A + static_cast<size_t>(i) * n;
            for (int j = k0; j < k1; ++j) {
                const double* Aj = A + static_cast<size_t>(j) * n;
                double s = Ai[j];
                for (int p = k0; p < j; ++p) s -= Ai[p] * Aj[p];
                Ai[j] = s / Aj[j];
            }
        }
#pragma omp parallel for schedule(dynamic, 8) num_threads(nthreads)
        for (int i = k1; i < n; ++i) {
            double* Ai = A + static_cast<size_t>(i) * n;
            for (int j = k1; j <= i; ++j) {
                const double* Aj = A + static_cast<size_t>(j) * n;
                double s = 0.0;
#pragma omp simd reduction(+ : s)
                for (int p = k0; p < k1; ++p) s += Ai[p] * Aj[p];
                Ai[j] -= s;
            }
        }
    }
    for (int i = 0; i < n; ++i)
        for (int j = i + 1; j < n; ++j) A[static_cast<size_t>(i) * n + j] = 0.0;
    return -1;
}

// Inverse of a lower-triangular L (row-major) into Linv (row-major, lower).
void lower_inverse(const double* L, int n, double* Linv, int nthreads) {
    std::fill(Linv, Linv + static_cast<size_t>(n) * n, 0.0);
#pragma omp parallel num_threads(nthreads)
    {
        std::vector<double> x(n);
#pragma omp for schedule(dynamic, 4)
        for (int j = 0; j < n; ++j) {
            x[j] = 1.0 / L[static_cast<size_t>(j) * n + j];
            for (int i = j + 1; i < n; ++i) {
                const double* Li = L + static_cast<size_t>(i) * n;
                double s = 0.0;
#pragma omp simd reduction(+ : s)
                for (int k = j; k < i; ++k) s += Li[k] * x[k];
                x[i] = -s / Li[i];
            }
            for (int i = j; i < n; ++i) Linv[static_cast<size_t>(i) * n + j] = x[i];
        }
    }
}

}  // namespace p4a
