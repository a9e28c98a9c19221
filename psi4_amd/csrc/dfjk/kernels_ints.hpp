// Device three-centre Coulomb integrals (P|mn) (see kernels_ints.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>

namespace p4a {

struct GpuShell {
    int l, nprim, p0, foff;  // angular momentum, primitives, first primitive, first function
    double A[3];             // centre (bohr)
};

// The tabulated Boys grid of the host engine (F_n(i h) for i < nt, n < ncol).
struct BoysGrid {
    const double* table = nullptr;
    int ncol = 0;
    double h = 0.0, inv_h = 0.0, tmax = 0.0;
};

struct ThreeIndexArgs {
    const GpuShell* prim = nullptr;   // primary shells
    const GpuShell* aux = nullptr;    // auxiliary shells
    const double* pexp = nullptr;     // primary exponents / normalised contraction coefficients
    const double* pcoef = nullptr;
    const double* aexp = nullptr;
    const double* acoef = nullptr;
    const double* c2s = nullptr;      // Cartesian -> solid harmonic rows, l = 0..4
    int c2s_off[5] = {0, 0, 0, 0, 0};
    int prim_pure = 1, aux_pure = 1;
    BoysGrid boys;
    const int64_t* dst = nullptr;     // [nbf*nbf] tiled-layout position of (m,n) or -1; null: the value
                                      // of (m, n) goes to column m (pairs of a shell with a unit s
                                      // shell, as for the two-centre metric (P|Q))
    int64_t p_stride = 0;             // elements per auxiliary function in out
    int nbf = 0;
    double* out = nullptr;            // [naux][p_stride]
    double omega = 0.0;               // > 0: erf(omega r)/r
};

// One (la, lb | lc) class, la >= lb (la, lb <= 3, lc <= 4): pairs[i] = (shell with la, shell with lb);
// afun[j] = (auxiliary shell, function within the shell).  Throws for unsupported classes.
void launch_three_index_class(const ThreeIndexArgs& a, int la, int lb, int lc, const int2* pairs, int npair,
                              const int2* afun, int nafun, hipStream_t s);

}  // namespace p4a
