// Host-side Gaussian integral engine for the DF-JK path.
//
// psi4 takes these integrals from Libint2 (psi4/src/psi4/libmints/eri.cc:44-84,
// eribase.cc:220-245); Libint2 is a third-party dependency that is not
// available here, so this engine computes the same quantities from first
// principles with the McMurchie-Davidson scheme:
//   * (P|mn)  three-centre Coulomb integrals       (dfhelper.cc:1284-1347)
//   * (P|Q)   two-centre Coulomb metric            (fittingmetric.cc:115-155)
//   * (mn|mn) diagonal quartets for Schwarz masks  (dfhelper.cc:299-350)
//   * S, T, V one-electron integrals for the RHF caller.
// Conventions follow psi4: contraction normalisation as in
// libmints/gshell.cc:60-145 (x^l component of each contracted shell has unit
// norm), Cartesian order xx,xy,xz,yy,yz,zz, and real solid harmonics in the
// "gaussian" order m = 0,+1,-1,+2,-2,... (CMakeLists.txt:208, core.cc:1118-1125).
#pragma once
#include <cstdint>
#include <vector>

namespace p4a {

constexpr int kMaxShellL = 6;                 // up to i functions
constexpr int kMaxTotalL = 4 * kMaxShellL;    // (ii|ii)

struct Shell {
    int l = 0;
    bool pure = true;
    double A[3] = {0, 0, 0};
    std::vector<double> exps;
    std::vector<double> coefs;  // normalised: primitive norm (x^l convention) x contraction norm
    int ncart() const { return (l + 1) * (l + 2) / 2; }
    int nfunc() const { return pure ? 2 * l + 1 : ncart(); }
};

struct Basis {
    std::vector<Shell> shells;
    std::vector<int> offset;  // first function of each shell
    int nbf = 0;
    int max_l = 0;
    int max_nfunc = 0;
    void finalize();
};

// Normalise original contraction coefficients in place (gshell.cc:75-145).
void normalize_shell(Shell& s);

// Cartesian -> real solid harmonic matrix for angular momentum l:
// row-major (2l+1) x ncart, rows in gaussian order m = 0,+1,-1,...
const std::vector<double>& cart2sph(int l);

// Boys function F_n(T) for n = 0..nmax written into F[0..nmax].
void boys(int nmax, double T, double* F);
// The tabulated grid behind boys(): F_n(i h) at [i * ncol + n] for i < nt; Taylor expansion
// of order 6 about the nearest grid point below tmax, asymptotic form above.
const double* boys_grid(int* nt, int* ncol, double* h, double* tmax);

// ---- integral drivers (all outputs row-major, FP64) ----
void overlap(const Basis& b, double* S);
void kinetic(const Basis& b, double* T);
// <mu| r_d - origin_d |nu> for d = x, y, z into M[3][nbf][nbf] (no electron charge).
void dipole(const Basis& b, const double* origin, double* M);
// <mu| (r - origin) x nabla |nu> for x, y, z into M[3][nbf][nbf] (antisymmetric; no factor -i).
void angular_momentum(const Basis& b, const double* origin, double* M);
void potential(const Basis& b, int natom, const double* Z, const double* xyz, double* V);
// Two-electron drivers take omega: 0 for the Coulomb operator 1/r, > 0 for the attenuated
// erf(omega r)/r of range-separated exchange (psi4's erf_eri, libmints/integral.h).
// (P|Q), naux x naux.
void metric(const Basis& aux, double* J, int nthreads, double omega = 0.0);
// (P|mn): out[P * p_stride + dst[m * nbf + n]] for every (m,n) whose dst >= 0.
// Shell pairs (MU,NU) with shell_mask[MU*nshell+NU] == 0 are skipped (mask may be null).
void three_index(const Basis& prim, const Basis& aux, const int64_t* dst, int64_t p_stride,
                 const uint8_t* shell_mask, double* out, int nthreads, double omega = 0.0);
// Schwarz diagonal (mn|mn) for all function pairs: fun_max nbf*nbf (both triangles),
// shell_max nshell*nshell, following dfhelper.cc:322-366. Returns max value.
double schwarz_diagonal(const Basis& prim, double* fun_max, double* shell_max, int nthreads);
// ---- gradient contractions (derivatives with respect to shell centres) ----
// g_shell[I][xyz] += sum_mn W_mn dS_mn/dR_I  (W symmetric), likewise T with D.
void overlap_grad(const Basis& b, const double* W, double* g_shell);
void kinetic_grad(const Basis& b, const double* D, double* g_shell);
// V: basis-centre part into g_shell, operator (nuclear position) part into g_nuc[natom][3].
void potential_grad(const Basis& b, const double* D, int natom, const double* Z, const double* xyz, double* g_shell,
                    double* g_nuc);
// sum_{P,m,n} G[P][m][n] d(P|mn)/dR, G symmetric in (m,n): into g_prim[nshell][3], g_aux[naux_shell][3].
void three_index_grad(const Basis& prim, const Basis& aux, const double* G, double* g_prim, double* g_aux,
                      int nthreads);
// sum_{P,Q} W[P][Q] d(P|Q)/dR, W symmetric: into g_aux[naux_shell][3].
void metric_grad(const Basis& aux, const double* W, double* g_aux, int nthreads);
// Full (mn|ls) over a basis, nbf^4, for brute-force checks of tiny systems.
void four_index(const Basis& b, double* out, int nthreads, double omega = 0.0);

}  // namespace p4a
