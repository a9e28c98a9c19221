/*
 * psi4amd_ints.h -- host integral engine behind a C ABI.
 *
 * In psi4 these integrals come from Libint2 through libmints
 * (psi4/src/psi4/libmints/eri.cc:44-84, twobody.h:65-333) and are consumed by
 * DFHelper (lib3index/dfhelper.cc:299-420 for Schwarz, 1284-1347 for (P|mn))
 * and FittingMetric (lib3index/fittingmetric.cc:115-155 for (P|Q)).  The DF-JK
 * library (psi4amd_dfjk.h) calls the same engine internally; these entry
 * points exist so that callers and tests can obtain the host-side inputs.
 *
 * Conventions: FP64, row-major, Cartesian coordinates in bohr, original
 * (unnormalised) contraction coefficients as listed in a .gbs file;
 * normalisation follows libmints/gshell.cc:60-145.  Every function returns 0
 * on success and a non-zero status on failure (message via p4a_ints_last_error).
 */
#ifndef PSI4AMD_INTS_H
#define PSI4AMD_INTS_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* A basis set as flat shell arrays (psi4: BasisSet / GaussianShell). */
typedef struct p4a_basis {
    int nshell;
    const int* l;          /* [nshell] angular momentum */
    const int* nprim;      /* [nshell] primitives per shell */
    const double* center;  /* [nshell*3] bohr */
    const double* exps;    /* [sum nprim] exponents */
    const double* coefs;   /* [sum nprim] original contraction coefficients */
    int puream;            /* 1: real solid harmonics (gaussian order), 0: Cartesian */
} p4a_basis;

const char* p4a_ints_last_error(void);
int p4a_ints_nbf(const p4a_basis* b, int* nbf);
/* S, T, V (nbf x nbf).  V uses point charges Z at xyz [natom*3]. */
int p4a_ints_overlap(const p4a_basis* b, double* S);
int p4a_ints_kinetic(const p4a_basis* b, double* T);
int p4a_ints_potential(const p4a_basis* b, int natom, const double* Z, const double* xyz, double* V);
/* <mu| r_d - origin_d |nu>, d = x, y, z, into M [3][nbf][nbf]; origin NULL = (0,0,0).
 * psi4's MintsHelper::ao_dipole (libmints/mintshelper.cc:1627-1645) returns these times the
 * electron charge, -1. */
int p4a_ints_dipole(const p4a_basis* b, const double* origin, double* M);
/* <mu| (r - origin) x nabla |nu>, x, y, z, into M [3][nbf][nbf] (antisymmetric, no factor -i);
 * the quantity of MintsHelper::ao_angular_momentum (libmints/mintshelper.cc:1613-1625). */
int p4a_ints_angular_momentum(const p4a_basis* b, const double* origin, double* M);
/* (P|Q) metric, naux x naux (fittingmetric.cc:115-155). */
int p4a_ints_metric(const p4a_basis* aux, double* J, int nthreads);
/* (P|mn) written to out[P*p_stride + dst[m*nbf+n]] for every dst >= 0;
 * shell pairs with shell_mask[M*nshell+N]==0 are skipped (mask may be NULL). */
int p4a_ints_three_index(const p4a_basis* prim, const p4a_basis* aux, const int64_t* dst, int64_t p_stride,
                         const uint8_t* shell_mask, double* out, int nthreads);
/* Gradient contractions for the DF-SCF gradient (scfgrad/scf_grad.cc:103-262,
 * jk_grad.cc:834-1000): derivatives with respect to the centre of each shell, added to
 * g_shell [nshell*3] (and, for the nuclear attraction operator, to g_nuc [natom*3]).
 *   overlap_grad:     sum_mn W_mn dS_mn/dR          (W symmetric, e.g. the energy-weighted density)
 *   kinetic_grad:     sum_mn D_mn dT_mn/dR
 *   potential_grad:   sum_mn D_mn dV_mn/dR, including the moving nuclei
 *   three_index_grad: sum_{P,mn} G[P][m][n] d(P|mn)/dR  (G [naux][nbf][nbf], symmetric in m,n)
 *   metric_grad:      sum_{PQ} W[P][Q] d(P|Q)/dR        (W symmetric) */
int p4a_ints_overlap_grad(const p4a_basis* b, const double* W, double* g_shell);
int p4a_ints_kinetic_grad(const p4a_basis* b, const double* D, double* g_shell);
int p4a_ints_potential_grad(const p4a_basis* b, const double* D, int natom, const double* Z, const double* xyz,
                            double* g_shell, double* g_nuc);
int p4a_ints_three_index_grad(const p4a_basis* prim, const p4a_basis* aux, const double* G, double* g_prim,
                              double* g_aux, int nthreads);
int p4a_ints_metric_grad(const p4a_basis* aux, const double* W, double* g_aux, int nthreads);
/* Schwarz diagonal (mn|mn): fun_max [nbf*nbf], shell_max [nshell*nshell], max_val
 * (dfhelper.cc:322-366). */
int p4a_ints_schwarz(const p4a_basis* prim, double* fun_max, double* shell_max, double* max_val, int nthreads);
/* Full (mn|ls), nbf^4, for brute-force checks on tiny systems. */
int p4a_ints_four_index(const p4a_basis* b, double* out, int nthreads);
/* The same integrals over the attenuated operator erf(omega r)/r, omega > 0
 * (IntegralFactory::erf_eri, libmints/integral.h:498, integral.cc:228): the metric,
 * the three-index (P|mn) with the layout arguments of p4a_ints_three_index, and the
 * four-index brute-force set. */
int p4a_ints_metric_erf(const p4a_basis* aux, double omega, double* J, int nthreads);
int p4a_ints_three_index_erf(const p4a_basis* prim, const p4a_basis* aux, double omega, const int64_t* dst,
                             int64_t p_stride, const uint8_t* shell_mask, double* out, int nthreads);
int p4a_ints_four_index_erf(const p4a_basis* b, double omega, double* out, int nthreads);
/* Boys function F_0..F_nmax at T. */
int p4a_ints_boys(int nmax, double T, double* F);

#ifdef __cplusplus
}
#endif
#endif
