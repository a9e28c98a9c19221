/*
 * psi4amd_dfjk.h -- MI355X density-fitted J/K engine behind a C ABI.
 *
 * Drop-in for psi4's in-core DF-JK: the JK object API of
 * psi4/src/psi4/libfock/jk.h:232-589 with the MemDFJK algorithm
 * (libfock/MemDFJK.cc:56-162) and DFHelper::build_JK
 * (lib3index/dfhelper.h:329-331, dfhelper.cc:3015-3377).  Each entry point
 * names the reference interface it replaces.
 *
 * Conventions
 *   - Every function returns DFJK_OK (0) or a non-zero status; no C++
 *     exception crosses this boundary.  dfjk_last_error() gives the message
 *     of the last failure on the calling thread (psi4 raises PsiException,
 *     libpsi4util/exception.h:54, which its pybind layer turns into RuntimeError).
 *   - Matrices are row-major FP64 (psi4 Matrix::pointer()[0] layout).
 *   - Host-pointer entry points copy through device buffers owned by the
 *     handle; *_device entry points take device pointers and never copy to the host.
 *   - Calls on one handle are not reentrant; separate handles are independent.
 *   - All device work of a handle runs on one HIP stream (dfjk_set_stream).  The
 *     handle's own stream is a blocking stream, so it is ordered with work on the
 *     legacy default (null) stream; a caller on a non-blocking stream passes that
 *     stream.  *_device calls return with their work queued (host-pointer calls
 *     return after it finished).  Two cases wait for the stream: the structure test
 *     of a non-symmetric pair when the pair form is DFJK_FORM_DETECT (the launch
 *     shapes depend on its outcome; declare the form with dfjk_set_pair_form to
 *     skip it), and the adjoint, which returns E on the host.
 */
#ifndef PSI4AMD_DFJK_H
#define PSI4AMD_DFJK_H

#include <stddef.h>
#include <stdint.h>

#include "psi4amd_ints.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct dfjk_handle dfjk_handle;

enum dfjk_status {
    DFJK_OK = 0,
    DFJK_ERR_ARG = 1,      /* bad argument or shape (psi4: "JK: ... mismatch") */
    DFJK_ERR_STATE = 2,    /* call out of lifecycle order */
    DFJK_ERR_MEMORY = 3,   /* memory budget or device allocation */
    DFJK_ERR_DEVICE = 4,   /* HIP runtime error */
    DFJK_ERR_INTERNAL = 5
};

enum dfjk_precision {
    DFJK_PREC_FP64 = 0,    /* FP64 MFMA everywhere */
    DFJK_PREC_FP32 = 1,    /* K: FP32 MFMA, FP32 accumulate; J: FP32 copy of B, FP64 accumulate */
    DFJK_PREC_BF16X3 = 2   /* K: 3-term BF16 split products, FP32 accumulate; J: FP32 copy of B, FP64 accumulate */
};

enum dfjk_fitting {
    DFJK_FIT_CHOLESKY = 0, /* B = L^-1 (P|mn), J = L L^T (FP64 blocked Cholesky, default) */
    DFJK_FIT_EIGEN = 1     /* B = J^-1/2 (P|mn), J^-1/2 = V f(a) V^T with psi4's condition cut: eigenvalues
                            * |a| < condition * max|a| and non-finite powers give 0 (libmints/matrix.cc:2370-2424,
                            * applied by lib3index/dfhelper.cc:1491-1517); FP64 device eigensolver */
};

/* Structure of a non-symmetric density pair (no psi4 counterpart: psi4 always runs the general
 * two-transform path).  DETECT tests the columns on the device (one host round trip); a caller
 * that knows the structure declares it and the test is skipped. */
enum dfjk_pair_form {
    DFJK_FORM_DETECT = 0,   /* default: signed (Cr = Cl diag(+-1)), cross or general, found on the device */
    DFJK_FORM_GENERAL = 1,  /* no structure assumed: two half-transforms, every output tile */
    DFJK_FORM_CROSS = 2     /* Cl = [A, V], Cr = [V, A] (ncol even): dD = A V^T + V A^T, K symmetric */
};

enum dfjk_subalgorithm {      /* psi4 SCF_SUBTYPE (DFHelper subalgo_, dfhelper.h:370-376; jk.h:1039 set_subalgo) */
    DFJK_SUBALGO_AUTO = 0,        /* out of core when B (+ Bw) + metric + 4 GiB do not fit the free HBM */
    DFJK_SUBALGO_INCORE = 1,      /* B resident in HBM */
    DFJK_SUBALGO_OUT_OF_CORE = 2  /* B in pinned host memory, streamed through the device per compute */
};

const char* dfjk_last_error(void);

/* JK::build_JK(primary, aux, ...) for MEM_DF (libfock/jk.cc:206-230) -> MemDFJK ctor (MemDFJK.cc:56-64). */
int dfjk_create(dfjk_handle** out, const p4a_basis* primary, const p4a_basis* aux);
/* Tensor-level handle: nbf x nbf x naux fitted tensor supplied by dfjk_set_B_* or dfjk_fill_synthetic
 * (DFHelper with a caller-owned Ppq_, dfhelper.h:51-597). */
int dfjk_create_tensor(dfjk_handle** out, int nbf, int naux);
int dfjk_destroy(dfjk_handle* h);

/* Knobs (jk.h:375-420; MemDFJK::set_condition jk.h:1181-1187). */
int dfjk_set_memory(dfjk_handle* h, size_t doubles);       /* JK::set_memory: device workspace budget */
int dfjk_set_cutoff(dfjk_handle* h, double cutoff);        /* JK::set_cutoff: Schwarz cutoff (default 1e-12) */
int dfjk_set_condition(dfjk_handle* h, double condition);  /* MemDFJK::set_condition (default 1e-10) */
int dfjk_set_do_J(dfjk_handle* h, int do_J);               /* JK::set_do_J */
int dfjk_set_do_K(dfjk_handle* h, int do_K);               /* JK::set_do_K */
int dfjk_set_omp_nthread(dfjk_handle* h, int n);           /* JK::set_omp_nthread: host integral threads */
int dfjk_set_do_wK(dfjk_handle* h, int do_wK);             /* JK::set_do_wK (jk.h:459) */
int dfjk_set_subalgorithm(dfjk_handle* h, int subalgo);    /* enum dfjk_subalgorithm (DFHelper subalgo_) */
int dfjk_set_omega(dfjk_handle* h, double omega);          /* JK::set_omega (jk.h:473): erf(omega r)/r */
/* JK::set_wcombine (jk.h:463-467; MemDFJK.cc:83-88, 149-154): combined exchange.  With do_wK,
 * K is not built (its output is zeroed unless accumulating, dfhelper.cc:3022) and wK returns
 * omega_alpha * K + omega_beta * wK(omega) (dfhelper.cc:659-667, 1411, 3378-3438).  Effective at
 * the next wK compute; a tensor-level handle must set the attenuated tensor after it (and after B). */
int dfjk_set_wcombine(dfjk_handle* h, int wcombine);
/* Wall-clock ms of the set-up phases of the last dfjk_initialize, the counterpart of DFHelper's
 * timers (dfhelper.cc:659-690, "DFH: AO Construction", "DFH: AO-Met. Contraction", ...):
 * ms[0] Schwarz screening and tiles, [1] metric (P|Q), [2] metric factor, [3] (P|mn),
 * [4] fit, [5] attenuated tensor (wK); the first min(n, 6) are written. */
int dfjk_get_setup_timings(dfjk_handle* h, double* ms, int n);
int dfjk_set_omega_alpha(dfjk_handle* h, double alpha);   /* JK::set_omega_alpha (jk.h:478-481), default 1 */
int dfjk_set_omega_beta(dfjk_handle* h, double beta);     /* JK::set_omega_beta (jk.h:485-488), default 0 */
int dfjk_set_precision(dfjk_handle* h, int precision);     /* enum dfjk_precision (no psi4 counterpart) */
int dfjk_set_fitting(dfjk_handle* h, int fitting);         /* enum dfjk_fitting */
int dfjk_set_stream(dfjk_handle* h, void* hip_stream);     /* NULL: the handle's own stream */
int dfjk_set_profiling(dfjk_handle* h, int on);            /* per-kernel HIP-event timing */

/* JK::memory_estimate (MemDFJK.cc:65-69) -> DFHelper::get_core_size (dfhelper.h:95-98, dfhelper.cc:216-236):
 * naux * nsig + naux^2 + omp_nthread * nbf^2 + 3 * nbf^2 * Qshell_max doubles, where nsig is the number
 * of Schwarz-significant function pairs and Qshell_max the largest auxiliary shell; this is the figure
 * psi4's JK::build_JK compares with its memory to choose MEM_DF or DISK_DF (jk.cc:206-230).  As in
 * MemDFJK, wK is not counted (the helper's do_wK is set later, in preiterations).  Needs no device. */
int dfjk_memory_estimate(dfjk_handle* h, size_t* doubles);
/* Doubles of device memory this handle keeps resident: B (and Bw with wK), the FP32 copies of B
 * a compute at the current precision makes (half of B for FP32; for BF16X3 that and a
 * tile-transposed one), the metric during the fit, D/J/K staging and the packed C panels (no
 * psi4 counterpart).  A mixed-precision compute whose copies do not fit the free device memory
 * fails with DFJK_ERR_MEMORY and says so. */
int dfjk_device_footprint(dfjk_handle* h, size_t* doubles);
/* The Schwarz function-pair mask of DFHelper::prepare_sparsity (dfhelper.cc:299-420):
 * mask[m*nbf + n] = 1 when (mn|mn) >= cutoff^2 / max (mn|mn), else 0 (nbf x nbf bytes), and the
 * cutoff it was built with.  The (mn|mn) are computed on the device when one is present, by
 * the host engine otherwise (so this also works without a GPU); either output may be NULL. */
int dfjk_get_pair_mask(dfjk_handle* h, uint8_t* mask, double* cutoff);
/* The screening integrals behind that mask, as psi4's TwoBodyAOInt keeps them
 * (libmints/twobody.h:128-142, 232-238: function_pair_values_, shell_pair_value(), max_integral();
 * filled by dfhelper.cc:329-356): fun_values[m*nbf + n] = |(mn|mn)| (nbf x nbf),
 * shell_values[M*nshell + N] = its maximum over the shell pair (nshell x nshell), the overall
 * maximum, and 1 if the device computed them.  Basis-built handles only; any output may be NULL. */
int dfjk_get_schwarz(dfjk_handle* h, double* fun_values, double* shell_values, double* max_integral,
                     int* on_device);
/* The metric fit in use: enum dfjk_fitting, the condition, and the number of metric
 * eigenvalues kept (psi4's Matrix::power return value; naux for the Cholesky fit, 0 before
 * initialize).  Any output may be NULL. */
int dfjk_get_fit_info(dfjk_handle* h, int* fitting, double* condition, int* naux_kept);
/* Declared structure of non-symmetric pairs (enum dfjk_pair_form); applies to later computes. */
int dfjk_set_pair_form(dfjk_handle* h, int form);
/* The auxiliary range set by dfjk_set_aux_range (all of [0, naux) by default). */
int dfjk_get_aux_range(dfjk_handle* h, int* q0, int* q1);
/* JK::initialize -> MemDFJK::preiterations -> DFHelper::initialize (dfhelper.cc:149-215):
 * Schwarz screening, metric, host FP64 factorisation, (P|mn), GPU fit into HBM. */
int dfjk_initialize(dfjk_handle* h);
/* Tensor-level inputs: dense B [naux][nbf][nbf] (host or device, must be symmetric in m,n). */
int dfjk_set_B_dense(dfjk_handle* h, const double* B, int on_device);
/* Seeded synthetic B_Qmn = s_Q exp(-|m-n|/64) (G_Qmn+G_Qnm)/2, s_Q=(Q+1)^-1/2, G~N(0,1) (Philox4x32-10). */
int dfjk_fill_synthetic(dfjk_handle* h, uint64_t seed);
/* Read back B_Q for Q in [q0,q1) as dense [q1-q0][nbf][nbf] host array (tests). */
int dfjk_get_B_dense(dfjk_handle* h, int q0, int q1, double* out);
/* The attenuated tensor Bw = M (Q|erf(omega r)/r|mn) of range-separated exchange, fitted with
 * the Coulomb metric factor M of B (J^-1 = M^T M).  With a basis it is built by
 * dfjk_initialize when do_wK is set, or at the first wK compute after an omega change;
 * tensor-level handles take it from the caller.  Same dense layout as dfjk_set_B_dense. */
int dfjk_set_Bw_dense(dfjk_handle* h, const double* Bw, int on_device);
int dfjk_get_Bw_dense(dfjk_handle* h, int q0, int q1, double* out);

/* JK::compute (jk.cc:595-681) -> MemDFJK::compute_JK (MemDFJK.cc:97-111) -> DFHelper::build_JK.
 * n density pairs; Cr == NULL means lr_symmetric (C_right = C_left, jk.cc:597-602).
 * Cl[i], Cr[i]: nbf x ncol[i] row-major.  D/J/K outputs: nbf x nbf (any may be NULL).
 * J and K are overwritten unless accumulate != 0 (then J += ..., K += ...). */
int dfjk_compute(dfjk_handle* h, int n, const double* const* Cl, const double* const* Cr, const int* ncol,
                 double* const* D, double* const* J, double* const* K, int accumulate);
int dfjk_compute_device(dfjk_handle* h, int n, const double* const* Cl, const double* const* Cr, const int* ncol,
                        double* const* D, double* const* J, double* const* K, int accumulate);
/* No counterpart in the reference.  For callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc): dfjk_compute_device
 * with the same arguments and results, but every workspace buffer comes from one static
 * 128 MiB device array per GPU, held from the library's load until the process ends; no
 * hipMalloc, hipFree or allocating library routine is called.  Limits: B in core; for
 * DFJK_PREC_FP32 / DFJK_PREC_BF16X3 one dfjk_compute_device at that precision must have made
 * the lower-precision copy of B first; a system whose fixed-size buffers (D, J partials, one
 * K split, one auxiliary function's half-transform) exceed the array is refused with
 * DFJK_ERR_MEMORY before any work is queued (from about nbf 1150 at 300 columns).  Every handle on
 * a GPU shares that GPU's array, which keeps the last call's contents: calls that overlap on
 * one GPU (other threads, other streams) race. */
int dfjk_compute_device_noalloc(dfjk_handle* h, int n, const double* const* Cl, const double* const* Cr,
                                const int* ncol, double* const* D, double* const* J, double* const* K,
                                int accumulate);
/* JK::wK() of JK::compute with do_wK (jk.cc:382-423) -> DFHelper::compute_wK
 * (dfhelper.cc:3378-3438), MemDFJK's non-combined path (wcombine off, MemDFJK.cc:81-90):
 *   wK = sum_Q (B_Q Cl)(Bw_Q Cr)^T = sum_{PQ} ((P|mn) Cl) [J^-1]_PQ ((Q|erf(omega r)/r|mn) Cr)^T,
 * or omega_alpha K + omega_beta of that with wcombine.  FP64 whatever the precision setting.
 * Cr == NULL means Cr = Cl (lr_symmetric): wK is then hermitivitized, (wK + wK^T)/2, as
 * MemDFJK::compute_JK does (MemDFJK.cc:101-107); otherwise it is not symmetric in general.
 * wK[i]: nbf x nbf, overwritten unless accumulate != 0. */
int dfjk_compute_wK(dfjk_handle* h, int n, const double* const* Cl, const double* const* Cr, const int* ncol,
                    double* const* wK, int accumulate);
int dfjk_compute_wK_device(dfjk_handle* h, int n, const double* const* Cl, const double* const* Cr, const int* ncol,
                           double* const* wK, int accumulate);
/* Restrict the auxiliary index range used by compute (J and K are additive over Q). */
int dfjk_set_aux_range(dfjk_handle* h, int q0, int q1);

/* Adjoint of E = tr(D(2J - K)), D = C C^T, in FP64 whatever the precision setting.  psi4 has
 * no single counterpart: its DF-SCF gradient (scfgrad/jk_grad.cc:294-475, 634-832) contracts
 * the same intermediates, c_Q = (Q|mn) D_mn and (Q|ij) = C^T (Q|mn) C.
 *   E    = sum_mn D_mn (2J - K)_mn                       (host scalar; may be NULL)
 *   dEdC = 4 (2J - K) C                                  (nbf x nocc; may be NULL)
 *   dEdB[Q][m][n] = 4 d_Q D_mn - 2 (D B_Q D)_mn          (may be NULL)
 * E and dEdC use every auxiliary function; dEdB covers Q in the range set by
 * dfjk_set_aux_range (all by default) and has (q1 - q0) x nbf x nbf elements.  dEdB is the
 * derivative with respect to each element of the dense symmetric B: a perturbation of the
 * stored pair (m,n),(n,m) together changes E by dEdB[Q][m][n] + dEdB[Q][n][m].
 * E and dEdC also work out of core (they come from J and K); dEdB needs B in core and is
 * refused (DFJK_ERR_STATE) out of core.
 * dfjk_energy_adjoint takes host arrays, the _device variant device arrays (E stays host). */
int dfjk_energy_adjoint(dfjk_handle* h, const double* C, int nocc, double* E, double* dEdC, double* dEdB);
int dfjk_energy_adjoint_device(dfjk_handle* h, const double* C, int nocc, double* E, double* dEdC, double* dEdB);

/* Timing of the last compute when profiling is on: milliseconds per kernel class and launch counts. */
int dfjk_get_timings(dfjk_handle* h, double* ms_j, double* ms_ht, double* ms_kgemm, double* ms_other, int* n_ht,
                     int* n_kgemm);
/* Sizes: nbf, naux, stored tile pairs, significant function pairs. */
int dfjk_get_info(dfjk_handle* h, int* nbf, int* naux, int* ntile_pairs, int64_t* nsig_pairs);

/* Whether initialize chose the out-of-core subalgorithm (MemDFJK print_header "Algorithm: Core/Disk",
 * MemDFJK.cc:124, from DFHelper::get_AO_core). */
int dfjk_get_out_of_core(dfjk_handle* h, int* ooc);

/* JK::finalize (jk.cc) -> postiterations: frees the workspace, keeps B. */
int dfjk_finalize(dfjk_handle* h);

#ifdef __cplusplus
}
#endif
#endif
