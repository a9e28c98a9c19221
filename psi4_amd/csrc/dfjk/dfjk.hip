// MI355X DF-JK engine: host orchestration and the C ABI (include/psi4amd_dfjk.h).
//
// Lifecycle mirrors psi4's JK object (libfock/jk.h:232-589) with the MemDFJK
// algorithm: create (JK::build_JK, jk.cc:206-230) -> knobs -> initialize
// (MemDFJK::preiterations -> DFHelper::initialize, dfhelper.cc:149-215) ->
// repeated compute (JK::compute, jk.cc:595-681 -> DFHelper::build_JK,
// dfhelper.cc:3015-3377) -> finalize.
#include <hip/hip_runtime.h>
#include <thread>

#include <algorithm>
#include <chrono>
#include <cstdlib>
#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include <rocsolver/rocsolver.h>

#include "../ints/basis_build.hpp"
#include "../ints/ints.hpp"
#include "../ints/hostlinalg.hpp"
#include "kernels.hpp"
#include "kernels_ints.hpp"
#include "kernels_k.hpp"
#include "kernels_schwarz.hpp"
#include "psi4amd_dfjk.h"

// Workspace of dfjk_compute_device_noalloc: one array per GPU, held from the library's load
// until the process ends.  The name below is a host-side stand-in, never a device address:
// host code uses only the address hipGetSymbolAddress returns for each GPU (see
// scratch_of_device), kept per device.  Every handle on a GPU shares that GPU's copy.
constexpr size_t kScratchBytes = size_t(128) << 20;
static __device__ __align__(256) char scratch[kScratchBytes];

namespace p4a {
namespace {

thread_local std::string g_err;

struct StatusError : std::runtime_error {
    int code;
    StatusError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

[[noreturn]] void fail(int code, const std::string& msg) { throw StatusError(code, msg); }

template <class F>
int guarded(F&& f) {
    try {
        f();
        return DFJK_OK;
    } catch (const StatusError& e) {
        g_err = e.what();
        return e.code;
    } catch (const std::bad_alloc&) {
        g_err = "host allocation failed";
        return DFJK_ERR_MEMORY;
    } catch (const std::exception& e) {
        g_err = e.what();
        return DFJK_ERR_INTERNAL;
    }
}

void hcheck(hipError_t e, const char* what) {
    if (e != hipSuccess) {
        if (e == hipErrorOutOfMemory) fail(DFJK_ERR_MEMORY, std::string(what) + ": device out of memory");
        fail(DFJK_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
    }
}

// A copy between device memory and pageable host memory (std::vector, caller arrays) that
// neither overtakes the work queued on s nor outlives the call: s is drained first, the copy
// is synchronous, and the device is idle afterwards, whatever the flags of s (a torch stream
// does not synchronise with the legacy stream) and however the runtime stages pageable memory.
// An asynchronous copy on s was seen to read a block of the metric before the upload queued
// ahead of it had landed when kernel dispatches were traced.
void copy_settled(void* dst, const void* src, size_t bytes, hipMemcpyKind kind, hipStream_t s, const char* what) {
    hcheck(hipStreamSynchronize(s), what);
    hcheck(hipMemcpy(dst, src, bytes, kind), what);
    hcheck(hipDeviceSynchronize(), what);
}
void copy2d_settled(void* dst, size_t dpitch, const void* src, size_t spitch, size_t width, size_t rows,
                    hipMemcpyKind kind, hipStream_t s, const char* what) {
    hcheck(hipStreamSynchronize(s), what);
    hcheck(hipMemcpy2D(dst, dpitch, src, spitch, width, rows, kind), what);
    hcheck(hipDeviceSynchronize(), what);
}

// Growable device buffer.
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    template <class T>
    T* get(size_t count) {
        const size_t need = std::max<size_t>(count * sizeof(T), 256);
        if (need > bytes) {
            release();
            hcheck(hipMalloc(&p, need), "hipMalloc");
            bytes = need;
        }
        return static_cast<T*>(p);
    }
    template <class T>
    T* as() const {
        return static_cast<T*>(p);
    }
};

inline int round_up(int a, int b) { return (a + b - 1) / b * b; }

struct EventTimer {
    bool on = false;
    hipStream_t s = nullptr;
    struct Rec {
        int cls;
        hipEvent_t a, b;
    };
    std::vector<Rec> recs;
    std::vector<hipEvent_t> pool;
    hipEvent_t take() {
        if (!pool.empty()) {
            hipEvent_t e = pool.back();
            pool.pop_back();
            return e;
        }
        hipEvent_t e;
        hcheck(hipEventCreate(&e), "hipEventCreate");
        return e;
    }
    hipEvent_t begin() {
        if (!on) return nullptr;
        hipEvent_t e = take();
        hcheck(hipEventRecord(e, s), "hipEventRecord");
        return e;
    }
    void end(int cls, hipEvent_t a) {
        if (!on) return;
        hipEvent_t b = take();
        hcheck(hipEventRecord(b, s), "hipEventRecord");
        recs.push_back({cls, a, b});
    }
    ~EventTimer() {
        for (auto& r : recs) {
            (void)hipEventDestroy(r.a);
            (void)hipEventDestroy(r.b);
        }
        for (auto e : pool) (void)hipEventDestroy(e);
    }
};

enum TimerClass { T_J = 0, T_HT = 1, T_KGEMM = 2, T_OTHER = 3 };

}  // namespace
}  // namespace p4a

using namespace p4a;

struct dfjk_handle {
    // configuration
    bool tensor_mode = false;
    Basis prim, aux;
    int nbf = 0, naux = 0;
    size_t memory_doubles = 0;  // 0: no explicit budget
    double cutoff = 1e-12, condition = 1e-10;
    bool do_J = true, do_K = true;
    bool do_wK = false;
    double omega = 0.0;
    int nthreads = 1;      // host integral threads
    int omp_nthread = 1;   // psi4's omp_nthread_ (jk.cc:254-257, default one thread): memory estimate
    int precision = DFJK_PREC_FP64;
    int fitting = DFJK_FIT_CHOLESKY;
    int pair_form = DFJK_FORM_DETECT;
    bool fit_lower = true;  // the metric factor is lower triangular (Cholesky): the fit skips its upper part
    int fit_kept = 0;       // metric eigenvalues kept by the fit (naux for Cholesky)
    bool initialized = false;
    int aux_q0 = 0, aux_q1 = 0;
    int device = -1;              // the HIP device current at initialize (its buffers live there)
    char* scratch_dev = nullptr;  // that device's address of the static scratch array
    // stream
    hipStream_t own_stream = nullptr;
    hipStream_t stream = nullptr;
    // screening and tile geometry
    std::vector<uint8_t> fun_mask;  // nbf*nbf
    std::vector<uint8_t> shell_mask;
    int64_t nsig = 0;
    int nT = 0, nbfp = 0, ntp = 0;
    std::vector<int> tidx;
    std::vector<int2> tiles;
    DevBuf d_tidx, d_tiles, d_mask;
    // output 128-tile lists
    int nB = 0;
    DevBuf d_otiles_sym, d_otiles_all, d_otile_of_sym, d_otile_of_all;
    std::vector<int2> otiles_sym, otiles_all;
    std::vector<int> otile_of_sym, otile_of_all;
    int ntile_sym = 0, ntile_all = 0;
    // resident tensor and its lower-precision copies
    DevBuf dB, dB32, dB32T;
    DevBuf dZero;  // one zero TILE x TILE FP64 block, read in place of absent blocks
    // out of core (DFHelper AO_core false, dfhelper.cc:216-260): B lives in pinned host memory
    // and each compute streams it through two device stage buffers, Q-block by Q-block
    int subalgo = DFJK_SUBALGO_AUTO;
    bool ooc = false;
    double* hB = nullptr;               // pinned host B, b_doubles
    hipStream_t cstream = nullptr;      // copy stream (default priority, ordered by events)
    hipEvent_t ev_copied[2] = {nullptr, nullptr}, ev_free[2] = {nullptr, nullptr};
    DevBuf ws_stage[2], ws_stage32, ws_stage32t;
    // range-separated exchange: Bw = M (P|erf(omega r)/r|mn), same layout and fit as B
    DevBuf dBw;
    double bw_omega = -1.0;         // omega dBw was built for (< 0: none)
    // combined exchange (wcombine): dBw then holds alpha B + beta Bw, so that the wK build
    // returns alpha K + beta wK; bw_* record what dBw holds
    bool wcombine = false;
    double omega_alpha = 1.0, omega_beta = 0.0;
    // wall-clock ms of the set-up phases of the last initialize (psi4's DFH timers):
    // screening + tiles, metric (P|Q), metric factor, (P|mn), fit, attenuated tensor
    double setup_ms[6] = {0, 0, 0, 0, 0, 0};
    bool dM_ready = false;  // ws_G holds the metric factor M (left by the device factorisation)
    bool metric_on_device = false;  // the last initialize computed (P|Q) on the device
    // the Schwarz diagonal the masks were built from: |(mn|mn)| [nbf][nbf], its shell maxima
    // [nshell][nshell] and overall maximum, and whether the device computed it
    std::vector<double> schwarz_fun, schwarz_shell;
    double schwarz_max = 0.0;
    bool schwarz_on_device = false;
    bool bw_comb = false;
    double bw_alpha = 1.0, bw_beta = 0.0;
    std::vector<double> fit_M;      // host metric factor, kept for rebuilding Bw after an omega change
    // workspace
    DevBuf ws_craw, ws_cl, ws_cr, ws_D, ws_W, ws_dpart, ws_d, ws_Jpart, ws_XL, ws_XR, ws_P, ws_Pd, ws_J, ws_K, ws_G, ws_sign, ws_c32, ws_cbf, ws_cmap;
    DevBuf ws_ct, ws_red, ws_Y, ws_PC, ws_gam, ws_dc;
    // column map of the half-transform: the device copy is re-uploaded (from pinned memory,
    // asynchronously) only when it changes
    std::vector<int> cmap_dev;          // what ws_cmap holds
    int* cmap_pin = nullptr;            // pinned staging of the last upload
    size_t cmap_pin_cap = 0;
    hipEvent_t ev_cmap = nullptr;       // that upload has been read
    int8_t* sign_pin = nullptr;         // pinned copy of the column-structure test
    size_t sign_pin_cap = 0;
    // timing
    EventTimer timer;
    double t_ms[4] = {0, 0, 0, 0};
    int n_ht = 0, n_kgemm = 0;

    TileGeom geom() const {
        TileGeom g;
        g.nbf = nbf;
        g.nbfp = nbfp;
        g.nT = nT;
        g.ntp = ntp;
        g.d_tidx = d_tidx.as<int>();
        g.d_tiles = d_tiles.as<int2>();
        return g;
    }
};

namespace {

void require_state(dfjk_handle* h, bool init_needed) {
    if (!h) fail(DFJK_ERR_ARG, "null handle");
    if (init_needed && !h->initialized) fail(DFJK_ERR_STATE, "dfjk: call initialize() before compute()");
}

// The handle's own stream is created on first device use, so that creating a
// handle, setting knobs and asking for the memory estimate need no device.
// The library's own stream is a blocking stream: work a caller queues on the legacy
// default (null) stream -- torch's default stream -- is ordered before and after it.
void ensure_stream(dfjk_handle* h) {
    if (!h->own_stream) hcheck(hipStreamCreateWithFlags(&h->own_stream, hipStreamDefault), "hipStreamCreate");
    if (!h->stream) h->stream = h->own_stream;
    h->timer.s = h->stream;
}

// Tile geometry from the function-pair mask (tile-granular screening).
void build_tiles(dfjk_handle* h) {
    const int n = h->nbf;
    h->nbfp = round_up(n, TILE);
    h->nT = h->nbfp / TILE;
    const int nT = h->nT;
    h->tidx.assign(static_cast<size_t>(nT) * nT, -1);
    h->tiles.clear();
    for (int I = 0; I < nT; ++I)
        for (int J = I; J < nT; ++J) {
            bool any = false;
            for (int m = I * TILE; m < std::min(n, (I + 1) * TILE) && !any; ++m)
                for (int q = J * TILE; q < std::min(n, (J + 1) * TILE); ++q)
                    if (h->fun_mask[static_cast<size_t>(m) * n + q]) {
                        any = true;
                        break;
                    }
            if (any) {
                h->tidx[I * nT + J] = static_cast<int>(h->tiles.size());
                h->tiles.push_back(make_int2(I, J));
            }
        }
    h->ntp = static_cast<int>(h->tiles.size());
    if (h->ntp == 0) fail(DFJK_ERR_ARG, "no significant basis-function pairs");
    // 128x128 output tiles for K/D products
    h->nB = (n + 127) / 128;
    const int nB = h->nB;
    std::vector<int2> sym, all;
    std::vector<int> of_sym(static_cast<size_t>(nB) * nB, -1), of_all(static_cast<size_t>(nB) * nB, -1);
    for (int I = 0; I < nB; ++I)
        for (int J = 0; J < nB; ++J) {
            of_all[I * nB + J] = static_cast<int>(all.size());
            all.push_back(make_int2(I, J));
            if (I <= J) {
                of_sym[I * nB + J] = static_cast<int>(sym.size());
                sym.push_back(make_int2(I, J));
            }
        }
    h->ntile_sym = static_cast<int>(sym.size());
    h->ntile_all = static_cast<int>(all.size());
    h->otiles_sym = sym;
    h->otiles_all = all;
    h->otile_of_sym = of_sym;
    h->otile_of_all = of_all;
}

int g_kgemm_wgs = 2048;  // workgroups the Q split of the contraction aims at (tuning: PSI4AMD_KGEMM_WGS)
size_t g_x_budget = size_t(16) << 30;       // FP64 X workspace bytes without a memory setting
size_t g_x_budget_mixed = size_t(4) << 30;  // the same for FP32 / BF16x3 (both: PSI4AMD_X_BUDGET_GB)

// Workspace bytes under the memory setting: what it leaves beyond the resident bytes, capped
// by want.  A setting below the resident bytes sizes host memory (psi4 passes its own memory
// option here), not this device, and leaves want unchanged.
size_t memory_limited(const dfjk_handle* h, size_t want, size_t resident) {
    if (!h->memory_doubles) return want;
    const size_t mem = h->memory_doubles * sizeof(double);
    return mem > resident ? std::min(want, mem - resident) : want;
}

void upload_tiles(dfjk_handle* h) {
    hcheck(hipMemset(h->dZero.get<double>(TILE2), 0, TILE2 * sizeof(double)), "zero tile");
    if (const char* e = std::getenv("PSI4AMD_HT_ROWWAVES")) set_ht_rowwaves(std::atoi(e));
    if (const char* e = std::getenv("PSI4AMD_KGEMM_DC")) set_kgemm_direct(std::atoi(e));
    if (const char* e = std::getenv("PSI4AMD_KGEMM_WGS")) g_kgemm_wgs = std::max(256, std::atoi(e));
    if (const char* e = std::getenv("PSI4AMD_X_BUDGET_GB"))
        g_x_budget = g_x_budget_mixed = static_cast<size_t>(std::max(1, std::atoi(e))) << 30;
    const auto& sym = h->otiles_sym;
    const auto& all = h->otiles_all;
    const auto& of_sym = h->otile_of_sym;
    const auto& of_all = h->otile_of_all;
    hcheck(hipMemcpy(h->d_tidx.get<int>(h->tidx.size()), h->tidx.data(), h->tidx.size() * sizeof(int),
                     hipMemcpyHostToDevice),
           "upload tidx");
    hcheck(hipMemcpy(h->d_tiles.get<int2>(h->tiles.size()), h->tiles.data(), h->tiles.size() * sizeof(int2),
                     hipMemcpyHostToDevice),
           "upload tiles");
    hcheck(hipMemcpy(h->d_mask.get<uint8_t>(h->fun_mask.size()), h->fun_mask.data(), h->fun_mask.size(),
                     hipMemcpyHostToDevice),
           "upload mask");
    hcheck(hipMemcpy(h->d_otiles_sym.get<int2>(sym.size()), sym.data(), sym.size() * sizeof(int2),
                     hipMemcpyHostToDevice),
           "upload");
    hcheck(hipMemcpy(h->d_otiles_all.get<int2>(all.size()), all.data(), all.size() * sizeof(int2),
                     hipMemcpyHostToDevice),
           "upload");
    hcheck(hipMemcpy(h->d_otile_of_sym.get<int>(of_sym.size()), of_sym.data(), of_sym.size() * sizeof(int),
                     hipMemcpyHostToDevice),
           "upload");
    hcheck(hipMemcpy(h->d_otile_of_all.get<int>(of_all.size()), of_all.data(), of_all.size() * sizeof(int),
                     hipMemcpyHostToDevice),
           "upload");
    // the kernels that read these tables run on h->stream, which need not wait for the legacy stream
    hcheck(hipDeviceSynchronize(), "upload tile tables");
}

size_t b_doubles(const dfjk_handle* h) { return static_cast<size_t>(h->naux) * h->ntp * TILE2; }

bool gpu_schwarz(dfjk_handle* h, double* fun_max);

// Schwarz screening exactly as dfhelper.cc:299-420: tolerance = cutoff^2 / max (mn|mn).  The
// diagonal (mn|mn) comes from the device when one is present (the host engine otherwise); the
// shell maxima, the global maximum and the masks are formed here from it.
void prepare_sparsity(dfjk_handle* h) {
    const int n = h->nbf;
    const int ns = static_cast<int>(h->prim.shells.size());
    std::vector<double>& fun_max = h->schwarz_fun;
    std::vector<double>& shell_max = h->schwarz_shell;
    fun_max.assign(static_cast<size_t>(n) * n, 0.0);
    shell_max.assign(static_cast<size_t>(ns) * ns, 0.0);
    double max_val = 0.0;
    // the device path only speeds the same integrals up: a device without room for the
    // nbf x nbf values leaves them to the host engine
    try {
        h->schwarz_on_device = gpu_schwarz(h, fun_max.data());
    } catch (const StatusError& e) {
        if (e.code != DFJK_ERR_MEMORY) throw;
        (void)hipGetLastError();  // the failed allocation's status
        h->schwarz_on_device = false;
    }
    if (h->schwarz_on_device) {
        for (int M = 0; M < ns; ++M)
            for (int N = 0; N < ns; ++N) {
                double v = 0.0;
                for (int m = h->prim.offset[M]; m < h->prim.offset[M] + h->prim.shells[M].nfunc(); ++m)
                    for (int k = h->prim.offset[N]; k < h->prim.offset[N] + h->prim.shells[N].nfunc(); ++k)
                        v = std::max(v, fun_max[static_cast<size_t>(m) * n + k]);
                shell_max[static_cast<size_t>(M) * ns + N] = v;
                max_val = std::max(max_val, v);
            }
    } else {
        max_val = schwarz_diagonal(h->prim, fun_max.data(), shell_max.data(), h->nthreads);
    }
    h->schwarz_max = max_val;
    const double tol = h->cutoff * h->cutoff / max_val;
    h->fun_mask.assign(static_cast<size_t>(n) * n, 0);
    h->shell_mask.assign(static_cast<size_t>(ns) * ns, 0);
    h->nsig = 0;
    for (size_t i = 0; i < fun_max.size(); ++i)
        if (fun_max[i] >= tol) {
            h->fun_mask[i] = 1;
            ++h->nsig;
        }
    for (size_t i = 0; i < shell_max.size(); ++i) h->shell_mask[i] = shell_max[i] >= tol;
}

// Screening and tile geometry on the host, once (no device needed).  A tensor-level handle has
// no screening: every pair is significant.
void build_tiles(dfjk_handle* h);
void ensure_sparsity(dfjk_handle* h) {
    if (!h->tiles.empty()) return;
    if (h->tensor_mode) {
        h->fun_mask.assign(static_cast<size_t>(h->nbf) * h->nbf, 1);
        h->nsig = static_cast<int64_t>(h->nbf) * h->nbf;
    } else {
        prepare_sparsity(h);
    }
    build_tiles(h);
}

// Functions in the largest auxiliary shell (BasisSet::max_function_per_shell, DFHelper's Qshell_max_)
int qshell_max(const dfjk_handle* h) {
    int m = 1;
    for (const Shell& sh : h->aux.shells) m = std::max(m, sh.nfunc());
    return m;
}

// Metric factor on the host (FP64): returns M with B = M (P|mn).
// Wall-clock laps of the set-up phases.
struct Lap {
    std::chrono::steady_clock::time_point t = std::chrono::steady_clock::now();
    double operator()() {
        const auto n = std::chrono::steady_clock::now();
        const double ms = std::chrono::duration<double, std::milli>(n - t).count();
        t = n;
        return ms;
    }
};

[[noreturn]] void cholesky_failed(int pivot) {
    fail(DFJK_ERR_ARG, "metric Cholesky failed at pivot " + std::to_string(pivot) +
                           " (metric not positive definite: set DFJK_FIT_EIGEN, psi4's truncating J^-1/2)");
}

// M = L^-1 with J = L L^T, on the device.  Right-looking blocked Cholesky: each NB x NB
// diagonal block is factorised and inverted on the host (cholesky_lower, lower_inverse), the
// panel below it becomes L_ik = A_ik L_kk^-T and the trailing block A_22 -= L_i L_i^T (FP64
// MFMA GEMMs); then X = L^-1 by block forward substitution, X_i = -L_ii^-1 (L_i,<i X_<i), with
// the host inverses of the diagonal blocks.  M stays in ws_G for the fit (dM_ready); it is
// returned on the host (row-major, lower) only when want_host.
std::vector<double> device_metric_factor(dfjk_handle* h, const std::vector<double>& Jm, bool want_host,
                                         const double* dJm = nullptr) {
    const int n = h->naux;
    constexpr int NB = 256;
    const hipStream_t s = h->stream;
    const size_t nn = static_cast<size_t>(n) * n;
    DevBuf dA, dP, dPT, dT, dL;
    h->dM_ready = false;
    double* A = dA.get<double>(nn);
    double* X = h->ws_G.get<double>(nn);
    double* P = dP.get<double>(static_cast<size_t>(n) * NB);
    double* PT = dPT.get<double>(static_cast<size_t>(n) * NB);
    double* T = dT.get<double>(static_cast<size_t>(n) * NB);
    double* Li = dL.get<double>(static_cast<size_t>(NB) * NB);
    const size_t ld = static_cast<size_t>(n) * sizeof(double);
    if (dJm)
        hcheck(hipMemcpyAsync(A, dJm, nn * sizeof(double), hipMemcpyDeviceToDevice, s), "copy metric");
    else
        copy_settled(A, Jm.data(), nn * sizeof(double), hipMemcpyHostToDevice, s, "upload metric");
    hcheck(hipMemsetAsync(X, 0, nn * sizeof(double), s), "memset");
    std::vector<double> blk(NB * NB), invT(NB * NB);
    std::vector<std::vector<double>> dinv;  // host inverses of the diagonal blocks
    for (int k0 = 0; k0 < n; k0 += NB) {
        const int nb = std::min(NB, n - k0), k1 = k0 + nb, m = n - k1;
        const size_t bw = static_cast<size_t>(nb) * sizeof(double);
        double* Akk = A + static_cast<size_t>(k0) * n + k0;
        copy2d_settled(blk.data(), bw, Akk, ld, bw, nb, hipMemcpyDeviceToHost, s, "diagonal block");
        const int bad = cholesky_lower(blk.data(), nb, h->nthreads);
        if (bad >= 0) cholesky_failed(k0 + bad);
        dinv.emplace_back(static_cast<size_t>(nb) * nb);
        std::vector<double>& inv = dinv.back();
        lower_inverse(blk.data(), nb, inv.data(), h->nthreads);
        for (int r = 0; r < nb; ++r)
            for (int c = 0; c < nb; ++c) invT[static_cast<size_t>(c) * nb + r] = inv[static_cast<size_t>(r) * nb + c];
        copy2d_settled(Akk, ld, blk.data(), bw, bw, nb, hipMemcpyHostToDevice, s, "L_kk");
        copy2d_settled(X + static_cast<size_t>(k0) * n + k0, ld, inv.data(), bw, bw, nb, hipMemcpyHostToDevice, s,
                       "X_kk");
        if (m > 0) {
            copy_settled(Li, invT.data(), bw * nb, hipMemcpyHostToDevice, s, "L_kk^-T");
            double* Aik = A + static_cast<size_t>(k1) * n + k0;
            launch_gemm_nn_f64(Aik, n, Li, nb, P, nb, m, nb, nb, 1.0, 0.0, s);
            hcheck(hipMemcpy2DAsync(Aik, ld, P, bw, bw, m, hipMemcpyDeviceToDevice, s), "panel");
            launch_transpose(P, m, nb, nb, PT, m, s);
            launch_gemm_nn_f64(P, nb, PT, m, A + static_cast<size_t>(k1) * n + k1, n, m, m, nb, -1.0, 1.0, s);
        }
        // the host blocks are rewritten in the next pass: the copies above must have run
        hcheck(hipStreamSynchronize(s), "panel");
    }
    for (int i0 = NB, bi = 1; i0 < n; i0 += NB, ++bi) {
        const int nb = std::min(NB, n - i0);
        launch_gemm_nn_f64(A + static_cast<size_t>(i0) * n, n, X, n, T, i0, nb, i0, i0, 1.0, 0.0, s);
        copy_settled(Li, dinv[bi].data(), static_cast<size_t>(nb) * nb * sizeof(double), hipMemcpyHostToDevice, s,
                     "L_ii^-1");
        launch_gemm_nn_f64(Li, nb, T, i0, X + static_cast<size_t>(i0) * n, n, nb, i0, nb, -1.0, 0.0, s, 1, 0, 0, 0,
                           1);
        hcheck(hipStreamSynchronize(s), "block row");
    }
    h->dM_ready = true;
    std::vector<double> M;
    if (want_host) {
        M.resize(nn);
        copy_settled(M.data(), X, nn * sizeof(double), hipMemcpyDeviceToHost, s, "download M");
    }
    return M;
}

// psi4's fit, J^-1/2 = Matrix::power(-0.5, condition) (libmints/matrix.cc:2370-2424, called from
// dfhelper.cc:1491-1517): J = V diag(a) V^T by an FP64 symmetric eigensolver on the device
// (rocSOLVER dsyevd, the counterpart of psi4's LAPACK DSYEV), f_i = a_i^-1/2 except that
// |a_i| < condition * max(|a_min|, |a_max|) and non-finite powers give 0, then
// M = V diag(f) V^T, symmetric and full (the fit runs over the whole k range).  M is left in
// ws_G (dM_ready) and returned on the host only when want_host.
std::vector<double> device_metric_eigen(dfjk_handle* h, const std::vector<double>& Jm, bool want_host,
                                        const double* dJm = nullptr) {
    const int n = h->naux;
    const size_t nn = static_cast<size_t>(n) * n;
    const hipStream_t s = h->stream;
    h->dM_ready = false;
    DevBuf dA, dS, dRT, dw, de, df, dinfo;
    double* A = dA.get<double>(nn);
    double* w = dw.get<double>(n);
    double* e = de.get<double>(n);
    rocblas_int* info = dinfo.get<rocblas_int>(1);
    if (dJm)
        hcheck(hipMemcpyAsync(A, dJm, nn * sizeof(double), hipMemcpyDeviceToDevice, s), "copy metric");
    else
        copy_settled(A, Jm.data(), nn * sizeof(double), hipMemcpyHostToDevice, s, "upload metric");
    rocblas_handle rb = nullptr;
    if (rocblas_create_handle(&rb) != rocblas_status_success) fail(DFJK_ERR_DEVICE, "rocblas_create_handle failed");
    struct RbGuard {
        rocblas_handle r;
        ~RbGuard() { rocblas_destroy_handle(r); }
    } guard{rb};
    if (rocblas_set_stream(rb, s) != rocblas_status_success) fail(DFJK_ERR_DEVICE, "rocblas_set_stream failed");
    // the metric is symmetric, so its row-major storage is also its column-major storage
    const rocblas_status st = rocsolver_dsyevd(rb, rocblas_evect_original, rocblas_fill_upper, n, A, n, w, e, info);
    if (st != rocblas_status_success)
        fail(DFJK_ERR_DEVICE, std::string("rocsolver_dsyevd: ") + rocblas_status_to_string(st));
    std::vector<double> a(n);
    rocblas_int hinfo = 0;
    copy_settled(a.data(), w, n * sizeof(double), hipMemcpyDeviceToHost, s, "eigenvalues");
    copy_settled(&hinfo, info, sizeof(hinfo), hipMemcpyDeviceToHost, s, "eigensolver status");
    if (hinfo != 0) fail(DFJK_ERR_INTERNAL, "metric eigensolver did not converge (info " + std::to_string(hinfo) + ")");
    // eigenvalues ascend: the largest magnitude is at one end
    const double max_a = std::max(std::fabs(a[n - 1]), std::fabs(a[0]));
    std::vector<double> f(n);
    int kept = 0;
    for (int i = 0; i < n; ++i) {
        if (std::fabs(a[i]) < h->condition * max_a) {
            f[i] = 0.0;
        } else {
            const double v = std::pow(a[i], -0.5);
            f[i] = std::isfinite(v) ? v : 0.0;
            kept += std::isfinite(v) ? 1 : 0;
        }
    }
    h->fit_kept = kept;
    double* fd = df.get<double>(n);
    copy_settled(fd, f.data(), n * sizeof(double), hipMemcpyHostToDevice, s, "upload powers");
    // column-major eigenvectors read row-major: row i of A is eigenvector i (R).  M = R^T diag(f) R.
    double* S = dS.get<double>(nn);
    double* RT = dRT.get<double>(nn);
    launch_scale_rows(A, n, n, fd, S, s);
    launch_transpose(A, n, n, n, RT, n, s);
    double* X = h->ws_G.get<double>(nn);
    launch_gemm_nn_f64(RT, n, S, n, X, n, n, n, n, 1.0, 0.0, s);
    h->dM_ready = true;
    std::vector<double> M;
    if (want_host) {
        M.resize(nn);
        copy_settled(M.data(), X, nn * sizeof(double), hipMemcpyDeviceToHost, s, "download M");
    }
    // the temporaries are freed on return: the products must have read them
    hcheck(hipStreamSynchronize(s), "metric power");
    return M;
}

// The metric factor M (L^-1, or psi4's truncated J^-1/2); on the host only when want_host
// (else it is left in ws_G).
bool gpu_metric(dfjk_handle* h, DevBuf& dJ);

std::vector<double> metric_factor(dfjk_handle* h, bool want_host) {
    const int na = h->naux;
    const size_t nn = static_cast<size_t>(na) * na;
    Lap lap;
    // (P|Q) on the device when the kernels cover the auxiliary shells (then it stays there for
    // the factorisation), else on the host
    const bool host_factor = h->fitting == DFJK_FIT_CHOLESKY && std::getenv("PSI4AMD_HOST_FACTOR");
    DevBuf dJ;
    std::vector<double> Jm;
    const bool on_dev = !host_factor && gpu_metric(h, dJ);
    if (!on_dev) {
        Jm.resize(nn);
        metric(h->aux, Jm.data(), h->nthreads);
    }
    h->metric_on_device = on_dev;
    h->setup_ms[1] = lap();
    std::vector<double> M;
    h->fit_lower = h->fitting == DFJK_FIT_CHOLESKY;
    h->fit_kept = na;
    const double* dJm = on_dev ? dJ.as<double>() : nullptr;
    if (h->fitting == DFJK_FIT_CHOLESKY) {
        if (host_factor) {
            h->dM_ready = false;
            M.resize(nn);
            const int bad = cholesky_lower(Jm.data(), na, h->nthreads);
            if (bad >= 0) cholesky_failed(bad);
            lower_inverse(Jm.data(), na, M.data(), h->nthreads);
        } else {
            M = device_metric_factor(h, Jm, want_host, dJm);
        }
    } else {
        M = device_metric_eigen(h, Jm, want_host, dJm);
    }
    h->setup_ms[2] = lap();
    return M;
}

// Upload the host A = (P|mn) (tiled layout) and fit M A on the device into dst, in column chunks.
// With host_dst, the fitted columns go to pinned host memory (out-of-core B) instead.
void fit_on_device(dfjk_handle* h, const std::vector<double>& M, const std::vector<double>& A, DevBuf& dst,
                   double* host_dst = nullptr) {
    const int na = h->naux;
    const size_t ncolB = static_cast<size_t>(h->ntp) * TILE2;
    double* dM = h->ws_G.get<double>(static_cast<size_t>(na) * na);
    if (!h->dM_ready)
        copy_settled(dM, M.data(), M.size() * sizeof(double), hipMemcpyHostToDevice, h->stream, "upload metric factor");
    // chunk of columns so that the staging buffer stays <= ~1 GiB
    size_t cw = std::max<size_t>(TILE2, (size_t(1) << 27) / std::max(1, na));
    cw = std::min(ncolB, cw / TILE2 * TILE2);
    double* dA = h->ws_XL.get<double>(static_cast<size_t>(na) * cw);
    double* dBp = host_dst ? h->ws_XR.get<double>(static_cast<size_t>(na) * cw) : dst.get<double>(b_doubles(h));
    for (size_t c0 = 0; c0 < ncolB; c0 += cw) {
        const size_t w = std::min(cw, ncolB - c0);
        hcheck(hipMemcpy2DAsync(dA, w * sizeof(double), A.data() + c0, ncolB * sizeof(double), w * sizeof(double), na,
                                hipMemcpyHostToDevice, h->stream),
               "upload (P|mn)");
        if (host_dst) {
            launch_gemm_nn_f64(dM, na, dA, static_cast<int>(w), dBp, static_cast<int>(w), na, static_cast<int>(w), na,
                               1.0, 0.0, h->stream, 1, 0, 0, 0, h->fit_lower ? 1 : 0);
            hcheck(hipMemcpy2DAsync(host_dst + c0, ncolB * sizeof(double), dBp, w * sizeof(double),
                                    w * sizeof(double), na, hipMemcpyDeviceToHost, h->stream),
                   "download fitted B");
        } else {
            launch_gemm_nn_f64(dM, na, dA, static_cast<int>(w), dBp + c0, static_cast<int>(ncolB), na,
                               static_cast<int>(w), na, 1.0, 0.0, h->stream, 1, 0, 0, 0, h->fit_lower ? 1 : 0);
        }
        hcheck(hipStreamSynchronize(h->stream), "fit");
    }
    h->ws_XL.release();
    h->ws_XR.release();
    h->ws_G.release();
    h->dM_ready = false;
}

// ---------------------------------------------------------------- device (P|mn)
// The three-index integrals on the device (kernels_ints*.hip) for bases with primary l <= 3
// and auxiliary l <= 4: written into A [naux][ntp*TILE2] in the tiled layout (A zeroed here).
// Returns false (nothing launched) when a shell is outside those classes, when A and B
// together do not fit the device, out of core, or with PSI4AMD_HOST_INTS set; the caller
// then uses the host engine.
std::vector<int64_t> tile_dst(const dfjk_handle* h);

// Upload the shell tables of a "primary" basis pb and the auxiliary basis ab and run every
// (la, lb | lc) class of the shell pairs pcls[la * 4 + lb] (la >= lb) against every auxiliary
// function: out[P * p_stride + position of (m, n)] = (P|mn), positions from dst ([nfun(pb)^2],
// or null: column m).  Waits for the stream.
// Device copies of the shell tables of a "primary" basis pb (and an auxiliary basis ab), the
// solid-harmonic rows l = 0..4 and the host engine's Boys grid; fills the table fields of a.
struct IntTables {
    DevBuf ps, as, pe, pc, ae, ac, c2s, boys;
};
void upload_int_tables(const Basis& pb, const Basis* ab, IntTables& t, ThreeIndexArgs& a) {
    auto shells = [](const Basis& b, std::vector<GpuShell>& sh, std::vector<double>& e, std::vector<double>& c) {
        for (size_t i = 0; i < b.shells.size(); ++i) {
            const Shell& x = b.shells[i];
            GpuShell g;
            g.l = x.l;
            g.nprim = static_cast<int>(x.exps.size());
            g.p0 = static_cast<int>(e.size());
            g.foff = b.offset[i];
            for (int d = 0; d < 3; ++d) g.A[d] = x.A[d];
            sh.push_back(g);
            e.insert(e.end(), x.exps.begin(), x.exps.end());
            c.insert(c.end(), x.coefs.begin(), x.coefs.end());
        }
    };
    std::vector<GpuShell> ps, as;
    std::vector<double> pe, pc, ae, ac;
    shells(pb, ps, pe, pc);
    if (ab) shells(*ab, as, ae, ac);
    std::vector<double> c2s;
    for (int l = 0; l <= 4; ++l) {
        a.c2s_off[l] = static_cast<int>(c2s.size());
        const std::vector<double>& tb = cart2sph(l);
        c2s.insert(c2s.end(), tb.begin(), tb.end());
    }
    int bnt = 0, bncol = 0;
    double bh = 0.0, btmax = 0.0;
    const double* btab = boys_grid(&bnt, &bncol, &bh, &btmax);
    auto up = [&](DevBuf& b, const void* src, size_t bytes) {
        b.get<char>(std::max<size_t>(bytes, 8));
        if (bytes) hcheck(hipMemcpy(b.p, src, bytes, hipMemcpyHostToDevice), "upload integral tables");
        // read by kernels on the handle's stream, which need not wait for the legacy stream
        hcheck(hipDeviceSynchronize(), "upload integral tables");
    };
    up(t.ps, ps.data(), ps.size() * sizeof(GpuShell));
    up(t.as, as.data(), as.size() * sizeof(GpuShell));
    up(t.pe, pe.data(), pe.size() * sizeof(double));
    up(t.pc, pc.data(), pc.size() * sizeof(double));
    up(t.ae, ae.data(), ae.size() * sizeof(double));
    up(t.ac, ac.data(), ac.size() * sizeof(double));
    up(t.c2s, c2s.data(), c2s.size() * sizeof(double));
    up(t.boys, btab, static_cast<size_t>(bnt) * bncol * sizeof(double));
    a.prim = t.ps.as<GpuShell>();
    a.aux = t.as.as<GpuShell>();
    a.pexp = t.pe.as<double>();
    a.pcoef = t.pc.as<double>();
    a.aexp = t.ae.as<double>();
    a.acoef = t.ac.as<double>();
    a.c2s = t.c2s.as<double>();
    a.prim_pure = pb.shells.empty() ? 1 : (pb.shells[0].pure ? 1 : 0);
    a.aux_pure = (!ab || ab->shells.empty()) ? 1 : (ab->shells[0].pure ? 1 : 0);
    a.boys.table = t.boys.as<double>();
    a.boys.ncol = bncol;
    a.boys.h = bh;
    a.boys.inv_h = 1.0 / bh;
    a.boys.tmax = btmax;
    a.nbf = pb.nbf;
}

// The Schwarz diagonal |(mn|mn)| [nbf][nbf] on the device, for every shell pair M >= N of
// dfhelper.cc:329-356 (grouped by class, the shell of higher l first).  Returns false (nothing
// launched) with no device, for primary l > 3, or with PSI4AMD_HOST_INTS or
// PSI4AMD_HOST_SCHWARZ set; the caller then uses the host engine.
bool gpu_schwarz(dfjk_handle* h, double* fun_max) {
    if (std::getenv("PSI4AMD_HOST_INTS") || std::getenv("PSI4AMD_HOST_SCHWARZ")) return false;
    const Basis& pb = h->prim;
    for (const Shell& sh : pb.shells)
        if (sh.l > 3) return false;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
        (void)hipGetLastError();
        return false;
    }
    ensure_stream(h);
    const int np = static_cast<int>(pb.shells.size());
    std::vector<std::vector<int2>> pcls(16);
    for (int M = 0; M < np; ++M)
        for (int N = 0; N <= M; ++N) {
            int X = M, Y = N;
            if (pb.shells[X].l < pb.shells[Y].l) std::swap(X, Y);
            pcls[pb.shells[X].l * 4 + pb.shells[Y].l].push_back(make_int2(X, Y));
        }
    std::vector<int2> pall;
    std::vector<size_t> poff(17, 0);
    for (int c = 0; c < 16; ++c) {
        poff[c] = pall.size();
        pall.insert(pall.end(), pcls[c].begin(), pcls[c].end());
    }
    poff[16] = pall.size();
    ThreeIndexArgs a;
    IntTables tabs;
    upload_int_tables(pb, nullptr, tabs, a);
    DevBuf d_pairs, d_out;
    d_pairs.get<int2>(std::max<size_t>(1, pall.size()));
    copy_settled(d_pairs.p, pall.data(), pall.size() * sizeof(int2), hipMemcpyHostToDevice, h->stream, "upload pairs");
    const size_t nn = static_cast<size_t>(pb.nbf) * pb.nbf;
    double* out = d_out.get<double>(nn);
    hcheck(hipMemsetAsync(out, 0, nn * sizeof(double), h->stream), "memset schwarz");
    const int2* dp = d_pairs.as<int2>();
    for (int la = 0; la <= 3; ++la)
        for (int lb = 0; lb <= la; ++lb) {
            const int c = la * 4 + lb;
            const int npair = static_cast<int>(poff[c + 1] - poff[c]);
            if (npair) launch_schwarz_class(a, la, lb, dp + poff[c], npair, out, h->stream);
        }
    hcheck(hipGetLastError(), "schwarz launch");
    copy_settled(fun_max, out, nn * sizeof(double), hipMemcpyDeviceToHost, h->stream, "schwarz download");
    return true;
}

void run_three_index(dfjk_handle* h, const Basis& pb, const Basis& ab, const std::vector<std::vector<int2>>& pcls,
                     const std::vector<int64_t>* dst, int64_t p_stride, double* out, double omega) {
    std::vector<std::vector<int2>> acls(5);
    for (size_t i = 0; i < ab.shells.size(); ++i)
        for (int k = 0; k < ab.shells[i].nfunc(); ++k) acls[ab.shells[i].l].push_back(make_int2(static_cast<int>(i), k));
    std::vector<int2> pall, aall;
    std::vector<size_t> poff(17, 0), aoff(6, 0);
    for (int c = 0; c < 16; ++c) {
        poff[c] = pall.size();
        pall.insert(pall.end(), pcls[c].begin(), pcls[c].end());
    }
    poff[16] = pall.size();
    for (int c = 0; c < 5; ++c) {
        aoff[c] = aall.size();
        aall.insert(aall.end(), acls[c].begin(), acls[c].end());
    }
    aoff[5] = aall.size();
    ThreeIndexArgs a;
    IntTables tabs;
    upload_int_tables(pb, &ab, tabs, a);
    DevBuf d_dst, d_pairs, d_afun;
    auto up = [&](DevBuf& b, const void* src, size_t bytes) {
        b.get<char>(bytes);
        copy_settled(b.p, src, bytes, hipMemcpyHostToDevice, h->stream, "upload integral tables");
    };
    if (dst) up(d_dst, dst->data(), dst->size() * sizeof(int64_t));
    up(d_pairs, pall.data(), std::max<size_t>(1, pall.size()) * sizeof(int2));
    up(d_afun, aall.data(), std::max<size_t>(1, aall.size()) * sizeof(int2));
    a.dst = dst ? d_dst.as<int64_t>() : nullptr;
    a.p_stride = p_stride;
    a.out = out;
    a.omega = omega;
    const int2* dp = d_pairs.as<int2>();
    const int2* da = d_afun.as<int2>();
    for (int la = 0; la <= 3; ++la)
        for (int lb = 0; lb <= la; ++lb) {
            const int c = la * 4 + lb;
            const int npair = static_cast<int>(poff[c + 1] - poff[c]);
            if (!npair) continue;
            for (int lc = 0; lc <= 4; ++lc) {
                const int nf = static_cast<int>(aoff[lc + 1] - aoff[lc]);
                if (!nf) continue;
                launch_three_index_class(a, la, lb, lc, dp + poff[c], npair, da + aoff[lc], nf, h->stream);
            }
        }
    hcheck(hipStreamSynchronize(h->stream), "three-index integrals");
}

bool gpu_three_index(dfjk_handle* h, double omega, DevBuf& Abuf) {
    if (std::getenv("PSI4AMD_HOST_INTS")) return false;
    if (h->ooc) return false;  // a device-resident (P|mn) would not respect the memory budget
    const Basis& pb = h->prim;
    const Basis& ab = h->aux;
    for (const Shell& sh : pb.shells)
        if (sh.l > 3) return false;
    for (const Shell& sh : ab.shells)
        if (sh.l > 4) return false;
    const size_t ncolB = static_cast<size_t>(h->ntp) * TILE2;
    const size_t abytes = static_cast<size_t>(h->naux) * ncolB * sizeof(double);
    size_t free_b = 0, total_b = 0;
    hcheck(hipMemGetInfo(&free_b, &total_b), "hipMemGetInfo");
    // the fit writes B (omega = 0) or the attenuated tensor: count it unless already allocated
    const DevBuf& fit_dst = omega > 0.0 ? h->dBw : h->dB;
    const size_t dst_bytes = fit_dst.bytes >= abytes ? 0 : abytes;
    if (abytes + dst_bytes + (size_t(2) << 30) > free_b) return false;
    // significant shell pairs (the host engine's M >= N set), higher l first, by class
    const int np = static_cast<int>(pb.shells.size());
    std::vector<std::vector<int2>> pcls(16);
    for (int M = 0; M < np; ++M)
        for (int N = 0; N <= M; ++N) {
            if (!h->shell_mask[static_cast<size_t>(M) * np + N]) continue;
            int X = M, Y = N;
            if (pb.shells[X].l < pb.shells[Y].l) std::swap(X, Y);
            pcls[pb.shells[X].l * 4 + pb.shells[Y].l].push_back(make_int2(X, Y));
        }
    const std::vector<int64_t> dst = tile_dst(h);
    double* A = Abuf.get<double>(static_cast<size_t>(h->naux) * ncolB);
    hcheck(hipMemsetAsync(A, 0, abytes, h->stream), "memset A");
    run_three_index(h, pb, ab, pcls, &dst, static_cast<int64_t>(ncolB), A, omega);
    return true;
}

// The Coulomb metric (P|Q) on the device (fittingmetric.cc:115-155, the Libint2 two-centre
// step), by the three-centre kernels with the pair (Q, unit s function): (P|Q 1) = (P|Q).
// dJ [naux][naux], both triangles.  Returns false (nothing launched) for auxiliary l > 3 (the
// pair slot of the kernels), with PSI4AMD_HOST_INTS or PSI4AMD_HOST_METRIC set; the caller then
// uses the host engine.
bool gpu_metric(dfjk_handle* h, DevBuf& dJ) {
    if (std::getenv("PSI4AMD_HOST_INTS") || std::getenv("PSI4AMD_HOST_METRIC")) return false;
    const Basis& ab = h->aux;
    for (const Shell& sh : ab.shells)
        if (sh.l > 3) return false;
    // the auxiliary shells at their own function offsets, plus one unit s shell (exponent 0,
    // coefficient 1: its product with any Gaussian is that Gaussian)
    Basis pb = ab;
    Shell unit;
    unit.l = 0;
    unit.pure = ab.shells.empty() ? true : ab.shells[0].pure;
    unit.exps = {0.0};
    unit.coefs = {1.0};
    pb.shells.push_back(unit);
    pb.finalize();
    const int iu = static_cast<int>(pb.shells.size()) - 1;
    std::vector<std::vector<int2>> pcls(16);
    for (int Q = 0; Q < iu; ++Q) pcls[pb.shells[Q].l * 4].push_back(make_int2(Q, iu));
    const size_t nn = static_cast<size_t>(h->naux) * h->naux;
    double* J = dJ.get<double>(nn);
    run_three_index(h, pb, ab, pcls, nullptr, h->naux, J, 0.0);
    return true;
}

// Fit B = M A from a device-resident A [naux][ncolB] into dst, or into pinned host memory.
void fit_device_A(dfjk_handle* h, const std::vector<double>& M, const double* A, DevBuf& dst,
                  double* host_dst = nullptr) {
    const int na = h->naux;
    const size_t ncolB = static_cast<size_t>(h->ntp) * TILE2;
    double* dM = h->ws_G.get<double>(static_cast<size_t>(na) * na);
    if (!h->dM_ready)
        copy_settled(dM, M.data(), M.size() * sizeof(double), hipMemcpyHostToDevice, h->stream, "upload metric factor");
    if (!host_dst) {
        double* dBp = dst.get<double>(b_doubles(h));
        size_t cw = std::max<size_t>(TILE2, (size_t(1) << 30) / (sizeof(double) * std::max(1, na)));
        cw = std::min(ncolB, cw / TILE2 * TILE2);
        for (size_t c0 = 0; c0 < ncolB; c0 += cw) {
            const size_t w = std::min(cw, ncolB - c0);
            launch_gemm_nn_f64(dM, na, A + c0, static_cast<int>(ncolB), dBp + c0, static_cast<int>(ncolB), na,
                               static_cast<int>(w), na, 1.0, 0.0, h->stream, 1, 0, 0, 0, h->fit_lower ? 1 : 0);
        }
    } else {
        size_t cw = std::max<size_t>(TILE2, (size_t(1) << 27) / std::max(1, na));
        cw = std::min(ncolB, cw / TILE2 * TILE2);
        double* tmp = h->ws_XR.get<double>(static_cast<size_t>(na) * cw);
        for (size_t c0 = 0; c0 < ncolB; c0 += cw) {
            const size_t w = std::min(cw, ncolB - c0);
            launch_gemm_nn_f64(dM, na, A + c0, static_cast<int>(ncolB), tmp, static_cast<int>(w), na,
                               static_cast<int>(w), na, 1.0, 0.0, h->stream, 1, 0, 0, 0, h->fit_lower ? 1 : 0);
            hcheck(hipMemcpy2DAsync(host_dst + c0, ncolB * sizeof(double), tmp, w * sizeof(double),
                                    w * sizeof(double), na, hipMemcpyDeviceToHost, h->stream),
                   "download fitted B");
            hcheck(hipStreamSynchronize(h->stream), "fit");
        }
        h->ws_XR.release();
    }
    hcheck(hipStreamSynchronize(h->stream), "fit");
    h->ws_G.release();
    h->dM_ready = false;
}

// Resident doubles of the in-core algorithm (B, Bw when wK is on, the metric).  AUTO decides
// like DFHelper::AO_core with required_core_size_ (dfhelper.cc:216-260), but against the
// device: B goes out of core only when it and 4 GiB of workspace do not fit the free HBM.
// (psi4's memory setting sizes host buffers; here it bounds the X workspace, not residency.)
// Doubles' worth of the lower-precision copies of B a compute at the handle's precision keeps
// resident: the FP32 copy (FP32 and BF16X3), and the tile-transposed FP32 copy (BF16X3).
size_t copy_doubles(const dfjk_handle* h) {
    if (h->precision == DFJK_PREC_FP32) return b_doubles(h) / 2;
    if (h->precision == DFJK_PREC_BF16X3) return b_doubles(h);
    return 0;
}

size_t resident_doubles(const dfjk_handle* h) {
    return b_doubles(h) * (h->do_wK ? 2 : 1) + static_cast<size_t>(h->naux) * h->naux + copy_doubles(h);
}

void choose_subalgorithm(dfjk_handle* h) {
    if (h->subalgo == DFJK_SUBALGO_OUT_OF_CORE) {
        h->ooc = true;
    } else if (h->subalgo == DFJK_SUBALGO_INCORE) {
        h->ooc = false;
    } else {
        size_t free_b = 0, total_b = 0;
        hcheck(hipMemGetInfo(&free_b, &total_b), "hipMemGetInfo");
        h->ooc = resident_doubles(h) * sizeof(double) + (size_t(4) << 30) > free_b;
    }
    if (h->ooc && h->do_wK)  // dfhelper.cc:201-205
        fail(DFJK_ERR_STATE, "DFJK: not equipped to do wK out of core; supply more memory or disable wK");
    if (h->ooc && h->tensor_mode) fail(DFJK_ERR_STATE, "DFJK: tensor-level handles are in core only");
    if (h->ooc) {
        if (!h->hB) hcheck(hipHostMalloc(reinterpret_cast<void**>(&h->hB), b_doubles(h) * sizeof(double), 0),
                           "hipHostMalloc B");
        if (!h->cstream) hcheck(hipStreamCreateWithFlags(&h->cstream, hipStreamNonBlocking), "copy stream");
        for (int i = 0; i < 2; ++i) {
            if (!h->ev_copied[i]) hcheck(hipEventCreateWithFlags(&h->ev_copied[i], hipEventDisableTiming), "event");
            if (!h->ev_free[i]) hcheck(hipEventCreateWithFlags(&h->ev_free[i], hipEventDisableTiming), "event");
        }
    }
}

// Destination of every significant function pair (m,n) in the tiled layout (-1: not stored).
std::vector<int64_t> tile_dst(const dfjk_handle* h) {
    const int n = h->nbf;
    std::vector<int64_t> dst(static_cast<size_t>(n) * n, -1);
    for (int m = 0; m < n; ++m)
        for (int q = 0; q < n; ++q) {
            if (!h->fun_mask[static_cast<size_t>(m) * n + q]) continue;
            const int I = m / TILE, J = q / TILE;
            if (I > J) continue;  // lower block-triangle is not stored
            const int t = h->tidx[I * h->nT + J];
            dst[static_cast<size_t>(m) * n + q] = static_cast<int64_t>(t) * TILE2 + (m % TILE) * TILE + q % TILE;
        }
    return dst;
}

// (P|erf(omega r)/r|mn) on the host, fitted with the Coulomb metric factor into dBw.  psi4 keeps
// the raw attenuated tensor and contracts the full inverse metric into the other side
// (dfhelper.cc:589-699, 3378-3438); with J^-1 = M^T M both sides carry one factor M here.
// dBw matches the current omega and exchange-combination settings
bool bw_current(const dfjk_handle* h) {
    if (h->bw_omega != h->omega || h->bw_comb != h->wcombine) return false;
    return !h->wcombine || (h->bw_alpha == h->omega_alpha && h->bw_beta == h->omega_beta);
}

// Combined exchange (dfhelper.cc:659-667, 1411): psi4 pairs J^-1 (Q|mn) with
// alpha (Q|mn) + beta (Q|w|mn); with one factor M on each side that is alpha B + beta Bw,
// formed in place in dBw.
void combine_Bw(dfjk_handle* h) {
    h->bw_comb = false;
    if (h->wcombine) {
        if (h->ooc) fail(DFJK_ERR_STATE, "wK is not available with B out of core");
        launch_axpby_mat(h->dB.as<double>(), h->omega_alpha, h->dBw.as<double>(), h->omega_beta,
                         h->dBw.as<double>(), b_doubles(h), h->stream);
        hcheck(hipStreamSynchronize(h->stream), "combine wK tensor");
        h->bw_comb = true;
        h->bw_alpha = h->omega_alpha;
        h->bw_beta = h->omega_beta;
    }
}

// B resident on the device (in core): every path that launches on dB checks this first
void require_device_B(const dfjk_handle* h, const char* what) {
    if (h->ooc || !h->dB.p)
        fail(DFJK_ERR_STATE, std::string(what) + " needs B resident on the device (subalgorithm INCORE); this handle "
                                                 "streams B from host memory");
}

void build_Bw(dfjk_handle* h) {
    require_device_B(h, "wK");  // dfhelper.cc:201-205: psi4 is not equipped to do wK out of core
    if (h->tensor_mode) fail(DFJK_ERR_STATE, "wK needs the attenuated tensor: dfjk_set_Bw_dense");
    if (!(h->omega > 0.0)) fail(DFJK_ERR_ARG, "wK needs omega > 0 (dfjk_set_omega)");
    if (h->fit_M.empty()) h->fit_M = metric_factor(h, true);
    if (gpu_three_index(h, h->omega, h->ws_XL)) {
        fit_device_A(h, h->fit_M, h->ws_XL.as<double>(), h->dBw);
        h->ws_XL.release();
    } else {
        const std::vector<int64_t> dst = tile_dst(h);
        const size_t ncolB = static_cast<size_t>(h->ntp) * TILE2;
        std::vector<double> A(static_cast<size_t>(h->naux) * ncolB, 0.0);
        three_index(h->prim, h->aux, dst.data(), static_cast<int64_t>(ncolB), h->shell_mask.data(), A.data(),
                    h->nthreads, h->omega);
        fit_on_device(h, h->fit_M, A, h->dBw);
    }
    h->bw_omega = h->omega;
    combine_Bw(h);
}

void initialize_basis(dfjk_handle* h) {
    for (double& t : h->setup_ms) t = 0.0;
    Lap lap;
    ensure_sparsity(h);
    upload_tiles(h);
    h->setup_ms[0] = lap();
    std::vector<double> M = metric_factor(h, h->do_wK);  // the host copy rebuilds Bw after an omega change
    choose_subalgorithm(h);
    lap();
    // (P|mn) directly in the tiled layout: on the device when the classes and memory allow,
    // else on the host
    if (gpu_three_index(h, 0.0, h->ws_XL)) {
        h->setup_ms[3] = lap();
        fit_device_A(h, M, h->ws_XL.as<double>(), h->dB, h->ooc ? h->hB : nullptr);
        h->ws_XL.release();
    } else {
        const std::vector<int64_t> dst = tile_dst(h);
        const size_t ncolB = static_cast<size_t>(h->ntp) * TILE2;
        std::vector<double> A(static_cast<size_t>(h->naux) * ncolB, 0.0);
        three_index(h->prim, h->aux, dst.data(), static_cast<int64_t>(ncolB), h->shell_mask.data(), A.data(),
                    h->nthreads);
        h->setup_ms[3] = lap();
        fit_on_device(h, M, A, h->dB, h->ooc ? h->hB : nullptr);
    }
    hcheck(hipStreamSynchronize(h->stream), "fit");
    h->setup_ms[4] = lap();
    if (h->do_wK) {
        h->fit_M = std::move(M);
        build_Bw(h);
        h->setup_ms[5] = lap();
    }
}

// ---------------------------------------------------------------- compute
// Q-splits of the contraction: at least about 2048 workgroups, a multiple of the 8 XCDs
// when the batch allows (the contraction maps each split to one XCD).  Among up to three
// times that many splits that keep >= 32 Q per split, the one whose workgroups fill each
// XCD's 64 slots (32 CUs x 2 workgroups) in the most nearly whole rounds is taken: the
// last round of a launch otherwise runs part-empty (e.g. 78 tiles x 32 splits leave 2.5 %
// of the rounds idle, 72 splits 0.3 %; measured 70.2 -> 68.8 ms for the FP64 contraction
// at nbf 1500 / naux 6000).
int choose_nsplit(int ntile, int nbatch) {
    const int target = g_kgemm_wgs;
    int ns0 = std::max(1, (target + ntile - 1) / ntile);
    ns0 = (ns0 + 7) / 8 * 8;
    auto fill = [&](int ns) {
        const long w = static_cast<long>(ns / 8) * ntile;  // workgroups per XCD
        return static_cast<double>(w) / static_cast<double>((w + 63) / 64 * 64);
    };
    int best = ns0;
    for (int ns = ns0; ns <= 3 * ns0 && nbatch / ns >= 32; ns += 8)
        if (fill(ns) > fill(best) + 0.005) best = ns;
    return std::min(best, std::max(1, nbatch));
}

// The lower-precision copy of B, made once per tensor on first use: FP32, read by J in both
// mixed modes, by the FP32 half-transform, and by the BF16 half-transform, which splits it into
// hi/lo as it stages each block.
// A copy that does not fit the free device memory is refused with a message that says so.
void require_copy_room(size_t bytes) {
    size_t free_b = 0, total_b = 0;
    hcheck(hipMemGetInfo(&free_b, &total_b), "hipMemGetInfo");
    if (bytes > free_b)
        fail(DFJK_ERR_MEMORY, "DFJK: the FP32 copies of B for the mixed precisions need " +
                                  std::to_string(bytes >> 20) + " MiB more device memory, " +
                                  std::to_string(free_b >> 20) + " MiB are free; use FP64 or OUT_OF_CORE");
}
const float* ensure_B32(dfjk_handle* h) {
    if (!h->dB32.p) {
        require_copy_room(b_doubles(h) * sizeof(float));
        float* p = h->dB32.get<float>(b_doubles(h));
        launch_to_f32(h->dB.as<double>(), p, b_doubles(h), h->stream);
    }
    return h->dB32.as<float>();
}
// tile-transposed FP32 copy, read by the BF16 half-transform for blocks used as (J, I)
void ensure_B32T(dfjk_handle* h) {
    if (!h->dB32T.p) {
        require_copy_room(b_doubles(h) * sizeof(float));
        float* p = h->dB32T.get<float>(b_doubles(h));
        launch_to_f32_t(h->dB.as<double>(), p, b_doubles(h), h->stream);
    }
}
void invalidate_copies(dfjk_handle* h) {
    h->dB32.release();
    h->dB32T.release();
}

// Device address of the scratch array on the current device, looked up once per device with
// that device current (each GPU holds its own copy at its own address).
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
char* g_scratch_addr[kMaxDevices] = {};
char* scratch_of_device(int dev) {
    if (dev < 0 || dev >= kMaxDevices) fail(DFJK_ERR_DEVICE, "device index beyond the scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_addr[dev]) {
        void* p = nullptr;
        hcheck(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
        if (!p) fail(DFJK_ERR_DEVICE, "hipGetSymbolAddress(scratch) returned no address");
        g_scratch_addr[dev] = static_cast<char*>(p);
    }
    return g_scratch_addr[dev];
}

// Device workspace of one compute.  Allocating mode: the handle's growable buffers.  Fixed
// mode (dfjk_compute_device_noalloc): a bump arena over this device's static scratch array;
// nothing is allocated, and a request past the arena's end is an error.
struct Work {
    char* base = nullptr;
    size_t cap = 0, used = 0;
    bool fixed() const { return base != nullptr; }
    template <class T>
    T* get(DevBuf& b, size_t n) {
        if (!base) return b.get<T>(n);
        const size_t bytes = (std::max<size_t>(n * sizeof(T), 256) + 255) / 256 * 256;
        if (bytes > cap - used)
            fail(DFJK_ERR_MEMORY, "dfjk_compute_device_noalloc: this system needs more than the " +
                                      std::to_string(cap >> 20) + " MiB scratch array");
        T* p = reinterpret_cast<T*>(base + used);
        used += bytes;
        return p;
    }
    size_t left() const { return cap - used; }
};

// Column map of the half-transform on the device.  Allocating mode: uploaded only when it
// differs from what the device buffer holds.  Either way the copy is made from pinned memory on
// the handle's stream (no host wait; the pinned staging is rewritten only after the event of
// its previous upload has passed).
int* upload_cmap(dfjk_handle* h, const std::vector<int>& cmap, Work& wk) {
    const size_t n = cmap.size();
    int* dmap;
    if (wk.fixed()) {
        dmap = wk.get<int>(h->ws_cmap, n);  // the scratch array is shared: always upload
        if (h->cmap_pin_cap < n) fail(DFJK_ERR_MEMORY, "dfjk_compute_device_noalloc: column map too long");
    } else {
        void* before = h->ws_cmap.p;
        dmap = h->ws_cmap.get<int>(n);
        if (h->ws_cmap.p != before) h->cmap_dev.clear();  // reallocated: contents gone
        if (h->cmap_dev == cmap) return dmap;
    }
    if (!h->ev_cmap) hcheck(hipEventCreateWithFlags(&h->ev_cmap, hipEventDisableTiming), "event");
    hcheck(hipEventSynchronize(h->ev_cmap), "column map staging");
    if (h->cmap_pin_cap < n) {
        if (h->cmap_pin) hcheck(hipHostFree(h->cmap_pin), "hipHostFree");
        h->cmap_pin = nullptr;
        h->cmap_pin_cap = 0;
        hcheck(hipHostMalloc(reinterpret_cast<void**>(&h->cmap_pin), std::max<size_t>(n, 1024) * sizeof(int), 0),
               "hipHostMalloc");
        h->cmap_pin_cap = std::max<size_t>(n, 1024);
    }
    std::memcpy(h->cmap_pin, cmap.data(), n * sizeof(int));
    hcheck(hipMemcpyAsync(dmap, h->cmap_pin, n * sizeof(int), hipMemcpyHostToDevice, h->stream), "column map");
    hcheck(hipEventRecord(h->ev_cmap, h->stream), "record");
    h->cmap_dev = wk.fixed() ? std::vector<int>() : cmap;
    return dmap;
}

// One density pair: Cl/Cr are device [nbf][ncol] row-major (ld = ncol).
// Cr == nullptr: lr_symmetric.  Two structured non-symmetric forms are recognised
// (bitwise column tests on the device) and kept symmetric:
//  * signed: every column of Cr is +/- the same column of Cl (the incremental
//    C_left=[C_new,C_old], C_right=[C_new,-C_old] form).  K = sum_Q X_Q diag(s) X_Q^T:
//    the columns are regrouped (+ first, then -), one half-transform is run and the
//    upper block-triangle is contracted twice, with +1 over the + group and -1 over the -.
//  * cross: Cl = [A, V], Cr = [V, A] (ncol even; dD = A V^T + V A^T, the form
//    A = (C_new + C_old)/2, V = C_new - C_old of a density change, in which every
//    product carries the small factor V).  K = sum_Q X_A X_V^T + X_V X_A^T: one
//    half-transform of [A, V], two contractions of the upper block-triangle with the
//    column halves swapped between the operands.
// The cross form can also be declared (dfjk_set_pair_form), which skips the test.
void compute_pair(dfjk_handle* h, const double* dCl, const double* dCr, int ncol, double* dD, double* dJ, double* dK,
                  int accumulate, Work& wk) {
    const TileGeom g = h->geom();
    const hipStream_t s = h->stream;
    const int q0 = h->aux_q0, q1 = h->aux_q1;
    const int nq_total = q1 - q0;
    const int nbf = h->nbf, nbfp = h->nbfp;
    const int xrows = (h->nT + 3) / 4 * 128;
    EventTimer& tm = h->timer;
    if (wk.fixed() && h->ooc) fail(DFJK_ERR_STATE, "dfjk_compute_device_noalloc needs B in core");
    if (wk.fixed()) {
        // Before any launch: the buffers of fixed size, one K split and one auxiliary
        // function's X must fit the scratch array (X blocks and splits then adapt to the rest).
        auto r = [](size_t b) { return (std::max<size_t>(b, 256) + 255) / 256 * 256; };
        int nbw_b, np_b, ldc_b;
        ht_panel_shape(ncol + 64, &nbw_b, &np_b, &ldc_b);  // column groups are padded to 32 each
        const size_t ldcb = static_cast<size_t>(ldc_b) + 320;
        const size_t qsb = static_cast<size_t>(h->ntp) * TILE2, cszb = static_cast<size_t>(nbfp) * ldcb;
        const size_t tiles = static_cast<size_t>(h->ntile_all) * 16384 * sizeof(double);
        const size_t need = r(2 * static_cast<size_t>(ncol)) + r(ldcb * sizeof(int)) + 2 * r(cszb * 8) +
                            r(static_cast<size_t>(nbf) * nbf * 8) + r(tiles) +                       // D
                            r(qsb * 8) + r(static_cast<size_t>(h->naux) * 8) +                       // J: W, d
                            r(static_cast<size_t>(h->ntp) * h->naux * 8) + r(qsb * 8) +             // J partials
                            r(tiles) + r(8 * cszb) +                                                  // K split, C copies
                            2 * r(static_cast<size_t>(xrows) * ldcb * 8) + 1024;                     // one Q of X
        if (need > wk.cap)
            fail(DFJK_ERR_MEMORY, "dfjk_compute_device_noalloc: nbf " + std::to_string(nbf) + ", naux " +
                                      std::to_string(h->naux) + " and " + std::to_string(ncol) +
                                      " columns need about " + std::to_string(need >> 20) + " MiB, beyond the " +
                                      std::to_string(wk.cap >> 20) + " MiB scratch array");
    }

    // classify the pair (or take the declared form) and build the column map
    bool sym = (dCr == nullptr), cross = false;
    std::vector<int> hs(ncol, 1);
    if (!sym && h->pair_form == DFJK_FORM_CROSS) {
        if (ncol % 2) fail(DFJK_ERR_ARG, "JK: the cross pair form needs an even column count");
        cross = true;
    } else if (!sym && h->pair_form == DFJK_FORM_DETECT) {
        // both column tests in one round trip: shift 0 (signed form), shift ncol/2 (cross form)
        const int ntest = ncol % 2 == 0 ? 2 : 1;
        int8_t* sg = wk.get<int8_t>(h->ws_sign, 2 * static_cast<size_t>(ncol));
        launch_column_signs(dCl, dCr, ncol, nbf, ncol, 0, sg, s);
        if (ntest == 2) launch_column_signs(dCl, dCr, ncol, nbf, ncol, ncol / 2, sg + ncol, s);
        if (h->sign_pin_cap < 2 * static_cast<size_t>(ncol)) {
            if (wk.fixed()) fail(DFJK_ERR_MEMORY, "dfjk_compute_device_noalloc: too many columns for the structure test");
            if (h->sign_pin) hcheck(hipHostFree(h->sign_pin), "hipHostFree");
            h->sign_pin = nullptr;
            h->sign_pin_cap = 0;
            const size_t cap = std::max<size_t>(2 * static_cast<size_t>(ncol), 4096);
            hcheck(hipHostMalloc(reinterpret_cast<void**>(&h->sign_pin), cap, 0), "hipHostMalloc");
            h->sign_pin_cap = cap;
        }
        hcheck(hipMemcpyAsync(h->sign_pin, sg, static_cast<size_t>(ntest) * ncol, hipMemcpyDeviceToHost, s), "signs");
        hcheck(hipStreamSynchronize(s), "signs");
        const int8_t* h8 = h->sign_pin;
        bool allpm = true;
        for (int i = 0; i < ncol; ++i) {
            hs[i] = h8[i];
            allpm &= h8[i] != 0;
        }
        sym = allpm;
        if (!sym && ntest == 2) {
            cross = true;
            for (int i = 0; i < ncol; ++i) cross &= h8[ncol + i] == 1;
            if (cross) std::fill(hs.begin(), hs.end(), 1);
        }
    }
    const bool general = !sym && !cross;
    std::vector<int> plus, minus;
    for (int i = 0; i < ncol; ++i) (hs[i] >= 0 ? plus : minus).push_back(i);
    const bool signed_sym = sym && !minus.empty();
    int ncol_eff = ncol, kplus = 0, half = 0, wcross = 0;
    std::vector<int> cmap;
    if (signed_sym) {
        kplus = round_up(static_cast<int>(plus.size()), 32);
        ncol_eff = kplus + round_up(static_cast<int>(minus.size()), 32);
    } else if (cross) {
        half = ncol / 2;
        wcross = round_up(half, 32);
        ncol_eff = 2 * wcross;
    }
    int nbw, npanel, ldc;
    ht_panel_shape(ncol_eff, &nbw, &npanel, &ldc);
    cmap.assign(ldc, -1);
    if (signed_sym) {
        for (size_t i = 0; i < plus.size(); ++i) cmap[i] = plus[i];
        for (size_t i = 0; i < minus.size(); ++i) cmap[kplus + i] = minus[i];
    } else if (cross) {
        for (int i = 0; i < half; ++i) {
            cmap[i] = i;
            cmap[wcross + i] = half + i;
        }
    } else {
        for (int i = 0; i < ncol; ++i) cmap[i] = i;
    }
    int* dmap = upload_cmap(h, cmap, wk);

    double* cl = wk.get<double>(h->ws_cl, static_cast<size_t>(nbfp) * ldc);
    launch_pack_cmap_f64(dCl, ncol, nbf, dmap, cl, nbfp, ldc, s);
    double* cr = cl;
    if (dCr) {
        cr = wk.get<double>(h->ws_cr, static_cast<size_t>(nbfp) * ldc);
        launch_pack_cmap_f64(dCr, ncol, nbf, dmap, cr, nbfp, ldc, s);
    }

    // D = Cl Cr^T (jk.cc:314-354), FP64
    const bool need_D = (dD != nullptr) || (h->do_J && dJ != nullptr);
    double* Dm = dD;
    if (need_D) {
        hipEvent_t e0 = tm.begin();
        if (!Dm) Dm = wk.get<double>(h->ws_D, static_cast<size_t>(nbf) * nbf);
        double* P = wk.get<double>(h->ws_Pd, static_cast<size_t>(h->ntile_all) * 16384);
        launch_gemm_nt_f64(cl, 0, ldc, cr, 0, ldc, nbf, nbf, ldc, 1, 1, h->d_otiles_all.as<int2>(), h->ntile_all, P, 0,
                           s);
        launch_tile_reduce(P, 1, h->ntile_all, h->d_otile_of_all.as<int>(), h->nB, 0, nbf, nbf, Dm, nbf, 1.0, 0, s);
        tm.end(T_OTHER, e0);
    }

    // J (dfhelper.cc:3162-3285) and K (dfhelper.cc:2162-2186, 3350-3377) over Q-blocks.
    // J_Q-block: d_Q = <B_Q, D>, J += sum_Q B_Q d_Q (additive over Q, FP64 accumulation);
    // K_Q-block: X_Q = B_Q C, K += sum_Q X_Q X'_Q^T.  In core the blocks are views of the
    // resident tensor (or of its FP32 / BF16 copies in the mixed modes); out of core each
    // block is copied from pinned host memory into one of two stage buffers on the copy
    // stream while the previous block computes, and converted there for the mixed modes.
    const bool doJ = h->do_J && dJ;
    const bool doK = h->do_K && dK;
    if (!doJ && !doK) return;
    const int prec = h->precision;
    const size_t qs = static_cast<size_t>(h->ntp) * TILE2;  // elements per Q
    const size_t xq = static_cast<size_t>(xrows) * ldc;      // X elements per Q
    const size_t esz = prec == DFJK_PREC_FP64 ? 8 : 4;       // FP64, FP32, or BF16 hi+lo

    // resident lower-precision copies (in core); the fixed-workspace entry point allocates
    // none, so an earlier allocating compute at this precision must have made them
    if (!h->ooc) {
        const bool miss =
            (prec != DFJK_PREC_FP64 && !h->dB32.p) || (prec == DFJK_PREC_BF16X3 && doK && !h->dB32T.p);
        if (wk.fixed() && miss)
            fail(DFJK_ERR_STATE, "dfjk_compute_device_noalloc: the lower-precision copy of B for this precision does "
                                 "not exist yet; run one dfjk_compute_device at this precision first");
        // J reads the FP32 copy in both mixed modes (2^-24 relative; a BF16 hi + lo pair's
        // 2^-17 would make the J error of a density change dominate its K error)
        if (prec != DFJK_PREC_FP64) ensure_B32(h);
        if (prec == DFJK_PREC_BF16X3 && doK) ensure_B32T(h);
    }

    // J workspace in core: the symmetrised tile-packed density, d_Q, the d_Q partials per
    // tile pair and the J partials per Q-chunk (J is one pass over the range, below)
    double* W = nullptr;
    double *dpart = nullptr, *dvec = nullptr, *Jpart = nullptr;
    int jchunk = 0, jnchunk = 0, qchunk = 0;
    if (doJ) {
        W = wk.get<double>(h->ws_W, qs);
        dvec = wk.get<double>(h->ws_d, h->naux);  // d_Q at its absolute Q (the adjoint reads it afterwards)
        if (!h->ooc) {
            dpart = wk.get<double>(h->ws_dpart, static_cast<size_t>(h->ntp) * h->naux);
            // about 4096 workgroups for the second pass (fewer chunks in a fixed workspace)
            int want_wg = 4096;
            if (wk.fixed())
                want_wg = static_cast<int>(std::max<size_t>(
                    h->ntp, std::min<size_t>(4096, wk.left() / 8 / (TILE2 * sizeof(double)))));
            jnchunk = std::max(1, std::min(nq_total, (want_wg + h->ntp - 1) / h->ntp));
            qchunk = (h->naux + jnchunk - 1) / jnchunk;
            jnchunk = (h->naux + qchunk - 1) / qchunk;
            Jpart = wk.get<double>(h->ws_Jpart, static_cast<size_t>(jnchunk) * qs);
        }
    }

    // K set-up: output tiles, Q-splits, contraction passes, packed C
    // symmetric K: contract the upper block-triangle only; the cross form contracts
    // M = X_A X_V^T over every tile in one pass and the reduction forms K = M + M^T
    const bool upper = sym;
    const int ntile = upper ? h->ntile_sym : h->ntile_all;
    const int2* otiles = upper ? h->d_otiles_sym.as<int2>() : h->d_otiles_all.as<int2>();
    const size_t per_q = xq * esz * (general ? 2 : 1);  // X bytes per Q
    const size_t csz = static_cast<size_t>(nbfp) * ldc;
    int qblk = nq_total, nsplit = 1;
    double* P = nullptr;
    if (!wk.fixed()) {
        // bytes of X workspace (16 GiB FP64, 4 GiB in the mixed modes, whose contraction runs
        // best on smaller Q-blocks), or what the memory setting leaves beyond the resident tensor
        size_t budget = prec == DFJK_PREC_FP64 ? g_x_budget : g_x_budget_mixed;
        budget = memory_limited(h, budget,
                                (h->ooc ? 0 : b_doubles(h)) * sizeof(double) +
                                    static_cast<size_t>(h->naux) * h->naux * sizeof(double));
        if (doK) {
            qblk = static_cast<int>(std::min<size_t>(nq_total, std::max<size_t>(1, budget / per_q)));
            qblk = std::min(qblk, 65535);
        }
        if (h->ooc) {
            // two FP64 stage buffers (plus their FP32 copies in the mixed modes) share the budget with X
            const size_t conv = prec == DFJK_PREC_FP64 ? 0 : prec == DFJK_PREC_FP32 ? 4 : 8;
            const size_t pq = qs * (2 * sizeof(double) + conv);
            const size_t stage_budget = std::max<size_t>(pq, std::min<size_t>(size_t(8) << 30, budget / 2));
            qblk = std::min<int>(qblk, static_cast<int>(std::max<size_t>(1, stage_budget / pq)));
        }
        const int nblocks0 = (nq_total + qblk - 1) / qblk;
        qblk = (nq_total + nblocks0 - 1) / nblocks0;
        if (doK) {
            nsplit = choose_nsplit(ntile, qblk);
            P = wk.get<double>(h->ws_P, static_cast<size_t>(nsplit) * ntile * 16384);
        }
    } else if (doK) {
        // fixed workspace: the K partials take at most a quarter of what is left, the packed C
        // copies their share, and X the rest
        const size_t tile_bytes = static_cast<size_t>(ntile) * 16384 * sizeof(double);
        const int cap_split = static_cast<int>(std::max<size_t>(1, wk.left() / 4 / tile_bytes));
        nsplit = std::min(choose_nsplit(ntile, std::min(nq_total, 65535)), cap_split);
        P = wk.get<double>(h->ws_P, static_cast<size_t>(nsplit) * ntile * 16384);
    }
    // contraction passes over k: (kbeg, kend, alpha, column offsets of the two operands)
    struct Pass {
        int kbeg, kend;
        double alpha;
        size_t offl, offr;
    };
    // padding columns of X are zero: each pass stops at the contraction's k-step after
    // the last real column of its group
    const int kstep = prec == DFJK_PREC_BF16X3 ? 32 : 16;
    std::vector<Pass> passes;
    if (signed_sym) {
        if (!plus.empty()) passes.push_back({0, round_up(static_cast<int>(plus.size()), kstep), 1.0, 0, 0});
        passes.push_back({kplus, kplus + round_up(static_cast<int>(minus.size()), kstep), -1.0, 0, 0});
    } else if (cross) {
        passes.push_back({0, round_up(half, kstep), 1.0, 0, static_cast<size_t>(wcross)});
    } else {
        passes.push_back({0, std::min(ldc, round_up(ncol, kstep)), 1.0, 0, 0});
    }
    const float *c32l = nullptr, *c32r = nullptr;
    uint16_t *cth_l = nullptr, *ctl_l = nullptr, *cth_r = nullptr, *ctl_r = nullptr;
    char *XL = nullptr, *XR = nullptr;
    if (doK) {
        if (prec == DFJK_PREC_FP32) {
            float* a = wk.get<float>(h->ws_c32, 2 * csz);
            launch_pack_cmap_f32(dCl, ncol, nbf, dmap, a, nbfp, ldc, s);
            c32l = c32r = a;
            if (general) {
                launch_pack_cmap_f32(dCr, ncol, nbf, dmap, a + csz, nbfp, ldc, s);
                c32r = a + csz;
            }
        } else if (prec == DFJK_PREC_BF16X3) {
            uint16_t* a = wk.get<uint16_t>(h->ws_cbf, 4 * csz);
            cth_l = a;
            ctl_l = a + csz;
            launch_pack_cT_bf(dCl, ncol, nbf, dmap, cth_l, ctl_l, nbfp, ldc, s);
            cth_r = cth_l;
            ctl_r = ctl_l;
            if (general) {
                cth_r = a + 2 * csz;
                ctl_r = a + 3 * csz;
                launch_pack_cT_bf(dCr, ncol, nbf, dmap, cth_r, ctl_r, nbfp, ldc, s);
            }
        }
        if (wk.fixed()) {
            const size_t room = wk.left() > 1024 ? wk.left() - 1024 : 0;  // alignment of the two X arrays
            qblk = static_cast<int>(std::min<size_t>(std::min(nq_total, 65535), room / per_q));
            if (qblk < 1)
                fail(DFJK_ERR_MEMORY, "dfjk_compute_device_noalloc: one auxiliary function's half-transform does not "
                                      "fit the " + std::to_string(wk.cap >> 20) + " MiB scratch array");
            const int nb0 = (nq_total + qblk - 1) / qblk;
            qblk = (nq_total + nb0 - 1) / nb0;
        }
        XL = wk.get<char>(h->ws_XL, xq * esz * qblk);
        XR = general ? wk.get<char>(h->ws_XR, xq * esz * qblk) : XL;
    }
    const int nblocks = (nq_total + qblk - 1) / qblk;
    if (doJ && h->ooc) {
        // out of core: d_Q partials and J partials per streamed block
        dpart = wk.get<double>(h->ws_dpart, static_cast<size_t>(h->ntp) * qblk);
        const int want_wg = 4096;
        jnchunk = std::max(1, std::min(qblk, (want_wg + h->ntp - 1) / h->ntp));
        jchunk = (qblk + jnchunk - 1) / jnchunk;
        jnchunk = (qblk + jchunk - 1) / jchunk;
        Jpart = wk.get<double>(h->ws_Jpart, static_cast<size_t>(jnchunk) * qs);
    }
    if (doJ) {
        hipEvent_t e0 = tm.begin();
        launch_pack_w(g, Dm, nbf, W, s);
        tm.end(T_J, e0);
    }

    // In core, J is one pass over the whole range (two streaming kernels over the resident
    // tensor, or its lower-precision copy); out of core it is accumulated block by block below.
    if (doJ && !h->ooc) {
        JSource js;
        if (prec != DFJK_PREC_FP64) {
            js.kind = JSource::F32;
            js.a = h->dB32.p;
        } else {
            js.a = h->dB.p;
        }
        hipEvent_t e0 = tm.begin();
        if (q0 > 0 || q1 < h->naux) launch_zero(dpart, static_cast<size_t>(h->ntp) * h->naux, s);
        launch_j1(g, js, W, h->naux, q0, q1, dpart, s);
        launch_j1_reduce(g, dpart, h->naux, dvec, s);
        launch_j2(g, js, dvec, h->naux, qchunk, Jpart, s);
        launch_j_unpack(g, Jpart, jnchunk, dJ, nbf, 1.0, accumulate, s);
        tm.end(T_J, e0);
    }
    auto stage_copy = [&](int blk) {  // out of core: host block -> stage buffer blk & 1
        const int qa = q0 + blk * qblk;
        const int nq = std::min(qblk, q1 - qa);
        double* dst = wk.get<double>(h->ws_stage[blk & 1], qs * qblk);
        hcheck(hipStreamWaitEvent(h->cstream, h->ev_free[blk & 1], 0), "wait stage");
        hcheck(hipMemcpyAsync(dst, h->hB + static_cast<size_t>(qa) * qs, qs * nq * sizeof(double),
                              hipMemcpyHostToDevice, h->cstream),
               "stage B");
        hcheck(hipEventRecord(h->ev_copied[blk & 1], h->cstream), "record");
    };
    if (h->ooc) {
        // both stage buffers start free
        for (int i = 0; i < 2; ++i) hcheck(hipEventRecord(h->ev_free[i], s), "record");
        stage_copy(0);
    }

    for (int blk = 0; blk < nblocks; ++blk) {
        const int qa = q0 + blk * qblk;
        const int nq = std::min(qblk, q1 - qa);
        const size_t nel = qs * nq;
        // this block's views of the tensor in every precision the kernels read
        const double* b64;
        const float *b32 = nullptr, *b32t = nullptr;
        if (!h->ooc) {
            const size_t off = static_cast<size_t>(qa) * qs;
            b64 = h->dB.as<double>() + off;
            if (prec != DFJK_PREC_FP64) b32 = h->dB32.as<float>() + off;
            if (prec == DFJK_PREC_BF16X3 && doK) b32t = h->dB32T.as<float>() + off;
        } else {
            if (blk + 1 < nblocks) stage_copy(blk + 1);
            hcheck(hipStreamWaitEvent(s, h->ev_copied[blk & 1], 0), "wait copy");
            double* st = h->ws_stage[blk & 1].as<double>();
            b64 = st;
            hipEvent_t e0 = tm.begin();
            if (prec != DFJK_PREC_FP64) {
                float* c = wk.get<float>(h->ws_stage32, qs * qblk);
                launch_to_f32(st, c, nel, s);
                b32 = c;
            }
            if (prec == DFJK_PREC_BF16X3 && doK) {
                float* c = wk.get<float>(h->ws_stage32t, qs * qblk);
                launch_to_f32_t(st, c, nel, s);
                b32t = c;
            }
            tm.end(T_OTHER, e0);
        }

        if (doJ && h->ooc) {
            JSource js;
            if (prec != DFJK_PREC_FP64) {
                js.kind = JSource::F32;
                js.a = b32;
            } else {
                js.a = b64;
            }
            hipEvent_t e0 = tm.begin();
            const int nch = (nq + jchunk - 1) / jchunk;
            launch_j1(g, js, W, nq, 0, nq, dpart, s);
            launch_j1_reduce(g, dpart, nq, dvec + qa, s);
            launch_j2(g, js, dvec + qa, nq, jchunk, Jpart, s);
            launch_j_unpack(g, Jpart, nch, dJ, nbf, 1.0, (blk > 0 || accumulate) ? 1 : 0, s);
            tm.end(T_J, e0);
        }

        if (doK) {
            const int nsp = std::min(nsplit, nq);
            hipEvent_t e0 = tm.begin();
            if (prec == DFJK_PREC_FP64) {
                launch_ht2_f64(g, b64, 0, nq, cl, ldc, nbw, npanel, reinterpret_cast<double*>(XL), xrows, h->dZero.p,
                               s);
                if (general)
                    launch_ht2_f64(g, b64, 0, nq, cr, ldc, nbw, npanel, reinterpret_cast<double*>(XR), xrows,
                                   h->dZero.p, s);
            } else if (prec == DFJK_PREC_FP32) {
                launch_ht2_f32(g, b32, 0, nq, c32l, ldc, nbw, npanel, reinterpret_cast<float*>(XL), xrows, h->dZero.p,
                               s);
                if (general)
                    launch_ht2_f32(g, b32, 0, nq, c32r, ldc, nbw, npanel, reinterpret_cast<float*>(XR), xrows,
                                   h->dZero.p, s);
            } else {
                uint16_t* xl = reinterpret_cast<uint16_t*>(XL);
                uint16_t* xr = reinterpret_cast<uint16_t*>(XR);
                launch_ht_bf3(g, b32, b32t, 0, nq, cth_l, ctl_l, ldc, nbw, npanel, xl, xl + xq * qblk, xrows, h->dZero.p, s);
                if (general)
                    launch_ht_bf3(g, b32, b32t, 0, nq, cth_r, ctl_r, ldc, nbw, npanel, xr, xr + xq * qblk, xrows,
                                  h->dZero.p, s);
            }
            tm.end(T_HT, e0);
            h->n_ht += general ? 2 : 1;
            hipEvent_t e1 = tm.begin();
            for (size_t ip = 0; ip < passes.size(); ++ip) {
                const Pass& pp = passes[ip];
                const int acc = (blk > 0 || ip > 0) ? 1 : 0;
                if (prec == DFJK_PREC_FP64) {
                    double* xl = reinterpret_cast<double*>(XL) + pp.offl;
                    double* xr = reinterpret_cast<double*>(XR) + pp.offr;
                    launch_kgemm2_f64(xl, xr, xq, xrows, nbf, ldc, pp.kbeg, pp.kend, nq, nsp, otiles, ntile, P, pp.alpha,
                                      acc, upper ? 1 : 0, s);
                } else if (prec == DFJK_PREC_FP32) {
                    float* xl = reinterpret_cast<float*>(XL) + pp.offl;
                    float* xr = reinterpret_cast<float*>(XR) + pp.offr;
                    launch_kgemm2_f32(xl, xr, xq, xrows, nbf, ldc, pp.kbeg, pp.kend, nq, nsp, otiles, ntile, P, pp.alpha,
                                      acc, upper ? 1 : 0, s);
                } else {
                    uint16_t* xl = reinterpret_cast<uint16_t*>(XL) + pp.offl;
                    uint16_t* xr = reinterpret_cast<uint16_t*>(XR) + pp.offr;
                    launch_kgemm_bf3(xl, xl + xq * qblk, xr, xr + xq * qblk, xq, xrows, nbf, ldc, pp.kbeg, pp.kend, nq, nsp,
                                     otiles, ntile, P, pp.alpha, acc, upper ? 1 : 0, s);
                }
                h->n_kgemm += 1;
            }
            if (nsp < nsplit && blk == 0)
                launch_zero(P + static_cast<size_t>(nsp) * ntile * 16384,
                            static_cast<size_t>(nsplit - nsp) * ntile * 16384, s);
            tm.end(T_KGEMM, e1);
        }
        if (h->ooc) hcheck(hipEventRecord(h->ev_free[blk & 1], s), "record");
    }
    if (doK) {
        hipEvent_t e2 = tm.begin();
        launch_tile_reduce(P, nsplit, ntile, upper ? h->d_otile_of_sym.as<int>() : h->d_otile_of_all.as<int>(), h->nB,
                           upper ? 1 : (cross ? 2 : 0), nbf, nbf, dK, nbf, 1.0, accumulate, s);
        tm.end(T_OTHER, e2);
    }
}

// wK of one pair (dfhelper.cc:3378-3438): wK = sum_Q (B_Q C_left)(Bw_Q C_right)^T with
// Bw = M (Q|erf(omega r)/r|mn), so that wK_mn = sum_{PQ,i} ((P|mi) C_l) [J^-1]_PQ ((Q|w|ni) C_r) as
// psi4 forms it from J^-1 (Q|mn) and the raw attenuated tensor.  FP64; every output tile is
// contracted (the product of two different tensors is not symmetric).
void compute_wk_pair(dfjk_handle* h, const double* dCl, const double* dCr, int ncol, double* dwK, int accumulate) {
    const TileGeom g = h->geom();
    const hipStream_t s = h->stream;
    const int q0 = h->aux_q0, q1 = h->aux_q1;
    const int nq_total = q1 - q0;
    const int nbf = h->nbf, nbfp = h->nbfp;
    const int xrows = (h->nT + 3) / 4 * 128;
    EventTimer& tm = h->timer;
    int nbw, npanel, ldc;
    ht_panel_shape(ncol, &nbw, &npanel, &ldc);
    std::vector<int> cmap(ldc, -1);
    for (int i = 0; i < ncol; ++i) cmap[i] = i;
    Work alloc_wk;
    int* dmap = upload_cmap(h, cmap, alloc_wk);
    double* cl = h->ws_cl.get<double>(static_cast<size_t>(nbfp) * ldc);
    launch_pack_cmap_f64(dCl, ncol, nbf, dmap, cl, nbfp, ldc, s);
    double* cr = cl;
    if (dCr) {
        cr = h->ws_cr.get<double>(static_cast<size_t>(nbfp) * ldc);
        launch_pack_cmap_f64(dCr, ncol, nbf, dmap, cr, nbfp, ldc, s);
    }
    const size_t xq = static_cast<size_t>(xrows) * ldc;
    const size_t budget = memory_limited(
        h, g_x_budget, (2 * b_doubles(h) + static_cast<size_t>(h->naux) * h->naux) * sizeof(double));
    const size_t per_q = 2 * xq * sizeof(double);
    int qblk = static_cast<int>(std::min<size_t>(nq_total, std::max<size_t>(1, budget / per_q)));
    qblk = std::min(qblk, 65535);
    const int nblocks = (nq_total + qblk - 1) / qblk;
    qblk = (nq_total + nblocks - 1) / nblocks;
    const int ntile = h->ntile_all;
    const int2* otiles = h->d_otiles_all.as<int2>();
    const int nsplit = choose_nsplit(ntile, qblk);
    double* P = h->ws_P.get<double>(static_cast<size_t>(nsplit) * ntile * 16384);
    const int kend = std::min(ldc, round_up(ncol, 16));
    double* XL = h->ws_XL.get<double>(xq * qblk);
    double* XR = h->ws_XR.get<double>(xq * qblk);
    for (int blk = 0; blk < nblocks; ++blk) {
        const int qa = q0 + blk * qblk;
        const int nq = std::min(qblk, q1 - qa);
        const int nsp = std::min(nsplit, nq);
        hipEvent_t e0 = tm.begin();
        launch_ht2_f64(g, h->dB.as<double>(), qa, nq, cl, ldc, nbw, npanel, XL, xrows, h->dZero.p, s);
        launch_ht2_f64(g, h->dBw.as<double>(), qa, nq, cr, ldc, nbw, npanel, XR, xrows, h->dZero.p, s);
        tm.end(T_HT, e0);
        h->n_ht += 2;
        hipEvent_t e1 = tm.begin();
        launch_kgemm2_f64(XL, XR, xq, xrows, nbf, ldc, 0, kend, nq, nsp, otiles, ntile, P, 1.0, blk > 0 ? 1 : 0, 0, s);
        h->n_kgemm += 1;
        if (nsp < nsplit && blk == 0)
            hcheck(hipMemsetAsync(P + static_cast<size_t>(nsp) * ntile * 16384, 0,
                                  sizeof(double) * static_cast<size_t>(nsplit - nsp) * ntile * 16384, s),
                   "memset");
        tm.end(T_KGEMM, e1);
    }
    hipEvent_t e2 = tm.begin();
    launch_tile_reduce(P, nsplit, ntile, h->d_otile_of_all.as<int>(), h->nB, 0, nbf, nbf, dwK, nbf, 1.0, accumulate,
                       s);
    tm.end(T_OTHER, e2);
}

void finish_timers(dfjk_handle* h) {
    if (!h->timer.on) return;
    hcheck(hipStreamSynchronize(h->stream), "sync");
    for (auto& r : h->timer.recs) {
        float ms = 0.f;
        hcheck(hipEventElapsedTime(&ms, r.a, r.b), "hipEventElapsedTime");
        h->t_ms[r.cls] += ms;
        h->timer.pool.push_back(r.a);
        h->timer.pool.push_back(r.b);
    }
    h->timer.recs.clear();
}

void reset_timers(dfjk_handle* h) {
    for (double& t : h->t_ms) t = 0.0;
    h->n_ht = h->n_kgemm = 0;
    ensure_stream(h);
}

// arena: null for the allocating entry points, else the device address of this device's
// static scratch array (dfjk_compute_device_noalloc; device pointers only, no timing events)
void compute_impl(dfjk_handle* h, int n, const double* const* Cl, const double* const* Cr, const int* ncol,
                  double* const* D, double* const* J, double* const* K, int accumulate, bool device,
                  char* arena = nullptr, size_t arena_bytes = 0) {
    require_state(h, true);
    if (n < 0 || (n > 0 && (!Cl || !ncol))) fail(DFJK_ERR_ARG, "JK: bad C_left arguments");
    const bool timing = h->timer.on;
    if (arena) h->timer.on = false;
    struct TimerRestore {
        dfjk_handle* h;
        bool on;
        ~TimerRestore() { h->timer.on = on; }
    } restore{h, timing};
    reset_timers(h);
    const size_t nn = static_cast<size_t>(h->nbf) * h->nbf;
    for (int i = 0; i < n; ++i) {
        if (ncol[i] < 0) fail(DFJK_ERR_ARG, "JK: negative column count");
        if (!Cl[i] && ncol[i] > 0) fail(DFJK_ERR_ARG, "JK: null C_left matrix");
        if (Cr && !Cr[i] && ncol[i] > 0) fail(DFJK_ERR_ARG, "JK: null C_right matrix");
        double* dD = D ? D[i] : nullptr;
        double* dJ = J ? J[i] : nullptr;
        double* dK = K ? K[i] : nullptr;
        if (dK && h->do_wK && h->wcombine) {
            // combined exchange: K is not built (dfhelper.cc:3022) and stays as MemDFJK::compute_JK's
            // zero() left it; the exchange comes back in wK
            if (!accumulate) {
                if (device)
                    launch_zero(dK, nn, h->stream);
                else
                    std::memset(dK, 0, nn * sizeof(double));
            }
            dK = nullptr;
        }
        if (ncol[i] == 0) {
            // psi4: D = 0, J += 0, K skipped for nocc == 0 (dfhelper.cc:3355-3358)
            if (device) {
                if (dD) launch_zero(dD, nn, h->stream);
                if (dJ && !accumulate) launch_zero(dJ, nn, h->stream);
                if (dK && !accumulate) launch_zero(dK, nn, h->stream);
            } else {
                if (dD) std::memset(dD, 0, nn * sizeof(double));
                if (dJ && !accumulate) std::memset(dJ, 0, nn * sizeof(double));
                if (dK && !accumulate) std::memset(dK, 0, nn * sizeof(double));
            }
            continue;
        }
        const size_t csz = static_cast<size_t>(h->nbf) * ncol[i];
        const double* cl = Cl[i];
        const double* cr = Cr ? Cr[i] : nullptr;
        double *gD = dD, *gJ = dJ, *gK = dK;
        if (!device) {
            double* raw = h->ws_craw.get<double>(2 * csz);
            copy_settled(raw, cl, csz * sizeof(double), hipMemcpyHostToDevice, h->stream, "upload C_left");
            cl = raw;
            if (cr) {
                copy_settled(raw + csz, cr, csz * sizeof(double), hipMemcpyHostToDevice, h->stream, "upload C_right");
                cr = raw + csz;
            }
            gD = dD ? h->ws_D.get<double>(nn) : nullptr;
            gJ = dJ ? h->ws_J.get<double>(nn) : nullptr;
            gK = dK ? h->ws_K.get<double>(nn) : nullptr;
            if (accumulate) {
                if (gJ) copy_settled(gJ, dJ, nn * sizeof(double), hipMemcpyHostToDevice, h->stream, "up");
                if (gK) copy_settled(gK, dK, nn * sizeof(double), hipMemcpyHostToDevice, h->stream, "up");
            }
        }
        Work wk;  // per pair: the pairs run in stream order, so a fixed arena is reused
        wk.base = arena;
        wk.cap = arena_bytes;
        compute_pair(h, cl, cr, ncol[i], gD, gJ, gK, accumulate, wk);
        if (!device) {
            if (dD) copy_settled(dD, gD, nn * sizeof(double), hipMemcpyDeviceToHost, h->stream, "download D");
            if (dJ && h->do_J)
                copy_settled(dJ, gJ, nn * sizeof(double), hipMemcpyDeviceToHost, h->stream, "download J");
            if (dK && h->do_K)
                copy_settled(dK, gK, nn * sizeof(double), hipMemcpyDeviceToHost, h->stream, "download K");
            hcheck(hipStreamSynchronize(h->stream), "compute");
        }
    }
    finish_timers(h);
}


void compute_wk_impl(dfjk_handle* h, int n, const double* const* Cl, const double* const* Cr, const int* ncol,
                     double* const* wK, int accumulate, bool device) {
    require_state(h, true);
    if (n < 0 || (n > 0 && (!Cl || !ncol || !wK))) fail(DFJK_ERR_ARG, "JK: bad wK arguments");
    require_device_B(h, "wK");
    if (h->tensor_mode) {
        if (!h->dBw.p) fail(DFJK_ERR_STATE, "wK needs the attenuated tensor: dfjk_set_Bw_dense");
        if (h->bw_comb != h->wcombine || (h->wcombine && (h->bw_alpha != h->omega_alpha || h->bw_beta != h->omega_beta)))
            fail(DFJK_ERR_STATE, "wcombine / omega_alpha / omega_beta changed after dfjk_set_Bw_dense: set it again");
    } else if (!bw_current(h)) {
        build_Bw(h);
    }
    reset_timers(h);
    const size_t nn = static_cast<size_t>(h->nbf) * h->nbf;
    for (int i = 0; i < n; ++i) {
        if (ncol[i] < 0) fail(DFJK_ERR_ARG, "JK: negative column count");
        if (!Cl[i] && ncol[i] > 0) fail(DFJK_ERR_ARG, "JK: null C_left matrix");
        if (Cr && !Cr[i] && ncol[i] > 0) fail(DFJK_ERR_ARG, "JK: null C_right matrix");
        double* dW = wK[i];
        if (!dW) continue;
        if (ncol[i] == 0) {
            if (!accumulate) {
                if (device)
                    hcheck(hipMemsetAsync(dW, 0, nn * sizeof(double), h->stream), "memset");
                else
                    std::memset(dW, 0, nn * sizeof(double));
            }
            continue;
        }
        const size_t csz = static_cast<size_t>(h->nbf) * ncol[i];
        const double* cl = Cl[i];
        const double* cr = Cr ? Cr[i] : nullptr;
        double* gW = dW;
        if (!device) {
            double* raw = h->ws_craw.get<double>(2 * csz);
            copy_settled(raw, cl, csz * sizeof(double), hipMemcpyHostToDevice, h->stream, "upload C_left");
            cl = raw;
            if (cr) {
                copy_settled(raw + csz, cr, csz * sizeof(double), hipMemcpyHostToDevice, h->stream, "upload C_right");
                cr = raw + csz;
            }
            gW = h->ws_K.get<double>(nn);
            if (accumulate) copy_settled(gW, dW, nn * sizeof(double), hipMemcpyHostToDevice, h->stream, "up");
        }
        compute_wk_pair(h, cl, cr, ncol[i], gW, accumulate);
        // lr_symmetric: MemDFJK::compute_JK hermitivitizes wK (MemDFJK.cc:101-107)
        if (!cr) launch_hermitivitize(gW, h->nbf, h->nbf, h->stream);
        if (!device) {
            copy_settled(dW, gW, nn * sizeof(double), hipMemcpyDeviceToHost, h->stream, "download wK");
            hcheck(hipStreamSynchronize(h->stream), "compute wK");
        }
    }
    finish_timers(h);
}

// Adjoint of E = tr(D(2J - K)), D = C C^T, at FP64 (the DF gradient terms of
// scfgrad/jk_grad.cc:294-475 contract the same c_Q = d_Q and (Q|ij) = C^T B_Q C):
//   dE/dC = 4 (2J - K) C,   Gamma_Qmn = dE/dB_Qmn = 4 d_Q D_mn - 2 (C (C^T X_Q) C^T)_mn, X_Q = B_Q C.
// E and dE/dC use the whole auxiliary range; Gamma is produced for the configured range
// [q0, q1), in blocks of Q.  dC is a device [nbf][nocc] matrix.  Gamma blocks go to
// gam_dev (device, [q1-q0][nbf][nbf]) or, when gam_host is set, are copied to host memory.
void adjoint_impl(dfjk_handle* h, const double* dC, int nocc, double* E, double* dEdC, double* gam_dev,
                  double* gam_host) {
    // E and dE/dC come from J and K, which stream B out of core too; Gamma needs B in core
    if (h->ooc && (gam_dev || gam_host))
        fail(DFJK_ERR_STATE, "DFJK: dE/dB of the energy adjoint needs B in core (subalgorithm INCORE)");
    const hipStream_t s = h->stream;
    const int nbf = h->nbf, nbfp = h->nbfp;
    const size_t nn = static_cast<size_t>(nbf) * nbf;
    const int uq0 = h->aux_q0, uq1 = h->aux_q1;
    const int prec = h->precision;
    const bool dj = h->do_J, dk = h->do_K;
    double* D = h->ws_D.get<double>(nn);
    double* Jm = h->ws_J.get<double>(nn);
    double* Km = h->ws_K.get<double>(nn);
    // full-range FP64 J and K (this also leaves d_Q for every Q in ws_d)
    h->aux_q0 = 0;
    h->aux_q1 = h->naux;
    h->precision = DFJK_PREC_FP64;
    h->do_J = h->do_K = true;
    try {
        Work wk;
        compute_pair(h, dC, nullptr, nocc, D, Jm, Km, 0, wk);
    } catch (...) {
        h->aux_q0 = uq0, h->aux_q1 = uq1, h->precision = prec, h->do_J = dj, h->do_K = dk;
        throw;
    }
    h->aux_q0 = uq0, h->aux_q1 = uq1, h->precision = prec, h->do_J = dj, h->do_K = dk;
    const double* dvec = h->ws_d.as<double>();

    // G = 2J - K; E = <D, G> (fixed-order reduction); dE/dC = 4 G C
    h->dM_ready = false;
    double* G = h->ws_G.get<double>(nn);
    launch_axpby_mat(Jm, 2.0, Km, -1.0, G, nn, s);
    if (E) {
        constexpr int kRed = 256;
        double* part = h->ws_red.get<double>(kRed);
        launch_dot_partial(D, G, nn, part, kRed, s);
        std::vector<double> hp(kRed);
        copy_settled(hp.data(), part, kRed * sizeof(double), hipMemcpyDeviceToHost, s, "energy");
        double e = 0.0;
        for (double v : hp) e += v;
        *E = e;
    }
    if (dEdC) launch_gemm_nn_f64(G, nbf, dC, nocc, dEdC, nocc, nbf, nocc, nbf, 4.0, 0.0, s);
    if (!gam_dev && !gam_host) return;

    // Gamma over [uq0, uq1) in blocks: X_Q = B_Q C (half-transform), Y_Q = C^T X_Q,
    // P_Q = C Y_Q, Gamma_Q = 4 d_Q D - 2 P_Q C^T.
    const TileGeom g = h->geom();
    const int xrows = (h->nT + 3) / 4 * 128;
    int nbw, npanel, ldc;
    ht_panel_shape(nocc, &nbw, &npanel, &ldc);
    std::vector<int> cmap(ldc, -1);
    for (int i = 0; i < nocc; ++i) cmap[i] = i;
    Work alloc_wk;
    int* dmap = upload_cmap(h, cmap, alloc_wk);
    double* cl = h->ws_cl.get<double>(static_cast<size_t>(nbfp) * ldc);
    launch_pack_cmap_f64(dC, nocc, nbf, dmap, cl, nbfp, ldc, s);
    double* CT = h->ws_ct.get<double>(static_cast<size_t>(nocc) * nbf);
    launch_transpose(dC, nbf, nocc, nocc, CT, nbf, s);

    const size_t xq = static_cast<size_t>(xrows) * ldc;
    const size_t yq = static_cast<size_t>(nocc) * nocc, pq = static_cast<size_t>(nbf) * nocc;
    const size_t per_q = (xq + yq + pq + (gam_host ? nn : 0)) * sizeof(double);
    const size_t budget = memory_limited(h, size_t(2) << 30, (b_doubles(h) + 4 * nn) * sizeof(double));
    const int nq_total = uq1 - uq0;
    int qblk = static_cast<int>(std::min<size_t>(nq_total, std::max<size_t>(1, budget / per_q)));
    qblk = std::min(qblk, 65535);
    double* X = h->ws_XL.get<double>(xq * qblk);
    double* Y = h->ws_Y.get<double>(yq * qblk);
    double* P = h->ws_PC.get<double>(pq * qblk);
    double* Gb = gam_host ? h->ws_gam.get<double>(nn * qblk) : nullptr;
    for (int qa = uq0; qa < uq1; qa += qblk) {
        const int nq = std::min(qblk, uq1 - qa);
        double* out = gam_host ? Gb : gam_dev + static_cast<size_t>(qa - uq0) * nn;
        launch_ht2_f64(g, h->dB.as<double>(), qa, nq, cl, ldc, nbw, npanel, X, xrows, h->dZero.p, s);
        launch_gemm_nn_f64(CT, nbf, X, ldc, Y, nocc, nocc, nocc, nbf, 1.0, 0.0, s, nq, 0, xq, yq);
        launch_gemm_nn_f64(dC, nocc, Y, nocc, P, nocc, nbf, nocc, nocc, 1.0, 0.0, s, nq, 0, yq, pq);
        launch_fill_scaled(dvec, qa, nq, 4.0, D, nbf, out, s);
        launch_gemm_nn_f64(P, nocc, CT, nbf, out, nbf, nbf, nbf, nocc, -2.0, 1.0, s, nq, pq, 0, nn);
        if (gam_host) {
            copy_settled(gam_host + static_cast<size_t>(qa - uq0) * nn, Gb, nn * nq * sizeof(double),
                         hipMemcpyDeviceToHost, s, "download dE/dB");
        }
    }
}

void adjoint_entry(dfjk_handle* h, const double* C, int nocc, double* E, double* dEdC, double* dEdB, bool device) {
    require_state(h, true);
    if (nocc <= 0 || !C) fail(DFJK_ERR_ARG, "energy adjoint: C must be nbf x nocc with nocc > 0");
    const size_t csz = static_cast<size_t>(h->nbf) * nocc;
    reset_timers(h);
    if (device) {
        adjoint_impl(h, C, nocc, E, dEdC, dEdB, nullptr);
        hcheck(hipStreamSynchronize(h->stream), "adjoint");
    } else {
        double* dc = h->ws_craw.get<double>(csz);
        copy_settled(dc, C, csz * sizeof(double), hipMemcpyHostToDevice, h->stream, "upload C");
        double* g = dEdC ? h->ws_dc.get<double>(csz) : nullptr;
        adjoint_impl(h, dc, nocc, E, g, nullptr, dEdB);
        if (dEdC) copy_settled(dEdC, g, csz * sizeof(double), hipMemcpyDeviceToHost, h->stream, "dE/dC");
        hcheck(hipStreamSynchronize(h->stream), "adjoint");
    }
    finish_timers(h);
}

}  // namespace

// ============================================================== C ABI
extern "C" {

const char* dfjk_last_error(void) { return g_err.c_str(); }

int dfjk_create(dfjk_handle** out, const p4a_basis* primary, const p4a_basis* aux) {
    return guarded([&] {
        if (!out || !primary || !aux) fail(DFJK_ERR_ARG, "dfjk_create: null argument");
        auto h = std::make_unique<dfjk_handle>();
        h->prim = basis_from_desc(primary);
        h->aux = basis_from_desc(aux);
        h->nbf = h->prim.nbf;
        h->naux = h->aux.nbf;
        h->aux_q1 = h->naux;
        h->nthreads = std::max(1, static_cast<int>(std::thread::hardware_concurrency()));
        *out = h.release();
    });
}

int dfjk_create_tensor(dfjk_handle** out, int nbf, int naux) {
    return guarded([&] {
        if (!out || nbf <= 0 || naux <= 0) fail(DFJK_ERR_ARG, "dfjk_create_tensor: bad sizes");
        auto h = std::make_unique<dfjk_handle>();
        h->tensor_mode = true;
        h->nbf = nbf;
        h->naux = naux;
        h->aux_q1 = naux;
        h->nthreads = std::max(1, static_cast<int>(std::thread::hardware_concurrency()));
        *out = h.release();
    });
}

int dfjk_destroy(dfjk_handle* h) {
    return guarded([&] {
        if (!h) return;
        if (h->own_stream) (void)hipStreamSynchronize(h->own_stream);
        if (h->cstream) (void)hipStreamSynchronize(h->cstream);
        hipStream_t s = h->own_stream, cs = h->cstream;
        for (int i = 0; i < 2; ++i) {
            if (h->ev_copied[i]) (void)hipEventDestroy(h->ev_copied[i]);
            if (h->ev_free[i]) (void)hipEventDestroy(h->ev_free[i]);
        }
        if (h->hB) (void)hipHostFree(h->hB);
        if (h->cmap_pin) (void)hipHostFree(h->cmap_pin);
        if (h->sign_pin) (void)hipHostFree(h->sign_pin);
        if (h->ev_cmap) (void)hipEventDestroy(h->ev_cmap);
        delete h;
        if (cs) (void)hipStreamDestroy(cs);
        if (s) (void)hipStreamDestroy(s);
    });
}

int dfjk_set_memory(dfjk_handle* h, size_t doubles) {
    return guarded([&] {
        require_state(h, false);
        h->memory_doubles = doubles;
    });
}
int dfjk_set_cutoff(dfjk_handle* h, double cutoff) {
    return guarded([&] {
        require_state(h, false);
        if (!(cutoff >= 0.0)) fail(DFJK_ERR_ARG, "cutoff must be >= 0");
        if (!h->initialized && cutoff != h->cutoff) {
            // a mask built for an earlier estimate is rebuilt with the new cutoff
            h->tiles.clear();
            h->fun_mask.clear();
        }
        h->cutoff = cutoff;
    });
}
int dfjk_set_condition(dfjk_handle* h, double condition) {
    return guarded([&] {
        require_state(h, false);
        h->condition = condition;
    });
}
int dfjk_set_do_J(dfjk_handle* h, int v) {
    return guarded([&] {
        require_state(h, false);
        h->do_J = v != 0;
    });
}
int dfjk_set_do_K(dfjk_handle* h, int v) {
    return guarded([&] {
        require_state(h, false);
        h->do_K = v != 0;
    });
}
int dfjk_set_do_wK(dfjk_handle* h, int v) {
    return guarded([&] {
        require_state(h, false);
        if (v && h->initialized && h->ooc)  // dfhelper.cc:201-205
            fail(DFJK_ERR_STATE, "DFJK: not equipped to do wK out of core; supply more memory or disable wK");
        h->do_wK = v != 0;
    });
}
int dfjk_set_subalgorithm(dfjk_handle* h, int v) {
    return guarded([&] {
        require_state(h, false);
        if (v < DFJK_SUBALGO_AUTO || v > DFJK_SUBALGO_OUT_OF_CORE) fail(DFJK_ERR_ARG, "unknown subalgorithm");
        if (h->initialized) fail(DFJK_ERR_STATE, "the subalgorithm must be chosen before initialize()");
        h->subalgo = v;
    });
}
int dfjk_get_setup_timings(dfjk_handle* h, double* ms, int n) {
    return guarded([&] {
        require_state(h, false);
        if (n < 0 || (n > 0 && !ms)) fail(DFJK_ERR_ARG, "bad timing buffer");
        for (int i = 0; i < n && i < 6; ++i) ms[i] = h->setup_ms[i];
    });
}
int dfjk_set_wcombine(dfjk_handle* h, int wcombine) {
    return guarded([&] {
        require_state(h, false);
        h->wcombine = wcombine != 0;  // takes effect at the next wK compute
    });
}
int dfjk_set_omega_alpha(dfjk_handle* h, double alpha) {
    return guarded([&] {
        require_state(h, false);
        if (!std::isfinite(alpha)) fail(DFJK_ERR_ARG, "omega_alpha must be finite");
        h->omega_alpha = alpha;
    });
}
int dfjk_set_omega_beta(dfjk_handle* h, double beta) {
    return guarded([&] {
        require_state(h, false);
        if (!std::isfinite(beta)) fail(DFJK_ERR_ARG, "omega_beta must be finite");
        h->omega_beta = beta;
    });
}
int dfjk_set_omega(dfjk_handle* h, double omega) {
    return guarded([&] {
        require_state(h, false);
        if (!(omega >= 0.0) || !std::isfinite(omega)) fail(DFJK_ERR_ARG, "omega must be finite and >= 0");
        h->omega = omega;  // a built attenuated tensor is rebuilt at the next wK compute
    });
}
int dfjk_set_omp_nthread(dfjk_handle* h, int n) {
    return guarded([&] {
        require_state(h, false);
        if (n < 1) fail(DFJK_ERR_ARG, "omp_nthread must be >= 1");
        h->nthreads = n;
        h->omp_nthread = n;
    });
}
int dfjk_set_precision(dfjk_handle* h, int p) {
    return guarded([&] {
        require_state(h, false);
        if (p < DFJK_PREC_FP64 || p > DFJK_PREC_BF16X3) fail(DFJK_ERR_ARG, "unknown precision");
        h->precision = p;
    });
}
int dfjk_set_fitting(dfjk_handle* h, int f) {
    return guarded([&] {
        require_state(h, false);
        if (f != DFJK_FIT_CHOLESKY && f != DFJK_FIT_EIGEN) fail(DFJK_ERR_ARG, "unknown fitting");
        if (h->initialized) fail(DFJK_ERR_STATE, "fitting must be chosen before initialize()");
        h->fitting = f;
    });
}
int dfjk_set_stream(dfjk_handle* h, void* stream) {
    return guarded([&] {
        require_state(h, false);
        h->stream = stream ? static_cast<hipStream_t>(stream) : h->own_stream;
        h->timer.s = h->stream;
    });
}
int dfjk_set_profiling(dfjk_handle* h, int on) {
    return guarded([&] {
        require_state(h, false);
        h->timer.on = on != 0;
    });
}
int dfjk_set_aux_range(dfjk_handle* h, int q0, int q1) {
    return guarded([&] {
        require_state(h, false);
        if (q0 < 0 || q1 > h->naux || q0 >= q1) fail(DFJK_ERR_ARG, "aux range out of bounds");
        h->aux_q0 = q0;
        h->aux_q1 = q1;
    });
}

int dfjk_memory_estimate(dfjk_handle* h, size_t* doubles) {
    return guarded([&] {
        require_state(h, false);
        if (!doubles) fail(DFJK_ERR_ARG, "null output");
        ensure_sparsity(h);
        // DFHelper::AO_core (dfhelper.cc:216-236) with do_wK unset, as MemDFJK::memory_estimate
        // calls it: the pQq tensor naux * nsig, the metric, the C buffers, the Q-shell temporaries
        const size_t nbf = static_cast<size_t>(h->nbf), naux = static_cast<size_t>(h->naux);
        size_t est = naux * static_cast<size_t>(h->nsig);
        est += naux * naux;
        est += static_cast<size_t>(h->omp_nthread) * nbf * nbf;
        est += 3 * nbf * nbf * static_cast<size_t>(qshell_max(h));
        *doubles = est;
    });
}

int dfjk_device_footprint(dfjk_handle* h, size_t* doubles) {
    return guarded([&] {
        require_state(h, false);
        if (!doubles) fail(DFJK_ERR_ARG, "null output");
        ensure_sparsity(h);
        const size_t nbfp = static_cast<size_t>(round_up(h->nbf, TILE));
        size_t est = static_cast<size_t>(h->naux) * h->ntp * TILE2;  // resident B
        if (h->do_wK) est *= 2;                                      // and the attenuated Bw
        est += copy_doubles(h);  // FP32 copies of B at the current precision
        est += static_cast<size_t>(h->naux) * h->naux;              // metric (fit only)
        est += 4 * static_cast<size_t>(h->nbf) * h->nbf;            // D, J, K, staging
        est += 2 * nbfp * 320;                                      // padded C panels
        *doubles = est;
    });
}

int dfjk_get_pair_mask(dfjk_handle* h, uint8_t* mask, double* cutoff) {
    return guarded([&] {
        require_state(h, false);
        ensure_sparsity(h);
        if (mask) std::memcpy(mask, h->fun_mask.data(), h->fun_mask.size());
        if (cutoff) *cutoff = h->cutoff;
    });
}

int dfjk_get_schwarz(dfjk_handle* h, double* fun_values, double* shell_values, double* max_integral,
                     int* on_device) {
    return guarded([&] {
        require_state(h, false);
        if (h->tensor_mode) fail(DFJK_ERR_STATE, "DFJK: a tensor-level handle has no screening integrals");
        ensure_sparsity(h);
        if (fun_values) std::memcpy(fun_values, h->schwarz_fun.data(), h->schwarz_fun.size() * sizeof(double));
        if (shell_values)
            std::memcpy(shell_values, h->schwarz_shell.data(), h->schwarz_shell.size() * sizeof(double));
        if (max_integral) *max_integral = h->schwarz_max;
        if (on_device) *on_device = h->schwarz_on_device ? 1 : 0;
    });
}

int dfjk_get_fit_info(dfjk_handle* h, int* fitting, double* condition, int* naux_kept) {
    return guarded([&] {
        require_state(h, false);
        if (fitting) *fitting = h->fitting;
        if (condition) *condition = h->condition;
        if (naux_kept) *naux_kept = h->initialized && !h->tensor_mode ? h->fit_kept : 0;
    });
}

int dfjk_set_pair_form(dfjk_handle* h, int form) {
    return guarded([&] {
        require_state(h, false);
        if (form < DFJK_FORM_DETECT || form > DFJK_FORM_CROSS) fail(DFJK_ERR_ARG, "unknown pair form");
        h->pair_form = form;
    });
}

int dfjk_get_aux_range(dfjk_handle* h, int* q0, int* q1) {
    return guarded([&] {
        require_state(h, false);
        if (q0) *q0 = h->aux_q0;
        if (q1) *q1 = h->aux_q1;
    });
}

int dfjk_initialize(dfjk_handle* h) {
    return guarded([&] {
        require_state(h, false);
        if (h->initialized) return;
        ensure_stream(h);
        // pinned staging of the column map and the structure test, and the staging event
        if (!h->cmap_pin) {
            hcheck(hipHostMalloc(reinterpret_cast<void**>(&h->cmap_pin), 16384 * sizeof(int), 0), "hipHostMalloc");
            h->cmap_pin_cap = 16384;
        }
        if (!h->sign_pin) {
            hcheck(hipHostMalloc(reinterpret_cast<void**>(&h->sign_pin), 16384, 0), "hipHostMalloc");
            h->sign_pin_cap = 16384;
        }
        if (!h->ev_cmap) hcheck(hipEventCreateWithFlags(&h->ev_cmap, hipEventDisableTiming), "event");
        // the device this handle's buffers live on, and its copy of the scratch array
        hcheck(hipGetDevice(&h->device), "hipGetDevice");
        h->scratch_dev = scratch_of_device(h->device);
        if (h->tensor_mode) {
            if (h->subalgo == DFJK_SUBALGO_OUT_OF_CORE)
                fail(DFJK_ERR_STATE, "DFJK: tensor-level handles are in core only");
            ensure_sparsity(h);
            upload_tiles(h);
            h->dB.get<double>(b_doubles(h));
            hcheck(hipMemsetAsync(h->dB.p, 0, b_doubles(h) * sizeof(double), h->stream), "memset B");
            hcheck(hipStreamSynchronize(h->stream), "init");
        } else {
            initialize_basis(h);
        }
        h->initialized = true;
    });
}

namespace {
// Pack a dense symmetric [naux][nbf][nbf] tensor (host or device) into the tiled layout of dst.
void set_dense(dfjk_handle* h, const double* B, int on_device, DevBuf& dst) {
    require_state(h, true);
    if (h->ooc) fail(DFJK_ERR_STATE, "DFJK: dense tensors are set in core only");
    if (!B) fail(DFJK_ERR_ARG, "null B");
    const size_t nn = static_cast<size_t>(h->nbf) * h->nbf;
    if (h->fun_mask.size() != nn) fail(DFJK_ERR_INTERNAL, "screening mask missing");
    const TileGeom g = h->geom();
    const hipStream_t s = h->stream;
    double* out = dst.get<double>(b_doubles(h));
    if (on_device) {
        for (int q0 = 0; q0 < h->naux; q0 += 32768) {
            const int nq = std::min(h->naux - q0, 32768);
            launch_pack_dense(g, B + static_cast<size_t>(q0) * nn, h->d_mask.as<uint8_t>(), q0, nq, out, s);
        }
    } else {
        // stage host Q-slices of about 256 MiB through a device buffer
        const int qc = static_cast<int>(std::max<size_t>(1, std::min<size_t>(32768, (size_t(32) << 20) / nn)));
        double* stage = h->ws_XL.get<double>(nn * qc);
        for (int q0 = 0; q0 < h->naux; q0 += qc) {
            const int nq = std::min(h->naux - q0, qc);
            copy_settled(stage, B + static_cast<size_t>(q0) * nn, nn * nq * sizeof(double), hipMemcpyHostToDevice, s,
                         "upload B");
            launch_pack_dense(g, stage, h->d_mask.as<uint8_t>(), q0, nq, out, s);
        }
    }
    hcheck(hipStreamSynchronize(s), "set B");
}

// Unpack Q in [q0,q1) of a tiled tensor into dense [q1-q0][nbf][nbf] host memory.
void get_dense(dfjk_handle* h, const DevBuf& src, int q0, int q1, double* out) {
    require_state(h, true);
    if (!out || q0 < 0 || q1 > h->naux || q0 >= q1) fail(DFJK_ERR_ARG, "bad range");
    if (!src.p && !((&src == &h->dB) && h->ooc)) fail(DFJK_ERR_STATE, "tensor not built");
    const int n = h->nbf;
    const size_t nn = static_cast<size_t>(n) * n;
    const size_t qs = static_cast<size_t>(h->ntp) * TILE2;
    std::vector<double> packed(qs);
    const bool host = (&src == &h->dB) && h->ooc;
    for (int Q = q0; Q < q1; ++Q) {
        if (host)
            std::memcpy(packed.data(), h->hB + Q * qs, qs * sizeof(double));
        else
            copy_settled(packed.data(), src.as<double>() + Q * qs, qs * sizeof(double), hipMemcpyDeviceToHost,
                         h->stream, "download B");
        double* o = out + (Q - q0) * nn;
        std::fill(o, o + nn, 0.0);
        for (int t = 0; t < h->ntp; ++t) {
            const int2 IJ = h->tiles[t];
            for (int r = 0; r < TILE; ++r)
                for (int c = 0; c < TILE; ++c) {
                    const int m = IJ.x * TILE + r, q = IJ.y * TILE + c;
                    if (m >= n || q >= n) continue;
                    const double v = packed[static_cast<size_t>(t) * TILE2 + r * TILE + c];
                    o[static_cast<size_t>(m) * n + q] = v;
                    o[static_cast<size_t>(q) * n + m] = v;
                }
        }
    }
}
}  // namespace

int dfjk_set_B_dense(dfjk_handle* h, const double* B, int on_device) {
    return guarded([&] {
        require_state(h, true);
        invalidate_copies(h);
        set_dense(h, B, on_device, h->dB);
    });
}

int dfjk_set_Bw_dense(dfjk_handle* h, const double* Bw, int on_device) {
    return guarded([&] {
        if (h && h->wcombine) require_state(h, true);  // combined with B on arrival
        set_dense(h, Bw, on_device, h->dBw);
        h->bw_omega = h->omega;
        combine_Bw(h);
    });
}

int dfjk_fill_synthetic(dfjk_handle* h, uint64_t seed) {
    return guarded([&] {
        require_state(h, true);
        require_device_B(h, "dfjk_fill_synthetic");
        invalidate_copies(h);
        std::vector<double> decay(h->nbf);
        for (int k = 0; k < h->nbf; ++k) decay[k] = std::exp(-static_cast<double>(k) / 64.0);
        double* dd = h->ws_d.get<double>(std::max(h->nbf, h->naux));
        copy_settled(dd, decay.data(), decay.size() * sizeof(double), hipMemcpyHostToDevice, h->stream, "upload decay");
        const TileGeom g = h->geom();
        for (int q0 = 0; q0 < h->naux; q0 += 32768) {
            const int q1 = std::min(h->naux, q0 + 32768);
            launch_synth_b(g, dd, seed, q0, q1, h->dB.as<double>(), h->stream);
        }
        hcheck(hipStreamSynchronize(h->stream), "synthetic B");
    });
}

int dfjk_get_B_dense(dfjk_handle* h, int q0, int q1, double* out) {
    return guarded([&] { get_dense(h, h->dB, q0, q1, out); });
}

int dfjk_get_Bw_dense(dfjk_handle* h, int q0, int q1, double* out) {
    return guarded([&] {
        require_state(h, true);
        if (!h->tensor_mode && !bw_current(h)) build_Bw(h);
        get_dense(h, h->dBw, q0, q1, out);
    });
}

int dfjk_compute(dfjk_handle* h, int n, const double* const* Cl, const double* const* Cr, const int* ncol,
                 double* const* D, double* const* J, double* const* K, int accumulate) {
    return guarded([&] { compute_impl(h, n, Cl, Cr, ncol, D, J, K, accumulate, false); });
}

int dfjk_compute_device(dfjk_handle* h, int n, const double* const* Cl, const double* const* Cr, const int* ncol,
                        double* const* D, double* const* J, double* const* K, int accumulate) {
    return guarded([&] { compute_impl(h, n, Cl, Cr, ncol, D, J, K, accumulate, true); });
}

int dfjk_compute_device_noalloc(dfjk_handle* h, int n, const double* const* Cl, const double* const* Cr,
                                const int* ncol, double* const* D, double* const* J, double* const* K, int accumulate) {
    return guarded([&] {
        require_state(h, true);
        if (!h->scratch_dev) fail(DFJK_ERR_STATE, "dfjk_compute_device_noalloc: no scratch address for this device");
        compute_impl(h, n, Cl, Cr, ncol, D, J, K, accumulate, true, h->scratch_dev, kScratchBytes);
    });
}

int dfjk_compute_wK(dfjk_handle* h, int n, const double* const* Cl, const double* const* Cr, const int* ncol,
                    double* const* wK, int accumulate) {
    return guarded([&] { compute_wk_impl(h, n, Cl, Cr, ncol, wK, accumulate, false); });
}

int dfjk_compute_wK_device(dfjk_handle* h, int n, const double* const* Cl, const double* const* Cr, const int* ncol,
                           double* const* wK, int accumulate) {
    return guarded([&] { compute_wk_impl(h, n, Cl, Cr, ncol, wK, accumulate, true); });
}

int dfjk_energy_adjoint(dfjk_handle* h, const double* C, int nocc, double* E, double* dEdC, double* dEdB) {
    return guarded([&] { adjoint_entry(h, C, nocc, E, dEdC, dEdB, false); });
}

int dfjk_energy_adjoint_device(dfjk_handle* h, const double* C, int nocc, double* E, double* dEdC, double* dEdB) {
    return guarded([&] { adjoint_entry(h, C, nocc, E, dEdC, dEdB, true); });
}

int dfjk_get_timings(dfjk_handle* h, double* ms_j, double* ms_ht, double* ms_kgemm, double* ms_other, int* n_ht,
                     int* n_kgemm) {
    return guarded([&] {
        require_state(h, false);
        if (ms_j) *ms_j = h->t_ms[T_J];
        if (ms_ht) *ms_ht = h->t_ms[T_HT];
        if (ms_kgemm) *ms_kgemm = h->t_ms[T_KGEMM];
        if (ms_other) *ms_other = h->t_ms[T_OTHER];
        if (n_ht) *n_ht = h->n_ht;
        if (n_kgemm) *n_kgemm = h->n_kgemm;
    });
}

int dfjk_get_info(dfjk_handle* h, int* nbf, int* naux, int* ntile_pairs, int64_t* nsig_pairs) {
    return guarded([&] {
        require_state(h, false);
        if (nbf) *nbf = h->nbf;
        if (naux) *naux = h->naux;
        if (ntile_pairs) *ntile_pairs = h->ntp;
        if (nsig_pairs) *nsig_pairs = h->nsig;
    });
}

int dfjk_get_out_of_core(dfjk_handle* h, int* ooc) {
    return guarded([&] {
        require_state(h, false);
        if (!ooc) fail(DFJK_ERR_ARG, "null output");
        *ooc = h->ooc ? 1 : 0;
    });
}

int dfjk_finalize(dfjk_handle* h) {
    return guarded([&] {
        require_state(h, false);
        if (h->initialized && h->stream) hcheck(hipStreamSynchronize(h->stream), "finalize");
        DevBuf* ws[] = {&h->ws_craw, &h->ws_cl,  &h->ws_cr,  &h->ws_D,    &h->ws_W,  &h->ws_dpart, &h->ws_d,
                        &h->ws_Jpart, &h->ws_XL, &h->ws_XR,  &h->ws_P,    &h->ws_Pd, &h->ws_J,     &h->ws_K,
                        &h->ws_G,    &h->ws_sign, &h->ws_c32, &h->ws_cbf, &h->ws_cmap, &h->ws_ct,   &h->ws_red,
                        &h->ws_Y,    &h->ws_PC,  &h->ws_gam, &h->ws_dc, &h->ws_stage[0], &h->ws_stage[1],
                        &h->ws_stage32, &h->ws_stage32t};
        for (DevBuf* b : ws) b->release();
        h->cmap_dev.clear();
        h->dM_ready = false;
    });
}

}  // extern "C"
