// K-path kernels for gfx950: half-transform X_Q = B_Q C and the contraction
// K = sum_Q X_Q S X_Q^T (S = diag of column signs, or a second operand X'),
// in FP64, FP32 and 3-term BF16 split precision.
//
// All kernels are software pipelined: the next K-step's global loads are
// issued into registers before the current step's MFMAs, then written to the
// other half of a double-buffered LDS stage, with one barrier per step.
//
// MFMA lane layouts (checked on the device):
//   f64 16x16x4 : A[i][k] lane i+16k; B[k][j] lane j+16k; D[i][j] lane j+16(i%4), reg i/4
//   f32 16x16x4 : A, B as above;                          D[i][j] lane j+16(i/4), reg i%4
//   bf16 16x16x32: A[i][8g+e] lane i+16g elem e; B[8g+e][j] lane j+16g elem e; D as f32
#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

#include "kernels.hpp"
#include "kernels_k.hpp"

namespace p4a {

namespace {

__device__ inline uint16_t f2bf(float x) {
    uint32_t u = __float_as_uint(x);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return static_cast<uint16_t>(u >> 16);
}
__device__ inline float bf2f(uint16_t h) { return __uint_as_float(static_cast<uint32_t>(h) << 16); }

// Element e of a 128 x KS chunk of A: row tile e / (32*KS), then (r, c) in the stored
// tile's own orientation so that global reads are contiguous for both orientations.
struct ACoord {
    int sub, r, c;
};
template <int KS>
__device__ inline ACoord a_coord(int e, int I0, int J) {
    ACoord a;
    a.sub = e / (32 * KS);
    const int f = e % (32 * KS);
    if (I0 + a.sub <= J) {
        a.r = f / KS;
        a.c = f % KS;
    } else {
        a.c = f / 32;
        a.r = f % 32;
    }
    return a;
}

// Workgroups are dispatched round-robin over the XCDs (flat id w runs on XCD w % 8,
// each XCD with its own L2).  xcd_map turns the flat id into (group, item) so that
// all items of one group -- the row blocks of one Q in the half-transform, the
// output tiles of one Q-split in the contraction -- run on the same XCD and share
// its L2.  The launch uses 8 * ceil(ngroups/8) * items workgroups.
constexpr int kXCD = 8;
struct XcdItem {
    int group, item;
    bool valid;
};
__device__ inline XcdItem xcd_map(int wg, int items, int ngroups) {
    const int xcd = wg % kXCD, local = wg / kXCD;
    XcdItem d;
    d.group = (local / items) * kXCD + xcd;
    d.item = local % items;
    d.valid = d.group < ngroups;
    return d;
}
inline unsigned xcd_grid(int ngroups, int items) {
    return static_cast<unsigned>(kXCD * ((ngroups + kXCD - 1) / kXCD) * items);
}

// Per-workgroup list of the column tiles J that meet at least one of the
// workgroup's RT row tiles, with the element offset of each (row tile, J) block
// in B_Q (or -1), built once in LDS at kernel start (parallel presence test and
// an inclusive scan), so the K loop does no dependent tile-table lookups.
constexpr int kMaxColTiles = 256;  // nbf <= 8192
// k-terms an FP32 MFMA accumulator of the K contraction holds before it is folded into FP64
constexpr int kFold32 = 512;

template <int RT>
struct StepList {
    int off[kMaxColTiles * RT];
    int J[kMaxColTiles];
    int cnt[kMaxColTiles];
    int n;
};

template <int RT>
__device__ inline void build_steps(StepList<RT>& sl, const int* __restrict__ tidx, int nT, int I0, int tid, int nthr) {
    for (int j = tid; j < kMaxColTiles; j += nthr) {
        int present = 0;
#pragma unroll
        for (int s = 0; s < RT; ++s) {
            const int I = I0 + s;
            int t = -1;
            if (j < nT && I < nT) t = tidx[min(I, j) * nT + max(I, j)];
            sl.off[j * RT + s] = t >= 0 ? t * TILE2 : -1;
            present |= (t >= 0);
        }
        sl.cnt[j] = present;
    }
    __syncthreads();
    // inclusive scan of the presence flags (Hillis-Steele over kMaxColTiles entries)
    for (int d = 1; d < kMaxColTiles; d <<= 1) {
        int v = 0;
        if (tid < kMaxColTiles && tid >= d) v = sl.cnt[tid - d];
        __syncthreads();
        if (tid < kMaxColTiles) sl.cnt[tid] += v;
        __syncthreads();
    }
    // compact: step k -> J (offsets are kept indexed by J)
    if (tid < kMaxColTiles) {
        const int prev = tid ? sl.cnt[tid - 1] : 0;
        if (sl.cnt[tid] != prev) sl.J[prev] = tid;
        if (tid == kMaxColTiles - 1) sl.n = sl.cnt[tid];
    }
    __syncthreads();
}

// =====================================================================================
// Half-transform, FP64 / FP32 (16x16x4 MFMA).  WM*256 threads = WM*4 waves (WM rows x
// 4 cols), workgroup tile 64*WM rows x 64*NBW columns, wave tile 64 x 16*NBW.
// K-step 8 (FP64) or 16 (FP32).  Absent tiles read a device zero tile (no branches).
// T = double or float; B is the FP64 tensor (T=double) or its FP32 copy (T=float).
// grid = (xrows / (64*WM), nq, npanel)
// =====================================================================================
template <typename T, int NBW, int WM>
__global__ __launch_bounds__(WM * 256) __attribute__((amdgpu_waves_per_eu(2, 2))) void k_ht_mfma(const T* __restrict__ B, const int* __restrict__ tidx, int nT,
                                                      int ntp, int q0, int nq, int npanel, const T* __restrict__ C,
                                                      int ldc, T* __restrict__ X, int xrows,
                                                      const T* __restrict__ ztile) {
    constexpr bool F64 = sizeof(T) == 8;
    constexpr int NTHR = WM * 256;
    constexpr int RT = 2 * WM;            // row tiles per workgroup
    constexpr int ROWS = 64 * WM;
    constexpr int PAN = 64 * NBW;
    constexpr int KS = F64 ? (WM == 1 ? 4 : 8) : 16;
    constexpr int AP = KS + 1;
    constexpr int CP = PAN + 4;
    constexpr int NA = ROWS * KS / NTHR;  // A elements per thread per step
    constexpr int NC = KS * PAN / NTHR;   // C elements per thread per step
    static_assert((ROWS * KS) % NTHR == 0 && (KS * PAN) % NTHR == 0, "staging shape");
    __shared__ T As[2][ROWS * AP];
    __shared__ T Cs[2][KS * CP];
    __shared__ StepList<RT> sl;
    typedef typename std::conditional<F64, d4_t, f4_t>::type acc_t;

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave >> 2, wn = wave & 3;
    const int nrb = xrows / ROWS;
    const XcdItem xi = xcd_map(blockIdx.x, nrb * npanel, nq);
    if (!xi.valid) return;
    const int rb = xi.item / npanel;
    const int I0 = RT * rb;
    const int q = xi.group;
    const int col0 = (xi.item % npanel) * PAN;
    const T* BQ = B + static_cast<size_t>(q0 + q) * ntp * TILE2;

    build_steps<RT>(sl, tidx, nT, I0, tid, NTHR);
    const int nsteps = sl.n * (TILE / KS);

    acc_t acc[4][NBW];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < NBW; ++j) acc[i][j] = acc_t{0, 0, 0, 0};

    T ra[NA], rc[NC];
    // Column tile of the next step to load, read from the LDS step list one step
    // ahead, so a step's loads wait on one LDS lookup (the block offsets) instead of two.
    // The FP64 5-column-block shape looks the tile up in place (measured faster there).
    constexpr bool PF = !(F64 && NBW == 5);
    int nJ = 0;
    auto fetch = [&](int st) {
        if constexpr (PF) nJ = sl.J[st / (TILE / KS)];
    };
    int lJ = 0;  // column tile of the step held in ra/rc
    auto gload = [&](int st) {
        int J;
        if constexpr (PF)
            J = nJ;
        else
            J = sl.J[st / (TILE / KS)];
        lJ = J;
        const int kh = (st % (TILE / KS)) * KS;
#pragma unroll
        for (int it = 0; it < NA; ++it) {
            const ACoord ac = a_coord<KS>(tid + NTHR * it, I0, J);
            const int off = sl.off[J * RT + ac.sub];
            const T* tp = off >= 0 ? BQ + off : ztile;
            // B is streamed (non-temporal) so that the C panel every workgroup re-reads stays in L2
            ra[it] = __builtin_nontemporal_load((I0 + ac.sub <= J) ? tp + ac.r * TILE + kh + ac.c
                                                                   : tp + (kh + ac.c) * TILE + ac.r);
        }
        const T* Cj = C + static_cast<size_t>(J * TILE + kh) * ldc + col0;
#pragma unroll
        for (int it = 0; it < NC; ++it) {
            const int e = tid + NTHR * it;
            const int kk = e / PAN, cc = e - kk * PAN;
            rc[it] = Cj[static_cast<size_t>(kk) * ldc + cc];
        }
        if (st + 1 < nsteps) fetch(st + 1);
    };
    auto lstore = [&](int buf, int st) {
        int J;
        if constexpr (PF)
            J = lJ;
        else
            J = sl.J[st / (TILE / KS)];
#pragma unroll
        for (int it = 0; it < NA; ++it) {
            const ACoord ac = a_coord<KS>(tid + NTHR * it, I0, J);
            As[buf][(ac.sub * 32 + ac.r) * AP + ac.c] = ra[it];
        }
#pragma unroll
        for (int it = 0; it < NC; ++it) {
            const int e = tid + NTHR * it;
            const int kk = e / PAN, cc = e - kk * PAN;
            Cs[buf][kk * CP + cc] = rc[it];
        }
    };
    // One step's products: per 4-k sub-step the A fragments of the next sub-step and the C
    // fragment of the next column block are read while the current block's MFMAs run, so
    // the LDS reads overlap the MFMAs (consecutive updates of one accumulator are 4*NBW
    // MFMAs apart).
    // The last row block can reach past the last row tile: a wave's 16-row fragments there
    // keep zero accumulators, and their A reads and MFMAs are skipped (NI = fragments kept).
    constexpr int NSUB = KS / 4;
    const int nfr = __builtin_amdgcn_readfirstlane(min(4, max(0, 2 * (nT - I0 - 2 * wm))));
    auto compute = [&](int buf, auto ni) {
        constexpr int NI = decltype(ni)::value;
        T a[2][4], b[2];
        auto aread = [&](int u, int slot) {
#pragma unroll
            for (int i = 0; i < NI; ++i)
                a[slot][i] = As[buf][(wm * 64 + i * 16 + (lane & 15)) * AP + 4 * u + (lane >> 4)];
        };
        auto bread = [&](int u, int j, int slot) {
            b[slot] = Cs[buf][(4 * u + (lane >> 4)) * CP + wn * NBW * 16 + j * 16 + (lane & 15)];
        };
        aread(0, 0);
        bread(0, 0, 0);
#pragma unroll
        for (int u = 0; u < NSUB; ++u) {
#pragma unroll
            for (int j = 0; j < NBW; ++j) {
                const int t = u * NBW + j;  // running block index: C fragment slot t & 1
                if (j + 1 < NBW)
                    bread(u, j + 1, (t + 1) & 1);
                else if (u + 1 < NSUB)
                    bread(u + 1, 0, (t + 1) & 1);
                if (j == 0 && u + 1 < NSUB) aread(u + 1, (u + 1) & 1);
                __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                for (int i = 0; i < NI; ++i) {
                    if constexpr (F64)
                        acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[u & 1][i], b[t & 1], acc[i][j], 0, 0, 0);
                    else
                        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[u & 1][i], b[t & 1], acc[i][j], 0, 0, 0);
                }
                __builtin_amdgcn_sched_barrier(0);
            }
        }
    };

    auto mainloop = [&](auto ni) {
        fetch(0);
        gload(0);
        lstore(0, 0);
        __syncthreads();
        int buf = 0;
        for (int st = 0; st < nsteps; ++st) {
            const bool more = st + 1 < nsteps;
            if (more) gload(st + 1);
            __builtin_amdgcn_sched_barrier(0);
            if constexpr (decltype(ni)::value > 0) compute(buf, ni);
            if (more) lstore(buf ^ 1, st + 1);
            __syncthreads();
            buf ^= 1;
        }
    };
    if (nsteps > 0) {
        if (nfr == 4)
            mainloop(std::integral_constant<int, 4>{});
        else if (nfr == 2)
            mainloop(std::integral_constant<int, 2>{});
        else
            mainloop(std::integral_constant<int, 0>{});
    }
    // every row of the workgroup's block is written (absent row tiles give zeros)
    T* Xq = X + static_cast<size_t>(q) * xrows * ldc;
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < NBW; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int rr = F64 ? (4 * r + (lane >> 4)) : (4 * (lane >> 4) + r);
                const int m = ROWS * rb + wm * 64 + i * 16 + rr;
                Xq[static_cast<size_t>(m) * ldc + col0 + wn * NBW * 16 + j * 16 + (lane & 15)] = acc[i][j][r];
            }
}

// BF16 LDS tiles of 32 k per row (64 bytes): the four 16-byte chunks of row r are stored at
// chunk c ^ ((r>>1)&3), so fragment reads (16 rows x one chunk per 16 lanes) and row-segment
// stores (4 lanes per row) are free of bank conflicts.
__device__ inline int kg_swz(int row, int chunk) { return row * 32 + ((chunk ^ ((row >> 1) & 3)) << 3); }

// Split 8 FP32 values into BF16 hi = RNE(x) and lo = RNE(x - hi), packed two per word
// (v_cvt_pk_bf16_f32 rounds to nearest even, as f2bf does).
typedef float v4f_t __attribute__((ext_vector_type(4)));
typedef float f2v_t __attribute__((ext_vector_type(2)));
typedef __bf16 bf2v_t __attribute__((ext_vector_type(2)));
typedef unsigned int v4u_t __attribute__((ext_vector_type(4)));
__device__ inline void split_bf16x8(const v4f_t& x, const v4f_t& y, v4u_t& h, v4u_t& l) {
    auto hi = [](float a, float b) { return __builtin_convertvector(f2v_t{a, b}, bf2v_t); };
    auto lo = [](float a, float b, bf2v_t hb) {
        return __builtin_convertvector(f2v_t{a, b} - __builtin_convertvector(hb, f2v_t), bf2v_t);
    };
    const bf2v_t h0 = hi(x[0], x[1]), h1 = hi(x[2], x[3]), h2 = hi(y[0], y[1]), h3 = hi(y[2], y[3]);
    const bf2v_t l0 = lo(x[0], x[1], h0), l1 = lo(x[2], x[3], h1), l2 = lo(y[0], y[1], h2), l3 = lo(y[2], y[3], h3);
    h = v4u_t{__builtin_bit_cast(unsigned int, h0), __builtin_bit_cast(unsigned int, h1),
              __builtin_bit_cast(unsigned int, h2), __builtin_bit_cast(unsigned int, h3)};
    l = v4u_t{__builtin_bit_cast(unsigned int, l0), __builtin_bit_cast(unsigned int, l1),
              __builtin_bit_cast(unsigned int, l2), __builtin_bit_cast(unsigned int, l3)};
}

// =====================================================================================
// Half-transform, 3-term BF16 split: X = Bh Ch + Bh Cl + Bl Ch with FP32 accumulation.
// B is read from its FP32 copies and split into hi/lo in registers before the LDS store (the
// same 4 bytes per element as stored hi/lo copies); a block used in the transposed
// orientation is read from the tile-transposed FP32 copy, so every A vector is a 32-byte
// row segment.
// C transposed ([ldc][nbfp], k contiguous) hi/lo.  K-step = one column tile (32).  WM x WN
// waves, workgroup tile 64*WM rows x 16*NBW*WN columns (WN = 4: one 8-wave workgroup per CU;
// WN = 2: two 4-wave workgroups per CU, whose barriers are independent).  DC: the C panel
// goes global -> LDS by direct loads (no staging registers, no LDS store instructions)
// instead of through registers.  Output X split again into hi/lo BF16 ([nq][xrows][ldc]).
// =====================================================================================
// s_waitcnt immediate (gfx9 encoding: vmcnt bits 3:0 and 15:14, expcnt 6:4, lgkmcnt 11:8)
// that waits until at most n vector-memory operations of this wave are outstanding
constexpr int wait_vm(int n) { return 0x0F70 | (n & 15) | ((n >> 4) << 14); }
__device__ inline void fence_compiler() { __asm__ __volatile__("" ::: "memory"); }
// (the target builtins below exist only in the device pass; the host pass of a kernel
// template that calls them directly would drop the kernel's launch stub)
template <int N>
__device__ inline void wait_vmem() {
#if defined(__HIP_DEVICE_COMPILE__)
    __builtin_amdgcn_s_waitcnt(wait_vm(N));
#endif
}
// 16 bytes per lane from global memory straight into LDS: lane L of the wave writes
// lds + 16 L bytes (lds is wave-uniform)
__device__ inline void load_lds16(const uint16_t* src, uint16_t* lds) {
#if defined(__HIP_DEVICE_COMPILE__)
    __builtin_amdgcn_global_load_lds(reinterpret_cast<const uint4*>(src), lds, 16, 0, 0);
#endif
}

template <int NBW, int WM, int WN, bool DC>
__global__ __launch_bounds__(WM * WN * 64) __attribute__((amdgpu_waves_per_eu(2, 2))) void k_ht_bf3(const float* __restrict__ B,
                                                     const float* __restrict__ BT,
                                                     const int* __restrict__ tidx, int nT, int ntp, int q0, int nq,
                                                     int npanel, const uint16_t* __restrict__ CTh,
                                                     const uint16_t* __restrict__ CTl,
                                                     int nbfp, int ldc, uint16_t* __restrict__ Xh,
                                                     uint16_t* __restrict__ Xl, int xrows,
                                                     const float* __restrict__ ztile) {
    constexpr int NTHR = WM * WN * 64;
    constexpr int RT = 2 * WM;
    constexpr int ROWS = 64 * WM;
    constexpr int PAN = 16 * NBW * WN;
    constexpr int A_SZ = ROWS * 32;          // swizzled 64-byte rows (kg_swz)
    constexpr int C_SZ = PAN * 32;
    constexpr int NAV = ROWS * 4 / NTHR;     // 16-byte A vectors per thread per array
    constexpr int NCV = PAN * 4 / NTHR;      // full rounds of 16-byte C vectors per thread per array
    constexpr int NCV_R = (PAN * 4) % NTHR;  // vectors in the last partial round
    static_assert((NAV == 1 || NAV == 2) && NAV * NTHR == ROWS * 4, "one or two A vectors per thread");
    static_assert(NCV + (NCV_R ? 1 : 0) <= 5, "C staging slots");
    static_assert(NCV_R % 4 == 0, "whole C rows per partial round");
    static_assert(NTHR >= kMaxColTiles, "build_steps scans with one thread per column tile");
    constexpr int NSLOT = NCV + (NCV_R ? 1 : 0);  // C vector rounds per thread
    // the two stage buffers are separate arrays, so the compiler can tell a direct load into
    // one from the fragment reads of the other and need not wait for it before them
    __shared__ __align__(16) uint16_t Ls0[2 * A_SZ + 2 * C_SZ];
    __shared__ __align__(16) uint16_t Ls1[2 * A_SZ + 2 * C_SZ];
    auto lbuf = [&](int buf) { return buf ? Ls1 : Ls0; };
    __shared__ StepList<RT> sl;

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave / WN, wn = wave % WN;
    const int nrb = xrows / ROWS;
    const XcdItem xi = xcd_map(blockIdx.x, nrb * npanel, nq);
    if (!xi.valid) return;
    const int rb = xi.item / npanel;
    const int I0 = RT * rb;
    const int q = xi.group;
    const int col0 = (xi.item % npanel) * PAN;
    const size_t qoff = static_cast<size_t>(q0 + q) * ntp * TILE2;

    build_steps<RT>(sl, tidx, nT, I0, tid, NTHR);
    const int nsteps = sl.n;

    f4_t acc[4][NBW];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < NBW; ++j) acc[i][j] = f4_t{0, 0, 0, 0};

    // A: per row tile 32x32 = 128 vectors of 8; RT row tiles -> NAV vectors per thread per
    // array, vector v = tid + NTHR * n in row tile v >> 7.
    // Staging registers.  The A vectors (B_Q blocks, streamed from HBM) are loaded two steps
    // before their LDS store, in two sets; the C vectors (a panel every workgroup re-reads,
    // mostly L2 hits) one step before, either into registers (C loads issued first in the
    // step, so the store's wait for them leaves the younger A loads in flight) or, with DC,
    // straight into the other LDS buffer (then the step's closing wait for the direct loads
    // waits for every vector-memory load, and the A loads have one step to land).
    typedef v4u_t v4u;
    v4f_t a0[2 * NAV], a1[2 * NAV];  // 8 FP32 of one A row segment per vector, per set
    v4u c0h, c0l, c1h, c1l, c2h, c2l, c3h, c3l, c4h, c4l;
    // column tile of step st; steps past the end re-read the last one (their data is never used)
    auto jof = [&](int st) { return sl.J[min(st, nsteps - 1)]; };
    // direct C loads: the panel's 16-row groups (PAN/16 per array, hi then lo) are dealt to
    // the waves; a load writes 16 rows of 64 bytes, lane L into 16-byte slot L, so lane L
    // fetches the logical chunk that the swizzle of kg_swz places in that slot
    constexpr int NWAVE = WM * WN;
    constexpr int NRG = PAN / 16;
    constexpr int NLD = 2 * NRG / NWAVE;
    static_assert(!DC || NLD * NWAVE == 2 * NRG, "whole direct loads per wave");
    auto issueC = [&](int st, int buf) {
        const size_t kb = static_cast<size_t>(jof(st)) * TILE;
        const int w = __builtin_amdgcn_readfirstlane(wave);
        const int r0 = lane >> 2, pc = lane & 3;
#pragma unroll
        for (int i = 0; i < NLD; ++i) {
            const int idx = w + NWAVE * i, arr = idx / NRG, rg = idx % NRG;
            const int r = rg * 16 + r0;
            const uint16_t* src = (arr ? CTl : CTh) + static_cast<size_t>(col0 + r) * nbfp + kb + 8 * (pc ^ ((r >> 1) & 3));
            load_lds16(src, lbuf(buf) + 2 * A_SZ + arr * C_SZ + rg * 512);
        }
    };
    // vector f = v & 127 covers row (f>>2) of its row tile's block, k = 8*(f&3)..+7
    auto gloadA = [&](int st, v4f_t(&a)[2 * NAV]) {
        const int J = jof(st);
#pragma unroll
        for (int n = 0; n < NAV; ++n) {
            const int v = tid + NTHR * n, sub = v >> 7, f = v & 127;
            const int off = sl.off[J * RT + sub];
            // a block stored as (J, I) (below the diagonal) is read from the tile-transposed
            // copy, so the vector is always a row segment of this row tile's block
            const bool tr = (I0 + sub > J);
            const float* t = off >= 0 ? (tr ? BT : B) + qoff + off : ztile;
            a[2 * n] = *reinterpret_cast<const v4f_t*>(t + (f >> 2) * TILE + 8 * (f & 3));
            a[2 * n + 1] = *reinterpret_cast<const v4f_t*>(t + (f >> 2) * TILE + 8 * (f & 3) + 4);
        }
    };
    // the A vectors split into hi/lo (done while the previous step's MFMAs run)
    v4u_t sah[NAV], sal[NAV];
    auto splitA = [&](const v4f_t(&a)[2 * NAV]) {
#pragma unroll
        for (int n = 0; n < NAV; ++n) split_bf16x8(a[2 * n], a[2 * n + 1], sah[n], sal[n]);
    };
    auto gloadC = [&](int st) {
        const size_t kb = static_cast<size_t>(jof(st)) * TILE;
        // every lane loads (idle lanes of the partial round re-read the last vector and do
        // not store it), so every step issues the same loads
        auto cload = [&](int it, v4u& h, v4u& l) {
            const int v = min(tid + NTHR * it, PAN * 4 - 1);
            const int col = v >> 2, k8 = 8 * (v & 3);
            const size_t co = static_cast<size_t>(col0 + col) * nbfp + kb + k8;
            h = *reinterpret_cast<const v4u*>(CTh + co);
            l = *reinterpret_cast<const v4u*>(CTl + co);
        };
        if constexpr (DC) return;
        if (NSLOT > 0) cload(0, c0h, c0l);
        if (NSLOT > 1) cload(1, c1h, c1l);
        if (NSLOT > 2) cload(2, c2h, c2l);
        if (NSLOT > 3) cload(3, c3h, c3l);
        if (NSLOT > 4) cload(4, c4h, c4l);
    };
    auto lstore = [&](int buf) {
        uint16_t* Ah = lbuf(buf);
        uint16_t* Al = Ah + A_SZ;
        uint16_t* Ch = Al + A_SZ;
        uint16_t* Cl = Ch + C_SZ;
#pragma unroll
        for (int n = 0; n < NAV; ++n) {
            const int v = tid + NTHR * n, f = v & 127;
            const int ao = kg_swz((v >> 7) * 32 + (f >> 2), f & 3);
            *reinterpret_cast<v4u*>(Ah + ao) = sah[n];
            *reinterpret_cast<v4u*>(Al + ao) = sal[n];
        }
        // every loaded value has an unconditional use, so the loads stay where they are issued:
        // lanes past the end of the partial round loaded the last vector and store it again
        // to the same place (the same bytes as its own lane's store)
        auto cstore = [&](int it, const v4u& h, const v4u& l) {
            const int v = min(tid + NTHR * it, PAN * 4 - 1);
            const int o = kg_swz(v >> 2, v & 3);
            *reinterpret_cast<v4u*>(Ch + o) = h;
            *reinterpret_cast<v4u*>(Cl + o) = l;
        };
        if constexpr (DC) return;
        if (NSLOT > 0) cstore(0, c0h, c0l);
        if (NSLOT > 1) cstore(1, c1h, c1l);
        if (NSLOT > 2) cstore(2, c2h, c2l);
        if (NSLOT > 3) cstore(3, c3h, c3l);
        if (NSLOT > 4) cstore(4, c4h, c4l);
    };
    // One step's products: all fragments first (one LDS wait), then the MFMAs product by
    // product (consecutive updates of one accumulator are 4*NBW MFMAs apart), then the next
    // step's staging into the other buffer.
    // A wave whose 16-row fragments lie past the last row tile (last row block only) skips
    // their fragment reads and MFMAs; their accumulators stay zero.
    const int nfr = __builtin_amdgcn_readfirstlane(min(4, max(0, 2 * (nT - I0 - 2 * wm))));
    auto compute = [&](int buf, auto ni, auto&& pre, auto&& mid, auto&& store_next) {
        constexpr int NI = decltype(ni)::value;
        const uint16_t* Ah = lbuf(buf);
        const uint16_t* Al = Ah + A_SZ;
        const uint16_t* Ch = Al + A_SZ;
        const uint16_t* Cl = Ch + C_SZ;
        const int fc = lane >> 4;
        bf8_t ah[4], al[4], bh[NBW], bl[NBW];
        // hi fragments first; the lo fragments are read while the hi x hi MFMAs run
#pragma unroll
        for (int i = 0; i < NI; ++i)
            ah[i] = __builtin_bit_cast(bf8_t, *reinterpret_cast<const uint4*>(Ah + kg_swz(wm * 64 + i * 16 + (lane & 15), fc)));
#pragma unroll
        for (int j = 0; j < NBW; ++j)
            bh[j] = __builtin_bit_cast(bf8_t, *reinterpret_cast<const uint4*>(Ch + kg_swz(wn * NBW * 16 + j * 16 + (lane & 15), fc)));
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int j = 0; j < NBW; ++j)
            bl[j] = __builtin_bit_cast(bf8_t, *reinterpret_cast<const uint4*>(Cl + kg_swz(wn * NBW * 16 + j * 16 + (lane & 15), fc)));
#pragma unroll
        for (int i = 0; i < NI; ++i)
            al[i] = __builtin_bit_cast(bf8_t, *reinterpret_cast<const uint4*>(Al + kg_swz(wm * 64 + i * 16 + (lane & 15), fc)));
        __builtin_amdgcn_sched_barrier(0);
        pre();
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int j = 0; j < NBW; ++j)
#pragma unroll
            for (int i = 0; i < NI; ++i) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bh[j], acc[i][j], 0, 0, 0);
        // the next A vectors' hi/lo split (24 VALU ops per vector) runs between the MFMAs below
        __builtin_amdgcn_sched_barrier(0);
        mid();
#pragma unroll
        for (int j = 0; j < NBW; ++j)
#pragma unroll
            for (int i = 0; i < NI; ++i) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bl[j], acc[i][j], 0, 0, 0);
#pragma unroll
        for (int j = 0; j < NBW; ++j)
#pragma unroll
            for (int i = 0; i < NI; ++i) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[i], bh[j], acc[i][j], 0, 0, 0);
        if constexpr (NI > 0) {
            // one MFMA, then two VALU ops, 12 times per A vector; then the remaining MFMAs
            constexpr int NG = 12 * NAV < 2 * NI * NBW ? 12 * NAV : 2 * NI * NBW;
#pragma unroll
            for (int k = 0; k < NG; ++k) {
                __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
                __builtin_amdgcn_sched_group_barrier(0x002, 2, 0);
            }
            if constexpr (2 * NI * NBW > NG) __builtin_amdgcn_sched_group_barrier(0x008, 2 * NI * NBW - NG, 0);
        }
        __builtin_amdgcn_sched_barrier(0);
        store_next();
    };
    // Steps run in pairs so the two A sets are named statically; the loop body has no
    // branches around its loads and stores (loads past the last step re-read it, and their
    // stores land in a buffer no later compute reads).
    auto mainloop = [&](auto ni) {
        // Staged C: the step's C and A loads are issued before its fragment reads, and the A
        // split runs between its MFMAs.  Direct C: the A loads first; the split and the direct
        // loads after the fragment reads (a fragment read or split that follows a direct load
        // in program order is made to wait for it), and the step ends by waiting for all its
        // loads, which have had the step's MFMAs to land.
        auto loads = [&](int st, v4f_t(&a)[2 * NAV]) {
            if constexpr (!DC) gloadC(st);
            gloadA(st + 1, a);
        };
        auto waitC = [&] {
            if constexpr (DC) wait_vmem<0>();
        };
        if constexpr (DC) {
            issueC(0, 0);
            fence_compiler();
        } else {
            gloadC(0);
        }
        gloadA(0, a0);
        splitA(a0);
        lstore(0);
        gloadA(1, a1);
        waitC();
        __syncthreads();
        int st = 0;
        // (scheduling fences keep the loads ahead of the MFMAs they are meant to overlap)
        for (; st + 1 < nsteps; st += 2) {
            loads(st + 1, a0);
            __builtin_amdgcn_sched_barrier(0);
            if constexpr (DC)
                compute(0, ni, [&] { splitA(a1); __builtin_amdgcn_sched_barrier(0); issueC(st + 1, 1); }, [] {}, [&] { lstore(1); });
            else
                compute(0, ni, [] {}, [&] { splitA(a1); }, [&] { lstore(1); });
            waitC();
            __syncthreads();
            loads(st + 2, a1);
            __builtin_amdgcn_sched_barrier(0);
            if constexpr (DC)
                compute(1, ni, [&] { splitA(a0); __builtin_amdgcn_sched_barrier(0); issueC(st + 2, 0); }, [] {}, [&] { lstore(0); });
            else
                compute(1, ni, [] {}, [&] { splitA(a0); }, [&] { lstore(0); });
            waitC();
            __syncthreads();
        }
        if (st < nsteps) compute(0, ni, [] {}, [] {}, [] {});  // odd step count: the last step is in buffer 0
        // direct loads still in flight (steps past the end) land before the LDS is released
        if constexpr (DC) wait_vmem<0>();
    };
    if (nsteps > 0) {
        if (nfr == 4)
            mainloop(std::integral_constant<int, 4>{});
        else if (nfr == 2)
            mainloop(std::integral_constant<int, 2>{});
        else
            mainloop(std::integral_constant<int, 0>{});
    }
    const size_t xq = static_cast<size_t>(q) * xrows * ldc;
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < NBW; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int m = ROWS * rb + wm * 64 + i * 16 + 4 * (lane >> 4) + r;
                const size_t o = xq + static_cast<size_t>(m) * ldc + col0 + wn * NBW * 16 + j * 16 + (lane & 15);
                const float x = acc[i][j][r];
                const uint16_t hi = f2bf(x);
                Xh[o] = hi;
                Xl[o] = f2bf(x - bf2f(hi));
            }
}

// =====================================================================================
// K contraction, FP64 / FP32 (16x16x4 MFMA): P[split][tile] (+)= sum_b A_b S B_b^T over
// 128x128 output tiles, 4 waves (2x2) of 64x64, K-step 16, double-buffered LDS.
// S = diag(sign) if sign != nullptr.  Accumulators are stored to FP64 partials.
// FP32: the MFMA accumulators hold at most kFold32 k-terms before they are added into FP64
// totals, so the rounding error of a sum does not grow with the number of auxiliary
// functions a split covers.
// =====================================================================================
template <typename T>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2))) void k_kgemm_mfma(const T* __restrict__ A, size_t strideA, const T* __restrict__ Bm,
                                                    size_t strideB, int ld, int kbeg, int kend, int nbatch,
                                                    int nsplit, const int2* __restrict__ tiles, int ntile,
                                                    double* __restrict__ P, double alpha, int accumulate,
                                                    int upper, int nvalid) {
    constexpr int KS = 16, AP = KS + 1;
    constexpr bool F64 = sizeof(T) == 8;
    typedef typename std::conditional<F64, d4_t, f4_t>::type acc_t;
    __shared__ T As[2][128 * AP];
    __shared__ T Bs[2][128 * AP];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wr = wave >> 1, wc = wave & 1;
    const XcdItem xi = xcd_map(blockIdx.x, ntile, nsplit);
    if (!xi.valid) return;
    const int tile = xi.item, split = xi.group;
    const int2 tb = tiles[tile];
    // a wave does the staging but no MFMAs when its 64x64 block lies below the diagonal of a
    // diagonal tile of a symmetric result (only the upper triangle is read) or entirely in the
    // zero padding rows / columns of X past nvalid
    const bool idle = (upper && tb.x == tb.y && wr > wc) || tb.x * 128 + wr * 64 >= nvalid ||
                      tb.y * 128 + wc * 64 >= nvalid;
    const int b0 = static_cast<int>((static_cast<long long>(nbatch) * split) / nsplit);
    const int b1 = static_cast<int>((static_cast<long long>(nbatch) * (split + 1)) / nsplit);
    const int nks = (kend - kbeg) / KS;
    const int nsteps = (b1 - b0) * nks;
    // thread's fixed (row, k) slots within a 128 x 16 chunk: rows r0 + 16*it, k = tid & 15
    const int r0 = tid >> 4, kk = tid & 15;
    const T* Ab = A + static_cast<size_t>(tb.x * 128 + r0) * ld + kbeg + kk;
    const T* Bb = Bm + static_cast<size_t>(tb.y * 128 + r0) * ld + kbeg + kk;

    acc_t acc[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) acc[i] = acc_t{0, 0, 0, 0};
    // FP64 totals of the folded FP32 accumulators (unused for FP64)
    constexpr int FOLD = F64 ? 0 : kFold32 / KS;
    d4_t tot[F64 ? 1 : 16];
#pragma unroll
    for (int i = 0; i < (F64 ? 1 : 16); ++i) tot[i] = d4_t{0, 0, 0, 0};
    auto fold = [&]() {
        if constexpr (!F64) {
#pragma unroll
            for (int i = 0; i < 16; ++i) {
#pragma unroll
                for (int r = 0; r < 4; ++r) tot[i][r] += static_cast<double>(acc[i][r]);
                acc[i] = acc_t{0, 0, 0, 0};
            }
        }
    };

    T ra[8], rb[8];
    int b = b0, ks0 = 0;
    auto gload = [&]() {
        const size_t oa = static_cast<size_t>(b) * strideA + ks0;
        const size_t ob = static_cast<size_t>(b) * strideB + ks0;
#pragma unroll
        for (int it = 0; it < 8; ++it) {
            ra[it] = Ab[oa + static_cast<size_t>(16 * it) * ld];
            rb[it] = Bb[ob + static_cast<size_t>(16 * it) * ld];
        }
        ks0 += KS;
        if (ks0 == nks * KS) {
            ks0 = 0;
            ++b;
        }
    };
    auto lstore = [&](int buf) {
#pragma unroll
        for (int it = 0; it < 8; ++it) {
            As[buf][(r0 + 16 * it) * AP + kk] = ra[it];
            Bs[buf][(r0 + 16 * it) * AP + kk] = rb[it];
        }
    };
    auto compute = [&](int buf) {
#pragma unroll
        for (int ks = 0; ks < KS; ks += 4) {
            T a[4], bv[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) a[i] = As[buf][(wr * 64 + i * 16 + (lane & 15)) * AP + ks + (lane >> 4)];
#pragma unroll
            for (int j = 0; j < 4; ++j) bv[j] = Bs[buf][(wc * 64 + j * 16 + (lane & 15)) * AP + ks + (lane >> 4)];
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    if constexpr (F64)
                        acc[i * 4 + j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i], bv[j], acc[i * 4 + j], 0, 0, 0);
                    else
                        acc[i * 4 + j] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[i], bv[j], acc[i * 4 + j], 0, 0, 0);
                }
        }
    };

    if constexpr (F64) {
        // FP64: a step's operands are loaded two steps before their LDS store (two register
        // sets named statically in a loop over step pairs; loads past the last step re-read it)
        T sa0[8], sb0[8], sa1[8], sb1[8];
        auto gload_at = [&](int st, T* xa, T* xb) {
            st = min(st, nsteps - 1);
            const size_t oa = static_cast<size_t>(b0 + st / nks) * strideA + (st % nks) * KS;
            const size_t ob = static_cast<size_t>(b0 + st / nks) * strideB + (st % nks) * KS;
#pragma unroll
            for (int it = 0; it < 8; ++it) {
                xa[it] = Ab[oa + static_cast<size_t>(16 * it) * ld];
                xb[it] = Bb[ob + static_cast<size_t>(16 * it) * ld];
            }
        };
        auto lstore_from = [&](int buf, const T* xa, const T* xb) {
#pragma unroll
            for (int it = 0; it < 8; ++it) {
                As[buf][(r0 + 16 * it) * AP + kk] = xa[it];
                Bs[buf][(r0 + 16 * it) * AP + kk] = xb[it];
            }
        };
        if (nsteps > 0) {
            gload_at(0, sa0, sb0);
            gload_at(1, sa1, sb1);
            lstore_from(0, sa0, sb0);
            __syncthreads();
            int st = 0;
            for (; st + 1 < nsteps; st += 2) {
                gload_at(st + 2, sa0, sb0);
                __builtin_amdgcn_sched_barrier(0);
                if (!idle) compute(0);
                lstore_from(1, sa1, sb1);
                __syncthreads();
                gload_at(st + 3, sa1, sb1);
                __builtin_amdgcn_sched_barrier(0);
                if (!idle) compute(1);
                lstore_from(0, sa0, sb0);
                __syncthreads();
            }
            if (st < nsteps && !idle) compute(0);  // odd step count: the last step is in buffer 0
        }
    } else if (nsteps > 0) {
        gload();
        lstore(0);
        __syncthreads();
        int buf = 0, nf = 0;
        for (int s = 0; s < nsteps; ++s) {
            const bool more = s + 1 < nsteps;
            if (more) gload();
            if (!idle) {
                compute(buf);
                if (++nf == FOLD) {
                    fold();
                    nf = 0;
                }
            }
            if (more) lstore(buf ^ 1);
            __syncthreads();
            buf ^= 1;
        }
    }
    if (idle) return;
    fold();
    double* Pt = P + (static_cast<size_t>(split) * ntile + tile) * 16384;
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int rr = F64 ? (4 * r + (lane >> 4)) : (4 * (lane >> 4) + r);
                const int row = wr * 64 + i * 16 + rr;
                const int col = wc * 64 + j * 16 + (lane & 15);
                double* o = Pt + row * 128 + col;
                double v;
                if constexpr (F64) v = acc[i * 4 + j][r];
                else v = tot[i * 4 + j][r];
                *o = alpha * v + (accumulate ? *o : 0.0);
            }
}

// =====================================================================================
// K contraction, 3-term BF16 split on hi/lo X: A S B^T ~ Ah Bh + Ah Bl + Al Bh.
// 128x128 tiles, 4 waves of 64x64, K-step 32, double-buffered LDS.  LDS rows are
// 64 bytes (32 k) with the four 16-byte chunks of row r stored at chunk c ^ ((r>>1)&3):
// the fragment reads (16 rows x one chunk per 16 lanes) and the staging writes are
// bank-conflict free, and the 64 KB footprint leaves room for two workgroups per CU.
// The MFMA accumulators are folded into a second set of FP32 totals every kFold32 k-terms
// (FP64 totals would not fit beside the fragments in 256 registers): a sum of N terms then
// rounds like two chains of about sqrt(N) additions instead of one chain of N.
// DC: the operands go from global memory straight into LDS (global_load_lds_dwordx4, each
// wave 8 loads of 16 rows per step, issued after the step's fragment reads) instead of
// through staging registers.
// =====================================================================================

template <bool DC>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2))) void k_kgemm_bf3(
    const uint16_t* __restrict__ Ah_, const uint16_t* __restrict__ Al_, const uint16_t* __restrict__ Bh_,
    const uint16_t* __restrict__ Bl_, size_t stride, int ld, int kbeg, int kend, int nbatch, int nsplit,
    const int2* __restrict__ tiles, int ntile, double* __restrict__ P, double alpha, int accumulate, int upper,
    int nvalid) {
    constexpr int KS = 32, SZ = 128 * 32;
    __shared__ __align__(16) uint16_t Ls0[4 * SZ];
    __shared__ __align__(16) uint16_t Ls1[4 * SZ];
    auto lbuf = [&](int b) { return b ? Ls1 : Ls0; };
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wr = wave >> 1, wc = wave & 1;
    const XcdItem xi = xcd_map(blockIdx.x, ntile, nsplit);
    if (!xi.valid) return;
    const int tile = xi.item, split = xi.group;
    const int2 tb = tiles[tile];
    // no MFMAs for the wave below the diagonal of a diagonal tile of a symmetric result, or
    // for one entirely in the zero padding rows / columns of X past nvalid
    const bool idle = (upper && tb.x == tb.y && wr > wc) || tb.x * 128 + wr * 64 >= nvalid ||
                      tb.y * 128 + wc * 64 >= nvalid;
    const int m0 = tb.x * 128, n0 = tb.y * 128;
    const int b0 = static_cast<int>((static_cast<long long>(nbatch) * split) / nsplit);
    const int b1 = static_cast<int>((static_cast<long long>(nbatch) * (split + 1)) / nsplit);
    const int nks = (kend - kbeg) / KS;
    const int nsteps = (b1 - b0) * nks;

    f4_t acc[16], tot[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        acc[i] = f4_t{0, 0, 0, 0};
        tot[i] = f4_t{0, 0, 0, 0};
    }
    auto fold = [&]() {
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            tot[i] += acc[i];
            acc[i] = f4_t{0, 0, 0, 0};
        }
    };

    // per step: 4 arrays x 128 rows x 32 k = 4 x 512 vectors of 8; 256 threads -> 2 vectors per array
    // thread's two 8-element vectors per array: rows (tid>>2) and (tid>>2)+64, chunk tid&3
    const int vrow = tid >> 2, vc = tid & 3, vk8 = 8 * vc;
    uint4 a0h, a0l, b0h, b0l, a1h, a1l, b1h, b1l;
    auto gload = [&](int s) {
        const int b = b0 + s / nks;
        const int k0 = kbeg + (s % nks) * KS;
        const size_t base = static_cast<size_t>(b) * stride + k0 + vk8;
        const size_t oa0 = base + static_cast<size_t>(m0 + vrow) * ld;
        const size_t oa1 = oa0 + static_cast<size_t>(64) * ld;
        const size_t ob0 = base + static_cast<size_t>(n0 + vrow) * ld;
        const size_t ob1 = ob0 + static_cast<size_t>(64) * ld;
        a0h = *reinterpret_cast<const uint4*>(Ah_ + oa0);
        a0l = *reinterpret_cast<const uint4*>(Al_ + oa0);
        a1h = *reinterpret_cast<const uint4*>(Ah_ + oa1);
        a1l = *reinterpret_cast<const uint4*>(Al_ + oa1);
        b0h = *reinterpret_cast<const uint4*>(Bh_ + ob0);
        b0l = *reinterpret_cast<const uint4*>(Bl_ + ob0);
        b1h = *reinterpret_cast<const uint4*>(Bh_ + ob1);
        b1l = *reinterpret_cast<const uint4*>(Bl_ + ob1);
    };
    const int o0 = kg_swz(vrow, vc), o1 = kg_swz(vrow + 64, vc);
    auto lstore = [&](int buf) {
        uint16_t* L = lbuf(buf);
        *reinterpret_cast<uint4*>(L + o0) = a0h;
        *reinterpret_cast<uint4*>(L + SZ + o0) = a0l;
        *reinterpret_cast<uint4*>(L + 2 * SZ + o0) = b0h;
        *reinterpret_cast<uint4*>(L + 3 * SZ + o0) = b0l;
        *reinterpret_cast<uint4*>(L + o1) = a1h;
        *reinterpret_cast<uint4*>(L + SZ + o1) = a1l;
        *reinterpret_cast<uint4*>(L + 2 * SZ + o1) = b1h;
        *reinterpret_cast<uint4*>(L + 3 * SZ + o1) = b1l;
    };
    // direct loads of step s into buf: the 32 groups of 16 LDS rows (Ah, Al, Bh, Bl x 8) are
    // dealt to the waves; lane L lands in 16-byte slot L and fetches the logical chunk that
    // the swizzle of kg_swz places there
    auto issue = [&](int s, int buf) {
        const int b = b0 + s / nks;
        const int k0 = kbeg + (s % nks) * KS;
        const int w = __builtin_amdgcn_readfirstlane(wave);
        const int r0 = lane >> 2, pc = lane & 3;
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const int arr = i >> 1, rg = w + 4 * (i & 1);  // group w + 4 i of the 32
            const int r = rg * 16 + r0;
            const uint16_t* base = arr == 0 ? Ah_ : arr == 1 ? Al_ : arr == 2 ? Bh_ : Bl_;
            const size_t row = static_cast<size_t>((arr < 2 ? m0 : n0) + r);
            load_lds16(base + static_cast<size_t>(b) * stride + row * ld + k0 + 8 * (pc ^ ((r >> 1) & 3)),
                       lbuf(buf) + arr * SZ + rg * 512);
        }
    };
    const int fc = lane >> 4;
    auto compute = [&](int buf, auto&& pre) {
        const uint16_t* L = lbuf(buf);
        bf8_t ah[4], al[4], bh[4], bl[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int ra = kg_swz(wr * 64 + i * 16 + (lane & 15), fc);
            const int rb = kg_swz(wc * 64 + i * 16 + (lane & 15), fc);
            ah[i] = __builtin_bit_cast(bf8_t, *reinterpret_cast<const uint4*>(L + ra));
            al[i] = __builtin_bit_cast(bf8_t, *reinterpret_cast<const uint4*>(L + SZ + ra));
            bh[i] = __builtin_bit_cast(bf8_t, *reinterpret_cast<const uint4*>(L + 2 * SZ + rb));
            bl[i] = __builtin_bit_cast(bf8_t, *reinterpret_cast<const uint4*>(L + 3 * SZ + rb));
        }
        if constexpr (DC) {
            __builtin_amdgcn_sched_barrier(0);
            pre();
            __builtin_amdgcn_sched_barrier(0);
        }
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
            for (int i = 0; i < 4; ++i)
                acc[i * 4 + j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bh[j], acc[i * 4 + j], 0, 0, 0);
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
            for (int i = 0; i < 4; ++i)
                acc[i * 4 + j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bl[j], acc[i * 4 + j], 0, 0, 0);
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
            for (int i = 0; i < 4; ++i)
                acc[i * 4 + j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[i], bh[j], acc[i * 4 + j], 0, 0, 0);
    };

    if (nsteps > 0) {
        if constexpr (DC) {
            issue(0, 0);
            wait_vmem<0>();
        } else {
            gload(0);
            lstore(0);
        }
        __syncthreads();
        int buf = 0, nf = 0;
        for (int s = 0; s < nsteps; ++s) {
            const bool more = s + 1 < nsteps;
            auto next = [&] {
                if constexpr (DC)
                    if (more) issue(s + 1, buf ^ 1);
            };
            if constexpr (!DC)
                if (more) gload(s + 1);
            if (!idle) {
                compute(buf, next);
                if (++nf == kFold32 / KS) {
                    fold();
                    nf = 0;
                }
            } else {
                next();
            }
            if constexpr (DC)
                wait_vmem<0>();
            else if (more)
                lstore(buf ^ 1);
            __syncthreads();
            buf ^= 1;
        }
    }
    if (idle) return;
    fold();
    double* Pt = P + (static_cast<size_t>(split) * ntile + tile) * 16384;
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
            for (int rr = 0; rr < 4; ++rr) {
                const int row = wr * 64 + i * 16 + 4 * (lane >> 4) + rr;
                const int col = wc * 64 + j * 16 + (lane & 15);
                double* o = Pt + row * 128 + col;
                *o = alpha * static_cast<double>(tot[i * 4 + j][rr]) + (accumulate ? *o : 0.0);
            }
}

// ------------------------------------------------------------------ conversions
__global__ void k_to_f32(const double* __restrict__ src, float* __restrict__ dst, size_t n) {
    for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += static_cast<size_t>(gridDim.x) * blockDim.x)
        dst[i] = static_cast<float>(src[i]);
}
// Same, each 32x32 tile transposed: dst[t][r][c] = src[t][c][r] (the half-transform reads these
// for the blocks it uses in the transposed orientation).
__global__ void k_to_f32_t(const double* __restrict__ src, float* __restrict__ dst, size_t n) {
    for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += static_cast<size_t>(gridDim.x) * blockDim.x) {
        const size_t e = i % TILE2, base = i - e;
        dst[i] = static_cast<float>(src[base + (e % TILE) * TILE + e / TILE]);
    }
}

// C [nbf][ld] FP64 -> padded [nbfp][ldc] in T, column c taken from source column cmap[c] (< 0: zero)
template <typename T>
__global__ void k_pack_cmap(const double* __restrict__ src, int ldsrc, int nbf, const int* __restrict__ cmap,
                            T* __restrict__ dst, int nbfp, int ldc) {
    const size_t idx = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (idx >= static_cast<size_t>(nbfp) * ldc) return;
    const int m = static_cast<int>(idx / ldc), c = static_cast<int>(idx % ldc);
    const int i = cmap[c];
    dst[idx] = (m < nbf && i >= 0) ? static_cast<T>(src[static_cast<size_t>(m) * ldsrc + i]) : T(0);
}

// C [nbf][ld] FP64 -> transposed padded [ldc][nbfp] BF16 hi/lo, with the column map
__global__ void k_pack_cT_bf(const double* __restrict__ src, int ldsrc, int nbf, const int* __restrict__ cmap,
                             uint16_t* __restrict__ hi, uint16_t* __restrict__ lo, int nbfp, int ldc) {
    const size_t idx = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (idx >= static_cast<size_t>(nbfp) * ldc) return;
    const int c = static_cast<int>(idx / nbfp), m = static_cast<int>(idx % nbfp);
    const int i = cmap[c];
    const float x = (m < nbf && i >= 0) ? static_cast<float>(src[static_cast<size_t>(m) * ldsrc + i]) : 0.0f;
    const uint16_t h = f2bf(x);
    hi[idx] = h;
    lo[idx] = f2bf(x - bf2f(h));
}

// Column signs: s[i] = +1 if Cr[:,i] == Cl[:,i], -1 if Cr[:,i] == -Cl[:,i], 0 otherwise
// (only bitwise-exact matches count).  One block per column.
// s[i] = +1 / -1 when column i of Cr is bitwise +/- column (i + shift) % ncol of Cl, else 0.
__global__ void k_column_signs(const double* __restrict__ Cl, const double* __restrict__ Cr, int ld, int nbf,
                               int ncol, int shift, int8_t* __restrict__ s) {
    const int i = blockIdx.x;
    const int il = (i + shift) % ncol;
    __shared__ int plus, minus;
    if (threadIdx.x == 0) {
        plus = 1;
        minus = 1;
    }
    __syncthreads();
    int p = 1, mi = 1;
    for (int m = threadIdx.x; m < nbf; m += blockDim.x) {
        const double a = Cl[static_cast<size_t>(m) * ld + il], b = Cr[static_cast<size_t>(m) * ld + i];
        p &= (a == b);
        mi &= (a == -b);
    }
    if (!p) atomicAnd(&plus, 0);
    if (!mi) atomicAnd(&minus, 0);
    __syncthreads();
    if (threadIdx.x == 0) s[i] = plus ? 1 : (minus ? -1 : 0);
}

inline unsigned nblocks(size_t n, int bs) {
    const size_t b = (n + bs - 1) / bs;
    return static_cast<unsigned>(b > 65535u * 16 ? 65535u * 16 : b);
}

}  // namespace

// ------------------------------------------------------------------ launchers
void ht_panel_shape(int ncol, int* nbw, int* npanel, int* ldc) {
    const int nbt = (ncol + 63) / 64;  // 64-column units
    const int np = (nbt + 4) / 5;
    const int w = (nbt + np - 1) / np;
    *nbw = w;
    *npanel = np;
    *ldc = np * 64 * w;
}

template <typename T, int NBW, int WM>
static void ht_mfma_nbw(const TileGeom& g, const T* B, int q0, int nq, const T* C, int ldc, int npanel, T* X,
                        int xrows, const void* ztile, hipStream_t s) {
    dim3 grid(xcd_grid(nq, xrows / (64 * WM) * npanel));
    hipLaunchKernelGGL((k_ht_mfma<T, NBW, WM>), grid, dim3(WM * 256), 0, s, B, g.d_tidx, g.nT, g.ntp, q0, nq, npanel, C,
                       ldc, X, xrows, static_cast<const T*>(ztile));
}

template <typename T, int WM>
static void ht_mfma_wm(const TileGeom& g, const T* B, int q0, int nq, const T* C, int ldc, int nbw, int npanel, T* X,
                       int xrows, const void* zt, hipStream_t s) {
    switch (nbw) {
        case 1: ht_mfma_nbw<T, 1, WM>(g, B, q0, nq, C, ldc, npanel, X, xrows, zt, s); break;
        case 2: ht_mfma_nbw<T, 2, WM>(g, B, q0, nq, C, ldc, npanel, X, xrows, zt, s); break;
        case 3: ht_mfma_nbw<T, 3, WM>(g, B, q0, nq, C, ldc, npanel, X, xrows, zt, s); break;
        case 4: ht_mfma_nbw<T, 4, WM>(g, B, q0, nq, C, ldc, npanel, X, xrows, zt, s); break;
        default: ht_mfma_nbw<T, 5, WM>(g, B, q0, nq, C, ldc, npanel, X, xrows, zt, s); break;
    }
}

int g_ht_wm = 0;       // 0: default shape; 1, 2 (and 4 for BF16x3) select others (tuning sweeps)

template <typename T>
static void ht_mfma(const TileGeom& g, const T* B, int q0, int nq, const T* C, int ldc, int nbw, int npanel, T* X,
                    int xrows, const void* zt, hipStream_t s) {
    require(nbw >= 1 && nbw <= 5 && ldc == npanel * 64 * nbw && nq > 0 && nq <= 65535 &&
                xrows == (g.nT + 3) / 4 * 128 && g.nT <= kMaxColTiles && zt != nullptr,
            "ht_mfma shape");
    const int wm = g_ht_wm ? g_ht_wm : 2;
    if (wm == 1)
        ht_mfma_wm<T, 1>(g, B, q0, nq, C, ldc, nbw, npanel, X, xrows, zt, s);
    else
        ht_mfma_wm<T, 2>(g, B, q0, nq, C, ldc, nbw, npanel, X, xrows, zt, s);
}

void launch_ht2_f64(const TileGeom& g, const double* B, int q0, int nq, const double* C, int ldc, int nbw, int npanel,
                    double* X, int xrows, const void* zt, hipStream_t s) {
    ht_mfma<double>(g, B, q0, nq, C, ldc, nbw, npanel, X, xrows, zt, s);
}
void launch_ht2_f32(const TileGeom& g, const float* B, int q0, int nq, const float* C, int ldc, int nbw, int npanel,
                    float* X, int xrows, const void* zt, hipStream_t s) {
    ht_mfma<float>(g, B, q0, nq, C, ldc, nbw, npanel, X, xrows, zt, s);
}
void set_ht_rowwaves(int wm) { g_ht_wm = (wm >= 1 && wm <= 4) ? wm : 0; }
int g_kgemm_dc = 1;    // BF16x3 contraction: operands loaded straight into LDS (1) or staged (0)
void set_kgemm_direct(int on) { g_kgemm_dc = on ? 1 : 0; }

// npanel counts panels of 64*nbw columns; the WN = 2 shape covers each with two workgroups
template <int WM, int WN, bool DC>
static void ht_bf3_wm(const TileGeom& g, const float* B, const float* BT, int q0, int nq, const uint16_t* CTh, const uint16_t* CTl,
                      int ldc, int nbw, int npanel, uint16_t* Xh, uint16_t* Xl, int xrows, const void* zt,
                      hipStream_t s) {
    const int np = npanel * 4 / WN;
    dim3 grid(xcd_grid(nq, xrows / (64 * WM) * np));
    const float* z = static_cast<const float*>(zt);
    switch (nbw) {
#define P4A_BF3(n)                                                                                                  \
    case n:                                                                                                         \
        hipLaunchKernelGGL((k_ht_bf3<n, WM, WN, DC>), grid, dim3(WM * WN * 64), 0, s, B, BT, g.d_tidx, g.nT, g.ntp, q0, nq, \
                           np, CTh, CTl, g.nbfp, ldc, Xh, Xl, xrows, z);                                           \
        break;
        P4A_BF3(1) P4A_BF3(2) P4A_BF3(3) P4A_BF3(4) P4A_BF3(5)
#undef P4A_BF3
    }
}

void launch_ht_bf3(const TileGeom& g, const float* B, const float* BT, int q0, int nq, const uint16_t* CTh, const uint16_t* CTl,
                   int ldc, int nbw, int npanel, uint16_t* Xh, uint16_t* Xl, int xrows, const void* zt,
                   hipStream_t s) {
    require(nbw >= 1 && nbw <= 5 && ldc == npanel * 64 * nbw && nq > 0 && nq <= 65535 && g.nbfp % 32 == 0 &&
                xrows == (g.nT + 3) / 4 * 128 && g.nT <= kMaxColTiles && zt != nullptr,
            "ht_bf3 shape");
    // default: one 8-wave workgroup of 128 x 64*nbw per CU, C loaded straight into LDS.
    // Kept for comparison: 1 (4 waves over 64 rows), 3 (4-wave workgroups of 128 x 32*nbw,
    // two per CU: as fast, but each B block is read by twice as many workgroups, 1.3x the
    // HBM reads), 4 (the 8-wave shape with C staged through registers).
    if (g_ht_wm == 1)
        ht_bf3_wm<1, 4, false>(g, B, BT, q0, nq, CTh, CTl, ldc, nbw, npanel, Xh, Xl, xrows, zt, s);
    else if (g_ht_wm == 3)
        ht_bf3_wm<2, 2, true>(g, B, BT, q0, nq, CTh, CTl, ldc, nbw, npanel, Xh, Xl, xrows, zt, s);
    else if (g_ht_wm == 4)
        ht_bf3_wm<2, 4, false>(g, B, BT, q0, nq, CTh, CTl, ldc, nbw, npanel, Xh, Xl, xrows, zt, s);
    else
        ht_bf3_wm<2, 4, true>(g, B, BT, q0, nq, CTh, CTl, ldc, nbw, npanel, Xh, Xl, xrows, zt, s);
}

static void kgemm_check(int xrows, int ld, int kbeg, int kend, int nbatch, int nsplit, int ntile, int ks) {
    require(ntile > 0 && nsplit > 0 && nbatch >= nsplit && xrows % 128 == 0 && kbeg >= 0 && kend <= ld &&
                kbeg < kend && kbeg % ks == 0 && kend % ks == 0,
            "kgemm shape");
}

void launch_kgemm2_f64(const double* A, const double* B, size_t stride, int xrows, int nvalid, int ld, int kbeg,
                       int kend, int nbatch, int nsplit, const int2* tiles, int ntile, double* P, double alpha,
                       int accumulate, int upper, hipStream_t s) {
    kgemm_check(xrows, ld, kbeg, kend, nbatch, nsplit, ntile, 16);
    require(nvalid > 0 && nvalid <= xrows, "kgemm valid rows");
    hipLaunchKernelGGL((k_kgemm_mfma<double>), dim3(xcd_grid(nsplit, ntile)), dim3(256), 0, s, A, stride, B, stride, ld,
                       kbeg, kend, nbatch, nsplit, tiles, ntile, P, alpha, accumulate, upper, nvalid);
}
void launch_kgemm2_f32(const float* A, const float* B, size_t stride, int xrows, int nvalid, int ld, int kbeg, int kend,
                       int nbatch, int nsplit, const int2* tiles, int ntile, double* P, double alpha, int accumulate,
                       int upper, hipStream_t s) {
    kgemm_check(xrows, ld, kbeg, kend, nbatch, nsplit, ntile, 16);
    require(nvalid > 0 && nvalid <= xrows, "kgemm valid rows");
    hipLaunchKernelGGL((k_kgemm_mfma<float>), dim3(xcd_grid(nsplit, ntile)), dim3(256), 0, s, A, stride, B, stride, ld,
                       kbeg, kend, nbatch, nsplit, tiles, ntile, P, alpha, accumulate, upper, nvalid);
}
void launch_kgemm_bf3(const uint16_t* Ah, const uint16_t* Al, const uint16_t* Bh, const uint16_t* Bl, size_t stride,
                      int xrows, int nvalid, int ld, int kbeg, int kend, int nbatch, int nsplit, const int2* tiles,
                      int ntile, double* P, double alpha, int accumulate, int upper, hipStream_t s) {
    kgemm_check(xrows, ld, kbeg, kend, nbatch, nsplit, ntile, 32);
    require(nvalid > 0 && nvalid <= xrows, "kgemm valid rows");
    hipLaunchKernelGGL(g_kgemm_dc ? k_kgemm_bf3<true> : k_kgemm_bf3<false>, dim3(xcd_grid(nsplit, ntile)), dim3(256), 0, s, Ah, Al, Bh, Bl, stride, ld, kbeg,
                       kend, nbatch, nsplit, tiles, ntile, P, alpha, accumulate, upper, nvalid);
}

void launch_to_f32(const double* src, float* dst, size_t n, hipStream_t s) {
    hipLaunchKernelGGL(k_to_f32, dim3(nblocks(n, 256)), dim3(256), 0, s, src, dst, n);
}
void launch_to_f32_t(const double* src, float* dst, size_t n, hipStream_t s) {
    require(n % TILE2 == 0, "to_f32_t");
    hipLaunchKernelGGL(k_to_f32_t, dim3(nblocks(n, 256)), dim3(256), 0, s, src, dst, n);
}
void launch_pack_cmap_f64(const double* src, int ldsrc, int nbf, const int* cmap, double* dst, int nbfp, int ldc,
                          hipStream_t s) {
    const size_t n = static_cast<size_t>(nbfp) * ldc;
    hipLaunchKernelGGL((k_pack_cmap<double>), dim3(static_cast<unsigned>((n + 255) / 256)), dim3(256), 0, s, src, ldsrc,
                       nbf, cmap, dst, nbfp, ldc);
}
void launch_pack_cmap_f32(const double* src, int ldsrc, int nbf, const int* cmap, float* dst, int nbfp, int ldc,
                          hipStream_t s) {
    const size_t n = static_cast<size_t>(nbfp) * ldc;
    hipLaunchKernelGGL((k_pack_cmap<float>), dim3(static_cast<unsigned>((n + 255) / 256)), dim3(256), 0, s, src, ldsrc,
                       nbf, cmap, dst, nbfp, ldc);
}
void launch_pack_cT_bf(const double* src, int ldsrc, int nbf, const int* cmap, uint16_t* hi, uint16_t* lo, int nbfp,
                       int ldc, hipStream_t s) {
    const size_t n = static_cast<size_t>(nbfp) * ldc;
    hipLaunchKernelGGL(k_pack_cT_bf, dim3(static_cast<unsigned>((n + 255) / 256)), dim3(256), 0, s, src, ldsrc, nbf,
                       cmap, hi, lo, nbfp, ldc);
}
void launch_column_signs(const double* Cl, const double* Cr, int ld, int nbf, int ncol, int shift, int8_t* s,
                         hipStream_t st) {
    require(ncol > 0 && ld >= ncol && shift >= 0 && shift < ncol, "column_signs");
    hipLaunchKernelGGL(k_column_signs, dim3(ncol), dim3(256), 0, st, Cl, Cr, ld, nbf, ncol, shift, s);
}

}  // namespace p4a
