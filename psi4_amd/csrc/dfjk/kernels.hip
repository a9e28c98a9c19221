// DF-JK device kernels for gfx950 (CDNA4), FP64 path.
//
// Data layout in HBM (all row-major, FP64):
//   B      [naux][ntp][TILE][TILE]  fitted three-index tensor, stored as the
//          TILE x TILE blocks (I,J), I <= J, of each symmetric B_Q; diagonal
//          blocks hold both triangles.  Blocks with no Schwarz-significant pair
//          are not stored (tile-granular screening).
//   C      [nbfp][ldc]              orbital coefficients, zero-padded.
//
// MFMA lane layouts used here (verified on the device):
//   v_mfma_f64_16x16x4f64: A[i][k] in lane i+16k, B[k][j] in lane j+16k,
//                          D[i][j] in lane j+16*(i%4), register i/4.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "kernels.hpp"

namespace p4a {

namespace {

inline dim3 grid1(size_t n, int block) { return dim3(static_cast<unsigned>((n + block - 1) / block)); }

// ------------------------------------------------------------------ J path
__global__ void k_pack_w(const double* __restrict__ D, int ldD, int nbf, const int2* __restrict__ tiles,
                         double* __restrict__ W) {
    const int t = blockIdx.x;
    const int2 IJ = tiles[t];
    for (int e = threadIdx.x; e < TILE2; e += blockDim.x) {
        const int r = e / TILE, c = e % TILE;
        const int m = IJ.x * TILE + r, n = IJ.y * TILE + c;
        double v = 0.0;
        if (m < nbf && n < nbf) {
            v = D[static_cast<size_t>(m) * ldD + n];
            if (IJ.x != IJ.y) v += D[static_cast<size_t>(n) * ldD + m];
        }
        W[static_cast<size_t>(t) * TILE2 + e] = v;
    }
}

// Element loaders of the resident tensor for the J path (FP64 accumulation in every case).
// load4 / load8 read 4 / 8 consecutive elements (i a multiple of 4 / 8) with 16-byte vector
// loads per lane.
typedef double d2_t __attribute__((ext_vector_type(2)));
typedef float f4v_t __attribute__((ext_vector_type(4)));
struct BSrcF64 {
    const double* p;
    __device__ void load4(size_t i, double* v) const {
        const d2_t a = *reinterpret_cast<const d2_t*>(p + i), b = *reinterpret_cast<const d2_t*>(p + i + 2);
        v[0] = a.x; v[1] = a.y; v[2] = b.x; v[3] = b.y;
    }
    __device__ void load8(size_t i, double* v) const {
        load4(i, v);
        load4(i + 4, v + 4);
    }
};
struct BSrcF32 {
    const float* p;
    __device__ void load4(size_t i, double* v) const {
        const f4v_t a = *reinterpret_cast<const f4v_t*>(p + i);
        v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
    }
    __device__ void load8(size_t i, double* v) const {
        load4(i, v);
        load4(i + 4, v + 4);
    }
};
// d_Q partials: one wave per Q, the density tile of pair t held in registers (lane l holds
// elements 8l..8l+7 and 512+8l..512+8l+7).  Each wave takes two Q per pass, so two tiles
// are in flight.  grid = (ntp, ceil((q1-q0)/qblk)), 256 threads.
template <class BS>
__global__ __launch_bounds__(256) void k_j1(BS B, const double* __restrict__ W, int ntp, int naux, int q0, int q1,
                                            int qblk, double* __restrict__ dpart) {
    const int t = blockIdx.x;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const double* Wt = W + static_cast<size_t>(t) * TILE2;
    double w[16];
#pragma unroll
    for (int c = 0; c < 2; ++c)
#pragma unroll
        for (int e = 0; e < 8; ++e) w[8 * c + e] = Wt[512 * c + 8 * lane + e];
    const int qa = q0 + blockIdx.y * qblk;
    const int qb = min(q1, qa + qblk);
    auto dot = [&](int Q) {
        const size_t base = (static_cast<size_t>(Q) * ntp + t) * TILE2 + 8 * lane;
        double v[16];
        B.load8(base, v);
        B.load8(base + 512, v + 8);
        double s = 0.0;
#pragma unroll
        for (int k = 0; k < 16; ++k) s = fma(v[k], w[k], s);
        return s;
    };
    for (int Q = qa + wave; Q < qb; Q += 8) {
        const int Q2 = Q + 4;
        double s = dot(Q);
        double s2 = Q2 < qb ? dot(Q2) : 0.0;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            s += __shfl_xor(s, off);
            s2 += __shfl_xor(s2, off);
        }
        if (lane == 0) {
            dpart[static_cast<size_t>(t) * naux + Q] = s;
            if (Q2 < qb) dpart[static_cast<size_t>(t) * naux + Q2] = s2;
        }
    }
}

__global__ void k_j1_reduce(const double* __restrict__ dpart, int ntp, int naux, double* __restrict__ d) {
    const int Q = blockIdx.x * blockDim.x + threadIdx.x;
    if (Q >= naux) return;
    double s = 0.0;
    for (int t = 0; t < ntp; ++t) s += dpart[static_cast<size_t>(t) * naux + Q];
    d[Q] = s;
}

// J partials over a chunk of Q: grid = (ntp, nchunk), 256 threads, 4 consecutive elements per
// thread (vector loads), two Q per pass.
template <class BS>
__global__ __launch_bounds__(256) void k_j2(BS B, const double* __restrict__ d, int ntp, int naux, int qchunk,
                                            double* __restrict__ Jpart) {
    const int t = blockIdx.x, c = blockIdx.y;
    const int qa = c * qchunk, qb = min(naux, qa + qchunk);
    const int e0 = 4 * threadIdx.x;
    double a[4] = {0.0, 0.0, 0.0, 0.0};
    int Q = qa;
    for (; Q + 1 < qb; Q += 2) {
        const double d0 = d[Q], d1 = d[Q + 1];
        const size_t base = (static_cast<size_t>(Q) * ntp + t) * TILE2 + e0;
        double v0[4], v1[4];
        B.load4(base, v0);
        B.load4(base + static_cast<size_t>(ntp) * TILE2, v1);
#pragma unroll
        for (int k = 0; k < 4; ++k) a[k] = fma(v1[k], d1, fma(v0[k], d0, a[k]));
    }
    if (Q < qb) {
        const double d0 = d[Q];
        double v0[4];
        B.load4((static_cast<size_t>(Q) * ntp + t) * TILE2 + e0, v0);
#pragma unroll
        for (int k = 0; k < 4; ++k) a[k] = fma(v0[k], d0, a[k]);
    }
    double* o = Jpart + (static_cast<size_t>(c) * ntp + t) * TILE2 + e0;
#pragma unroll
    for (int k = 0; k < 4; ++k) o[k] = a[k];
}

__global__ void k_j_unpack(const double* __restrict__ Jpart, int nchunk, int ntp, const int* __restrict__ tidx,
                           int nT, int nbf, double* __restrict__ J, int ldJ, double alpha, int accumulate) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    const int m = blockIdx.y;
    if (n >= nbf) return;
    int I = m / TILE, Jt = n / TILE, r = m % TILE, c = n % TILE;
    if (I > Jt) {
        int tmp = I;
        I = Jt;
        Jt = tmp;
        tmp = r;
        r = c;
        c = tmp;
    }
    const int t = tidx[I * nT + Jt];
    double s = 0.0;
    if (t >= 0)
        for (int ch = 0; ch < nchunk; ++ch) s += Jpart[(static_cast<size_t>(ch) * ntp + t) * TILE2 + r * TILE + c];
    double* o = J + static_cast<size_t>(m) * ldJ + n;
    *o = alpha * s + (accumulate ? *o : 0.0);
}

// ------------------------------------------------------------------ K path
__global__ void k_pack_c(const double* __restrict__ src, int ldsrc, int nbf, int ncol, double scale,
                         double* __restrict__ dst, int nbfp, int ldc) {
    const size_t idx = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (idx >= static_cast<size_t>(nbfp) * ldc) return;
    const int m = static_cast<int>(idx / ldc), i = static_cast<int>(idx % ldc);
    dst[idx] = (m < nbf && i < ncol) ? scale * src[static_cast<size_t>(m) * ldsrc + i] : 0.0;
}

// Batched NT product into 128x128 tiles, FP64 MFMA.  4 waves in a 2x2 grid,
// each wave a 64x64 sub-tile (4x4 MFMA blocks).  grid = (ntile, nsplit).
__global__ __launch_bounds__(256) void k_gemm_nt_f64(const double* __restrict__ A, size_t strideA, int lda,
                                                     const double* __restrict__ Bm, size_t strideB, int ldb, int M,
                                                     int N, int K, int nbatch, const int2* __restrict__ tiles,
                                                     int ntile, double* __restrict__ P, int accumulate) {
    __shared__ double As[128][17];
    __shared__ double Bs[128][17];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wr = wave >> 1, wc = wave & 1;
    const int tile = blockIdx.x, split = blockIdx.y, nsplit = gridDim.y;
    const int2 tb = tiles[tile];
    const int m0 = tb.x * 128, n0 = tb.y * 128;
    const int b0 = static_cast<int>((static_cast<long long>(nbatch) * split) / nsplit);
    const int b1 = static_cast<int>((static_cast<long long>(nbatch) * (split + 1)) / nsplit);
    d4_t acc[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) acc[i] = d4_t{0.0, 0.0, 0.0, 0.0};
    for (int b = b0; b < b1; ++b) {
        const double* Ab = A + static_cast<size_t>(b) * strideA;
        const double* Bb = Bm + static_cast<size_t>(b) * strideB;
        for (int k0 = 0; k0 < K; k0 += 16) {
            __syncthreads();
#pragma unroll
            for (int it = 0; it < 8; ++it) {
                const int e = tid + 256 * it;
                const int r = e >> 4, kk = e & 15;
                const bool kin = (k0 + kk) < K;
                As[r][kk] = (kin && m0 + r < M) ? Ab[static_cast<size_t>(m0 + r) * lda + k0 + kk] : 0.0;
                Bs[r][kk] = (kin && n0 + r < N) ? Bb[static_cast<size_t>(n0 + r) * ldb + k0 + kk] : 0.0;
            }
            __syncthreads();
#pragma unroll
            for (int ks = 0; ks < 16; ks += 4) {
                double a[4], bv[4];
#pragma unroll
                for (int i = 0; i < 4; ++i) a[i] = As[wr * 64 + i * 16 + (lane & 15)][ks + (lane >> 4)];
#pragma unroll
                for (int j = 0; j < 4; ++j) bv[j] = Bs[wc * 64 + j * 16 + (lane & 15)][ks + (lane >> 4)];
#pragma unroll
                for (int i = 0; i < 4; ++i)
#pragma unroll
                    for (int j = 0; j < 4; ++j)
                        acc[i * 4 + j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i], bv[j], acc[i * 4 + j], 0, 0, 0);
            }
        }
    }
    double* Pt = P + (static_cast<size_t>(split) * ntile + tile) * 16384;
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = wr * 64 + i * 16 + 4 * r + (lane >> 4);
                const int col = wc * 64 + j * 16 + (lane & 15);
                double* o = Pt + row * 128 + col;
                *o = acc[i * 4 + j][r] + (accumulate ? *o : 0.0);
            }
}

__global__ void k_tile_reduce(const double* __restrict__ P, int nsplit, int ntile, const int* __restrict__ tile_of,
                              int nB, int sym, int M, int N, double* __restrict__ out, int ldo, double alpha,
                              int accumulate) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    const int m = blockIdx.y;
    if (n >= N) return;
    int Ib = m >> 7, Jb = n >> 7, r = m & 127, c = n & 127;
    if (sym == 2) {
        // out = M + M^T from the partials of every tile; (m,n) and (n,m) add the same two
        // split sums in either order, so the result is exactly symmetric
        const int t1 = tile_of[Ib * nB + Jb], t2 = tile_of[Jb * nB + Ib];
        double s1 = 0.0, s2 = 0.0;
        for (int sp = 0; sp < nsplit; ++sp) {
            if (t1 >= 0) s1 += P[(static_cast<size_t>(sp) * ntile + t1) * 16384 + r * 128 + c];
            if (t2 >= 0) s2 += P[(static_cast<size_t>(sp) * ntile + t2) * 16384 + c * 128 + r];
        }
        double* o = out + static_cast<size_t>(m) * ldo + n;
        const double v = m <= n ? s1 + s2 : s2 + s1;
        *o = alpha * v + (accumulate ? *o : 0.0);
        return;
    }
    // symmetric: read the upper triangle only (also inside diagonal tiles), so the
    // result is exactly symmetric whatever the accumulation order of the tile
    if (sym && (Ib > Jb || (Ib == Jb && r > c))) {
        int tmp = Ib;
        Ib = Jb;
        Jb = tmp;
        tmp = r;
        r = c;
        c = tmp;
    }
    const int t = tile_of[Ib * nB + Jb];
    double s = 0.0;
    if (t >= 0)
        for (int sp = 0; sp < nsplit; ++sp) s += P[(static_cast<size_t>(sp) * ntile + t) * 16384 + r * 128 + c];
    double* o = out + static_cast<size_t>(m) * ldo + n;
    *o = alpha * s + (accumulate ? *o : 0.0);
}

// Row-major C = alpha*A*B + beta*C, 128x128 output tiles, FP64 MFMA, 4 waves of 64x64.
// grid = (ceil(N/128), ceil(M/128), batch); operand z is offset by z * stride (0 = shared).
// lower_a: A is lower triangular (A[m][k] = 0 for k > m), so row block m0 stops at k = m0 + 128.
// K-step 16, double-buffered LDS: the next step's operands are loaded into registers before
// the current step's MFMAs and stored to the other buffer after them, one barrier per step.
// Out-of-range elements load as zero.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2))) void k_gemm_nn_f64(
    const double* __restrict__ A, int lda, const double* __restrict__ Bm, int ldb, double* __restrict__ Cm, int ldc,
    int M, int N, int K, double alpha, double beta, size_t sA, size_t sB, size_t sC, int lower_a) {
    constexpr int KS = 16;
    __shared__ double As[2][128][17];
    __shared__ double Bs[2][16][132];
    A += blockIdx.z * sA;
    Bm += blockIdx.z * sB;
    Cm += blockIdx.z * sC;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wr = wave >> 1, wc = wave & 1;
    const int m0 = blockIdx.y * 128, n0 = blockIdx.x * 128;
    const int kend = lower_a ? min(K, m0 + 128) : K;
    const int nsteps = (kend + KS - 1) / KS;
    d4_t acc[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) acc[i] = d4_t{0.0, 0.0, 0.0, 0.0};
    // thread slots: A element (r, kk) = (e >> 4, e & 15), B element (kb, cc) = (e >> 7, e & 127)
    double ra[8], rb[8];
    auto gload = [&](int st) {
        const int k0 = st * KS;
#pragma unroll
        for (int it = 0; it < 8; ++it) {
            const int e = tid + 256 * it;
            const int r = e >> 4, kk = e & 15;
            ra[it] = (m0 + r < M && k0 + kk < K) ? A[static_cast<size_t>(m0 + r) * lda + k0 + kk] : 0.0;
            const int kb = e >> 7, cc = e & 127;
            rb[it] = (k0 + kb < K && n0 + cc < N) ? Bm[static_cast<size_t>(k0 + kb) * ldb + n0 + cc] : 0.0;
        }
    };
    auto lstore = [&](int buf) {
#pragma unroll
        for (int it = 0; it < 8; ++it) {
            const int e = tid + 256 * it;
            As[buf][e >> 4][e & 15] = ra[it];
            Bs[buf][e >> 7][e & 127] = rb[it];
        }
    };
    auto compute = [&](int buf) {
#pragma unroll
        for (int ks = 0; ks < KS; ks += 4) {
            double a[4], bv[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) a[i] = As[buf][wr * 64 + i * 16 + (lane & 15)][ks + (lane >> 4)];
#pragma unroll
            for (int j = 0; j < 4; ++j) bv[j] = Bs[buf][ks + (lane >> 4)][wc * 64 + j * 16 + (lane & 15)];
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    acc[i * 4 + j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i], bv[j], acc[i * 4 + j], 0, 0, 0);
        }
    };
    if (nsteps > 0) {
        gload(0);
        lstore(0);
        __syncthreads();
        int buf = 0;
        for (int st = 0; st < nsteps; ++st) {
            const bool more = st + 1 < nsteps;
            if (more) gload(st + 1);
            __builtin_amdgcn_sched_barrier(0);
            compute(buf);
            if (more) lstore(buf ^ 1);
            __syncthreads();
            buf ^= 1;
        }
    }
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = m0 + wr * 64 + i * 16 + 4 * r + (lane >> 4);
                const int col = n0 + wc * 64 + j * 16 + (lane & 15);
                if (row < M && col < N) {
                    double* o = Cm + static_cast<size_t>(row) * ldc + col;
                    *o = alpha * acc[i * 4 + j][r] + (beta != 0.0 ? beta * *o : 0.0);
                }
            }
}

// ------------------------------------------------------------------ synthetic B
__device__ inline void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1) {
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = static_cast<uint64_t>(0xD2511F53u) * c[0];
        const uint64_t p1 = static_cast<uint64_t>(0xCD9E8D57u) * c[2];
        const uint32_t hi0 = static_cast<uint32_t>(p0 >> 32), lo0 = static_cast<uint32_t>(p0);
        const uint32_t hi1 = static_cast<uint32_t>(p1 >> 32), lo1 = static_cast<uint32_t>(p1);
        const uint32_t n0 = hi1 ^ c[1] ^ k0, n2 = hi0 ^ c[3] ^ k1;
        c[0] = n0;
        c[1] = lo1;
        c[2] = n2;
        c[3] = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
}

__device__ inline double synth_normal(uint32_t m, uint32_t n, uint32_t Q, uint32_t k0, uint32_t k1) {
    uint32_t c[4] = {m, n, Q, 0u};
    philox4x32_10(c, k0, k1);
    const uint64_t i1 = (static_cast<uint64_t>(c[0]) << 21) | (c[1] >> 11);
    const uint64_t i2 = (static_cast<uint64_t>(c[2]) << 21) | (c[3] >> 11);
    const double u1 = static_cast<double>(i1 + 1) * 0x1.0p-53;
    const double u2 = static_cast<double>(i2) * 0x1.0p-53;
    return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

// grid = (ntp, q1-q0), 256 threads.
__global__ void k_synth_b(const int2* __restrict__ tiles, int ntp, int nbf, const double* __restrict__ decay,
                          uint32_t k0, uint32_t k1, int q0, double* __restrict__ B) {
    const int t = blockIdx.x, Q = q0 + blockIdx.y;
    const int2 IJ = tiles[t];
    const double sQ = 1.0 / sqrt(static_cast<double>(Q) + 1.0);
    double* Bt = B + (static_cast<size_t>(Q) * ntp + t) * TILE2;
    for (int e = threadIdx.x; e < TILE2; e += blockDim.x) {
        const int m = IJ.x * TILE + e / TILE, n = IJ.y * TILE + e % TILE;
        double v = 0.0;
        if (m < nbf && n < nbf) {
            const double g = 0.5 * (synth_normal(m, n, Q, k0, k1) + synth_normal(n, m, Q, k0, k1));
            v = sQ * decay[m > n ? m - n : n - m] * g;
        }
        Bt[e] = v;
    }
}

}  // namespace

// ------------------------------------------------------------------ launchers
void launch_pack_w(const TileGeom& g, const double* D, int ldD, double* W, hipStream_t s) {
    require(g.ntp > 0 && ldD >= g.nbf, "pack_w");
    hipLaunchKernelGGL(k_pack_w, dim3(g.ntp), dim3(256), 0, s, D, ldD, g.nbf, g.d_tiles, W);
}

template <class F>
void with_bsrc(const JSource& b, F&& f) {
    if (b.kind == JSource::F64)
        f(BSrcF64{static_cast<const double*>(b.a)});
    else
        f(BSrcF32{static_cast<const float*>(b.a)});
}

void launch_j1(const TileGeom& g, const JSource& B, const double* W, int naux, int q0, int q1, double* dpart,
               hipStream_t s) {
    require(q0 >= 0 && q1 <= naux && q0 < q1, "j1 q range");
    require(B.a != nullptr, "j1 source");
    const int qblk = 64;
    dim3 grid(g.ntp, (q1 - q0 + qblk - 1) / qblk);
    with_bsrc(B, [&](auto src) {
        hipLaunchKernelGGL(k_j1<decltype(src)>, grid, dim3(256), 0, s, src, W, g.ntp, naux, q0, q1, qblk, dpart);
    });
}

void launch_j1_reduce(const TileGeom& g, const double* dpart, int naux, double* d, hipStream_t s) {
    hipLaunchKernelGGL(k_j1_reduce, grid1(naux, 256), dim3(256), 0, s, dpart, g.ntp, naux, d);
}

void launch_j2(const TileGeom& g, const JSource& B, const double* d, int naux, int qchunk, double* Jpart,
               hipStream_t s) {
    require(qchunk > 0, "j2 qchunk");
    require(B.a != nullptr, "j2 source");
    dim3 grid(g.ntp, (naux + qchunk - 1) / qchunk);
    with_bsrc(B, [&](auto src) {
        hipLaunchKernelGGL(k_j2<decltype(src)>, grid, dim3(256), 0, s, src, d, g.ntp, naux, qchunk, Jpart);
    });
}

void launch_j_unpack(const TileGeom& g, const double* Jpart, int nchunk, double* J, int ldJ, double alpha,
                     int accumulate, hipStream_t s) {
    require(ldJ >= g.nbf, "j_unpack ld");
    dim3 grid((g.nbf + 255) / 256, g.nbf);
    hipLaunchKernelGGL(k_j_unpack, grid, dim3(256), 0, s, Jpart, nchunk, g.ntp, g.d_tidx, g.nT, g.nbf, J, ldJ, alpha,
                       accumulate);
}

void launch_pack_c(const double* Csrc, int ldsrc, int nbf, int ncol, double scale, double* Cpad, int nbfp, int ldc,
                   hipStream_t s) {
    require(nbfp >= nbf && ldc >= ncol && ldsrc >= ncol, "pack_c");
    const size_t n = static_cast<size_t>(nbfp) * ldc;
    hipLaunchKernelGGL(k_pack_c, grid1(n, 256), dim3(256), 0, s, Csrc, ldsrc, nbf, ncol, scale, Cpad, nbfp, ldc);
}

void launch_gemm_nt_f64(const double* A, size_t strideA, int lda, const double* B, size_t strideB, int ldb, int M,
                        int N, int K, int nbatch, int nsplit, const int2* d_tiles, int ntile, double* P, int accumulate,
                        hipStream_t s) {
    require(ntile > 0 && nsplit > 0 && nbatch >= nsplit && lda >= K && ldb >= K, "gemm_nt_f64");
    dim3 grid(ntile, nsplit);
    hipLaunchKernelGGL(k_gemm_nt_f64, grid, dim3(256), 0, s, A, strideA, lda, B, strideB, ldb, M, N, K, nbatch,
                       d_tiles, ntile, P, accumulate);
}

void launch_tile_reduce(const double* P, int nsplit, int ntile, const int* d_tile_of, int nB, int sym, int M, int N,
                        double* out, int ldo, double alpha, int accumulate, hipStream_t s) {
    require(M <= nB * 128 && N <= nB * 128 && ldo >= N, "tile_reduce");
    dim3 grid((N + 255) / 256, M);
    hipLaunchKernelGGL(k_tile_reduce, grid, dim3(256), 0, s, P, nsplit, ntile, d_tile_of, nB, sym, M, N, out, ldo,
                       alpha, accumulate);
}

void launch_gemm_nn_f64(const double* A, int lda, const double* B, int ldb, double* C, int ldc, int M, int N, int K,
                        double alpha, double beta, hipStream_t s, int batch, size_t sA, size_t sB, size_t sC,
                        int lower_a) {
    require(lda >= K && ldb >= N && ldc >= N && batch >= 1 && batch <= 65535, "gemm_nn_f64");
    if (M <= 0 || N <= 0) return;
    dim3 grid((N + 127) / 128, (M + 127) / 128, batch);
    hipLaunchKernelGGL(k_gemm_nn_f64, grid, dim3(256), 0, s, A, lda, B, ldb, C, ldc, M, N, K, alpha, beta, sA, sB, sC,
                       lower_a);
}

// ---- adjoint helpers ----
__global__ void k_axpby_mat(const double* __restrict__ X, double a, const double* Y, double b, double* Z, size_t n) {
    for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += static_cast<size_t>(gridDim.x) * blockDim.x)
        Z[i] = a * X[i] + b * Y[i];
}
// A = (A + A^T) / 2 in place (Matrix::hermitivitize, matrix.cc:2681-2700): the thread of
// (r, c), c > r, owns the pair
__global__ void k_hermitivitize(double* __restrict__ A, int n, int lda) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x, r = blockIdx.y;
    if (c > r && c < n) {
        const size_t rc = static_cast<size_t>(r) * lda + c, cr = static_cast<size_t>(c) * lda + r;
        const double v = 0.5 * (A[rc] + A[cr]);
        A[rc] = v;
        A[cr] = v;
    }
}
// partial[b] = sum over the block's slice of X.*Y (fixed tree, deterministic)
__global__ __launch_bounds__(256) void k_dot_partial(const double* __restrict__ X, const double* __restrict__ Y,
                                                     size_t n, double* __restrict__ partial) {
    __shared__ double red[256];
    double s = 0.0;
    for (size_t i = static_cast<size_t>(blockIdx.x) * 256 + threadIdx.x; i < n; i += static_cast<size_t>(gridDim.x) * 256)
        s = fma(X[i], Y[i], s);
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}
__global__ void k_transpose(const double* __restrict__ A, int rows, int cols, int lda, double* __restrict__ AT,
                            int ldt) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x, r = blockIdx.y;
    if (c < cols) AT[static_cast<size_t>(c) * ldt + r] = A[static_cast<size_t>(r) * lda + c];
}
// G[q][m][n] = a * d[q0+q] * D[m][n]
__global__ void k_fill_scaled(const double* __restrict__ d, int q0, double a, const double* __restrict__ D, int n,
                              double* __restrict__ G) {
    const size_t nn = static_cast<size_t>(n) * n;
    const int q = blockIdx.y;
    const double f = a * d[q0 + q];
    for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < nn;
         i += static_cast<size_t>(gridDim.x) * blockDim.x)
        G[q * nn + i] = f * D[i];
}

// p[0..n) = 0 (a kernel, not hipMemsetAsync: a caller may capture this work in a graph)
__global__ void k_zero(double* __restrict__ p, size_t n) {
    for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += static_cast<size_t>(gridDim.x) * blockDim.x)
        p[i] = 0.0;
}
void launch_zero(double* p, size_t n, hipStream_t s) {
    if (n == 0) return;
    const size_t nb = std::min<size_t>((n + 255) / 256, size_t(1) << 16);
    hipLaunchKernelGGL(k_zero, dim3(static_cast<unsigned>(nb)), dim3(256), 0, s, p, n);
}

// out[r][c] = f[r] * A[r][c]
__global__ void k_scale_rows(const double* __restrict__ A, int rows, int cols, const double* __restrict__ f,
                             double* __restrict__ out) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x, r = blockIdx.y;
    if (c < cols) out[static_cast<size_t>(r) * cols + c] = f[r] * A[static_cast<size_t>(r) * cols + c];
}
void launch_scale_rows(const double* A, int rows, int cols, const double* f, double* out, hipStream_t s) {
    require(rows >= 1 && rows <= 65535 && cols >= 1, "scale_rows");
    hipLaunchKernelGGL(k_scale_rows, dim3((cols + 255) / 256, rows), dim3(256), 0, s, A, rows, cols, f, out);
}
void launch_axpby_mat(const double* X, double a, const double* Y, double b, double* Z, size_t n, hipStream_t s) {
    const size_t nb = std::min<size_t>((n + 255) / 256, size_t(1) << 20);
    if (nb == 0) return;
    hipLaunchKernelGGL(k_axpby_mat, dim3(static_cast<unsigned>(nb)), dim3(256), 0, s, X, a, Y, b, Z, n);
}
void launch_hermitivitize(double* A, int n, int lda, hipStream_t s) {
    require(n >= 1 && n <= 65535 && lda >= n, "hermitivitize");
    hipLaunchKernelGGL(k_hermitivitize, dim3((n + 255) / 256, n), dim3(256), 0, s, A, n, lda);
}
void launch_dot_partial(const double* X, const double* Y, size_t n, double* partial, int nblocks, hipStream_t s) {
    hipLaunchKernelGGL(k_dot_partial, dim3(nblocks), dim3(256), 0, s, X, Y, n, partial);
}
void launch_transpose(const double* A, int rows, int cols, int lda, double* AT, int ldt, hipStream_t s) {
    require(lda >= cols && ldt >= rows, "transpose");
    hipLaunchKernelGGL(k_transpose, dim3((cols + 255) / 256, rows), dim3(256), 0, s, A, rows, cols, lda, AT, ldt);
}
void launch_fill_scaled(const double* d, int q0, int nq, double a, const double* D, int n, double* G, hipStream_t s) {
    require(nq >= 1 && nq <= 65535, "fill_scaled");
    hipLaunchKernelGGL(k_fill_scaled, dim3(64, nq), dim3(256), 0, s, d, q0, a, D, n, G);
}

// Dense [nq][nbf][nbf] -> stored tiles of Q = q0 .. q0+nq-1; pairs outside the Schwarz mask are zero.
__global__ __launch_bounds__(256) void k_pack_dense(const double* __restrict__ src, const int2* __restrict__ tiles,
                                                    int ntp, int nbf, const uint8_t* __restrict__ mask, int q0,
                                                    double* __restrict__ B) {
    const int t = blockIdx.x, q = blockIdx.y;
    const int2 IJ = tiles[t];
    const double* S = src + static_cast<size_t>(q) * nbf * nbf;
    double* o = B + (static_cast<size_t>(q0 + q) * ntp + t) * TILE2;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int e = threadIdx.x + 256 * k;
        const int r = e >> 5, c = e & 31;
        const int m = IJ.x * TILE + r, n = IJ.y * TILE + c;
        double v = 0.0;
        if (m < nbf && n < nbf) {
            const size_t mn = static_cast<size_t>(m) * nbf + n;
            if (mask[mn]) v = S[mn];
        }
        o[e] = v;
    }
}

void launch_pack_dense(const TileGeom& g, const double* src, const uint8_t* mask, int q0, int nq, double* B,
                       hipStream_t s) {
    require(q0 >= 0 && nq >= 1 && nq <= 65535, "pack_dense");
    hipLaunchKernelGGL(k_pack_dense, dim3(g.ntp, nq), dim3(256), 0, s, src, g.d_tiles, g.ntp, g.nbf, mask, q0, B);
}

void launch_synth_b(const TileGeom& g, const double* decay, uint64_t seed, int q0, int q1, double* B, hipStream_t s) {
    require(q0 >= 0 && q1 > q0, "synth_b");
    dim3 grid(g.ntp, q1 - q0);
    hipLaunchKernelGGL(k_synth_b, grid, dim3(256), 0, s, g.d_tiles, g.ntp, g.nbf, decay,
                       static_cast<uint32_t>(seed & 0xffffffffu), static_cast<uint32_t>(seed >> 32), q0, B);
}

}  // namespace p4a
