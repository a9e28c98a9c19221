// Launchers for the DF-JK device kernels.  Every launcher validates the shapes
// it derives its grid from and throws std::runtime_error on a mismatch.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "common.hpp"

namespace p4a {

// Geometry of the tiled, block-upper-triangular B layout.
struct TileGeom {
    int nbf = 0;    // basis functions
    int nbfp = 0;   // padded to a multiple of TILE
    int nT = 0;     // tiles per dimension
    int ntp = 0;    // stored tile pairs (I <= J)
    const int* d_tidx = nullptr;   // [nT*nT] stored index of (min,max) or -1
    const int2* d_tiles = nullptr; // [ntp] (I, J), I <= J
    size_t q_stride() const { return static_cast<size_t>(ntp) * TILE2; }
};

// ---- J path (FP64, HBM-bound) ----
// W = tile-packed symmetrised density: I<J tiles hold D_mn + D_nm, diagonal tiles D_mn.
void launch_pack_w(const TileGeom& g, const double* D, int ldD, double* W, hipStream_t s);
// dpart[t*naux+Q] = <B_Q tile t, W tile t> for Q in [q0, q1)
// Which copy of the resident tensor the J kernels read; every variant accumulates in FP64.
struct JSource {
    enum Kind { F64 = 0, F32 = 1 } kind = F64;
    const void* a = nullptr;  // the FP64 tensor or its FP32 copy
};
void launch_j1(const TileGeom& g, const JSource& B, const double* W, int naux, int q0, int q1, double* dpart,
               hipStream_t s);
void launch_j1_reduce(const TileGeom& g, const double* dpart, int naux, double* d, hipStream_t s);
// Jpart[c][t][e] = sum_{Q in chunk c} B[Q][t][e] d[Q]   (chunks of qchunk over [0,naux))
void launch_j2(const TileGeom& g, const JSource& B, const double* d, int naux, int qchunk, double* Jpart,
               hipStream_t s);
// J[m][n] = alpha * sum_c Jpart[c][tile(m,n)] (+ J if accumulate)
void launch_j_unpack(const TileGeom& g, const double* Jpart, int nchunk, double* J, int ldJ, double alpha,
                     int accumulate, hipStream_t s);

// ---- K path ----
// Pack a dense row-major nbf x ncol matrix (any leading dim) into [nbfp][ldc] zero-padded.
void launch_pack_c(const double* Csrc, int ldsrc, int nbf, int ncol, double scale, double* Cpad, int nbfp, int ldc,
                   hipStream_t s);
// Batched NT product accumulated into 128x128 output tiles:
//   P[split][tile] (+)= sum_{b in split} A_b[m][:K] . B_b[n][:K]
// A_b = A + b*strideA (row-major, lda), rows < M; likewise B (rows < N).
// tiles: list of (Ib, Jb) 128-blocks, ntile entries.
void launch_gemm_nt_f64(const double* A, size_t strideA, int lda, const double* B, size_t strideB, int ldb, int M,
                        int N, int K, int nbatch, int nsplit, const int2* d_tiles, int ntile, double* P, int accumulate,
                        hipStream_t s);
// out[m][n] = alpha * sum_split P[split][tile(m,n)][..] (+ out if accumulate).  sym 1: only
// tiles Ib<=Jb exist and the lower triangle is read transposed; sym 2: every tile exists and
// out = alpha * (P + P^T), exactly symmetric.  tile_of: [nB*nB] -> tile index.
void launch_tile_reduce(const double* P, int nsplit, int ntile, const int* d_tile_of, int nB, int sym, int M, int N,
                        double* out, int ldo, double alpha, int accumulate, hipStream_t s);

// Plain FP64 GEMM, row-major, C = alpha*A*B + beta*C (A MxK lda, B KxN ldb).  MFMA, 128x128 tiles.
// batch > 1: operand z at A + z*sA, B + z*sB, C + z*sC (a zero stride shares the operand).
// lower_a: A lower triangular (its zero upper part is skipped)
void launch_gemm_nn_f64(const double* A, int lda, const double* B, int ldb, double* C, int ldc, int M, int N, int K,
                        double alpha, double beta, hipStream_t s, int batch = 1, size_t sA = 0, size_t sB = 0,
                        size_t sC = 0, int lower_a = 0);

// ---- adjoint helpers ----
// Z = a X + b Y (n elements)
void launch_axpby_mat(const double* X, double a, const double* Y, double b, double* Z, size_t n, hipStream_t s);
// A = (A + A^T) / 2 in place (n x n, leading dimension lda)
void launch_hermitivitize(double* A, int n, int lda, hipStream_t s);
// partial[0..nblocks) = block sums of X.*Y (deterministic; sum the partials on the host)
void launch_dot_partial(const double* X, const double* Y, size_t n, double* partial, int nblocks, hipStream_t s);
// p[0..n) = 0 by a kernel (graph-capturable, unlike hipMemsetAsync)
void launch_zero(double* p, size_t n, hipStream_t s);
// out[r][c] = f[r] * A[r][c] (rows x cols, dense row-major)
void launch_scale_rows(const double* A, int rows, int cols, const double* f, double* out, hipStream_t s);
// AT[c][r] = A[r][c]
void launch_transpose(const double* A, int rows, int cols, int lda, double* AT, int ldt, hipStream_t s);
// G[q][:][:] = a * d[q0+q] * D for q < nq (D n x n)
void launch_fill_scaled(const double* d, int q0, int nq, double a, const double* D, int n, double* G, hipStream_t s);

// ---- synthetic tensor-level workload ----
// B_Qmn = s_Q * decay[|m-n|] * (G_Qmn + G_Qnm)/2, G from Philox4x32-10 keyed by seed, Q in [q0,q1).
// Stored tiles of Q = q0 .. q0+nq-1 from a dense device [nq][nbf][nbf] tensor; mask is nbf x nbf (0 = screened).
void launch_pack_dense(const TileGeom& g, const double* src, const uint8_t* mask, int q0, int nq, double* B,
                       hipStream_t s);
void launch_synth_b(const TileGeom& g, const double* decay, uint64_t seed, int q0, int q1, double* B, hipStream_t s);

}  // namespace p4a
