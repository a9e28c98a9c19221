// Launchers for the pipelined K-path kernels (kernels_k.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "kernels.hpp"

namespace p4a {

// Column panels of the half-transform: 64*nbw columns per panel (nbw <= 5), npanel panels, ldc = npanel*64*nbw.
void ht_panel_shape(int ncol, int* nbw, int* npanel, int* ldc);

// X[q][m][i] = sum_n B_{q0+q}[m][n] C[n][i]; C [nbfp][ldc], X [nq][xrows][ldc] with
// xrows = 128*ceil(nT/4); every row of X is written (rows past nbf are zero).
// zt: device zero tile of at least TILE*TILE*8 bytes (read for absent blocks).
void launch_ht2_f64(const TileGeom& g, const double* B, int q0, int nq, const double* C, int ldc, int nbw, int npanel,
                    double* X, int xrows, const void* zt, hipStream_t s);
void launch_ht2_f32(const TileGeom& g, const float* B, int q0, int nq, const float* C, int ldc, int nbw, int npanel,
                    float* X, int xrows, const void* zt, hipStream_t s);
// Half-transform workgroup shape (tuning): 0 = default; 1 = one row-wave (256 threads);
// 2 = two row-waves (512 threads, the FP64/FP32 default); for BF16x3 also 3 = two 4-wave
// 128-row workgroups per CU and 4 = C staged through registers.
void set_ht_rowwaves(int wm);
// BF16x3 contraction: load its operands straight into LDS (1) or stage them in registers (0).
void set_kgemm_direct(int on);
// 3-term BF16: B FP32 [naux][ntp][32][32] and its tile-transposed copy BT (split into hi/lo
// as they are staged), C^T hi/lo [ldc][nbfp], X hi/lo [nq][xrows][ldc].
void launch_ht_bf3(const TileGeom& g, const float* B, const float* BT, int q0, int nq, const uint16_t* CTh, const uint16_t* CTl,
                   int ldc, int nbw, int npanel, uint16_t* Xh, uint16_t* Xl, int xrows, const void* zt,
                   hipStream_t s);

// P[split][tile] (+)= alpha * sum_b A_b[:, kbeg:kend] B_b[:, kbeg:kend]^T over 128x128 tiles;
// A_b = A + b*stride, rows of length ld, xrows rows (a multiple of 128).
// nvalid: rows of X that are not zero padding (nbf); fragments past it are skipped
void launch_kgemm2_f64(const double* A, const double* B, size_t stride, int xrows, int nvalid, int ld, int kbeg,
                       int kend, int nbatch, int nsplit, const int2* tiles, int ntile, double* P, double alpha,
                       int accumulate, int upper, hipStream_t s);
void launch_kgemm2_f32(const float* A, const float* B, size_t stride, int xrows, int nvalid, int ld, int kbeg, int kend,
                       int nbatch, int nsplit, const int2* tiles, int ntile, double* P, double alpha, int accumulate,
                       int upper, hipStream_t s);
void launch_kgemm_bf3(const uint16_t* Ah, const uint16_t* Al, const uint16_t* Bh, const uint16_t* Bl, size_t stride,
                      int xrows, int nvalid, int ld, int kbeg, int kend, int nbatch, int nsplit, const int2* tiles,
                      int ntile, double* P, double alpha, int accumulate, int upper, hipStream_t s);

// Precision copies and operand packing (cmap: destination column -> source column or -1).
void launch_to_f32(const double* src, float* dst, size_t n, hipStream_t s);
// FP32 with every 32x32 tile transposed (the blocks the BF16 half-transform reads as (J,I)).
void launch_to_f32_t(const double* src, float* dst, size_t n, hipStream_t s);
void launch_pack_cmap_f64(const double* src, int ldsrc, int nbf, const int* cmap, double* dst, int nbfp, int ldc,
                          hipStream_t s);
void launch_pack_cmap_f32(const double* src, int ldsrc, int nbf, const int* cmap, float* dst, int nbfp, int ldc,
                          hipStream_t s);
void launch_pack_cT_bf(const double* src, int ldsrc, int nbf, const int* cmap, uint16_t* hi, uint16_t* lo, int nbfp,
                       int ldc, hipStream_t s);
// s[i] = +1 / -1 if Cr[:,i] is bitwise +/- Cl[:,(i + shift) % ncol], else 0.
void launch_column_signs(const double* Cl, const double* Cr, int ld, int nbf, int ncol, int shift, int8_t* s,
                         hipStream_t st);

}  // namespace p4a
