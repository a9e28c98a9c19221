#pragma once
// Device code of the three-centre Coulomb integrals (P|mn), included by one translation unit per
// auxiliary angular momentum (kernels_ints_l<LC>.hip) so the classes compile in parallel.
//
// Three-centre Coulomb integrals (P|mn) on the device: the integral-generation step before
// the fit (dfhelper.cc:1284-1347, the Libint2 role), by the McMurchie-Davidson scheme of the
// host engine in psi4_amd/csrc/ints (same normalisation, Cartesian and solid-harmonic order).
//
// One launch per angular-momentum class (LA, LB | LC) with LA >= LB, so every loop bound is a
// compile-time constant.  A workgroup of 64 threads takes one primary shell pair and 64
// auxiliary functions of angular momentum LC; the pair's primitive data (exponent sum, centre,
// prefactor, Hermite expansion coefficients E) are built once per chunk of primitive pairs in
// LDS, and each thread accumulates the Cartesian (ab| block of its auxiliary function:
//   (ab|k) = sum_{pairs} pref sum_{tuv} E^ab_tuv W_tuv,   W_tuv = sum_{xyz} V^k_xyz R_{t+x,u+y,v+z}
// where V^k folds the auxiliary function's solid-harmonic combination and its own Hermite
// expansion (sign (-1)^(x+y+z) included), and R are the Hermite Coulomb integrals of the
// primitive pair and auxiliary primitive, built in place from F_n by downward levels.
// The block is transformed to solid harmonics and written to the tiled B layout through
// the same destination map as the host engine (both (m,n) and (n,m) positions).
#include <hip/hip_runtime.h>

#include "kernels_ints.hpp"

namespace p4a {
namespace {

constexpr double kPi = 3.14159265358979323846;
// primitive pairs per LDS fill (one per thread), at most 16 KB of pair data per workgroup
__host__ __device__ constexpr int chunk_for(int prec) { return prec * 8 * 64 <= 16384 ? 64 : (16384 / (prec * 8) > 0 ? 16384 / (prec * 8) : 1); }

__host__ __device__ constexpr int ncart(int l) { return (l + 1) * (l + 2) / 2; }
__host__ __device__ constexpr int nherm(int L) { return (L + 1) * (L + 2) * (L + 3) / 6; }
__host__ __device__ constexpr int hidx(int t, int u, int v) {
    return (t + u + v) * (t + u + v + 1) * (t + u + v + 2) / 6 + (u + v) * (u + v + 1) / 2 + v;
}
// Cartesian component i of degree l in the order xx, xy, xz, yy, yz, zz (lx descending, then ly)
__host__ __device__ constexpr void cart_comp(int l, int i, int& lx, int& ly, int& lz) {
    int k = 0;
    for (int a = l; a >= 0; --a)
        for (int b = l - a; b >= 0; --b) {
            if (k == i) {
                lx = a;
                ly = b;
                lz = l - a - b;
                return;
            }
            ++k;
        }
    lx = ly = lz = 0;
}

// F_0..F_L at T: Taylor expansion of the highest order from the tabulated grid, then downward
// recursion; above the table the asymptotic form with upward recursion (the host engine's
// scheme, psi4_amd/csrc/ints/ints.cpp boys()).
template <int L>
__device__ inline void boys_dev(double T, const BoysGrid& bg, double* F) {
    if (T < bg.tmax) {
        const int i = static_cast<int>(T * bg.inv_h + 0.5);
        const double dT = i * bg.h - T;
        const double* r = bg.table + static_cast<size_t>(i) * bg.ncol + L;
        double f = r[6];
        f = r[5] + dT * (1.0 / 6.0) * f;
        f = r[4] + dT * 0.2 * f;
        f = r[3] + dT * 0.25 * f;
        f = r[2] + dT * (1.0 / 3.0) * f;
        f = r[1] + dT * 0.5 * f;
        f = r[0] + dT * f;
        F[L] = f;
        if (L > 0) {
            const double e = exp(-T);
#pragma unroll
            for (int n = L - 1; n >= 0; --n) F[n] = (2.0 * T * F[n + 1] + e) / (2 * n + 1);
        }
    } else {
        const double e = exp(-T);
        F[0] = 0.5 * sqrt(kPi / T);
#pragma unroll
        for (int n = 1; n <= L; ++n) F[n] = ((2 * n - 1) * F[n - 1] - e) / (2.0 * T);
    }
}

// Hermite Coulomb integrals R^0_tuv(alpha, PQ) for t+u+v <= L, in place over n = L..0:
// level N of R^n is built from levels N-1, N-2 of R^(n+1) before those are overwritten.
template <int L>
__device__ inline void hermite_R_dev(double alpha, double X, double Y, double Z, const BoysGrid& bg, double* R) {
    double F[L + 1];
    boys_dev<L>(alpha * (X * X + Y * Y + Z * Z), bg, F);
    double pw[L + 1];
    pw[0] = 1.0;
#pragma unroll
    for (int n = 1; n <= L; ++n) pw[n] = pw[n - 1] * (-2.0 * alpha);
    R[0] = pw[L] * F[L];
#pragma unroll
    for (int n = L - 1; n >= 0; --n) {
#pragma unroll
        for (int N = L - n; N >= 1; --N) {
#pragma unroll
            for (int t = N; t >= 0; --t)
#pragma unroll
                for (int u = N - t; u >= 0; --u) {
                    const int v = N - t - u;
                    double r;
                    if (t > 0) {
                        r = X * R[hidx(t - 1, u, v)];
                        if (t > 1) r += (t - 1) * R[hidx(t - 2, u, v)];
                    } else if (u > 0) {
                        r = Y * R[hidx(0, u - 1, v)];
                        if (u > 1) r += (u - 1) * R[hidx(0, u - 2, v)];
                    } else {
                        r = Z * R[hidx(0, 0, v - 1)];
                        if (v > 1) r += (v - 1) * R[hidx(0, 0, v - 2)];
                    }
                    R[hidx(t, u, v)] = r;
                }
        }
        R[0] = pw[n] * F[n];
    }
}

// Hermite expansion coefficients of a primitive pair along one axis:
// E[(i*(LB+1)+j)*(LA+LB+1)+t] for x_A^i x_B^j (host engine build_E).
template <int LA, int LB>
__device__ inline void build_E_dev(double oo2p, double PA, double PB, double* e) {
    constexpr int st = LA + LB + 1;
    constexpr int esz = (LA + 1) * (LB + 1) * st;
#pragma unroll
    for (int k = 0; k < esz; ++k) e[k] = 0.0;
    e[0] = 1.0;
#pragma unroll
    for (int i = 0; i < LA; ++i)
#pragma unroll
        for (int t = 0; t <= i + 1; ++t) {
            double v = (t <= i) ? PA * e[(i * (LB + 1)) * st + t] : 0.0;
            if (t > 0) v += oo2p * e[(i * (LB + 1)) * st + t - 1];
            if (t + 1 <= i) v += (t + 1) * e[(i * (LB + 1)) * st + t + 1];
            e[((i + 1) * (LB + 1)) * st + t] = v;
        }
#pragma unroll
    for (int i = 0; i <= LA; ++i)
#pragma unroll
        for (int j = 0; j < LB; ++j)
#pragma unroll
            for (int t = 0; t <= i + j + 1; ++t) {
                double v = (t <= i + j) ? PB * e[(i * (LB + 1) + j) * st + t] : 0.0;
                if (t > 0) v += oo2p * e[(i * (LB + 1) + j) * st + t - 1];
                if (t + 1 <= i + j) v += (t + 1) * e[(i * (LB + 1) + j) * st + t + 1];
                e[(i * (LB + 1) + j + 1) * st + t] = v;
            }
}

template <int LA, int LB, int LC>
__global__ __launch_bounds__(64) void k_three_index(ThreeIndexArgs a, const int2* __restrict__ pairs, int npair,
                                                    const int2* __restrict__ afun, int nafun) {
    constexpr int LAB = LA + LB;
    constexpr int L = LAB + LC;
    constexpr int NCA = ncart(LA), NCB = ncart(LB), NCC = ncart(LC);
    constexpr int ST = LAB + 1;
    constexpr int ESZ = (LA + 1) * (LB + 1) * ST;
    constexpr int PREC = 5 + 3 * ESZ;  // p, Px, Py, Pz, pref, Ex, Ey, Ez
    constexpr int CH = chunk_for(PREC);
    __shared__ double lds[CH * PREC];

    const int ip = blockIdx.x;
    if (ip >= npair) return;
    const int2 pr = pairs[ip];
    const GpuShell sa = a.prim[pr.x], sb = a.prim[pr.y];
    const int f = blockIdx.y * 64 + threadIdx.x;
    const bool active = f < nafun;
    const int2 fk = afun[active ? f : nafun - 1];
    const GpuShell sc = a.aux[fk.x];

    // V^k: the auxiliary function's solid-harmonic row folded into its Hermite expansion per
    // Cartesian axis; built per auxiliary primitive below from the unit-exponent-free parts.
    const double* srow = a.aux_pure ? a.c2s + a.c2s_off[LC] + fk.y * NCC : nullptr;

    double out[NCA * NCB];
#pragma unroll
    for (int i = 0; i < NCA * NCB; ++i) out[i] = 0.0;

    const double ABx = sa.A[0] - sb.A[0], ABy = sa.A[1] - sb.A[1], ABz = sa.A[2] - sb.A[2];
    const double AB2 = ABx * ABx + ABy * ABy + ABz * ABz;
    const int npp = sa.nprim * sb.nprim;
    for (int c0 = 0; c0 < npp; c0 += CH) {
        const int nch = min(CH, npp - c0);
        __syncthreads();
        if (threadIdx.x < nch) {
            const int q = c0 + threadIdx.x;
            const int ia = q / sb.nprim, ib = q % sb.nprim;
            const double ea = a.pexp[sa.p0 + ia], eb = a.pexp[sb.p0 + ib];
            const double p = ea + eb, oo2p = 0.5 / p;
            double* d = lds + threadIdx.x * PREC;
            const double Px = (ea * sa.A[0] + eb * sb.A[0]) / p;
            const double Py = (ea * sa.A[1] + eb * sb.A[1]) / p;
            const double Pz = (ea * sa.A[2] + eb * sb.A[2]) / p;
            d[0] = p;
            d[1] = Px;
            d[2] = Py;
            d[3] = Pz;
            d[4] = a.pcoef[sa.p0 + ia] * a.pcoef[sb.p0 + ib] * exp(-ea * eb / p * AB2);
            build_E_dev<LA, LB>(oo2p, Px - sa.A[0], Px - sb.A[0], d + 5);
            build_E_dev<LA, LB>(oo2p, Py - sa.A[1], Py - sb.A[1], d + 5 + ESZ);
            build_E_dev<LA, LB>(oo2p, Pz - sa.A[2], Pz - sb.A[2], d + 5 + 2 * ESZ);
        }
        __syncthreads();
        if (!active) continue;
        for (int jc = 0; jc < sc.nprim; ++jc) {
            const double g = a.aexp[sc.p0 + jc], cg = a.acoef[sc.p0 + jc];
            // Hermite expansion of x^i exp(-g x^2) about the auxiliary centre, per axis
            double ec[(LC + 1) * (LC + 1)];
#pragma unroll
            for (int k = 0; k < (LC + 1) * (LC + 1); ++k) ec[k] = 0.0;
            ec[0] = 1.0;
#pragma unroll
            for (int i = 0; i < LC; ++i)
#pragma unroll
                for (int t = 0; t <= i + 1; ++t) {
                    double v = (t > 0) ? (0.5 / g) * ec[i * (LC + 1) + t - 1] : 0.0;
                    if (t + 1 <= i) v += (t + 1) * ec[i * (LC + 1) + t + 1];
                    ec[(i + 1) * (LC + 1) + t] = v;
                }
            double V[nherm(LC)];
#pragma unroll
            for (int k = 0; k < nherm(LC); ++k) V[k] = 0.0;
#pragma unroll
            for (int c = 0; c < NCC; ++c) {
                int cx = 0, cy = 0, cz = 0;
                cart_comp(LC, c, cx, cy, cz);
                const double w = srow ? srow[c] : (c == fk.y ? 1.0 : 0.0);
                if (w == 0.0) continue;
#pragma unroll
                for (int x = 0; x <= cx; ++x)
#pragma unroll
                    for (int y = 0; y <= cy; ++y)
#pragma unroll
                        for (int z = 0; z <= cz; ++z) {
                            const double s = ((x + y + z) & 1) ? -1.0 : 1.0;
                            V[hidx(x, y, z)] +=
                                w * s * ec[cx * (LC + 1) + x] * ec[cy * (LC + 1) + y] * ec[cz * (LC + 1) + z];
                        }
            }
            for (int q = 0; q < nch; ++q) {
                const double* d = lds + q * PREC;
                const double p = d[0];
                double alpha = p * g / (p + g);
                double pref = 2.0 * kPi * kPi * sqrt(kPi) / (p * g * sqrt(p + g)) * d[4] * cg;
                if (a.omega > 0.0) {
                    const double w2 = a.omega * a.omega;
                    const double aw = alpha * w2 / (alpha + w2);
                    pref *= sqrt(aw / alpha);
                    alpha = aw;
                }
                double R[nherm(L)];
                hermite_R_dev<L>(alpha, d[1] - sc.A[0], d[2] - sc.A[1], d[3] - sc.A[2], a.boys, R);
                double W[nherm(LAB)];
#pragma unroll
                for (int N = 0; N <= LAB; ++N)
#pragma unroll
                    for (int t = N; t >= 0; --t)
#pragma unroll
                        for (int u = N - t; u >= 0; --u) {
                            const int v = N - t - u;
                            double s = 0.0;
#pragma unroll
                            for (int M = 0; M <= LC; ++M)
#pragma unroll
                                for (int x = M; x >= 0; --x)
#pragma unroll
                                    for (int y = M - x; y >= 0; --y) {
                                        const int z = M - x - y;
                                        s += V[hidx(x, y, z)] * R[hidx(t + x, u + y, v + z)];
                                    }
                            W[hidx(t, u, v)] = pref * s;
                        }
                const double* Ex = d + 5;
                const double* Ey = d + 5 + ESZ;
                const double* Ez = d + 5 + 2 * ESZ;
#pragma unroll
                for (int ca = 0; ca < NCA; ++ca)
#pragma unroll
                    for (int cb = 0; cb < NCB; ++cb) {
                        int ax = 0, ay = 0, az = 0, bx = 0, by = 0, bz = 0;
                        cart_comp(LA, ca, ax, ay, az);
                        cart_comp(LB, cb, bx, by, bz);
                        const double* ex = Ex + (ax * (LB + 1) + bx) * ST;
                        const double* ey = Ey + (ay * (LB + 1) + by) * ST;
                        const double* ez = Ez + (az * (LB + 1) + bz) * ST;
                        double s = 0.0;
#pragma unroll
                        for (int t = 0; t <= ax + bx; ++t)
#pragma unroll
                            for (int u = 0; u <= ay + by; ++u)
#pragma unroll
                                for (int v = 0; v <= az + bz; ++v) s += ex[t] * ey[u] * ez[v] * W[hidx(t, u, v)];
                        out[ca * NCB + cb] += s;
                    }
            }
        }
    }
    if (!active) return;

    // solid harmonics on the primary pair (identity for Cartesian or s shells) and scatter
    constexpr int NFA = 2 * LA + 1, NFB = 2 * LB + 1;
    const bool pa = a.prim_pure && LA > 0, pb = a.prim_pure && LB > 0;
    const int nfa = pa ? NFA : NCA, nfb = pb ? NFB : NCB;
    const double* ta = a.c2s + a.c2s_off[LA];
    const double* tb = a.c2s + a.c2s_off[LB];
    const int64_t row = static_cast<int64_t>(sc.foff + fk.y) * a.p_stride;
    for (int ma = 0; ma < nfa; ++ma)
        for (int mb = 0; mb < nfb; ++mb) {
            double v = 0.0;
#pragma unroll
            for (int ca = 0; ca < NCA; ++ca) {
                const double wa = pa ? ta[ma * NCA + ca] : (ca == ma ? 1.0 : 0.0);
                if (wa == 0.0) continue;
#pragma unroll
                for (int cb = 0; cb < NCB; ++cb) {
                    const double wb = pb ? tb[mb * NCB + cb] : (cb == mb ? 1.0 : 0.0);
                    v += wa * wb * out[ca * NCB + cb];
                }
            }
            const int m = sa.foff + ma, n = sb.foff + mb;
            if (!a.dst) {
                a.out[row + m] = v;
                continue;
            }
            const int64_t d1 = a.dst[static_cast<int64_t>(m) * a.nbf + n];
            const int64_t d2 = a.dst[static_cast<int64_t>(n) * a.nbf + m];
            if (d1 >= 0) a.out[row + d1] = v;
            if (d2 >= 0 && d2 != d1) a.out[row + d2] = v;
        }
}

template <int LA, int LB, int LC>
void launch_class(const ThreeIndexArgs& a, const int2* pairs, int npair, const int2* afun, int nafun, hipStream_t s) {
    if (npair == 0 || nafun == 0) return;
    dim3 grid(npair, (nafun + 63) / 64);
    hipLaunchKernelGGL((k_three_index<LA, LB, LC>), grid, dim3(64), 0, s, a, pairs, npair, afun, nafun);
}

}  // namespace
}  // namespace p4a
