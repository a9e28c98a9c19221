#pragma once
// Device code of the Schwarz diagonal (mn|mn), included by kernels_schwarz.hip (s, p, d shells)
// and kernels_schwarz_f.hip (pairs with an f shell), so the classes compile in parallel.
//
// Schwarz diagonal (mn|mn) on the device: the screening integrals of DFHelper::prepare_sparsity
// (dfhelper.cc:299-360, which asks Libint2 for every shell quartet (MN|MN), M >= N), by the
// McMurchie-Davidson scheme of the three-centre kernels (kernels_ints_impl.hpp) and of the host
// engine's schwarz_diagonal (psi4_amd/csrc/ints/ints.cpp).
//
// One thread per (shell pair, function pair).  For the function pair (a, b) of the shell pair
// (A, B), with the solid-harmonic rows of a and b folded into the Hermite expansion of each
// primitive pair i,
//   G^i_tuv = sum_{cart a', b'} w_a a' w_b b' E^{a'b'}_t E^{a'b'}_u E^{a'b'}_v,
// the integral is
//   (ab|ab) = sum_{i,j} K_ij sum_{tuv} G^i_tuv sum_{t'u'v'} (-1)^(t'+u'+v') G^j_t'u'v' R_{t+t',u+u',v+v'}
// with K_ij = 2 pi^(5/2) / (p_i p_j sqrt(p_i + p_j)) times the pairs' prefactors and R the Hermite
// Coulomb integrals at alpha = p_i p_j / (p_i + p_j), P_i - P_j.  The (i, j) and (j, i) terms are
// equal (R changes sign with the parity of its index under P_i <-> P_j, the same parity the
// (-1) factor moves to the other side), so j runs from i with a weight 2 off the diagonal.
#include "kernels_ints_impl.hpp"
#include "kernels_schwarz.hpp"

namespace p4a {
namespace {

// Hermite coefficients of one function pair of the shell pair (sa, sb) for its primitive pair q:
// G[hidx(t,u,v)], the exponent sum p, centre P and prefactor c_a c_b exp(-ea eb / p |AB|^2).
// sign != 0 multiplies G_tuv by (-1)^(t+u+v).
template <int LA, int LB>
__device__ inline void pair_hermite(const ThreeIndexArgs& a, const GpuShell& sa, const GpuShell& sb, double AB2,
                                    int q, const double* wa, const double* wb, bool sign, double& p, double* P,
                                    double& pref, double* G) {
    constexpr int ST = LA + LB + 1;
    constexpr int ESZ = (LA + 1) * (LB + 1) * ST;
    constexpr int NCA = ncart(LA), NCB = ncart(LB);
    const int ia = q / sb.nprim, ib = q - ia * sb.nprim;
    const double ea = a.pexp[sa.p0 + ia], eb = a.pexp[sb.p0 + ib];
    p = ea + eb;
    const double oo2p = 0.5 / p;
#pragma unroll
    for (int d = 0; d < 3; ++d) P[d] = (ea * sa.A[d] + eb * sb.A[d]) / p;
    pref = a.pcoef[sa.p0 + ia] * a.pcoef[sb.p0 + ib] * exp(-ea * eb / p * AB2);
    double Ex[ESZ], Ey[ESZ], Ez[ESZ];
    build_E_dev<LA, LB>(oo2p, P[0] - sa.A[0], P[0] - sb.A[0], Ex);
    build_E_dev<LA, LB>(oo2p, P[1] - sa.A[1], P[1] - sb.A[1], Ey);
    build_E_dev<LA, LB>(oo2p, P[2] - sa.A[2], P[2] - sb.A[2], Ez);
#pragma unroll
    for (int k = 0; k < nherm(LA + LB); ++k) G[k] = 0.0;
#pragma unroll
    for (int ca = 0; ca < NCA; ++ca) {
        if (wa[ca] == 0.0) continue;
#pragma unroll
        for (int cb = 0; cb < NCB; ++cb) {
            const double w = wa[ca] * wb[cb];
            if (w == 0.0) continue;
            int ax = 0, ay = 0, az = 0, bx = 0, by = 0, bz = 0;
            cart_comp(LA, ca, ax, ay, az);
            cart_comp(LB, cb, bx, by, bz);
            const double* ex = Ex + (ax * (LB + 1) + bx) * ST;
            const double* ey = Ey + (ay * (LB + 1) + by) * ST;
            const double* ez = Ez + (az * (LB + 1) + bz) * ST;
#pragma unroll
            for (int t = 0; t <= ax + bx; ++t)
#pragma unroll
                for (int u = 0; u <= ay + by; ++u)
#pragma unroll
                    for (int v = 0; v <= az + bz; ++v) {
                        const double s = (sign && ((t + u + v) & 1)) ? -w : w;
                        G[hidx(t, u, v)] += s * ex[t] * ey[u] * ez[v];
                    }
        }
    }
}

// sum_{tuv} Gi_tuv sum_{xyz} Gj_xyz R_{t+x,u+y,v+z} over t+u+v, x+y+z <= LAB, fully unrolled
template <int LAB>
__device__ inline double hermite_contract(const double* Gi, const double* Gj, const double* R) {
    double s = 0.0;
#pragma unroll
    for (int N = 0; N <= LAB; ++N)
#pragma unroll
        for (int t = N; t >= 0; --t)
#pragma unroll
            for (int u = N - t; u >= 0; --u) {
                const int v = N - t - u;
                double w = 0.0;
#pragma unroll
                for (int M = 0; M <= LAB; ++M)
#pragma unroll
                    for (int x = M; x >= 0; --x)
#pragma unroll
                        for (int y = M - x; y >= 0; --y) {
                            const int z = M - x - y;
                            w += Gj[hidx(x, y, z)] * R[hidx(t + x, u + y, v + z)];
                        }
                s += Gi[hidx(t, u, v)] * w;
            }
    return s;
}

// The same sum with the outer (t, u, v) loop kept rolled: for the (ff|ff) class, whose R
// (455 values) lives in scratch memory anyway, this keeps the kernel small enough to compile
// in reasonable time.
template <int LAB>
__device__ inline double hermite_contract_rolled(const double* Gi, const double* Gj, const double* R) {
    double s = 0.0;
#pragma unroll 1
    for (int N = 0; N <= LAB; ++N)
#pragma unroll 1
        for (int t = N; t >= 0; --t)
#pragma unroll 1
            for (int u = N - t; u >= 0; --u) {
                const int v = N - t - u;
                double w = 0.0;
#pragma unroll
                for (int M = 0; M <= LAB; ++M)
#pragma unroll
                    for (int x = M; x >= 0; --x)
#pragma unroll
                        for (int y = M - x; y >= 0; --y) {
                            const int z = M - x - y;
                            w += Gj[hidx(x, y, z)] * R[hidx(t + x, u + y, v + z)];
                        }
                s += Gi[hidx(t, u, v)] * w;
            }
    return s;
}

template <int LA, int LB>
__global__ __launch_bounds__(64) void k_schwarz(ThreeIndexArgs a, const int2* __restrict__ pairs, int npair, int nfa,
                                                int nfb, double* __restrict__ fun_max) {
    constexpr int LAB = LA + LB;
    constexpr int NCA = ncart(LA), NCB = ncart(LB);
    constexpr int NH = nherm(LAB);
    const int nfp = nfa * nfb;
    const int64_t gid = static_cast<int64_t>(blockIdx.x) * 64 + threadIdx.x;
    if (gid >= static_cast<int64_t>(npair) * nfp) return;
    const int ip = static_cast<int>(gid / nfp), fp = static_cast<int>(gid - static_cast<int64_t>(ip) * nfp);
    const int2 pr = pairs[ip];
    const int ma = fp / nfb, mb = fp - ma * nfb;
    // within a diagonal shell pair only functions m >= n, as dfhelper.cc:344 (the mirror is written below)
    if (pr.x == pr.y && ma < mb) return;
    const GpuShell sa = a.prim[pr.x], sb = a.prim[pr.y];

    // the solid-harmonic rows of the two functions (identity for Cartesian or s shells)
    const bool pa = a.prim_pure && LA > 0, pb = a.prim_pure && LB > 0;
    double wa[NCA], wb[NCB];
#pragma unroll
    for (int c = 0; c < NCA; ++c) wa[c] = pa ? a.c2s[a.c2s_off[LA] + ma * NCA + c] : (c == ma ? 1.0 : 0.0);
#pragma unroll
    for (int c = 0; c < NCB; ++c) wb[c] = pb ? a.c2s[a.c2s_off[LB] + mb * NCB + c] : (c == mb ? 1.0 : 0.0);

    const double AB2 = (sa.A[0] - sb.A[0]) * (sa.A[0] - sb.A[0]) + (sa.A[1] - sb.A[1]) * (sa.A[1] - sb.A[1]) +
                       (sa.A[2] - sb.A[2]) * (sa.A[2] - sb.A[2]);
    const int npp = sa.nprim * sb.nprim;
    double acc = 0.0;
    for (int i = 0; i < npp; ++i) {
        double pi, Pi[3], ci, Gi[NH];
        pair_hermite<LA, LB>(a, sa, sb, AB2, i, wa, wb, false, pi, Pi, ci, Gi);
        for (int j = i; j < npp; ++j) {
            double pj, Pj[3], cj, Gj[NH];
            pair_hermite<LA, LB>(a, sa, sb, AB2, j, wa, wb, true, pj, Pj, cj, Gj);
            const double alpha = pi * pj / (pi + pj);
            double R[nherm(2 * LAB)];
            hermite_R_dev<2 * LAB>(alpha, Pi[0] - Pj[0], Pi[1] - Pj[1], Pi[2] - Pj[2], a.boys, R);
            const double s = LAB >= 6 ? hermite_contract_rolled<LAB>(Gi, Gj, R) : hermite_contract<LAB>(Gi, Gj, R);
            const double k = 2.0 * kPi * kPi * sqrt(kPi) / (pi * pj * sqrt(pi + pj)) * ci * cj;
            acc += (j == i ? 1.0 : 2.0) * k * s;
        }
    }
    const double val = fabs(acc);
    const int m = sa.foff + ma, n = sb.foff + mb;
    fun_max[static_cast<int64_t>(m) * a.nbf + n] = val;
    fun_max[static_cast<int64_t>(n) * a.nbf + m] = val;
}

template <int LA, int LB>
void launch(const ThreeIndexArgs& a, const int2* pairs, int npair, int nfa, int nfb, double* fun_max,
            hipStream_t s) {
    const int64_t nthr = static_cast<int64_t>(npair) * nfa * nfb;
    if (nthr == 0) return;
    const int64_t nblk = (nthr + 63) / 64;
    if (nblk > 0x7fffffff) throw std::runtime_error("schwarz: too many function pairs in a class");
    hipLaunchKernelGGL((k_schwarz<LA, LB>), dim3(static_cast<unsigned>(nblk)), dim3(64), 0, s, a, pairs, npair, nfa,
                       nfb, fun_max);
}

}  // namespace
}  // namespace p4a
