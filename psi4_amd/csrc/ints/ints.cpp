// McMurchie-Davidson integral engine (host, FP64).  See ints.hpp for scope.
//
// Algorithm (Helgaker, Jorgensen, Olsen, "Molecular Electronic-Structure
// Theory", ch. 9): Hermite expansion coefficients E^{ij}_t for each primitive
// pair, Hermite Coulomb integrals R^n_{tuv} by the standard recursion seeded
// with Boys functions, and
//   (ab|cd) = 2 pi^{5/2} / (p q sqrt(p+q)) sum_{tuv} E^{ab}_{tuv}
//             sum_{tau nu phi} (-1)^{tau+nu+phi} E^{cd}_{tau nu phi} R_{t+tau,u+nu,v+phi}.
// Three- and two-centre integrals use a dummy s shell of exponent zero.
#include "ints.hpp"

#include <omp.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <mutex>
#include <stdexcept>

namespace p4a {

namespace {

constexpr double kPi = 3.14159265358979323846;

// (n-1)!!, with (-1)!! = 0!! = 1; psi4's df[n] table (libmints/gshell.cc).
double dfact_m1(int n) {
    double r = 1.0;
    for (int k = n - 1; k > 1; k -= 2) r *= k;
    return r;
}

inline int ncart_of(int l) { return (l + 1) * (l + 2) / 2; }
inline int nherm(int L) { return (L + 1) * (L + 2) * (L + 3) / 6; }
// Linear Hermite index: shells of total order N, within a shell the Cartesian order.
inline int hidx(int t, int u, int v) {
    int N = t + u + v;
    return N * (N + 1) * (N + 2) / 6 + (N - t) * (N - t + 1) / 2 + v;
}

struct CartTable {
    std::array<std::vector<std::array<int, 3>>, kMaxTotalL + 1> comps;
    CartTable() {
        for (int l = 0; l <= kMaxTotalL; ++l)
            for (int lx = l; lx >= 0; --lx)
                for (int ly = l - lx; ly >= 0; --ly) comps[l].push_back({lx, ly, l - lx - ly});
    }
};
const CartTable& carts() {
    static CartTable t;
    return t;
}

// ---------------------------------------------------------------- Boys
constexpr int kBoysNCol = kMaxTotalL + 8;
constexpr double kBoysH = 1.0 / 32.0;
constexpr double kBoysTMax = 45.0;
constexpr int kBoysNT = static_cast<int>(kBoysTMax * 32) + 2;

struct BoysTable {
    std::vector<double> t;  // [kBoysNT][kBoysNCol]
    BoysTable() : t(static_cast<size_t>(kBoysNT) * kBoysNCol) {
        for (int i = 0; i < kBoysNT; ++i) {
            long double T = i * static_cast<long double>(kBoysH);
            int n = kBoysNCol - 1;
            // F_n(T) = e^{-T} sum_k (2T)^k / prod_{j=0..k} (2n+2j+1)
            long double term = 1.0L / (2 * n + 1), sum = term;
            for (int k = 1; k < 2000; ++k) {
                term *= 2.0L * T / (2 * n + 2 * k + 1);
                sum += term;
                if (term < sum * 1e-22L) break;
            }
            long double eT = std::exp(-T);
            long double F = eT * sum;
            t[static_cast<size_t>(i) * kBoysNCol + n] = static_cast<double>(F);
            for (int m = n - 1; m >= 0; --m) {
                F = (2.0L * T * F + eT) / (2 * m + 1);
                t[static_cast<size_t>(i) * kBoysNCol + m] = static_cast<double>(F);
            }
        }
    }
};
const BoysTable& boys_table() {
    static BoysTable b;
    return b;
}

// ---------------------------------------------------------------- solid harmonics
struct SphTable {
    std::array<std::vector<double>, kMaxShellL + 1> c2s;
    SphTable() {
        // Real solid harmonics by the recursion of Helgaker et al. eq. (6.4.70)-(6.4.72),
        // held as dense coefficient vectors over the Cartesian monomials of degree l.
        using Poly = std::vector<double>;
        auto cidx = [](int l, int lx, int ly) { return (l - lx) * (l - lx + 1) / 2 + (l - lx - ly); };
        // S[l][m + l]
        std::vector<std::vector<Poly>> S(kMaxShellL + 1);
        S[0].push_back(Poly{1.0});
        auto mulvar = [&](const Poly& p, int l, int var) {  // p (degree l) * x/y/z -> degree l+1
            Poly r(ncart_of(l + 1), 0.0);
            for (auto& c : carts().comps[l]) {
                double v = p[cidx(l, c[0], c[1])];
                if (v == 0.0) continue;
                int a = c[0] + (var == 0), b = c[1] + (var == 1);
                r[cidx(l + 1, a, b)] += v;
            }
            return r;
        };
        auto mulr2 = [&](const Poly& p, int l) {  // p (degree l) * r^2 -> degree l+2
            Poly r(ncart_of(l + 2), 0.0);
            for (auto& c : carts().comps[l]) {
                double v = p[cidx(l, c[0], c[1])];
                if (v == 0.0) continue;
                r[cidx(l + 2, c[0] + 2, c[1])] += v;
                r[cidx(l + 2, c[0], c[1] + 2)] += v;
                r[cidx(l + 2, c[0], c[1])] += v;
            }
            return r;
        };
        auto axpy = [](double a, const Poly& x, double b, const Poly& y) {
            Poly r(x.size());
            for (size_t i = 0; i < x.size(); ++i) r[i] = a * x[i] + b * y[i];
            return r;
        };
        for (int l = 0; l < kMaxShellL; ++l) {
            S[l + 1].assign(2 * l + 3, Poly());
            double d0 = (l == 0) ? 1.0 : 0.0;
            double f = std::sqrt(std::pow(2.0, d0) * (2 * l + 1) / (2.0 * l + 2.0));
            const Poly& Sll = S[l][2 * l];
            const Poly& Slml = S[l][0];
            S[l + 1][2 * l + 2] = axpy(f, mulvar(Sll, l, 0), -f * (1 - d0), mulvar(Slml, l, 1));
            S[l + 1][0] = axpy(f, mulvar(Sll, l, 1), f * (1 - d0), mulvar(Slml, l, 0));
            for (int m = -l; m <= l; ++m) {
                Poly zS = mulvar(S[l][m + l], l, 2);
                Poly r2S(ncart_of(l + 1), 0.0);
                if (l >= 1 && std::abs(m) <= l - 1) r2S = mulr2(S[l - 1][m + l - 1], l - 1);
                double a = (2 * l + 1) / std::sqrt((l + m + 1.0) * (l - m + 1.0));
                double b = -std::sqrt((l + m) * static_cast<double>(l - m)) / std::sqrt((l + m + 1.0) * (l - m + 1.0));
                S[l + 1][m + l + 1] = axpy(a, zS, b, r2S);
            }
        }
        for (int l = 0; l <= kMaxShellL; ++l) {
            int nc = ncart_of(l);
            c2s[l].assign(static_cast<size_t>(2 * l + 1) * nc, 0.0);
            for (int k = 0; k < 2 * l + 1; ++k) {
                int m = (k == 0) ? 0 : ((k % 2) ? (k + 1) / 2 : -(k / 2));  // 0,+1,-1,+2,-2,...
                const Poly& p = S[l][m + l];
                // norm^2 relative to the x^l Cartesian function (same radial part)
                double nrm = 0.0;
                for (auto& c1 : carts().comps[l])
                    for (auto& c2 : carts().comps[l]) {
                        int a = c1[0] + c2[0], b = c1[1] + c2[1], c = c1[2] + c2[2];
                        if ((a | b | c) & 1) continue;
                        nrm += p[cidx(l, c1[0], c1[1])] * p[cidx(l, c2[0], c2[1])] * dfact_m1(a) * dfact_m1(b) *
                               dfact_m1(c);
                    }
                nrm /= dfact_m1(2 * l);
                double s = 1.0 / std::sqrt(nrm);
                for (int i = 0; i < nc; ++i) c2s[l][static_cast<size_t>(k) * nc + i] = p[i] * s;
            }
        }
    }
};

// ---------------------------------------------------------------- pair data
struct PrimPair {
    double p, P[3], pref;  // pref = c_a c_b exp(-mu |AB|^2)
    double ea, eb;         // the two primitive exponents
};

struct PairData {
    int la = 0, lb = 0, lab = 0, esz = 0;
    std::vector<PrimPair> prims;
    std::vector<double> E;  // per prim: 3 * esz; index ((i*(lb+1)+j)*(lab+1)+t)
    const double* Eprim(size_t k) const { return E.data() + k * 3 * esz; }
};

void build_E(int la, int lb, double oo2p, const double* PA, const double* PB, double* E) {
    const int lab = la + lb, st = lab + 1;
    const int esz = (la + 1) * (lb + 1) * st;
    for (int d = 0; d < 3; ++d) {
        double* e = E + d * esz;
        std::fill(e, e + esz, 0.0);
        auto at = [&](int i, int j, int t) -> double& { return e[(i * (lb + 1) + j) * st + t]; };
        at(0, 0, 0) = 1.0;
        for (int i = 0; i < la; ++i)
            for (int t = 0; t <= i + 1; ++t) {
                double v = PA[d] * (t <= i ? at(i, 0, t) : 0.0);
                if (t > 0) v += oo2p * at(i, 0, t - 1);
                if (t + 1 <= i) v += (t + 1) * at(i, 0, t + 1);
                at(i + 1, 0, t) = v;
            }
        for (int i = 0; i <= la; ++i)
            for (int j = 0; j < lb; ++j)
                for (int t = 0; t <= i + j + 1; ++t) {
                    double v = PB[d] * (t <= i + j ? at(i, j, t) : 0.0);
                    if (t > 0) v += oo2p * at(i, j, t - 1);
                    if (t + 1 <= i + j) v += (t + 1) * at(i, j, t + 1);
                    at(i, j + 1, t) = v;
                }
    }
}

PairData make_pair(const Shell& a, const Shell& b) {
    PairData pd;
    pd.la = a.l;
    pd.lb = b.l;
    pd.lab = a.l + b.l;
    pd.esz = (a.l + 1) * (b.l + 1) * (pd.lab + 1);
    double AB2 = 0;
    for (int d = 0; d < 3; ++d) AB2 += (a.A[d] - b.A[d]) * (a.A[d] - b.A[d]);
    for (size_t i = 0; i < a.exps.size(); ++i)
        for (size_t j = 0; j < b.exps.size(); ++j) {
            double ai = a.exps[i], bj = b.exps[j], p = ai + bj;
            double mu = (p > 0) ? ai * bj / p : 0.0;
            PrimPair pp;
            pp.p = p;
            pp.ea = ai;
            pp.eb = bj;
            pp.pref = a.coefs[i] * b.coefs[j] * std::exp(-mu * AB2);
            if (std::fabs(pp.pref) < 1e-30) continue;
            double PA[3], PB[3];
            for (int d = 0; d < 3; ++d) {
                pp.P[d] = (ai * a.A[d] + bj * b.A[d]) / p;
                PA[d] = pp.P[d] - a.A[d];
                PB[d] = pp.P[d] - b.A[d];
            }
            pd.prims.push_back(pp);
            size_t off = pd.E.size();
            pd.E.resize(off + 3 * pd.esz);
            build_E(a.l, b.l, 0.5 / p, PA, PB, pd.E.data() + off);
        }
    return pd;
}

Shell dummy_s(const Shell& like) {
    Shell s;
    s.l = 0;
    s.pure = false;
    std::copy(like.A, like.A + 3, s.A);
    s.exps = {0.0};
    s.coefs = {1.0};
    return s;
}

// Per-thread scratch.  omega > 0 selects the attenuated operator erf(omega r)/r for the
// two-electron integrals (0: plain Coulomb).
struct Work {
    std::vector<double> tmp, R, W, cart, half;
    double omega = 0.0;
};

// R^0_{tuv}(alpha, PQ) for t+u+v <= L into R (size nherm(L)).
void hermite_R(int L, double alpha, const double* PQ, Work& w) {
    const int nh = nherm(L);
    w.tmp.resize(static_cast<size_t>(L + 1) * nh);
    w.R.resize(nh);
    double F[kMaxTotalL + 1];
    double T = alpha * (PQ[0] * PQ[0] + PQ[1] * PQ[1] + PQ[2] * PQ[2]);
    boys(L, T, F);
    double* t = w.tmp.data();
    double f = 1.0;
    for (int n = 0; n <= L; ++n) {
        t[n * nh] = f * F[n];
        f *= -2.0 * alpha;
    }
    for (int N = 1; N <= L; ++N)
        for (int n = 0; n <= L - N; ++n) {
            const double* up = t + (n + 1) * nh;
            double* cur = t + n * nh;
            for (int tt = N; tt >= 0; --tt)
                for (int uu = N - tt; uu >= 0; --uu) {
                    int vv = N - tt - uu;
                    double v;
                    if (tt > 0) {
                        v = PQ[0] * up[hidx(tt - 1, uu, vv)];
                        if (tt > 1) v += (tt - 1) * up[hidx(tt - 2, uu, vv)];
                    } else if (uu > 0) {
                        v = PQ[1] * up[hidx(0, uu - 1, vv)];
                        if (uu > 1) v += (uu - 1) * up[hidx(0, uu - 2, vv)];
                    } else {
                        v = PQ[2] * up[hidx(0, 0, vv - 1)];
                        if (vv > 1) v += (vv - 1) * up[hidx(0, 0, vv - 2)];
                    }
                    cur[hidx(tt, uu, vv)] = v;
                }
        }
    std::copy(t, t + nh, w.R.begin());
}

// Cartesian (ab|cd) for the contracted pairs, accumulated into out[(a*ncb+b)][(c*ncd+d)].
void eri_cart(const PairData& bra, const PairData& ket, double* out, Work& w) {
    const auto& ca = carts().comps[bra.la];
    const auto& cb = carts().comps[bra.lb];
    const auto& cc = carts().comps[ket.la];
    const auto& cd = carts().comps[ket.lb];
    const int nbra = static_cast<int>(ca.size() * cb.size());
    const int nket = static_cast<int>(cc.size() * cd.size());
    const int L = bra.lab + ket.lab;
    const int nhb = nherm(bra.lab);
    std::fill(out, out + static_cast<size_t>(nbra) * nket, 0.0);
    w.W.resize(static_cast<size_t>(nket) * nhb);
    const int stb = bra.lab + 1, stk = ket.lab + 1;
    for (size_t ib = 0; ib < bra.prims.size(); ++ib) {
        const PrimPair& pb = bra.prims[ib];
        const double* Eb = bra.Eprim(ib);
        for (size_t ik = 0; ik < ket.prims.size(); ++ik) {
            const PrimPair& pk = ket.prims[ik];
            const double* Ek = ket.Eprim(ik);
            double p = pb.p, q = pk.p;
            double alpha = p * q / (p + q);
            double pref = 2.0 * std::pow(kPi, 2.5) / (p * q * std::sqrt(p + q)) * pb.pref * pk.pref;
            if (w.omega > 0.0) {
                // erf(omega r)/r acts like 1/r between the two distributions with one of them
                // smeared by a Gaussian of exponent omega^2: 1/alpha' = 1/alpha + 1/omega^2, and
                // the prefactor scales by sqrt(alpha'/alpha)
                const double w2 = w.omega * w.omega;
                const double aw = alpha * w2 / (alpha + w2);
                pref *= std::sqrt(aw / alpha);
                alpha = aw;
            }
            double PQ[3] = {pb.P[0] - pk.P[0], pb.P[1] - pk.P[1], pb.P[2] - pk.P[2]};
            hermite_R(L, alpha, PQ, w);
            const double* R = w.R.data();
            // ket contraction
            std::fill(w.W.begin(), w.W.end(), 0.0);
            int kc = 0;
            for (auto& c : cc)
                for (auto& d : cd) {
                    double* Wk = w.W.data() + static_cast<size_t>(kc) * nhb;
                    const double* ex = Ek + (c[0] * (ket.lb + 1) + d[0]) * stk;
                    const double* ey = Ek + ket.esz + (c[1] * (ket.lb + 1) + d[1]) * stk;
                    const double* ez = Ek + 2 * ket.esz + (c[2] * (ket.lb + 1) + d[2]) * stk;
                    for (int tau = 0; tau <= c[0] + d[0]; ++tau) {
                        if (ex[tau] == 0.0) continue;
                        for (int nu = 0; nu <= c[1] + d[1]; ++nu) {
                            double exy = ex[tau] * ey[nu];
                            if (exy == 0.0) continue;
                            for (int phi = 0; phi <= c[2] + d[2]; ++phi) {
                                double e = exy * ez[phi];
                                if (e == 0.0) continue;
                                if ((tau + nu + phi) & 1) e = -e;
                                for (int N = 0; N <= bra.lab; ++N)
                                    for (int t = N; t >= 0; --t)
                                        for (int u = N - t; u >= 0; --u) {
                                            int v = N - t - u;
                                            Wk[hidx(t, u, v)] += e * R[hidx(t + tau, u + nu, v + phi)];
                                        }
                            }
                        }
                    }
                    ++kc;
                }
            // bra contraction
            int kb = 0;
            for (auto& a : ca)
                for (auto& b : cb) {
                    const double* ex = Eb + (a[0] * (bra.lb + 1) + b[0]) * stb;
                    const double* ey = Eb + bra.esz + (a[1] * (bra.lb + 1) + b[1]) * stb;
                    const double* ez = Eb + 2 * bra.esz + (a[2] * (bra.lb + 1) + b[2]) * stb;
                    double* o = out + static_cast<size_t>(kb) * nket;
                    for (int t = 0; t <= a[0] + b[0]; ++t) {
                        if (ex[t] == 0.0) continue;
                        for (int u = 0; u <= a[1] + b[1]; ++u) {
                            double exy = ex[t] * ey[u];
                            if (exy == 0.0) continue;
                            for (int v = 0; v <= a[2] + b[2]; ++v) {
                                double e = pref * exy * ez[v];
                                if (e == 0.0) continue;
                                const double* Wc = w.W.data() + hidx(t, u, v);
                                for (int k = 0; k < nket; ++k) o[k] += e * Wc[static_cast<size_t>(k) * nhb];
                            }
                        }
                    }
                    ++kb;
                }
        }
    }
}

// Transform a block [n0c][n1c] (Cartesian on both axes) to [n0][n1] with the
// shells' pure/Cartesian choices.  Both axes may be composite (pair) indices;
// the caller passes the per-axis matrices as (rows, cols, M) lists.
void transform_axis(const double* in, double* out, int nouter, int nin, int ninner, const double* M, int nout) {
    // out[o][r][i] = sum_c M[r][c] in[o][c][i]
    for (int o = 0; o < nouter; ++o)
        for (int r = 0; r < nout; ++r) {
            double* dst = out + (static_cast<size_t>(o) * nout + r) * ninner;
            std::fill(dst, dst + ninner, 0.0);
            for (int c = 0; c < nin; ++c) {
                double m = M[r * nin + c];
                if (m == 0.0) continue;
                const double* src = in + (static_cast<size_t>(o) * nin + c) * ninner;
                for (int i = 0; i < ninner; ++i) dst[i] += m * src[i];
            }
        }
}

// In-place spherical transform of a 4-index Cartesian block for shells s0..s3 (any may be a
// Cartesian/dummy shell).  Result is [f0][f1][f2][f3] in w.cart.
void to_spherical(const Shell* s[4], Work& w) {
    int n[4], nc[4];
    for (int k = 0; k < 4; ++k) {
        nc[k] = s[k]->ncart();
        n[k] = nc[k];
    }
    for (int ax = 0; ax < 4; ++ax) {
        if (!s[ax]->pure || s[ax]->l == 0) continue;
        int outer = 1, inner = 1;
        for (int k = 0; k < ax; ++k) outer *= n[k];
        for (int k = ax + 1; k < 4; ++k) inner *= n[k];
        int nout = 2 * s[ax]->l + 1;
        w.half.resize(static_cast<size_t>(outer) * nout * inner);
        transform_axis(w.cart.data(), w.half.data(), outer, nc[ax], inner, cart2sph(s[ax]->l).data(), nout);
        n[ax] = nout;
        w.cart.swap(w.half);
        w.cart.resize(static_cast<size_t>(outer) * nout * inner);
    }
}

}  // namespace

// ---------------------------------------------------------------- public
void Basis::finalize() {
    offset.clear();
    nbf = 0;
    max_l = 0;
    max_nfunc = 0;
    for (auto& s : shells) {
        offset.push_back(nbf);
        nbf += s.nfunc();
        max_l = std::max(max_l, s.l);
        max_nfunc = std::max(max_nfunc, s.nfunc());
    }
    if (max_l > kMaxShellL) throw std::runtime_error("angular momentum above the supported maximum");
}

void normalize_shell(Shell& s) {
    const int l = s.l;
    for (size_t i = 0; i < s.exps.size(); ++i) {
        double g = 2.0 * s.exps[i];
        double z = std::pow(g, l + 1.5);
        s.coefs[i] *= std::sqrt(std::pow(2.0, l) * z / (kPi * std::sqrt(kPi) * dfact_m1(2 * l)));
    }
    double e_sum = 0.0;
    for (size_t i = 0; i < s.exps.size(); ++i)
        for (size_t j = 0; j < s.exps.size(); ++j)
            e_sum += s.coefs[i] * s.coefs[j] / std::pow(s.exps[i] + s.exps[j], l + 1.5);
    double tmp = kPi * std::sqrt(kPi) * dfact_m1(2 * l) / std::pow(2.0, l);
    double norm = std::sqrt(1.0 / (tmp * e_sum));
    if (norm != norm) {
        std::fill(s.coefs.begin(), s.coefs.end(), 1.0);
    } else {
        for (auto& c : s.coefs) c *= norm;
    }
}

const std::vector<double>& cart2sph(int l) {
    static SphTable tab;
    return tab.c2s[l];
}

const double* boys_grid(int* nt, int* ncol, double* h, double* tmax) {
    *nt = kBoysNT;
    *ncol = kBoysNCol;
    *h = kBoysH;
    *tmax = kBoysTMax;
    return boys_table().t.data();
}

void boys(int nmax, double T, double* F) {
    if (T < kBoysTMax) {
        const BoysTable& bt = boys_table();
        int i = static_cast<int>(T * 32.0 + 0.5);
        double dT = i * kBoysH - T;
        const double* r = bt.t.data() + static_cast<size_t>(i) * kBoysNCol + nmax;
        double f = r[6];
        f = r[5] + dT / 6.0 * f;
        f = r[4] + dT / 5.0 * f;
        f = r[3] + dT / 4.0 * f;
        f = r[2] + dT / 3.0 * f;
        f = r[1] + dT / 2.0 * f;
        f = r[0] + dT * f;
        F[nmax] = f;
        if (nmax > 0) {
            double e = std::exp(-T);
            for (int n = nmax - 1; n >= 0; --n) F[n] = (2.0 * T * F[n + 1] + e) / (2 * n + 1);
        }
    } else {
        double e = std::exp(-T);
        F[0] = 0.5 * std::sqrt(kPi / T);
        for (int n = 1; n <= nmax; ++n) F[n] = ((2 * n - 1) * F[n - 1] - e) / (2.0 * T);
    }
}

namespace {

// Transform a [nca][ncb] Cartesian one-electron block to functions and scatter.
// transpose_sign = +1 for Hermitian operators, -1 for antisymmetric ones (r x nabla).
void scatter_1e(const Basis& b, int I, int J, const std::vector<double>& blk, double* M, double transpose_sign = 1.0) {
    const Shell& si = b.shells[I];
    const Shell& sj = b.shells[J];
    Work w;
    w.cart = blk;
    Shell one = dummy_s(si);
    const Shell* s4[4] = {&si, &sj, &one, &one};
    to_spherical(s4, w);
    int ni = si.nfunc(), nj = sj.nfunc();
    for (int a = 0; a < ni; ++a)
        for (int c = 0; c < nj; ++c) {
            double v = w.cart[static_cast<size_t>(a) * nj + c];
            // within a diagonal block the later (c,a) visit overwrites both elements, so the
            // matrix is exactly (anti)symmetric
            M[static_cast<size_t>(b.offset[I] + a) * b.nbf + b.offset[J] + c] = v;
            M[static_cast<size_t>(b.offset[J] + c) * b.nbf + b.offset[I] + a] = transpose_sign * v;
        }
}

enum class OneE { S, T, V };

// Cartesian block <A|O|B> ([ncart(A)][ncart(B)], coefficients as stored in the shells).
// V sums over the nuclei [c0, c1).
void onee_cart(const Shell& A, const Shell& B, OneE kind, int c0, int c1, const double* Z, const double* xyz,
               std::vector<double>& blk, Work& w) {
    Shell Bx = B;
    if (kind == OneE::T) Bx.l = B.l + 2;  // E tables up to lb+2 for the Laplacian
    PairData pd = make_pair(A, Bx);
    const auto& ca = carts().comps[A.l];
    const auto& cb = carts().comps[B.l];
    blk.assign(ca.size() * cb.size(), 0.0);
    const int st = pd.lab + 1;
    for (size_t k = 0; k < pd.prims.size(); ++k) {
        const PrimPair& pp = pd.prims[k];
        const double* E = pd.Eprim(k);
        double s0 = std::pow(kPi / pp.p, 1.5) * pp.pref;
        if (kind == OneE::V) {
            int L = pd.lab;
            for (int c = c0; c < c1; ++c) {
                double PC[3] = {pp.P[0] - xyz[3 * c], pp.P[1] - xyz[3 * c + 1], pp.P[2] - xyz[3 * c + 2]};
                hermite_R(L, pp.p, PC, w);
                double f = -Z[c] * 2.0 * kPi / pp.p * pp.pref;
                size_t kk = 0;
                for (auto& a : ca)
                    for (auto& bb : cb) {
                        const double* ex = E + (a[0] * (pd.lb + 1) + bb[0]) * st;
                        const double* ey = E + pd.esz + (a[1] * (pd.lb + 1) + bb[1]) * st;
                        const double* ez = E + 2 * pd.esz + (a[2] * (pd.lb + 1) + bb[2]) * st;
                        double acc = 0.0;
                        for (int t = 0; t <= a[0] + bb[0]; ++t)
                            for (int u = 0; u <= a[1] + bb[1]; ++u)
                                for (int v = 0; v <= a[2] + bb[2]; ++v)
                                    acc += ex[t] * ey[u] * ez[v] * w.R[hidx(t, u, v)];
                        blk[kk++] += f * acc;
                    }
            }
        } else if (kind == OneE::S) {
            size_t kk = 0;
            for (auto& a : ca)
                for (auto& bb : cb) {
                    double sx = E[(a[0] * (pd.lb + 1) + bb[0]) * st];
                    double sy = E[pd.esz + (a[1] * (pd.lb + 1) + bb[1]) * st];
                    double sz = E[2 * pd.esz + (a[2] * (pd.lb + 1) + bb[2]) * st];
                    blk[kk++] += s0 * sx * sy * sz;
                }
        } else {
            const double beta = pp.eb;
            size_t kk = 0;
            auto s1 = [&](int d, int i, int j) -> double {
                if (j < 0) return 0.0;
                return E[d * pd.esz + (i * (pd.lb + 1) + j) * st];
            };
            for (auto& a : ca)
                for (auto& bb : cb) {
                    double s[3], t[3];
                    for (int d = 0; d < 3; ++d) {
                        int i = a[d], j = bb[d];
                        s[d] = s1(d, i, j);
                        t[d] = -0.5 * (4.0 * beta * beta * s1(d, i, j + 2) - 2.0 * beta * (2 * j + 1) * s1(d, i, j) +
                                       j * (j - 1) * s1(d, i, j - 2));
                    }
                    blk[kk++] += s0 * (t[0] * s[1] * s[2] + s[0] * t[1] * s[2] + s[0] * s[1] * t[2]);
                }
        }
    }
}

void one_electron(const Basis& b, OneE kind, int natom, const double* Z, const double* xyz, double* M) {
    const int ns = static_cast<int>(b.shells.size());
#pragma omp parallel for schedule(dynamic)
    for (int I = 0; I < ns; ++I) {
        Work w;
        std::vector<double> blk;
        for (int J = 0; J <= I; ++J) {
            onee_cart(b.shells[I], b.shells[J], kind, 0, natom, Z, xyz, blk, w);
            scatter_1e(b, I, J, blk, M);
        }
    }
}

// ---------------------------------------------------------------- derivative integrals
// d/dA_x of a primitive x^a e^{-alpha r_A^2} is 2 alpha x^{a+1} e^{..} - a x^{a-1} e^{..}.
// For a contracted shell S of angular momentum l the derivative functions are therefore
// combinations of the Cartesian components of two auxiliary shells on the same centre:
// S+ (l+1, coefficients 2 alpha_k c_k) and S- (l-1, coefficients c_k), with c_k the
// normalised coefficients of S (one normalisation per shell, shared by its components).
struct ShiftedShells {
    Shell plus, minus;
    bool has_minus = false;
};
ShiftedShells shifted(const Shell& s) {
    ShiftedShells r;
    r.plus = s;
    r.plus.l = s.l + 1;
    r.plus.pure = false;
    for (size_t k = 0; k < s.exps.size(); ++k) r.plus.coefs[k] = 2.0 * s.exps[k] * s.coefs[k];
    if (s.l > 0) {
        r.has_minus = true;
        r.minus = s;
        r.minus.l = s.l - 1;
        r.minus.pure = false;
    }
    return r;
}
inline int cart_index(int l, int lx, int ly) { return (l - lx) * (l - lx + 1) / 2 + (l - lx - ly); }

// Given Cartesian blocks of S+ and S- (S's index first, `rest` trailing elements each),
// form the three derivative blocks d[dir][ncart(S)][rest] w.r.t. S's centre.
void combine_deriv(int l, const std::vector<double>& bp, const std::vector<double>* bm, size_t rest,
                   std::vector<double> d[3]) {
    const auto& cs = carts().comps[l];
    for (int dir = 0; dir < 3; ++dir) d[dir].assign(cs.size() * rest, 0.0);
    for (size_t a = 0; a < cs.size(); ++a) {
        const auto& c = cs[a];
        for (int dir = 0; dir < 3; ++dir) {
            int up[3] = {c[0], c[1], c[2]};
            up[dir] += 1;
            const double* sp = bp.data() + static_cast<size_t>(cart_index(l + 1, up[0], up[1])) * rest;
            double* o = d[dir].data() + a * rest;
            for (size_t r = 0; r < rest; ++r) o[r] = sp[r];
            if (c[dir] > 0 && bm) {
                int dn[3] = {c[0], c[1], c[2]};
                dn[dir] -= 1;
                const double* sm = bm->data() + static_cast<size_t>(cart_index(l - 1, dn[0], dn[1])) * rest;
                const double f = static_cast<double>(c[dir]);
                for (size_t r = 0; r < rest; ++r) o[r] -= f * sm[r];
            }
        }
    }
}

// Spherical transform of a derivative block and contraction with a weight block W
// (rows: functions of shell A at offset oa, cols: functions of shell B at offset ob).
double contract_1e(const Basis& b, const Shell& A, const Shell& B, int oa, int ob, std::vector<double>& cart,
                   const double* W, Work& w) {
    w.cart = cart;
    Shell one = dummy_s(A);
    const Shell* s4[4] = {&A, &B, &one, &one};
    to_spherical(s4, w);
    const int na = A.nfunc(), nb = B.nfunc();
    double s = 0.0;
    for (int a = 0; a < na; ++a)
        for (int c = 0; c < nb; ++c)
            s += W[static_cast<size_t>(oa + a) * b.nbf + ob + c] * w.cart[static_cast<size_t>(a) * nb + c];
    return s;
}

// g_shell[I][dir] += 2 sum_{m in I, n} W_mn d<m|O|n>/dA_I (W symmetric); for V also the
// derivative with respect to each nucleus (translational invariance per nucleus) into g_nuc.
void onee_grad(const Basis& b, OneE kind, const double* W, int natom, const double* Z, const double* xyz,
               double* g_shell, double* g_nuc) {
    const int ns = static_cast<int>(b.shells.size());
    const int nc = kind == OneE::V ? natom : 1;
    std::vector<double> gn(static_cast<size_t>(ns) * nc * 3, 0.0);
#pragma omp parallel for schedule(dynamic)
    for (int I = 0; I < ns; ++I) {
        Work w;
        const Shell& A = b.shells[I];
        const ShiftedShells sh = shifted(A);
        std::vector<double> bp, bm, d[3];
        for (int J = 0; J < ns; ++J) {
            const Shell& B = b.shells[J];
            const size_t rest = static_cast<size_t>(B.ncart());
            for (int c = 0; c < nc; ++c) {
                const int c0 = kind == OneE::V ? c : 0, c1 = kind == OneE::V ? c + 1 : 0;
                onee_cart(sh.plus, B, kind, c0, c1, Z, xyz, bp, w);
                if (sh.has_minus) onee_cart(sh.minus, B, kind, c0, c1, Z, xyz, bm, w);
                combine_deriv(A.l, bp, sh.has_minus ? &bm : nullptr, rest, d);
                for (int dir = 0; dir < 3; ++dir)
                    gn[(static_cast<size_t>(I) * nc + c) * 3 + dir] +=
                        2.0 * contract_1e(b, A, B, b.offset[I], b.offset[J], d[dir], W, w);
            }
        }
    }
    for (int I = 0; I < ns; ++I)
        for (int c = 0; c < nc; ++c)
            for (int dir = 0; dir < 3; ++dir) {
                const double v = gn[(static_cast<size_t>(I) * nc + c) * 3 + dir];
                g_shell[3 * I + dir] += v;
                if (kind == OneE::V) g_nuc[3 * c + dir] -= v;  // d/dC = -(d/dA + d/dB) for nucleus C
            }
}

// Dipole integrals <a| r_d - O_d |b>, d = x, y, z, into M[3][nbf][nbf].  With the
// Hermite expansion of the pair overlap distribution, x_C = X_PC + (x - X_P) and
// integral (x - X_P) Lambda_t = delta_t1 (pi/p)^1/2, so
// <a|x_C|b>_x = (E^{ij}_1 + X_PC E^{ij}_0) (pi/p)^1/2 (Helgaker, Jorgensen, Olsen 9.5.41).
void dipole_ints(const Basis& b, const double* O, double* M) {
    const int ns = static_cast<int>(b.shells.size());
    const size_t nn = static_cast<size_t>(b.nbf) * b.nbf;
#pragma omp parallel for schedule(dynamic)
    for (int I = 0; I < ns; ++I) {
        for (int J = 0; J <= I; ++J) {
            const Shell& A = b.shells[I];
            const Shell& B = b.shells[J];
            PairData pd = make_pair(A, B);
            const auto& ca = carts().comps[A.l];
            const auto& cb = carts().comps[B.l];
            std::vector<double> blk[3];
            for (auto& v : blk) v.assign(ca.size() * cb.size(), 0.0);
            const int st = pd.lab + 1;
            for (size_t k = 0; k < pd.prims.size(); ++k) {
                const PrimPair& pp = pd.prims[k];
                const double* E = pd.Eprim(k);
                const double s0 = std::pow(kPi / pp.p, 1.5) * pp.pref;
                size_t kk = 0;
                for (auto& a : ca)
                    for (auto& bb : cb) {
                        double e0[3], e1[3];
                        for (int d = 0; d < 3; ++d) {
                            const double* e = E + d * pd.esz + (a[d] * (pd.lb + 1) + bb[d]) * st;
                            e0[d] = e[0];
                            e1[d] = (st > 1 ? e[1] : 0.0) + (pp.P[d] - O[d]) * e[0];
                        }
                        blk[0][kk] += s0 * e1[0] * e0[1] * e0[2];
                        blk[1][kk] += s0 * e0[0] * e1[1] * e0[2];
                        blk[2][kk] += s0 * e0[0] * e0[1] * e1[2];
                        ++kk;
                    }
            }
            for (int d = 0; d < 3; ++d) scatter_1e(b, I, J, blk[d], M + d * nn);
        }
    }
}

// Angular momentum integrals <a| (r - O) x nabla |b> (no factor -i), d = x, y, z, into
// M[3][nbf][nbf] (antisymmetric).  Per Cartesian direction, with the 1D factors of the
// Hermite expansion (common (pi/p)^3/2 prefactor taken out):
//   S(i,j) = E^{ij}_0,   M(i,j) = E^{ij}_1 + (P - O) E^{ij}_0          (moment x - O_x)
//   D(i,j) = j S(i,j-1) - 2 beta S(i,j+1)                              (d/dx on the ket)
//   L_x = S_x (M_y D_z - D_y M_z), and cyclic.
void angmom_ints(const Basis& b, const double* O, double* M) {
    const int ns = static_cast<int>(b.shells.size());
    const size_t nn = static_cast<size_t>(b.nbf) * b.nbf;
#pragma omp parallel for schedule(dynamic)
    for (int I = 0; I < ns; ++I) {
        for (int J = 0; J <= I; ++J) {
            const Shell& A = b.shells[I];
            const Shell& B = b.shells[J];
            Shell Bx = B;
            Bx.l = B.l + 1;  // E tables up to lb+1 for the derivative
            PairData pd = make_pair(A, Bx);
            const auto& ca = carts().comps[A.l];
            const auto& cb = carts().comps[B.l];
            std::vector<double> blk[3];
            for (auto& v : blk) v.assign(ca.size() * cb.size(), 0.0);
            const int st = pd.lab + 1;
            for (size_t k = 0; k < pd.prims.size(); ++k) {
                const PrimPair& pp = pd.prims[k];
                const double* E = pd.Eprim(k);
                const double s0 = std::pow(kPi / pp.p, 1.5) * pp.pref;
                const double beta = pp.eb;
                auto e = [&](int d, int i, int j, int t) -> double {
                    if (j < 0 || t > i + j) return 0.0;
                    return E[d * pd.esz + (i * (pd.lb + 1) + j) * st + t];
                };
                size_t kk = 0;
                for (auto& a : ca)
                    for (auto& bb : cb) {
                        double S[3], Mo[3], Dv[3];
                        for (int d = 0; d < 3; ++d) {
                            const int i = a[d], j = bb[d];
                            S[d] = e(d, i, j, 0);
                            Mo[d] = e(d, i, j, 1) + (pp.P[d] - O[d]) * e(d, i, j, 0);
                            Dv[d] = j * e(d, i, j - 1, 0) - 2.0 * beta * e(d, i, j + 1, 0);
                        }
                        blk[0][kk] += s0 * S[0] * (Mo[1] * Dv[2] - Dv[1] * Mo[2]);
                        blk[1][kk] += s0 * S[1] * (Mo[2] * Dv[0] - Dv[2] * Mo[0]);
                        blk[2][kk] += s0 * S[2] * (Mo[0] * Dv[1] - Dv[0] * Mo[1]);
                        ++kk;
                    }
            }
            for (int d = 0; d < 3; ++d) scatter_1e(b, I, J, blk[d], M + d * nn, -1.0);
        }
    }
}

}  // namespace

void angular_momentum(const Basis& b, const double* origin, double* M) { angmom_ints(b, origin, M); }
void overlap(const Basis& b, double* S) { one_electron(b, OneE::S, 0, nullptr, nullptr, S); }
void overlap_grad(const Basis& b, const double* W, double* g_shell) {
    onee_grad(b, OneE::S, W, 0, nullptr, nullptr, g_shell, nullptr);
}
void kinetic_grad(const Basis& b, const double* D, double* g_shell) {
    onee_grad(b, OneE::T, D, 0, nullptr, nullptr, g_shell, nullptr);
}
void potential_grad(const Basis& b, const double* D, int natom, const double* Z, const double* xyz, double* g_shell,
                    double* g_nuc) {
    onee_grad(b, OneE::V, D, natom, Z, xyz, g_shell, g_nuc);
}
void dipole(const Basis& b, const double* origin, double* M) { dipole_ints(b, origin, M); }
void kinetic(const Basis& b, double* T) {
    one_electron(b, OneE::T, 0, nullptr, nullptr, T);
}
void potential(const Basis& b, int natom, const double* Z, const double* xyz, double* V) {
    one_electron(b, OneE::V, natom, Z, xyz, V);
}

void metric(const Basis& aux, double* J, int nthreads, double omega) {
    const int ns = static_cast<int>(aux.shells.size());
    std::vector<PairData> pd(ns);
    std::vector<Shell> dum(ns);
    for (int P = 0; P < ns; ++P) {
        dum[P] = dummy_s(aux.shells[P]);
        pd[P] = make_pair(aux.shells[P], dum[P]);
    }
    const int n = aux.nbf;
#pragma omp parallel for schedule(dynamic) num_threads(nthreads)
    for (int P = 0; P < ns; ++P) {
        Work w;
        w.omega = omega;
        for (int Q = 0; Q <= P; ++Q) {
            const Shell& sp = aux.shells[P];
            const Shell& sq = aux.shells[Q];
            w.cart.resize(static_cast<size_t>(sp.ncart()) * sq.ncart());
            eri_cart(pd[P], pd[Q], w.cart.data(), w);
            const Shell* s4[4] = {&sp, &dum[P], &sq, &dum[Q]};
            to_spherical(s4, w);
            int np = sp.nfunc(), nq = sq.nfunc();
            for (int a = 0; a < np; ++a)
                for (int c = 0; c < nq; ++c) {
                    double v = w.cart[static_cast<size_t>(a) * nq + c];
                    J[static_cast<size_t>(aux.offset[P] + a) * n + aux.offset[Q] + c] = v;
                    J[static_cast<size_t>(aux.offset[Q] + c) * n + aux.offset[P] + a] = v;
                }
        }
    }
}

void three_index(const Basis& prim, const Basis& aux, const int64_t* dst, int64_t p_stride, const uint8_t* shell_mask,
                 double* out, int nthreads, double omega) {
    const int na = static_cast<int>(aux.shells.size());
    const int np = static_cast<int>(prim.shells.size());
    const int nbf = prim.nbf;
    std::vector<PairData> apd(na);
    std::vector<Shell> dum(na);
    for (int P = 0; P < na; ++P) {
        dum[P] = dummy_s(aux.shells[P]);
        apd[P] = make_pair(aux.shells[P], dum[P]);
    }
    std::vector<std::pair<int, int>> pairs;
    for (int M = 0; M < np; ++M)
        for (int N = 0; N <= M; ++N)
            if (!shell_mask || shell_mask[static_cast<size_t>(M) * np + N]) pairs.emplace_back(M, N);
    const int npairs = static_cast<int>(pairs.size());
#pragma omp parallel for schedule(dynamic) num_threads(nthreads)
    for (int ip = 0; ip < npairs; ++ip) {
        Work w;
        w.omega = omega;
        const int M = pairs[ip].first, N = pairs[ip].second;
        const Shell& sm = prim.shells[M];
        const Shell& sn = prim.shells[N];
        PairData bra = make_pair(sm, sn);
        const int nm = sm.nfunc(), nn = sn.nfunc();
        for (int P = 0; P < na; ++P) {
            const Shell& sp = aux.shells[P];
            w.cart.resize(static_cast<size_t>(sm.ncart()) * sn.ncart() * sp.ncart());
            eri_cart(bra, apd[P], w.cart.data(), w);
            const Shell* s4[4] = {&sm, &sn, &sp, &dum[P]};
            to_spherical(s4, w);
            const int nP = sp.nfunc();
            for (int a = 0; a < nm; ++a)
                for (int b = 0; b < nn; ++b) {
                    int m = prim.offset[M] + a, nidx = prim.offset[N] + b;
                    int64_t d1 = dst[static_cast<int64_t>(m) * nbf + nidx];
                    int64_t d2 = dst[static_cast<int64_t>(nidx) * nbf + m];
                    const double* src = w.cart.data() + (static_cast<size_t>(a) * nn + b) * nP;
                    for (int c = 0; c < nP; ++c) {
                        int64_t row = static_cast<int64_t>(aux.offset[P] + c) * p_stride;
                        if (d1 >= 0) out[row + d1] = src[c];
                        if (d2 >= 0 && d2 != d1) out[row + d2] = src[c];
                    }
                }
        }
    }
}

// sum_{P,m,n} G[P][m][n] d(P|mn)/dR over the centres R of the shells.  G is symmetric in
// (m,n); with translational invariance d/dP = -(d/dm + d/dn), only the derivative with
// respect to the first primitive function is needed:
//   g_prim[M] += 2 sum_{m in M, n, P} G d_m(P|mn),   g_aux[P] -= the same sum over its P.
void three_index_grad(const Basis& prim, const Basis& aux, const double* G, double* g_prim, double* g_aux,
                      int nthreads) {
    const int na = static_cast<int>(aux.shells.size());
    const int np = static_cast<int>(prim.shells.size());
    const int nbf = prim.nbf;
    const size_t nn = static_cast<size_t>(nbf) * nbf;
    std::vector<PairData> apd(na);
    std::vector<Shell> dum(na);
    for (int P = 0; P < na; ++P) {
        dum[P] = dummy_s(aux.shells[P]);
        apd[P] = make_pair(aux.shells[P], dum[P]);
    }
    std::vector<double> gp(static_cast<size_t>(np) * 3, 0.0), ga(static_cast<size_t>(na) * 3, 0.0);
#pragma omp parallel num_threads(nthreads)
    {
        std::vector<double> lgp(gp.size(), 0.0), lga(ga.size(), 0.0);
        Work w;
        std::vector<double> bp, bm, d[3];
#pragma omp for schedule(dynamic)
        for (int ip = 0; ip < np * np; ++ip) {
            const int M = ip / np, N = ip % np;
            const Shell& sm = prim.shells[M];
            const Shell& sn = prim.shells[N];
            const ShiftedShells sh = shifted(sm);
            PairData pplus = make_pair(sh.plus, sn);
            PairData pminus;
            if (sh.has_minus) pminus = make_pair(sh.minus, sn);
            const int nm = sm.nfunc(), nnf = sn.nfunc();
            for (int P = 0; P < na; ++P) {
                const Shell& sp = aux.shells[P];
                const size_t rest = static_cast<size_t>(sn.ncart()) * sp.ncart();
                bp.resize(static_cast<size_t>(sh.plus.ncart()) * rest);
                eri_cart(pplus, apd[P], bp.data(), w);
                if (sh.has_minus) {
                    bm.resize(static_cast<size_t>(sh.minus.ncart()) * rest);
                    eri_cart(pminus, apd[P], bm.data(), w);
                }
                combine_deriv(sm.l, bp, sh.has_minus ? &bm : nullptr, rest, d);
                const int nP = sp.nfunc();
                for (int dir = 0; dir < 3; ++dir) {
                    w.cart = d[dir];
                    const Shell* s4[4] = {&sm, &sn, &sp, &dum[P]};
                    to_spherical(s4, w);
                    double v = 0.0;
                    for (int a = 0; a < nm; ++a)
                        for (int b = 0; b < nnf; ++b) {
                            const size_t mn = static_cast<size_t>(prim.offset[M] + a) * nbf + prim.offset[N] + b;
                            const double* src = w.cart.data() + (static_cast<size_t>(a) * nnf + b) * nP;
                            for (int c = 0; c < nP; ++c) v += G[static_cast<size_t>(aux.offset[P] + c) * nn + mn] * src[c];
                        }
                    lgp[3 * M + dir] += 2.0 * v;
                    lga[3 * P + dir] -= 2.0 * v;
                }
            }
        }
#pragma omp critical
        {
            for (size_t i = 0; i < gp.size(); ++i) gp[i] += lgp[i];
            for (size_t i = 0; i < ga.size(); ++i) ga[i] += lga[i];
        }
    }
    for (size_t i = 0; i < gp.size(); ++i) g_prim[i] += gp[i];
    for (size_t i = 0; i < ga.size(); ++i) g_aux[i] += ga[i];
}

// sum_{P,Q} W[P][Q] d(P|Q)/dR (W symmetric): g_aux[P] += 2 sum_{p in P, Q} W_pQ d_p(p|Q).
void metric_grad(const Basis& aux, const double* W, double* g_aux, int nthreads) {
    const int na = static_cast<int>(aux.shells.size());
    const int n = aux.nbf;
    std::vector<PairData> apd(na);
    std::vector<Shell> dum(na);
    for (int P = 0; P < na; ++P) {
        dum[P] = dummy_s(aux.shells[P]);
        apd[P] = make_pair(aux.shells[P], dum[P]);
    }
    std::vector<double> ga(static_cast<size_t>(na) * 3, 0.0);
#pragma omp parallel for schedule(dynamic) num_threads(nthreads)
    for (int P = 0; P < na; ++P) {
        Work w;
        std::vector<double> bp, bm, d[3];
        const Shell& sp = aux.shells[P];
        const ShiftedShells sh = shifted(sp);
        PairData pplus = make_pair(sh.plus, dum[P]);
        PairData pminus;
        if (sh.has_minus) pminus = make_pair(sh.minus, dum[P]);
        double acc[3] = {0, 0, 0};
        for (int Q = 0; Q < na; ++Q) {
            const Shell& sq = aux.shells[Q];
            const size_t rest = static_cast<size_t>(sq.ncart());
            bp.resize(static_cast<size_t>(sh.plus.ncart()) * rest);
            eri_cart(pplus, apd[Q], bp.data(), w);
            if (sh.has_minus) {
                bm.resize(static_cast<size_t>(sh.minus.ncart()) * rest);
                eri_cart(pminus, apd[Q], bm.data(), w);
            }
            combine_deriv(sp.l, bp, sh.has_minus ? &bm : nullptr, rest, d);
            const int nP = sp.nfunc(), nQ = sq.nfunc();
            for (int dir = 0; dir < 3; ++dir) {
                w.cart = d[dir];
                const Shell* s4[4] = {&sp, &dum[P], &sq, &dum[Q]};
                to_spherical(s4, w);
                double v = 0.0;
                for (int a = 0; a < nP; ++a)
                    for (int c = 0; c < nQ; ++c)
                        v += W[static_cast<size_t>(aux.offset[P] + a) * n + aux.offset[Q] + c] *
                             w.cart[static_cast<size_t>(a) * nQ + c];
                acc[dir] += 2.0 * v;
            }
        }
        for (int dir = 0; dir < 3; ++dir) ga[3 * P + dir] = acc[dir];
    }
    for (size_t i = 0; i < ga.size(); ++i) g_aux[i] += ga[i];
}

double schwarz_diagonal(const Basis& prim, double* fun_max, double* shell_max, int nthreads) {
    const int np = static_cast<int>(prim.shells.size());
    const int nbf = prim.nbf;
    std::fill(fun_max, fun_max + static_cast<size_t>(nbf) * nbf, 0.0);
    std::fill(shell_max, shell_max + static_cast<size_t>(np) * np, 0.0);
    double max_val = 0.0;
#pragma omp parallel for schedule(dynamic) num_threads(nthreads) reduction(max : max_val)
    for (int MU = 0; MU < np; ++MU) {
        Work w;
        const Shell& sm = prim.shells[MU];
        const int nmu = sm.nfunc();
        for (int NU = 0; NU <= MU; ++NU) {
            const Shell& sn = prim.shells[NU];
            const int nnu = sn.nfunc();
            PairData pd = make_pair(sm, sn);
            w.cart.resize(static_cast<size_t>(sm.ncart()) * sn.ncart() * sm.ncart() * sn.ncart());
            eri_cart(pd, pd, w.cart.data(), w);
            const Shell* s4[4] = {&sm, &sn, &sm, &sn};
            to_spherical(s4, w);
            for (int mu = 0; mu < nmu; ++mu) {
                int omu = prim.offset[MU] + mu;
                for (int nu = 0; nu < nnu; ++nu) {
                    int onu = prim.offset[NU] + nu;
                    if (omu < onu) continue;
                    size_t index = static_cast<size_t>(mu) * (nnu * nmu * nnu + nnu) + nu * (nmu * nnu + 1);
                    double val = std::fabs(w.cart[index]);
                    max_val = std::max(val, max_val);
                    // distinct (MU,NU) entries are owned by one thread (MU is the loop index)
                    double& sm1 = shell_max[static_cast<size_t>(MU) * np + NU];
                    if (sm1 <= val) {
                        sm1 = val;
                        shell_max[static_cast<size_t>(NU) * np + MU] = val;
                    }
                    double& f1 = fun_max[static_cast<size_t>(omu) * nbf + onu];
                    if (f1 <= val) {
                        f1 = val;
                        fun_max[static_cast<size_t>(onu) * nbf + omu] = val;
                    }
                }
            }
        }
    }
    return max_val;
}

void four_index(const Basis& b, double* out, int nthreads, double omega) {
    const int ns = static_cast<int>(b.shells.size());
    const size_t n = b.nbf;
    std::vector<std::pair<int, int>> pairs;
    for (int M = 0; M < ns; ++M)
        for (int N = 0; N < ns; ++N) pairs.emplace_back(M, N);
    std::vector<PairData> pd(pairs.size());
    for (size_t i = 0; i < pairs.size(); ++i) pd[i] = make_pair(b.shells[pairs[i].first], b.shells[pairs[i].second]);
    const int np = static_cast<int>(pairs.size());
#pragma omp parallel for schedule(dynamic) num_threads(nthreads)
    for (int i = 0; i < np; ++i) {
        Work w;
        w.omega = omega;
        for (int j = 0; j < np; ++j) {
            const Shell* s4[4] = {&b.shells[pairs[i].first], &b.shells[pairs[i].second], &b.shells[pairs[j].first],
                                  &b.shells[pairs[j].second]};
            w.cart.resize(static_cast<size_t>(s4[0]->ncart()) * s4[1]->ncart() * s4[2]->ncart() * s4[3]->ncart());
            eri_cart(pd[i], pd[j], w.cart.data(), w);
            to_spherical(s4, w);
            int n0 = s4[0]->nfunc(), n1 = s4[1]->nfunc(), n2 = s4[2]->nfunc(), n3 = s4[3]->nfunc();
            int o0 = b.offset[pairs[i].first], o1 = b.offset[pairs[i].second];
            int o2 = b.offset[pairs[j].first], o3 = b.offset[pairs[j].second];
            size_t k = 0;
            for (int a = 0; a < n0; ++a)
                for (int c = 0; c < n1; ++c)
                    for (int d = 0; d < n2; ++d)
                        for (int e = 0; e < n3; ++e)
                            out[(((o0 + a) * n + o1 + c) * n + o2 + d) * n + o3 + e] = w.cart[k++];
        }
    }
}

}  // namespace p4a
