// C ABI over the host integral engine (declarations in include/psi4amd_ints.h).
#include "psi4amd_ints.h"

#include <exception>
#include <string>

#include "basis_build.hpp"
#include "ints.hpp"

namespace {
thread_local std::string g_err;

template <class F>
int guarded(F&& f) {
    try {
        f();
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return 1;
    } catch (...) {
        g_err = "unknown error";
        return 1;
    }
}
}  // namespace

extern "C" {

const char* p4a_ints_last_error(void) { return g_err.c_str(); }

int p4a_ints_nbf(const p4a_basis* b, int* nbf) {
    return guarded([&] { *nbf = p4a::basis_from_desc(b).nbf; });
}

int p4a_ints_overlap(const p4a_basis* b, double* S) {
    return guarded([&] { p4a::overlap(p4a::basis_from_desc(b), S); });
}

int p4a_ints_kinetic(const p4a_basis* b, double* T) {
    return guarded([&] { p4a::kinetic(p4a::basis_from_desc(b), T); });
}

int p4a_ints_dipole(const p4a_basis* b, const double* origin, double* M) {
    return guarded([&] {
        const double zero[3] = {0.0, 0.0, 0.0};
        p4a::dipole(p4a::basis_from_desc(b), origin ? origin : zero, M);
    });
}

int p4a_ints_angular_momentum(const p4a_basis* b, const double* origin, double* M) {
    return guarded([&] {
        const double zero[3] = {0.0, 0.0, 0.0};
        p4a::angular_momentum(p4a::basis_from_desc(b), origin ? origin : zero, M);
    });
}

int p4a_ints_overlap_grad(const p4a_basis* b, const double* W, double* g_shell) {
    return guarded([&] { p4a::overlap_grad(p4a::basis_from_desc(b), W, g_shell); });
}
int p4a_ints_kinetic_grad(const p4a_basis* b, const double* D, double* g_shell) {
    return guarded([&] { p4a::kinetic_grad(p4a::basis_from_desc(b), D, g_shell); });
}
int p4a_ints_potential_grad(const p4a_basis* b, const double* D, int natom, const double* Z, const double* xyz,
                            double* g_shell, double* g_nuc) {
    return guarded([&] { p4a::potential_grad(p4a::basis_from_desc(b), D, natom, Z, xyz, g_shell, g_nuc); });
}
int p4a_ints_three_index_grad(const p4a_basis* prim, const p4a_basis* aux, const double* G, double* g_prim,
                              double* g_aux, int nthreads) {
    return guarded([&] {
        p4a::three_index_grad(p4a::basis_from_desc(prim), p4a::basis_from_desc(aux), G, g_prim, g_aux, nthreads);
    });
}
int p4a_ints_metric_grad(const p4a_basis* aux, const double* W, double* g_aux, int nthreads) {
    return guarded([&] { p4a::metric_grad(p4a::basis_from_desc(aux), W, g_aux, nthreads); });
}

int p4a_ints_potential(const p4a_basis* b, int natom, const double* Z, const double* xyz, double* V) {
    return guarded([&] { p4a::potential(p4a::basis_from_desc(b), natom, Z, xyz, V); });
}

int p4a_ints_metric(const p4a_basis* aux, double* J, int nthreads) {
    return guarded([&] { p4a::metric(p4a::basis_from_desc(aux), J, nthreads); });
}

int p4a_ints_three_index(const p4a_basis* prim, const p4a_basis* aux, const int64_t* dst, int64_t p_stride,
                         const uint8_t* shell_mask, double* out, int nthreads) {
    return guarded([&] {
        p4a::three_index(p4a::basis_from_desc(prim), p4a::basis_from_desc(aux), dst, p_stride, shell_mask, out,
                         nthreads);
    });
}

int p4a_ints_metric_erf(const p4a_basis* aux, double omega, double* J, int nthreads) {
    return guarded([&] {
        if (!(omega > 0.0)) throw std::runtime_error("omega must be positive");
        p4a::metric(p4a::basis_from_desc(aux), J, nthreads, omega);
    });
}

int p4a_ints_three_index_erf(const p4a_basis* prim, const p4a_basis* aux, double omega, const int64_t* dst,
                             int64_t p_stride, const uint8_t* shell_mask, double* out, int nthreads) {
    return guarded([&] {
        if (!(omega > 0.0)) throw std::runtime_error("omega must be positive");
        p4a::three_index(p4a::basis_from_desc(prim), p4a::basis_from_desc(aux), dst, p_stride, shell_mask, out,
                         nthreads, omega);
    });
}

int p4a_ints_four_index_erf(const p4a_basis* b, double omega, double* out, int nthreads) {
    return guarded([&] {
        if (!(omega > 0.0)) throw std::runtime_error("omega must be positive");
        p4a::four_index(p4a::basis_from_desc(b), out, nthreads, omega);
    });
}

int p4a_ints_schwarz(const p4a_basis* prim, double* fun_max, double* shell_max, double* max_val, int nthreads) {
    return guarded(
        [&] { *max_val = p4a::schwarz_diagonal(p4a::basis_from_desc(prim), fun_max, shell_max, nthreads); });
}

int p4a_ints_four_index(const p4a_basis* b, double* out, int nthreads) {
    return guarded([&] { p4a::four_index(p4a::basis_from_desc(b), out, nthreads); });
}

int p4a_ints_boys(int nmax, double T, double* F) {
    return guarded([&] {
        if (nmax < 0 || nmax > p4a::kMaxTotalL) throw std::runtime_error("boys: nmax out of range");
        p4a::boys(nmax, T, F);
    });
}

}  // extern "C"
