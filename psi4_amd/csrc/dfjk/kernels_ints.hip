// Three-centre Coulomb integrals (P|mn) on the device: dispatch over the auxiliary angular
// momentum to the per-class translation units (kernels_ints_l<LC>.hip, device code in
// kernels_ints_impl.hpp).
#include "kernels_ints.hpp"

namespace p4a {

void launch_three_index_lc0(const ThreeIndexArgs&, int, int, const int2*, int, const int2*, int, hipStream_t);
void launch_three_index_lc1(const ThreeIndexArgs&, int, int, const int2*, int, const int2*, int, hipStream_t);
void launch_three_index_lc2(const ThreeIndexArgs&, int, int, const int2*, int, const int2*, int, hipStream_t);
void launch_three_index_lc3(const ThreeIndexArgs&, int, int, const int2*, int, const int2*, int, hipStream_t);
void launch_three_index_lc4(const ThreeIndexArgs&, int, int, const int2*, int, const int2*, int, hipStream_t);

void launch_three_index_class(const ThreeIndexArgs& a, int la, int lb, int lc, const int2* pairs, int npair,
                              const int2* afun, int nafun, hipStream_t s) {
    if (la < lb) throw std::runtime_error("three_index: pair class must have la >= lb");
    if (npair == 0 || nafun == 0) return;
    if ((nafun + 63) / 64 > 65535) throw std::runtime_error("three_index: too many auxiliary functions in a class");
    switch (lc) {
        case 0: launch_three_index_lc0(a, la, lb, pairs, npair, afun, nafun, s); break;
        case 1: launch_three_index_lc1(a, la, lb, pairs, npair, afun, nafun, s); break;
        case 2: launch_three_index_lc2(a, la, lb, pairs, npair, afun, nafun, s); break;
        case 3: launch_three_index_lc3(a, la, lb, pairs, npair, afun, nafun, s); break;
        case 4: launch_three_index_lc4(a, la, lb, pairs, npair, afun, nafun, s); break;
        default: throw std::runtime_error("three_index: auxiliary angular momentum above 4");
    }
}

}  // namespace p4a
