// Three-centre integral classes (la, lb | 4) for every la >= lb <= 3.
#include "kernels_ints_impl.hpp"

namespace p4a {

void launch_three_index_lc4(const ThreeIndexArgs& a, int la, int lb, const int2* pairs, int npair, const int2* afun,
                              int nafun, hipStream_t s) {
    switch (la * 4 + lb) {
        case 0: launch_class<0, 0, 4>(a, pairs, npair, afun, nafun, s); break;
        case 4: launch_class<1, 0, 4>(a, pairs, npair, afun, nafun, s); break;
        case 5: launch_class<1, 1, 4>(a, pairs, npair, afun, nafun, s); break;
        case 8: launch_class<2, 0, 4>(a, pairs, npair, afun, nafun, s); break;
        case 9: launch_class<2, 1, 4>(a, pairs, npair, afun, nafun, s); break;
        case 10: launch_class<2, 2, 4>(a, pairs, npair, afun, nafun, s); break;
        case 12: launch_class<3, 0, 4>(a, pairs, npair, afun, nafun, s); break;
        case 13: launch_class<3, 1, 4>(a, pairs, npair, afun, nafun, s); break;
        case 14: launch_class<3, 2, 4>(a, pairs, npair, afun, nafun, s); break;
        case 15: launch_class<3, 3, 4>(a, pairs, npair, afun, nafun, s); break;
        default: throw std::runtime_error("three_index: primary angular momentum above 3");
    }
}

}  // namespace p4a
