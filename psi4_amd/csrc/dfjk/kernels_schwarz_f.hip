// Schwarz diagonal (mn|mn) on the device: the classes (f l | f l) of pairs with an f shell
// (device code in kernels_schwarz_impl.hpp).
#include "kernels_schwarz_impl.hpp"

namespace p4a {

void launch_schwarz_f(const ThreeIndexArgs& a, int lb, const int2* pairs, int npair, int nfa, int nfb,
                      double* fun_max, hipStream_t s) {
    switch (lb) {
        case 0: launch<3, 0>(a, pairs, npair, nfa, nfb, fun_max, s); break;
        case 1: launch<3, 1>(a, pairs, npair, nfa, nfb, fun_max, s); break;
        case 2: launch<3, 2>(a, pairs, npair, nfa, nfb, fun_max, s); break;
        case 3: launch<3, 3>(a, pairs, npair, nfa, nfb, fun_max, s); break;
        default: throw std::runtime_error("schwarz: primary angular momentum above 3");
    }
}

}  // namespace p4a
