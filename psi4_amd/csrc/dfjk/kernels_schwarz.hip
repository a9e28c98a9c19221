// Schwarz diagonal (mn|mn) on the device: the classes of s, p and d shell pairs and the
// dispatch (device code in kernels_schwarz_impl.hpp; pairs with an f shell in
// kernels_schwarz_f.hip).
#include "kernels_schwarz_impl.hpp"

namespace p4a {

void launch_schwarz_f(const ThreeIndexArgs& a, int lb, const int2* pairs, int npair, int nfa, int nfb,
                      double* fun_max, hipStream_t s);

void launch_schwarz_class(const ThreeIndexArgs& a, int la, int lb, const int2* pairs, int npair, double* fun_max,
                          hipStream_t s) {
    if (la < lb) throw std::runtime_error("schwarz: pair class must have la >= lb");
    if (npair == 0) return;
    const int nfa = (a.prim_pure && la > 0) ? 2 * la + 1 : ncart(la);
    const int nfb = (a.prim_pure && lb > 0) ? 2 * lb + 1 : ncart(lb);
    switch (la * 4 + lb) {
        case 0: launch<0, 0>(a, pairs, npair, nfa, nfb, fun_max, s); break;
        case 4: launch<1, 0>(a, pairs, npair, nfa, nfb, fun_max, s); break;
        case 5: launch<1, 1>(a, pairs, npair, nfa, nfb, fun_max, s); break;
        case 8: launch<2, 0>(a, pairs, npair, nfa, nfb, fun_max, s); break;
        case 9: launch<2, 1>(a, pairs, npair, nfa, nfb, fun_max, s); break;
        case 10: launch<2, 2>(a, pairs, npair, nfa, nfb, fun_max, s); break;
        case 12:
        case 13:
        case 14:
        case 15: launch_schwarz_f(a, lb, pairs, npair, nfa, nfb, fun_max, s); break;
        default: throw std::runtime_error("schwarz: primary angular momentum above 3");
    }
}

}  // namespace p4a
