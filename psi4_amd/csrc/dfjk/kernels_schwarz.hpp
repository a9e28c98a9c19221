// Schwarz diagonal (mn|mn) on the device (see kernels_schwarz.hip).
#pragma once
#include "kernels_ints.hpp"

namespace p4a {

// One (la lb | la lb) class, la >= lb <= 3: pairs[i] = (shell with la, shell with lb) of the
// primary basis in a.prim.  Writes |(mn|mn)| to fun_max[m * a.nbf + n] and fun_max[n * a.nbf + m]
// for every function pair of every listed shell pair (within a pair of a shell with itself, the
// pairs m >= n and their mirrors).  Uses a.prim, a.pexp, a.pcoef, a.c2s, a.prim_pure, a.boys, a.nbf.
void launch_schwarz_class(const ThreeIndexArgs& a, int la, int lb, const int2* pairs, int npair, double* fun_max,
                          hipStream_t s);

}  // namespace p4a
