// Build a normalised p4a::Basis from the flat C-ABI description.
#pragma once
#include <stdexcept>

#include "ints.hpp"
#include "psi4amd_ints.h"

namespace p4a {

inline Basis basis_from_desc(const p4a_basis* d) {
    if (!d || d->nshell < 0) throw std::runtime_error("basis description is null");
    Basis b;
    b.shells.resize(d->nshell);
    size_t off = 0;
    for (int i = 0; i < d->nshell; ++i) {
        Shell& s = b.shells[i];
        s.l = d->l[i];
        if (s.l < 0 || s.l > kMaxShellL) throw std::runtime_error("shell angular momentum out of range");
        if (d->nprim[i] <= 0) throw std::runtime_error("shell with no primitives");
        s.pure = d->puream != 0;
        for (int k = 0; k < 3; ++k) s.A[k] = d->center[3 * i + k];
        s.exps.assign(d->exps + off, d->exps + off + d->nprim[i]);
        s.coefs.assign(d->coefs + off, d->coefs + off + d->nprim[i]);
        off += d->nprim[i];
        normalize_shell(s);
    }
    b.finalize();
    return b;
}

}  // namespace p4a
