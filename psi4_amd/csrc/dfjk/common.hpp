// Shared device-side definitions for the DF-JK kernels (gfx950 / CDNA4).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>

namespace p4a {

// B is stored as TILE x TILE blocks of the upper block-triangle of each B_Q.
constexpr int TILE = 32;
constexpr int TILE2 = TILE * TILE;

typedef double d4_t __attribute__((ext_vector_type(4)));
typedef float f4_t __attribute__((ext_vector_type(4)));
typedef __bf16 bf8_t __attribute__((ext_vector_type(8)));

inline void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

// Launch-shape guard: every launcher checks the shapes it derives its grid
// from before the kernel runs (a mismatch is a host error, never a device fault).
inline void require(bool ok, const char* what) {
    if (!ok) throw std::runtime_error(std::string("shape check failed: ") + what);
}

}  // namespace p4a
