// oz_aie.h -- C++ drop-in for the reference API (include/oz_aie.h of atlarge-research/ozAIE).
//
// A consumer of the reference (namespace oz_aie, enum Mode, ozAIE_setup, ozAIE_DGEMM)
// compiles unchanged against this header and links against libozaie.so.  Unlike the
// reference header, every function here is inline, so including it from several
// translation units is fine.
#pragma once

#include <cstdint>

#include "ozaie.h"

namespace oz_aie {

// include/oz_aie.h:82 of the reference (accurate, fast), plus this library's b = 7 presets
// and its CRT mode (16 int8 residue GEMMs, exact integer product of the scaled operands;
// no reference counterpart).
enum Mode { accurate = 0, fast = 1, fast3 = 2, accurate7 = 3, crt = 4 };

// Replaces include/oz_aie.h:182 of the reference.  Same status codes: 0 OK, 1 rejected
// dimensions (only with OZAIE_OPT_STRICT, otherwise any M, N, K >= 1 is accepted); the
// reference's 2 (environment) and 4 (build) cannot occur; 5+ are HIP-side failures.
inline int32_t ozAIE_setup(int M, int N, int K, enum Mode mode) {
    return ::ozAIE_setup(M, N, K, static_cast<int>(mode));
}

// Replaces include/oz_aie.h:311 of the reference: C = A * B, row-major, blocking.
inline int32_t ozAIE_DGEMM(int M, int N, int K, const double *A, const double *B, double *C) {
    return ::ozAIE_DGEMM(M, N, K, A, B, C);
}

// Replaces cleanup_xrt (include/oz_aie.h:153 of the reference).
inline void cleanup_xrt() { (void)::ozaie_release(); }

}  // namespace oz_aie
