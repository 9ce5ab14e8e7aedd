// Argument checks and request decoding shared by the C ABI entry points.
#pragma once

#include <stdint.h>
#include <string.h>

#include "../../include/zynum_gpu.h"

namespace zg {

inline int32_t max1_i32(int32_t x) { return x > 1 ? x : 1; }

// fortran.zig:88-103 (trans already upper-cased: N, T or C)
inline int gemm_error(char ta, char tb, int32_t m, int32_t n, int32_t k, int32_t lda, int32_t ldb, int32_t ldc) {
    auto valid = [](char c) { return c == 'N' || c == 'T' || c == 'C'; };
    if (!valid(ta)) return 1;
    if (!valid(tb)) return 2;
    if (m < 0) return 3;
    if (n < 0) return 4;
    if (k < 0) return 5;
    const int32_t nrowa = (ta == 'N') ? m : k;
    const int32_t nrowb = (tb == 'N') ? k : n;
    if (lda < max1_i32(nrowa)) return 8;
    if (ldb < max1_i32(nrowb)) return 10;
    if (ldc < max1_i32(m)) return 13;
    return 0;
}

// Request transpose code to the Fortran character ('X' for an invalid code);
// a real conjugate transpose is the transpose.
inline char request_trans_char(uint8_t t, bool complex_scalar) {
    return t == ZYNUM_GPU_NO_TRANS ? 'N'
           : t == ZYNUM_GPU_TRANS  ? 'T'
           : t == ZYNUM_GPU_CONJ_TRANS ? (complex_scalar ? 'C' : 'T')
                                       : 'X';
}

// Scalar from the request's 64-bit pattern (a float sits in the low 32 bits).
template <typename T>
inline T scalar_from_bits(uint64_t bits);
template <>
inline double scalar_from_bits<double>(uint64_t bits) {
    double v;
    memcpy(&v, &bits, sizeof(v));
    return v;
}
template <>
inline float scalar_from_bits<float>(uint64_t bits) {
    const uint32_t lo = (uint32_t)bits;
    float v;
    memcpy(&v, &lo, sizeof(v));
    return v;
}

}  // namespace zg
