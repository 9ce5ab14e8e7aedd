// Split-precision (Ozaki) DGEMM: shared types of the INT8 GEMM core and the
// CRT reconstruction.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace zg {

constexpr int kMaxModuli = 20;
constexpr int kMaxDigits = 5;

// Chinese-remainder constants for the first `nmod` moduli.  M_l = P / m_l and
// P are split into 40-bit digits: M_l = sum_d h[l][d] 2^dig_exp[d].
struct CrtConsts {
    int nmod, ndig, alpha;
    int m[kMaxModuli];
    int y[kMaxModuli];  // (P / m_l)^-1 mod m_l
    double inv_m[kMaxModuli];
    double h[kMaxModuli][kMaxDigits];
    double p[kMaxDigits];
    double inv_p_dig[kMaxDigits];  // 2^dig_exp[d] / P
    int dig_exp[kMaxDigits];
    double log2p;
};

// Fills `cc` for `nmod` moduli and inner dimension k; false if unsupported.
bool crt_setup(int nmod, int64_t k, CrtConsts& cc);

struct I8Pair {
    int a, b;  // plane indices of op(A) and op(B)
};

struct I8Out {
    int npairs;
    I8Pair pairs[8];
    double scale;   // FP64-accumulate mode: weight of this output
    int modulus;    // residue mode: m_l
    int y;          // residue mode: (P / m_l)^-1 mod m_l
    double inv_m;
};

// Batched INT8 (or BF16) GEMM over K-contiguous planes [rows][kp], kp = bytes per row (a multiple of 128).
// Output z (grid.z) sums the products of its plane pairs in INT32, then
//   mode 0: stores u = (y * acc) mod m as one byte at u + z*u_plane + col*ldu + row
//   mode 1: w[col*ldw + row] (+)= acc * scale
struct I8Args {
    int m, n, kp;
    const int8_t* a;
    int64_t a_plane;
    const int8_t* b;
    int64_t b_plane;
    int tiles_m, tiles_n;
    int mode;
    uint8_t* u;
    int64_t u_plane;
    int ldu;
    double* w;
    int64_t ldw;
    int accumulate;
    I8Out outs[kMaxModuli];
};

// tile configurations 0..kI8Tiles-1 of i8_gemm_launch (0..5: 16x16x64 MFMA, 6..9: 32x32x32 MFMA,
// 10..13: 32x32x32 MFMA fed by direct global -> LDS loads, 14..17: unpadded XOR-swizzled LDS rows,
// 18..21 and 25: 16 waves of 64x64, 22..24: 8 waves of 64x64, two workgroups per CU)
constexpr int kI8Tiles = 26;
hipError_t i8_gemm_launch(const I8Args& args, int nout, hipStream_t s, int tile);
// tile configurations 0..kBf16Tiles-1 of the BF16 slice GEMM (same I8Args, bf16 planes, kp in bytes)
constexpr int kBf16Tiles = 4;
hipError_t bf16_gemm_launch(const I8Args& args, int nout, hipStream_t s, int tile);
int i8_pick_tile(int m, int n, int kp);



}  // namespace zg
