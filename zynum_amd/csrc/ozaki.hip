// Split-precision DGEMM on gfx950 INT8 matrix cores.
//
// Two forms share one batched INT8 MFMA GEMM (v_mfma_i32_16x16x64_i8, exact
// INT32 accumulation) over K-contiguous int8 "planes" of op(A) and op(B):
//
//  * Ozaki scheme II (mode OZAKI_CRT): every row of op(A) / column of op(B) is
//    scaled by a power of two so its 2-norm is at most 2^alpha, rounded to an
//    integer (a', b'), and reduced modulo N pairwise-coprime moduli m_l <= 256.
//    The N residue products A'_l B'_l are exact in INT32; their epilogue
//    stores u_l = (y_l * (A'_l B'_l mod m_l)) mod m_l as one byte, with
//    y_l = (P / m_l)^-1 mod m_l.  The reconstruction kernel rebuilds
//    C' = A'B' exactly from the u_l by the Chinese remainder theorem in FP64
//    digits of 40 bits (every partial sum an exact integer), then applies the
//    inverse scaling and the alpha/beta epilogue.  By Cauchy-Schwarz
//    |C'_ij| <= ||a'_i|| ||b'_j|| < P/2, so C' is recovered exactly and the
//    only error is the rounding of a and b to 2^-alpha of their row/column
//    norms.  (Ozaki, Uchino, Imamura: "Ozaki Scheme II: a GEMM-oriented
//    emulation of floating-point matrix multiplication using an integer
//    modular technique", 2025.)
//  * Ozaki scheme I (mode OZAKI_INT8): rows/columns are scaled by their max
//    exponent and split into s signed slices of 7 bits (round to nearest);
//    products of slice pairs with i + j <= s + 1 are grouped by level
//    (one INT32 GEMM over the concatenated pairs per level) and accumulated
//    level by level into an FP64 buffer.  (Ozaki, Ogita, Oishi, Rump 2012;
//    Ootomo, Ozaki, Yokota 2024 for the INT8 form.)
//
// FP64 is used only for the scaling exponents, the residue / slice
// extraction, the CRT reconstruction and the final level sums.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "common.hpp"
#include "gemm.hpp"
#include "ops.hpp"
#include "ozaki.hpp"
#include "runtime.hpp"

namespace zg {

// Pairwise-coprime moduli, largest first (256 = 2^8, 255 = 3*5*17,
// 253 = 11*23, 247 = 13*19, 217 = 7*31, the rest prime).
static const int kModuli[kMaxModuli] = {256, 255, 253, 251, 247, 241, 239, 233, 229, 227,
                                        223, 217, 211, 199, 197, 193, 191, 181, 179, 173};

// ---------------------------------------------------------------------------
// host bignum helpers (little-endian base 2^32), enough for P < 2^170
// ---------------------------------------------------------------------------
namespace {
struct Big {
    uint32_t w[8] = {0, 0, 0, 0, 0, 0, 0, 0};
};
void big_mul_small(Big& x, uint32_t s) {
    uint64_t carry = 0;
    for (int i = 0; i < 8; ++i) {
        uint64_t v = (uint64_t)x.w[i] * s + carry;
        x.w[i] = (uint32_t)v;
        carry = v >> 32;
    }
}
uint32_t big_div_small(Big& x, uint32_t d) {  // x /= d, returns remainder
    uint64_t rem = 0;
    for (int i = 7; i >= 0; --i) {
        uint64_t cur = (rem << 32) | x.w[i];
        x.w[i] = (uint32_t)(cur / d);
        rem = cur % d;
    }
    return (uint32_t)rem;
}
int big_bits(const Big& x) {
    for (int i = 7; i >= 0; --i)
        if (x.w[i]) return 32 * i + 32 - __builtin_clz(x.w[i]);
    return 0;
}
// bits [lo, lo + 40) of x as an exact double
double big_digit(const Big& x, int lo) {
    uint64_t v = 0;
    for (int b = 39; b >= 0; --b) {
        int bit = lo + b;
        uint32_t word = (bit >= 0 && bit < 256) ? x.w[bit / 32] : 0;
        v = (v << 1) | ((bit >= 0 && bit < 256) ? ((word >> (bit % 32)) & 1u) : 0u);
    }
    return (double)v;
}
int inv_mod(int a, int m) {  // a^-1 mod m (a, m coprime)
    int t = 0, nt = 1, r = m, nr = a % m;
    while (nr) {
        int q = r / nr;
        int tmp = t - q * nt; t = nt; nt = tmp;
        tmp = r - q * nr; r = nr; nr = tmp;
    }
    return t < 0 ? t + m : t;
}
}  // namespace

bool crt_setup(int nmod, int64_t k, CrtConsts& cc) {
    if (nmod < 2 || nmod > kMaxModuli) return false;
    memset(&cc, 0, sizeof(cc));
    Big P;
    P.w[0] = 1;
    for (int l = 0; l < nmod; ++l) big_mul_small(P, (uint32_t)kModuli[l]);
    const int pbits = big_bits(P);
    // digit grid: digits of 40 bits, the top digit at bit `top`
    int nd = (pbits + 8 + 39) / 40;  // X < nmod * P < 2^(pbits + 5)
    if (nd > kMaxDigits) return false;
    cc.nmod = nmod;
    cc.ndig = nd;
    for (int d = 0; d < nd; ++d) cc.dig_exp[d] = 40 * (nd - 1 - d);
    for (int l = 0; l < nmod; ++l) {
        Big Ml = P;
        big_div_small(Ml, (uint32_t)kModuli[l]);
        Big tmp = Ml;
        const uint32_t ml_mod = big_div_small(tmp, (uint32_t)kModuli[l]);
        cc.m[l] = kModuli[l];
        cc.y[l] = inv_mod((int)ml_mod, kModuli[l]);
        cc.inv_m[l] = 1.0 / kModuli[l];
        for (int d = 0; d < nd; ++d) cc.h[l][d] = big_digit(Ml, cc.dig_exp[d]);
    }
    for (int d = 0; d < nd; ++d) cc.p[d] = big_digit(P, cc.dig_exp[d]);
    // 1/P scaled per digit: q = round(sum_d S_d 2^e_d / P)
    double p_approx = 0.0;
    for (int d = 0; d < nd; ++d) p_approx += ldexp(cc.p[d], cc.dig_exp[d]);
    for (int d = 0; d < nd; ++d) cc.inv_p_dig[d] = ldexp(1.0, cc.dig_exp[d]) / p_approx;
    // Largest alpha with (2^alpha + sqrt(k)/2)^2 <= P/2 * (1 - 2^-30):
    // |C'| <= ||a'|| ||b'|| stays clear of P/2 so q = round(X/P) is exact.
    const double log2p = log2(p_approx);
    int alpha = (int)floor((log2p - 1.0) / 2.0);
    const double half_sqrt_k = 0.5 * sqrt((double)(k > 0 ? k : 1));
    while (alpha > 0) {
        const double lhs = 2.0 * log2(ldexp(1.0, alpha) + half_sqrt_k);
        if (lhs <= log2p - 1.0 - 1e-6) break;
        --alpha;
    }
    cc.alpha = alpha;
    cc.log2p = log2p;
    return alpha >= 8;
}

// ---------------------------------------------------------------------------
// operand statistics: per-row (of op(X)) max |x| and scaled sum of squares
// element (r, k) of op(X) is x[k + r*ld] when kcontig, x[r + k*ld] otherwise
// ---------------------------------------------------------------------------
__global__ void fill2_u64_kernel(uint64_t* p, int64_t n, uint64_t* q, int64_t nq, uint64_t v) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n + nq; i += (int64_t)gridDim.x * blockDim.x) {
        if (i < n) p[i] = v;
        else q[i - n] = v;
    }
}


// One pass over op(X): per-row max |x| (atomicMax on the bit pattern, which
// orders like the value for non-negative doubles and puts NaN above +inf) and
// the sum of squares of each 256-deep k chunk, stored per chunk so that the
// row sum is formed in a fixed order (run-to-run reproducible scaling).
// Strided rows: one thread per row and chunk.  Contiguous rows: one wave per
// row, lanes over k.
constexpr int kStatChunk = 256;
// Smallest N of a CRT product that runs in parts on two streams, and the
// number of parts.
constexpr int kCrtSplitMinN = 4096;
#ifndef ZG_CRT_PARTS
#define ZG_CRT_PARTS 2
#endif
constexpr int kCrtParts = ZG_CRT_PARTS;
// sigma of a row or column holding a NaN or an infinity: its results are NaN
constexpr int kNonFinite = 0x7fffffff;

// One operand of a statistics (or plane) launch: workgroups [first, first +
// blocks) of the grid take its rows.
struct StatOp {
    const double* x;
    int64_t ld;
    int rows, kcontig;
    uint64_t* maxbits;
    double* part;
    int blocks;
};

template <bool KCONTIG>
__device__ __forceinline__ void row_stats_tile(const StatOp& o, int bx, int k) {
    const double* __restrict__ x = o.x;
    const int64_t ld = o.ld;
    const int rows = o.rows;
    uint64_t* __restrict__ maxbits = o.maxbits;
    const int k0 = blockIdx.y * kStatChunk, k1 = min(k, k0 + kStatChunk);
    double* ps = o.part + (int64_t)blockIdx.y * rows;  // this chunk's partial sums
    if constexpr (!KCONTIG) {
        const int r = bx * 256 + threadIdx.x;
        if (r >= rows) return;
        double mx = 0.0, ss = 0.0;
        bool nan = false;  // fmax drops NaN operands: flag them separately
        const double* px = x + r + (int64_t)k0 * ld;
        int kk = k0;
        for (; kk + 4 <= k1; kk += 4, px += 4 * ld) {
            const double v0 = px[0], v1 = px[ld], v2 = px[2 * ld], v3 = px[3 * ld];
            mx = fmax(fmax(mx, fabs(v0)), fmax(fabs(v1), fmax(fabs(v2), fabs(v3))));
            nan |= (v0 != v0) | (v1 != v1) | (v2 != v2) | (v3 != v3);
            ss = fma(v0, v0, ss); ss = fma(v1, v1, ss); ss = fma(v2, v2, ss); ss = fma(v3, v3, ss);
        }
        for (; kk < k1; ++kk, px += ld) {
            const double v = *px;
            mx = fmax(mx, fabs(v));
            nan |= (v != v);
            ss = fma(v, v, ss);
        }
        if (nan) mx = __longlong_as_double(0x7ff8000000000000ll);
        atomicMax((unsigned long long*)&maxbits[r], (unsigned long long)__double_as_longlong(mx));
        ps[r] = ss;
    } else {
        // one wave per row, lanes over k; four rows at a time so that their
        // 16 loads per lane are in flight together (indices clamped into the
        // chunk and the row range, the extra values masked)
        const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
        constexpr int kRowsTogether = 4, kPer = kStatChunk / 64;
        for (int rr0 = wave * kRowsTogether; rr0 < 64; rr0 += 4 * kRowsTogether) {
            double v[kRowsTogether][kPer];
#pragma unroll
            for (int q = 0; q < kRowsTogether; ++q) {
                const int r = min(bx * 64 + rr0 + q, rows - 1);
#pragma unroll
                for (int t = 0; t < kPer; ++t) v[q][t] = x[min(k0 + lane + 64 * t, k1 - 1) + (int64_t)r * ld];
            }
#pragma unroll
            for (int q = 0; q < kRowsTogether; ++q) {
                const int r = bx * 64 + rr0 + q;
                double mx = 0.0, ss = 0.0;
                bool nan = false;
#pragma unroll
                for (int t = 0; t < kPer; ++t) {
                    if (k0 + lane + 64 * t < k1) {
                        const double w = v[q][t];
                        mx = fmax(mx, fabs(w));
                        nan |= (w != w);
                        ss = fma(w, w, ss);
                    }
                }
                for (int off = 32; off > 0; off >>= 1) {
                    mx = fmax(mx, __shfl_xor(mx, off));
                    nan |= (bool)__shfl_xor((int)nan, off);
                    ss += __shfl_xor(ss, off);
                }
                if (nan) mx = __longlong_as_double(0x7ff8000000000000ll);
                if (lane == 0 && r < rows) {
                    atomicMax((unsigned long long*)&maxbits[r], (unsigned long long)__double_as_longlong(mx));
                    ps[r] = ss;
                }
            }
        }
    }
}

// Both operands in one launch (their rows are independent): workgroups
// [0, a.blocks) take op(A)'s rows, the rest op(B)'s.
__global__ __launch_bounds__(256) void row_stats_kernel(StatOp a, StatOp b, int k) {
    const bool first = (int)blockIdx.x < a.blocks;
    const StatOp& o = first ? a : b;
    const int bx = first ? (int)blockIdx.x : (int)blockIdx.x - a.blocks;
    if (o.kcontig) row_stats_tile<true>(o, bx, k);
    else row_stats_tile<false>(o, bx, k);
}

// Per-row scaling exponent sigma_r.
//   CRT:    ||x_r 2^sigma||_2 <= 2^alpha.  The norm comes from the sum of
//           squares when that is trustworthy (finite, and the max square is a
//           normal number so no significant term underflowed); otherwise from
//           the bound sqrt(k) max |x|.
//   slices: max |x_r 2^sigma| < 1
//   rows holding a NaN or an infinity: kNonFinite
// Rows [0, a.rows) of the launch are op(A)'s, the next b.rows op(B)'s.
struct ExpOp {
    int rows;
    const uint64_t* maxbits;
    const double* part;
    int* sigma;
};

__global__ void row_exponent_kernel(ExpOp a, ExpOp b, int k, int nchunks, int alpha, int crt) {
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    const bool first = r < a.rows;
    const ExpOp& o = first ? a : b;
    if (!first) r -= a.rows;
    if (r >= o.rows) return;
    const double mx = __longlong_as_double((long long)o.maxbits[r]);
    if (!isfinite(mx)) {
        o.sigma[r] = kNonFinite;
        return;
    }
    if (mx == 0.0) {
        o.sigma[r] = 0;
        return;
    }
    if (crt) {
        // chunk sums added in chunk order (fixed, so the scaling is
        // reproducible); the loads are issued ahead of the additions
        double ss = 0.0;
#pragma unroll 8
        for (int c = 0; c < nchunks; ++c) ss += o.part[(int64_t)c * o.rows + r];
        const double m2 = mx * mx;
        const bool ok = isfinite(ss) && isfinite(m2) && m2 >= 0x1p-1000 && ss >= m2;
        // relative rounding of a k-term sum is below k 2^-53 <= 2^-36 for k <= 2^17
        const double nrm = ok ? sqrt(ss) * (1.0 + 0x1p-30) : sqrt((double)k) * mx * (1.0 + 0x1p-30);
        int ex;
        const double fr = frexp(nrm, &ex);  // nrm = fr 2^ex, fr in [0.5, 1)
        (void)fr;
        o.sigma[r] = alpha - ex;  // ||x|| 2^sigma < 2^ex 2^(alpha - ex) = 2^alpha
    } else {
        o.sigma[r] = -(ilogb(mx) + 1);  // |x| 2^sigma < 1
    }
}

// ---------------------------------------------------------------------------
// plane extraction: op(X) (rows x k, FP64) -> nplanes int8 (or bf16) planes
// [rows][kp] (K-contiguous, zero padded to kp).  32 x 128 tiles through LDS so
// the FP64 reads are coalesced whatever the transpose and every plane row
// segment written is 128 contiguous int8 (256 bf16) bytes.
// ---------------------------------------------------------------------------
// One operand of a plane launch: workgroups [first, first + blocks) along x
// take its rows.
struct PlaneOp {
    const double* x;
    int64_t ld;
    int rows, kcontig;
    const int* sigma;
    int8_t* out;
    int64_t plane_stride;  // bytes between planes
    int blocks;
};

struct PlaneArgs {
    PlaneOp op[2];         // op(A), op(B)
    int k, kp;
    int nplanes;
    int crt;               // 1: residues modulo m[l]; 0: 7-bit slices
    int bf16;              // slices stored as bf16 (2 bytes, exact: |s| <= 64) instead of int8
    int m[kMaxModuli];
    double inv_m[kMaxModuli];
    double c32[kMaxModuli];  // 2^32 mod m
    double c64[kMaxModuli];  // 2^64 mod m
};

// Rows of op(X) per block of the plane kernel; each of the 256 threads then
// converts EPT = TR * 128 / 256 consecutive k of one row.
#ifndef ZG_PLANE_ROWS
#define ZG_PLANE_ROWS 16
#endif

// Column of element k of a plane tile row: one pad double after every 8, so
// that the 8-element runs the threads read, and the columns the loads write,
// fall in distinct LDS banks (a row of 128 + 16 + 1 doubles).
__device__ __forceinline__ int plane_col(int kk) { return kk + (kk >> 3); }

// Residues of the CRT form.  a = x 2^sigma rounded is an exact integer with
// |a| <= 2^(alpha+1) < 2^80.  It is split once per element into parts of at
// most 32 bits, a = a2 2^64 + a1 2^32 + a0, each part with the sign of what it
// is taken from, so every difference is exact (PARTS = 2 when |a| < 2^63:
// a2 = 0).  The residue of a modulo m is then that of
// t = a2 (2^64 mod m) + a1 (2^32 mod m) + a0, an exact integer below 2^41.
// For an odd m one rounded quotient gives the balanced residue
// r = t - m rint(t inv_m) in [-(m-1)/2, (m-1)/2]: the computed t inv_m is
// within 2^-11 of t / m, which, m being odd, is at least 1 / (2m) >= 2^-9 away
// from every half-integer.  For m = 256 the residue is a0 mod 256 as a signed
// byte (the other parts are multiples of 2^32).
constexpr int kPlaneTK = 128;
template <bool KCONTIG, int PARTS>
__device__ __forceinline__ void plane_tile(const PlaneArgs& p, const PlaneOp& o, int bx,
                                           double (&tile)[ZG_PLANE_ROWS][kPlaneTK + kPlaneTK / 8 + 1]) {
    constexpr int TR = ZG_PLANE_ROWS, TK = kPlaneTK;
    constexpr int EPT = TR * TK / 256;  // elements per thread (8 or 16)
    constexpr int TPR = TK / EPT;       // threads per row
    static_assert((EPT == 8 || EPT == 16) && TPR * EPT == TK, "plane tile: 16 or 32 rows per block");
    const int r0 = bx * TR, k0 = blockIdx.y * TK;
    const int t = threadIdx.x;
    // load: coalesced along the contiguous dimension
#pragma unroll
    for (int i = 0; i < TR * TK / 256; ++i) {
        const int e = t + 256 * i;
        int r, kk;
        if constexpr (KCONTIG) { kk = e & (TK - 1); r = e / TK; }
        else { r = e & (TR - 1); kk = e / TR; }
        const int gr = r0 + r, gk = k0 + kk;
        double v = 0.0;
        if (gr < o.rows && gk < p.k) v = KCONTIG ? o.x[gk + (int64_t)gr * o.ld] : o.x[gr + (int64_t)gk * o.ld];
        tile[r][plane_col(kk)] = v;
    }
    __syncthreads();
    // each thread: one row, EPT consecutive k -> EPT (bf16: 2 EPT) bytes per plane
    const int r = t / TPR, kq = (t % TPR) * EPT;
    const int gr = r0 + r;
    if (gr >= o.rows || k0 + kq >= p.kp) return;
    const int sg = o.sigma[gr];
    const int eb = p.bf16 ? 2 : 1;
    int8_t* dst = o.out + ((int64_t)gr * p.kp + k0 + kq) * eb;
    const double* src = &tile[r][plane_col(kq)];  // EPT contiguous values (kq is a multiple of 8)
    constexpr int W = EPT / 4;  // 32-bit words of int8 output per plane
    if (sg == kNonFinite) {
        // the row's results are set to NaN by the final kernel; its planes are zero
        for (int l = 0; l < p.nplanes; ++l) {
            uint32_t* d = reinterpret_cast<uint32_t*>(dst + (int64_t)l * o.plane_stride);
            for (int w = 0; w < W * eb; ++w) d[w] = 0u;
        }
        return;
    }
    if (p.crt) {
        double a2[EPT], a1[EPT], a0[EPT];
#pragma unroll
        for (int b = 0; b < EPT; ++b) {
            const double a = rint(ldexp(src[(b >> 3) * 9 + (b & 7)], sg));
            double rest = a;
            a2[b] = 0.0;
            if constexpr (PARTS == 3) {
                a2[b] = trunc(a * 0x1p-64);
                rest = fma(-a2[b], 0x1p64, a);
            }
            a1[b] = trunc(rest * 0x1p-32);
            a0[b] = fma(-a1[b], 0x1p32, rest);
        }
        for (int l = 0; l < p.nplanes; ++l) {
            const int ml = p.m[l];
            uint32_t pw[W];
            if (ml == 256) {
#pragma unroll
                for (int w = 0; w < W; ++w) {
                    uint32_t word = 0;
#pragma unroll
                    for (int b = 0; b < 4; ++b) {
                        const double u = fma(-floor(a0[4 * w + b] * 0x1p-8), 256.0, a0[4 * w + b]);  // in [0, 256)
                        word |= ((uint32_t)u & 0xffu) << (8 * b);
                    }
                    pw[w] = word;
                }
            } else {
                const double dm = (double)ml, iml = p.inv_m[l], c32 = p.c32[l], c64 = p.c64[l];
#pragma unroll
                for (int w = 0; w < W; ++w) {
                    uint32_t word = 0;
#pragma unroll
                    for (int b = 0; b < 4; ++b) {
                        double tt = fma(a1[4 * w + b], c32, a0[4 * w + b]);
                        if constexpr (PARTS == 3) tt = fma(a2[4 * w + b], c64, tt);
                        const int res = (int)fma(-rint(tt * iml), dm, tt);
                        word |= ((uint32_t)(uint8_t)(int8_t)res) << (8 * b);
                    }
                    pw[w] = word;
                }
            }
            uint32_t* d = reinterpret_cast<uint32_t*>(dst + (int64_t)l * o.plane_stride);
            if constexpr (W == 4) *reinterpret_cast<uint4*>(d) = uint4{pw[0], pw[1], pw[2], pw[3]};
            else *reinterpret_cast<uint2*>(d) = uint2{pw[0], pw[1]};
        }
    } else {
        // slices: s_1 = rint(x 2^6), then 7 more bits per slice (|s_j| <= 64)
        double rem[EPT];
#pragma unroll
        for (int b = 0; b < EPT; ++b) rem[b] = ldexp(src[(b >> 3) * 9 + (b & 7)], sg);  // |rem| < 1
        for (int l = 0; l < p.nplanes; ++l) {
            const int shift = (l == 0) ? 6 : 6 + 7 * l;
            double v[EPT];
#pragma unroll
            for (int b = 0; b < EPT; ++b) {
                v[b] = rint(ldexp(rem[b], shift));
                rem[b] = fma(-v[b], ldexp(1.0, -shift), rem[b]);  // exact
            }
            uint32_t* d = reinterpret_cast<uint32_t*>(dst + (int64_t)l * o.plane_stride);
            if (p.bf16) {
                // |v| <= 64: the float has at most 7 significant bits, so its top 16 bits are the bf16
#pragma unroll
                for (int w = 0; w < EPT / 2; ++w)
                    d[w] = (__float_as_uint((float)v[2 * w]) >> 16) | (__float_as_uint((float)v[2 * w + 1]) & 0xffff0000u);
            } else {
                uint32_t pw[W];
#pragma unroll
                for (int w = 0; w < W; ++w) {
                    uint32_t word = 0;
#pragma unroll
                    for (int b = 0; b < 4; ++b) word |= ((uint32_t)(uint8_t)(int8_t)(int)v[4 * w + b]) << (8 * b);
                    pw[w] = word;
                }
                if constexpr (W == 4) *reinterpret_cast<uint4*>(d) = uint4{pw[0], pw[1], pw[2], pw[3]};
                else *reinterpret_cast<uint2*>(d) = uint2{pw[0], pw[1]};
            }
        }
    }
}

// Both operands in one launch: workgroups [0, op[0].blocks) along x take
// op(A)'s rows, the rest op(B)'s.
template <int PARTS>
__global__ __launch_bounds__(256) void plane_kernel(PlaneArgs p) {
    __shared__ double tile[ZG_PLANE_ROWS][kPlaneTK + kPlaneTK / 8 + 1];
    const bool first = (int)blockIdx.x < p.op[0].blocks;
    const PlaneOp& o = first ? p.op[0] : p.op[1];
    const int bx = first ? (int)blockIdx.x : (int)blockIdx.x - p.op[0].blocks;
    if (o.kcontig) plane_tile<true, PARTS>(p, o, bx, tile);
    else plane_tile<false, PARTS>(p, o, bx, tile);
}

// ---------------------------------------------------------------------------
// batched INT8 GEMM over K-contiguous planes: D_z = sum_{(a,b) in pairs(z)} A_a B_b^T
// ---------------------------------------------------------------------------
// MFMA shapes: MS = 16 -> v_mfma_i32_16x16x64_i8 (4 results per lane: rows
// 4*(lane/16)+r of column lane%16); MS = 32 -> v_mfma_i32_32x32x32_i8 (16
// results per lane: rows 8*(r/4)+4*(lane/32)+r%4 of column lane%32).  Each
// lane feeds 16 consecutive k of one row of A and of B; the k order inside
// the instruction is the same for both operands, so the sum is the plain dot
// product whatever it is.
template <int MS>
struct I8Mfma;
template <>
struct I8Mfma<16> {
    static constexpr int KD = 64, NR = 4, MSZ = 16;
    static constexpr bool kFloat = false;
    typedef v4i acc_t;
    static __device__ __forceinline__ acc_t zero() { return acc_t{0, 0, 0, 0}; }
    static __device__ __forceinline__ acc_t mma(v4i a, v4i b, acc_t c) {
        return __builtin_amdgcn_mfma_i32_16x16x64_i8(a, b, c, 0, 0, 0);
    }
    // row offset of result group g (4 consecutive rows) for this lane
    static __device__ __forceinline__ int group_row(int lane, int) { return (lane >> 4) * 4; }
};
template <>
struct I8Mfma<32> {
    static constexpr int KD = 32, NR = 16, MSZ = 32;
    static constexpr bool kFloat = false;
    typedef int acc_t __attribute__((ext_vector_type(16)));
    static __device__ __forceinline__ acc_t zero() {
        acc_t z;
#pragma unroll
        for (int r = 0; r < 16; ++r) z[r] = 0;
        return z;
    }
    static __device__ __forceinline__ acc_t mma(v4i a, v4i b, acc_t c) {
        return __builtin_amdgcn_mfma_i32_32x32x32_i8(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int group_row(int lane, int g) { return 8 * g + 4 * (lane >> 5); }
};

// BF16 forms over the same byte layout: 16 bytes per lane are 8 bf16 values,
// so v_mfma_f32_16x16x32_bf16 reads what v_mfma_i32_16x16x64_i8 reads (64
// bytes of k per row) and v_mfma_f32_32x32x16_bf16 what the 32x32x32 INT8
// form reads; the result layouts are the same as the INT8 forms'.
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
template <int MS>
struct Bf16Mfma;
template <>
struct Bf16Mfma<16> {
    static constexpr int KD = 64, NR = 4, MSZ = 16;
    static constexpr bool kFloat = true;
    typedef float acc_t __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ acc_t zero() { return acc_t{0.f, 0.f, 0.f, 0.f}; }
    static __device__ __forceinline__ acc_t mma(v4i a, v4i b, acc_t c) {
        return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(v8bf, a), __builtin_bit_cast(v8bf, b), c,
                                                       0, 0, 0);
    }
    static __device__ __forceinline__ int group_row(int lane, int) { return (lane >> 4) * 4; }
};
template <>
struct Bf16Mfma<32> {
    static constexpr int KD = 32, NR = 16, MSZ = 32;
    static constexpr bool kFloat = true;
    typedef float acc_t __attribute__((ext_vector_type(16)));
    static __device__ __forceinline__ acc_t zero() {
        acc_t z;
#pragma unroll
        for (int r = 0; r < 16; ++r) z[r] = 0.f;
        return z;
    }
    static __device__ __forceinline__ acc_t mma(v4i a, v4i b, acc_t c) {
        return __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(v8bf, a), __builtin_bit_cast(v8bf, b), c,
                                                       0, 0, 0);
    }
    static __device__ __forceinline__ int group_row(int lane, int g) { return 8 * g + 4 * (lane >> 5); }
};

// Epilogue shared by the INT8/BF16 GEMM kernels: the wave's FM x FN blocks of
// MS x MS results start at row r0, column c0; value(i, j, r) is result item r
// of block (i, j) as an exact FP64 integer.  Per block, result group g of this
// lane is rows group_row(g)+0..3 of column lane % MS.
template <class MF, int FM, int FN, class V>
__device__ __forceinline__ void i8_store(const I8Args& p, const I8Out& out, int r0, int c0, int lane, V value) {
    constexpr int MS = MF::MSZ;
    constexpr int NG = MF::NR / 4;
    const int lr = lane & (MS - 1);
    if (p.mode == 0) {
        const int m = out.modulus, y = out.y;
        const double inv_m = out.inv_m;
        uint8_t* u = p.u + (int64_t)blockIdx.z * p.u_plane;
#pragma unroll
        for (int i = 0; i < FM; ++i)
#pragma unroll
            for (int j = 0; j < FN; ++j) {
                const int col = c0 + j * MS + lr;
                if (col >= p.n) continue;
#pragma unroll
                for (int g = 0; g < NG; ++g) {
                    const int row = r0 + i * MS + MF::group_row(lane, g);
                    if (row >= p.m) continue;
                    uint32_t word = 0;
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const double v = (double)y * value(i, j, 4 * g + r);  // |v| < 2^40, exact
                        double q = floor(v * inv_m);
                        double rr = fma(-q, (double)m, v);  // exact small integer
                        if (rr >= m) rr -= m;
                        if (rr < 0) rr += m;
                        word |= ((uint32_t)(int)rr & 0xffu) << (8 * r);
                    }
                    // rows beyond m land in the padded part of the plane (ldu >= m rounded to 4)
                    *reinterpret_cast<uint32_t*>(u + (int64_t)col * p.ldu + row) = word;
                }
            }
    } else {
        const double sc = out.scale;
#pragma unroll
        for (int i = 0; i < FM; ++i)
#pragma unroll
            for (int j = 0; j < FN; ++j) {
                const int col = c0 + j * MS + lr;
                if (col >= p.n) continue;
#pragma unroll
                for (int g = 0; g < NG; ++g) {
                    const int row = r0 + i * MS + MF::group_row(lane, g);
                    double* w = p.w + (int64_t)col * p.ldw + row;
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        if (row + r >= p.m) break;
                        const double v = value(i, j, 4 * g + r) * sc;
                        w[r] = p.accumulate ? w[r] + v : v;
                    }
                }
            }
    }
}

// BK is in bytes (p.kp is the plane row length in bytes).  For the BF16 forms
// every product of two slices is an integer of at most 2^12 in magnitude, so
// FP32 sums of up to 4096 of them are exact; the FP32 accumulators are added
// into FP64 ones every kFlush steps (BK/2 products per step) and cleared.
// Padding (bytes) of each LDS row of the register-staged INT8 kernel; a build knob for tuning.
#ifndef ZG_I8_LDS_PAD
#define ZG_I8_LDS_PAD 16
#endif

// SWZ: rows are unpadded and the 16-byte chunk c of row r is stored at chunk
// c ^ (r % 8) (BK of 128 bytes: 8 chunks per row), so the 16-byte reads of 8
// consecutive rows at one k offset, and the writes of one row, cover all banks.
template <int BM, int BN, int BK, int WM, int WN, int OCC, class MF, bool SWZ = false>
__global__ __launch_bounds__(64 * WM * WN) __attribute__((amdgpu_waves_per_eu(OCC)))
void i8_gemm_kernel(I8Args p) {
    constexpr int MS = MF::MSZ;
    constexpr bool FL = MF::kFloat;
    constexpr int kFlush = 8192 / BK;
    typedef typename MF::acc_t acc_t;
    constexpr int NT = 64 * WM * WN;
    constexpr int RS = SWZ ? BK : BK + ZG_I8_LDS_PAD;  // LDS row (bytes)
    constexpr int CPR = BK / 16;     // 16-byte chunks per row
    static_assert(!SWZ || CPR == 8, "the swizzle is laid out for 128-byte rows");
    // byte offset of 16-byte chunk `c16` (a multiple of 16) of row `row`
    auto at = [](int row, int c16) { return row * RS + (SWZ ? (c16 ^ ((row & 7) << 4)) : c16); };
    constexpr int WTM = BM / WM, WTN = BN / WN;
    constexpr int FM = WTM / MS, FN = WTN / MS;
    constexpr int LA = BM * CPR / NT;  // 16-byte loads per thread (A)
    constexpr int LB = BN * CPR / NT;
    static_assert(LA >= 1 && LB >= 1 && BM * CPR % NT == 0 && BN * CPR % NT == 0, "tile/workgroup mismatch");
    static_assert(BK % 64 == 0 && WTM % MS == 0 && WTN % MS == 0, "tile is a multiple of the MFMA shape");

    extern __shared__ __attribute__((aligned(16))) int8_t lds[];
    // stage buf: A rows at lds + buf*(BM+BN)*RS, B rows right after them
    auto stage_a = [&](int buf) { return lds + buf * (BM + BN) * RS; };
    auto stage_b = [&](int buf) { return lds + buf * (BM + BN) * RS + BM * RS; };

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm0 = (wave % WM) * WTM, wn0 = (wave / WM) * WTN;
    const int total = p.tiles_m * p.tiles_n;
    int tm, tn;
    grouped_tile(xcd_remap(blockIdx.x, total), p.tiles_m, p.tiles_n, 8, tm, tn);
    const int m0 = tm * BM, n0 = tn * BN;
    const I8Out& out = p.outs[blockIdx.z];

    acc_t acc[FM][FN];
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j) acc[i][j] = MF::zero();
    double dacc[FL ? FM : 1][FL ? FN : 1][FL ? MF::NR : 1];
    if constexpr (FL) {
#pragma unroll
        for (int i = 0; i < FM; ++i)
#pragma unroll
            for (int j = 0; j < FN; ++j)
#pragma unroll
                for (int r = 0; r < MF::NR; ++r) dacc[i][j][r] = 0.0;
    }
    auto flush = [&]() {
        if constexpr (FL) {
#pragma unroll
            for (int i = 0; i < FM; ++i)
#pragma unroll
                for (int j = 0; j < FN; ++j) {
#pragma unroll
                    for (int r = 0; r < MF::NR; ++r) dacc[i][j][r] += (double)acc[i][j][r];
                    acc[i][j] = MF::zero();
                }
        }
    };

    const int kslices = p.kp / BK;
    const int nsteps = kslices * out.npairs;
    int4 ra[LA], rb[LB];

    auto load = [&](int s) {
        const int pr = s / kslices, ks = s % kslices;
        const int8_t* a = p.a + (int64_t)out.pairs[pr].a * p.a_plane + (int64_t)ks * BK;
        const int8_t* b = p.b + (int64_t)out.pairs[pr].b * p.b_plane + (int64_t)ks * BK;
#pragma unroll
        for (int i = 0; i < LA; ++i) {
            const int e = tid + NT * i, row = e / CPR, c = (e % CPR) * 16;
            const int gr = m0 + row;
            ra[i] = (gr < p.m) ? *reinterpret_cast<const int4*>(a + (int64_t)gr * p.kp + c) : int4{0, 0, 0, 0};
        }
#pragma unroll
        for (int i = 0; i < LB; ++i) {
            const int e = tid + NT * i, row = e / CPR, c = (e % CPR) * 16;
            const int gr = n0 + row;
            rb[i] = (gr < p.n) ? *reinterpret_cast<const int4*>(b + (int64_t)gr * p.kp + c) : int4{0, 0, 0, 0};
        }
    };
    auto store = [&](int buf) {
#pragma unroll
        for (int i = 0; i < LA; ++i) {
            const int e = tid + NT * i, row = e / CPR, c = (e % CPR) * 16;
            *reinterpret_cast<int4*>(stage_a(buf) + at(row, c)) = ra[i];
        }
#pragma unroll
        for (int i = 0; i < LB; ++i) {
            const int e = tid + NT * i, row = e / CPR, c = (e % CPR) * 16;
            *reinterpret_cast<int4*>(stage_b(buf) + at(row, c)) = rb[i];
        }
    };

    if (nsteps > 0) {
        load(0);
        store(0);
    }
    __syncthreads();
    const int lr = lane & (MS - 1), lk = (lane / MS) * 16;
    for (int s = 0; s < nsteps; ++s) {
        const int cur = s & 1;
        const bool more = s + 1 < nsteps;
        if (more) load(s + 1);
        const int8_t* la = stage_a(cur);
        const int8_t* lb = stage_b(cur);
#pragma unroll
        for (int kk = 0; kk < BK; kk += MF::KD) {
            // B fragments of this k step stay in registers; A fragments are
            // streamed one block row at a time, the next one read before the
            // MFMAs of the current one (the scheduling barriers keep that
            // order, so the LDS latency hides behind FN MFMAs)
            v4i bf[FN];
#pragma unroll
            for (int j = 0; j < FN; ++j) bf[j] = *reinterpret_cast<const v4i*>(&lb[at(wn0 + j * MS + lr, kk + lk)]);
            v4i a_cur = *reinterpret_cast<const v4i*>(&la[at(wm0 + lr, kk + lk)]);
#pragma unroll
            for (int i = 0; i < FM; ++i) {
                v4i a_next = a_cur;
                if (i + 1 < FM) a_next = *reinterpret_cast<const v4i*>(&la[at(wm0 + (i + 1) * MS + lr, kk + lk)]);
                __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                for (int j = 0; j < FN; ++j) acc[i][j] = MF::mma(a_cur, bf[j], acc[i][j]);
                __builtin_amdgcn_sched_barrier(0);
                a_cur = a_next;
            }
        }
        if (more) store(cur ^ 1);
        if constexpr (FL) {
            if ((s + 1) % kFlush == 0) flush();
        }
        __syncthreads();
    }
    flush();
    // value of result item r of block (i, j) as an exact FP64 integer
    auto value = [&](int i, int j, int r) -> double {
        if constexpr (FL) return dacc[i][j][r];
        else return (double)acc[i][j][r];
    };

    i8_store<MF, FM, FN>(p, out, m0 + wm0, n0 + wn0, lane, value);
}


// ---------------------------------------------------------------------------
// The same GEMM with global -> LDS direct loads (global_load_lds_dwordx4) in
// an NS-stage ring, so NS-1 stages of 64 k-bytes are in flight while one is
// multiplied and no registers stage the tiles.
//
// Stage layout: BM rows of A then BN rows of B, 64 bytes each.  A direct load
// writes 64 lanes x 16 bytes = 16 consecutive rows; lane l fills slot l % 4 of
// row l / 4 with chunk (slot ^ sw(row)) of that row, sw(row) = (row >> 3) & 3.
// A 32x32 fragment read (32 lanes: 32 consecutive rows, one chunk) then hits
// 32 distinct 16-byte bank groups of the 512-byte LDS row.
//
// Synchronisation per stage s: wait until this wave's loads of stage s have
// landed (vmcnt), barrier (every wave's have, and every wave is done reading
// stage s-1), refill stage s-1's buffer with stage s+NS-1, then read and
// multiply stage s.  The LDS reads are issued as inline ds_read_b128 so that
// the compiler does not wait for every load in flight before each of them;
// their completion is waited for explicitly (lgkmcnt) before use.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void wait_vmcnt_max(int n) {
    // s_waitcnt vmcnt(n), expcnt and lgkmcnt left alone (gfx9 encoding: vmcnt in bits 3:0 and 15:14)
    switch (n) {
        case 0: __builtin_amdgcn_s_waitcnt(0x0F70); break;
        case 1: __builtin_amdgcn_s_waitcnt(0x0F71); break;
        case 2: __builtin_amdgcn_s_waitcnt(0x0F72); break;
        case 3: __builtin_amdgcn_s_waitcnt(0x0F73); break;
        case 4: __builtin_amdgcn_s_waitcnt(0x0F74); break;
        case 5: __builtin_amdgcn_s_waitcnt(0x0F75); break;
        case 6: __builtin_amdgcn_s_waitcnt(0x0F76); break;
        case 7: __builtin_amdgcn_s_waitcnt(0x0F77); break;
        case 8: __builtin_amdgcn_s_waitcnt(0x0F78); break;
        case 9: __builtin_amdgcn_s_waitcnt(0x0F79); break;
        case 10: __builtin_amdgcn_s_waitcnt(0x0F7A); break;
        case 11: __builtin_amdgcn_s_waitcnt(0x0F7B); break;
        case 12: __builtin_amdgcn_s_waitcnt(0x0F7C); break;
        case 13: __builtin_amdgcn_s_waitcnt(0x0F7D); break;
        case 14: __builtin_amdgcn_s_waitcnt(0x0F7E); break;
        case 15: __builtin_amdgcn_s_waitcnt(0x0F7F); break;
        case 16: __builtin_amdgcn_s_waitcnt(0x4F70); break;
        case 17: __builtin_amdgcn_s_waitcnt(0x4F71); break;
        case 18: __builtin_amdgcn_s_waitcnt(0x4F72); break;
        case 19: __builtin_amdgcn_s_waitcnt(0x4F73); break;
        case 20: __builtin_amdgcn_s_waitcnt(0x4F74); break;
        case 21: __builtin_amdgcn_s_waitcnt(0x4F75); break;
        case 22: __builtin_amdgcn_s_waitcnt(0x4F76); break;
        case 23: __builtin_amdgcn_s_waitcnt(0x4F77); break;
        case 24: __builtin_amdgcn_s_waitcnt(0x4F78); break;
        default: __builtin_amdgcn_s_waitcnt(0x0F70); break;
    }
}

__device__ __forceinline__ v4i lds_read16(uint32_t addr) {
    v4i x;
    asm volatile("ds_read_b128 %0, %1" : "=v"(x) : "v"(addr) : "memory");
    return x;
}

template <int BM, int BN, int WM, int WN, int NS, int OCC, class MF>
__global__ __launch_bounds__(64 * WM * WN) __attribute__((amdgpu_waves_per_eu(OCC)))
void i8_dma_kernel(I8Args p) {
    constexpr int MS = MF::MSZ;
    static_assert(MS == 32, "the swizzle is laid out for 32-row fragment reads");
    constexpr bool FL = MF::kFloat;
    constexpr int BKS = 64;                   // k bytes per stage
    constexpr int kFlush = 8192 / BKS;        // BF16: stages per exact FP32 run (BKS/2 products each)
    typedef typename MF::acc_t acc_t;
    constexpr int NW = WM * WN;
    constexpr int WTM = BM / WM, WTN = BN / WN;
    constexpr int FM = WTM / MS, FN = WTN / MS;
    constexpr int SB = (BM + BN) * BKS;       // bytes per stage
    constexpr int NQ = (BM + BN) / 16;        // direct loads (16 rows each) per stage
    constexpr int PW = NQ / NW;               // per wave
    static_assert(NQ % NW == 0 && (BM / 16) % NW == 0 && WTM % MS == 0 && WTN % MS == 0, "tile/workgroup mismatch");
    static_assert(NS >= 3 && PW * (NS - 1) <= 24, "vmcnt immediates");

    extern __shared__ __attribute__((aligned(16))) int8_t lds[];
    const uint32_t lds0 = (uint32_t)(uintptr_t)(__attribute__((address_space(3))) int8_t*)lds;

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm0 = (wave % WM) * WTM, wn0 = (wave / WM) * WTN;
    const int total = p.tiles_m * p.tiles_n;
    int tm, tn;
    grouped_tile(xcd_remap(blockIdx.x, total), p.tiles_m, p.tiles_n, 8, tm, tn);
    const int m0 = tm * BM, n0 = tn * BN;
    const I8Out& out = p.outs[blockIdx.z];

    // per-lane source offsets of this wave's direct loads: load i of wave w is
    // number q = w + NW i, A rows for i < PA, B rows after (rows past m / n
    // read row m-1 / n-1: they only feed results that are never stored)
    constexpr int PA = (BM / 16) / NW;
    int64_t src[PW];
#pragma unroll
    for (int i = 0; i < PW; ++i) {
        const int q = wave + NW * i;
        const bool qa = i < PA;
        const int rl = (qa ? q : q - BM / 16) * 16 + (lane >> 2);  // row within the A or B part
        const int chunk = (lane & 3) ^ ((rl >> 3) & 3);
        const int gr = qa ? min(m0 + rl, p.m - 1) : min(n0 + rl, p.n - 1);
        src[i] = (int64_t)gr * p.kp + chunk * 16;
    }

    const int kslices = p.kp / BKS;
    const int nsteps = kslices * out.npairs;
    // next stage to issue: plane pair, k slice, ring buffer
    int iss_pr = 0, iss_ks = 0, iss_buf = 0;
    const int8_t* iss_a = p.a + (int64_t)out.pairs[0].a * p.a_plane;
    const int8_t* iss_b = p.b + (int64_t)out.pairs[0].b * p.b_plane;
    auto issue = [&]() {
        int8_t* dst = lds + iss_buf * SB;
        const int64_t ko = (int64_t)iss_ks * BKS;
#pragma unroll
        for (int i = 0; i < PW; ++i) {
            const int q = wave + NW * i;
            const int8_t* g = (i < PA ? iss_a : iss_b) + ko + src[i];
            __builtin_amdgcn_global_load_lds((const void __attribute__((address_space(1)))*)g,
                                             (void __attribute__((address_space(3)))*)(dst + q * 1024), 16, 0, 0);
        }
        iss_buf = (iss_buf + 1 == NS) ? 0 : iss_buf + 1;
        if (++iss_ks == kslices) {
            iss_ks = 0;
            if (++iss_pr < out.npairs) {
                iss_a = p.a + (int64_t)out.pairs[iss_pr].a * p.a_plane;
                iss_b = p.b + (int64_t)out.pairs[iss_pr].b * p.b_plane;
            }
        }
    };

    acc_t acc[FM][FN];
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j) acc[i][j] = MF::zero();
    double dacc[FL ? FM : 1][FL ? FN : 1][FL ? MF::NR : 1];
    if constexpr (FL) {
#pragma unroll
        for (int i = 0; i < FM; ++i)
#pragma unroll
            for (int j = 0; j < FN; ++j)
#pragma unroll
                for (int r = 0; r < MF::NR; ++r) dacc[i][j][r] = 0.0;
    }

    // fragment addresses within a stage (substep t adds chunk 2t)
    const int lr = lane & 31, lc = lane >> 5;
    uint32_t fa_off[FM], fb_off[FN];
#pragma unroll
    for (int i = 0; i < FM; ++i) {
        const int r = wm0 + i * MS + lr;
        fa_off[i] = (uint32_t)(r * BKS);
    }
#pragma unroll
    for (int j = 0; j < FN; ++j) {
        const int r = wn0 + j * MS + lr;
        fb_off[j] = (uint32_t)(BM * BKS + r * BKS);
    }
    // swizzle of row r: (r >> 3) & 3; every fragment row above has the same low 5 bits as lr
    const int sw = (lr >> 3) & 3;

    // Fragments of one stage (both 32-byte substeps).  The reads of stage s+1
    // are issued before the MFMAs of stage s (sched_barrier keeps that order),
    // so the LDS latency and bandwidth overlap the MFMA pipe.
    struct Frag {
        v4i a[2][FM];
        v4i b[2][FN];
    };
    auto read_stage = [&](Frag& f, int buf) {
        const uint32_t base = lds0 + (uint32_t)(buf * SB);
#pragma unroll
        for (int t = 0; t < 2; ++t) {
            const uint32_t co = (uint32_t)((((2 * t + lc) ^ sw)) << 4);
#pragma unroll
            for (int i = 0; i < FM; ++i) f.a[t][i] = lds_read16(base + fa_off[i] + co);
#pragma unroll
            for (int j = 0; j < FN; ++j) f.b[t][j] = lds_read16(base + fb_off[j] + co);
        }
    };
    auto tie = [&](Frag& f) {  // the values are used after the preceding lgkmcnt wait
#pragma unroll
        for (int t = 0; t < 2; ++t) {
#pragma unroll
            for (int i = 0; i < FM; ++i) asm volatile("" : "+v"(f.a[t][i]));
#pragma unroll
            for (int j = 0; j < FN; ++j) asm volatile("" : "+v"(f.b[t][j]));
        }
    };
    auto multiply = [&](const Frag& f) {
#pragma unroll
        for (int t = 0; t < 2; ++t)
#pragma unroll
            for (int i = 0; i < FM; ++i)
#pragma unroll
                for (int j = 0; j < FN; ++j) acc[i][j] = MF::mma(f.a[t][i], f.b[t][j], acc[i][j]);
    };

    // prologue: all NS buffers in flight, stage 0 read
#pragma unroll
    for (int s = 0; s < NS; ++s)
        if (s < nsteps) issue();
    Frag f0, f1;
    if (nsteps > 0) {
        if (nsteps >= NS) wait_vmcnt_max(PW * (NS - 1));
        else wait_vmcnt_max(0);
        __builtin_amdgcn_s_barrier();
        read_stage(f0, 0);
        asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
        tie(f0);
    }
    int rd_buf = (1 == NS) ? 0 : 1;  // buffer of stage s + 1
    auto step = [&](Frag& cur, Frag& nxt, int s) {
        // this wave's loads of stage s+1 have landed (outstanding: stages s+2 .. s+NS-1)
        if (nsteps - 2 - s >= NS - 2) wait_vmcnt_max(PW * (NS - 2));
        else wait_vmcnt_max(0);
        // every wave's stage s+1 is in LDS and every wave has read stage s
        __builtin_amdgcn_s_barrier();
        if (s + NS < nsteps) issue();  // into stage s's buffer
        if (s + 1 < nsteps) read_stage(nxt, rd_buf);
        __builtin_amdgcn_sched_barrier(0);
        multiply(cur);
        __builtin_amdgcn_sched_barrier(0);
        asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
        tie(nxt);
        rd_buf = (rd_buf + 1 == NS) ? 0 : rd_buf + 1;
        if constexpr (FL) {
            if ((s + 1) % kFlush == 0) {
#pragma unroll
                for (int i = 0; i < FM; ++i)
#pragma unroll
                    for (int j = 0; j < FN; ++j) {
#pragma unroll
                        for (int r = 0; r < MF::NR; ++r) dacc[i][j][r] += (double)acc[i][j][r];
                        acc[i][j] = MF::zero();
                    }
            }
        }
    };
    int s = 0;
    for (; s + 1 < nsteps; s += 2) {
        step(f0, f1, s);
        step(f1, f0, s + 1);
    }
    if (s < nsteps) step(f0, f1, s);
    if constexpr (FL) {
#pragma unroll
        for (int i = 0; i < FM; ++i)
#pragma unroll
            for (int j = 0; j < FN; ++j)
#pragma unroll
                for (int r = 0; r < MF::NR; ++r) dacc[i][j][r] += (double)acc[i][j][r];
    }
    auto value = [&](int i, int j, int r) -> double {
        if constexpr (FL) return dacc[i][j][r];
        else return (double)acc[i][j][r];
    };
    i8_store<MF, FM, FN>(p, out, m0 + wm0, n0 + wn0, lane, value);
}

// ---------------------------------------------------------------------------
// CRT reconstruction + inverse scaling + alpha/beta epilogue
// ---------------------------------------------------------------------------
struct CrtArgs {
    int m, n;
    const uint8_t* u;
    int64_t u_plane;
    int ldu;
    const int* sig_a;
    const int* sig_b;
    double alpha, beta;
    double* c;
    int64_t ldc;
    CrtConsts cc;
};

// One thread per kCrtRows consecutive rows of one column: the residue bytes of
// all moduli are loaded first (one 32- or 64-bit load per modulus, all in
// flight; 4 rows per thread, 68 VGPRs and 7 waves per SIMD, measured the same
// as 8 rows with 126 VGPRs and 4 waves), then
// summed into the CRT digits.  Rows m..ldu-1 of a residue plane are padding.
// The products go through LDS so that the block's 2048 rows of C are read (for
// beta) and written with consecutive lanes on consecutive rows.
#ifndef ZG_CRT_ROWS
#define ZG_CRT_ROWS 8
#endif
constexpr int kCrtRows = ZG_CRT_ROWS;  // 4 or 8
static_assert(kCrtRows == 4 || kCrtRows == 8, "rows per reconstruction thread");
constexpr int kCrtThreads = 256;
// ND: number of 40-bit digits (cc.ndig), a template constant so the digit
// loops carry no per-digit selects.
template <int ND>
__global__ __launch_bounds__(kCrtThreads) void crt_reconstruct_kernel(CrtArgs p) {
    __shared__ double prods[kCrtThreads * (kCrtRows + 1)];  // row stride 9: conflict-free 8-byte writes
    const int base = blockIdx.x * kCrtThreads * kCrtRows;
    const int t = threadIdx.x;
    const int i0 = base + t * kCrtRows;
    const int j = blockIdx.y;
    const int sb = p.sig_b[j];
    if (i0 < p.m) {
        const int nm = p.cc.nmod;
        const uint8_t* u = p.u + (int64_t)j * p.ldu + i0;
        uint32_t w[kMaxModuli][kCrtRows / 4];
#pragma unroll
        for (int l = 0; l < kMaxModuli; ++l) {
            if (l < nm) {
                if constexpr (kCrtRows == 8) {
                    const uint2 v = *reinterpret_cast<const uint2*>(u + (int64_t)l * p.u_plane);
                    w[l][0] = v.x;
                    w[l][kCrtRows / 4 - 1] = v.y;
                } else {
                    w[l][0] = *reinterpret_cast<const uint32_t*>(u + (int64_t)l * p.u_plane);
                }
            }
        }
        double S[kCrtRows][ND] = {};
#pragma unroll
        for (int l = 0; l < kMaxModuli; ++l) {
            if (l < nm) {
#pragma unroll
                for (int r = 0; r < kCrtRows; ++r) {
                    const uint32_t word = w[l][r >> 2];
                    const double ul = (double)((word >> (8 * (r & 3))) & 0xffu);
#pragma unroll
                    for (int d = 0; d < ND; ++d) S[r][d] = fma(p.cc.h[l][d], ul, S[r][d]);  // exact: integer < 2^53
                }
            }
        }
#pragma unroll
        for (int r = 0; r < kCrtRows; ++r) {
            const int i = i0 + r;
            double xq = 0.0;
#pragma unroll
            for (int d = 0; d < ND; ++d) xq = fma(S[r][d], p.cc.inv_p_dig[d], xq);
            const double q = rint(xq);
            // C' = sum_d D_d 2^(40 (nd-1-d)) with D_d = S_d - q P_d (exact, |D_d| < 2^53).
            // Normalize to balanced 40-bit digits with exact carries, so that every
            // lower part is below half an ulp of what is added on top of it; the sum
            // is then within one rounding of C' and exact when C' is representable.
            double D[ND];
#pragma unroll
            for (int d = 0; d < ND; ++d) D[d] = fma(-q, p.cc.p[d], S[r][d]);
#pragma unroll
            for (int d = ND - 1; d >= 1; --d) {
                const double carry = rint(D[d] * 0x1p-40);
                D[d] = fma(-carry, 0x1p40, D[d]);  // exact, |D_d| <= 2^39
                D[d - 1] += carry;                 // exact, integers well below 2^53
            }
            double cp = 0.0;
#pragma unroll
            for (int d = ND - 1; d >= 0; --d) cp = ldexp(D[d], p.cc.dig_exp[d]) + cp;
            const int sa = (i < p.m) ? p.sig_a[i] : 0;
            prods[t * (kCrtRows + 1) + r] = (sa == kNonFinite || sb == kNonFinite)
                                                ? __longlong_as_double(0x7ff8000000000000ll)
                                                : ldexp(cp, -(sa + sb));
        }
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kCrtRows; ++r) {
        const int e = t + kCrtThreads * r;  // row base + e of the block
        const int i = base + e;
        if (i < p.m) {
            double* cptr = p.c + (int64_t)j * p.ldc + i;
            *cptr = epilogue<double>(p.alpha, p.beta, prods[(e / kCrtRows) * (kCrtRows + 1) + e % kCrtRows], cptr);
        }
    }
}

// slices: C = alpha * 2^-(e_a + e_b) W + beta C
__global__ __launch_bounds__(256) void slices_finish_kernel(int m, int n, const double* __restrict__ w, int64_t ldw,
                                                            const int* __restrict__ sig_a,
                                                            const int* __restrict__ sig_b, double alpha, double beta,
                                                            double* __restrict__ c, int64_t ldc) {
    const int i = blockIdx.x * 256 + threadIdx.x, j = blockIdx.y;
    if (i >= m) return;
    const int sa = sig_a[i], sb = sig_b[j];
    const double prod = (sa == kNonFinite || sb == kNonFinite) ? __longlong_as_double(0x7ff8000000000000ll)
                                                               : ldexp(w[(int64_t)j * ldw + i], -(sa + sb));
    double* cptr = c + (int64_t)j * ldc + i;
    *cptr = epilogue<double>(alpha, beta, prod, cptr);
}

// ---------------------------------------------------------------------------
// host driver
// ---------------------------------------------------------------------------
namespace {
struct Workspace {
    char* base = nullptr;
    size_t used = 0;
    template <typename T>
    T* take(size_t count) {
        used = (used + 255) & ~size_t(255);
        T* p = reinterpret_cast<T*>(base + used);
        used += count * sizeof(T);
        return p;
    }
};

// Zeroes the row maxima of op(A) and op(B) (one launch).
hipError_t stats_zero(uint64_t* max_a, int m, uint64_t* max_b, int n, hipStream_t s) {
    const int fb = std::max(1, std::min(1024, (m + n + 255) / 256));
    fill2_u64_kernel<<<fb, 256, 0, s>>>(max_a, m, max_b, n, 0ull);
    return hipGetLastError();
}

// Statistics and scaling exponents of rows of op(A) (m of them, contiguous in
// k when kca) and of op(B) (its columns, contiguous in k when kcb), one
// launch each over both; either count may be 0.  The chunk sums of an operand
// are laid out [chunk][row] over the rows of this call.
hipError_t stats_pass(const double* a, int64_t lda, int m, bool kca, uint64_t* max_a, double* ss_a, int* sig_a,
                      const double* b, int64_t ldb, int n, bool kcb, uint64_t* max_b, double* ss_b, int* sig_b, int k,
                      int nchunks, int alpha, bool crt, hipStream_t s) {
    if (m + n <= 0) return hipSuccess;
    auto op = [&](const double* x, int64_t ld, int rows, bool kc, uint64_t* mx, double* ss) {
        StatOp o;
        o.x = x; o.ld = ld; o.rows = rows; o.kcontig = kc ? 1 : 0; o.maxbits = mx; o.part = ss;
        o.blocks = kc ? (rows + 63) / 64 : (rows + 255) / 256;
        return o;
    };
    const StatOp oa = op(a, lda, m, kca, max_a, ss_a), ob = op(b, ldb, n, kcb, max_b, ss_b);
    dim3 grid(oa.blocks + ob.blocks, nchunks);
    row_stats_kernel<<<grid, 256, 0, s>>>(oa, ob, k);
    const ExpOp ea{m, max_a, ss_a, sig_a}, eb{n, max_b, ss_b, sig_b};
    row_exponent_kernel<<<(m + n + 255) / 256, 256, 0, s>>>(ea, eb, k, nchunks, alpha, crt ? 1 : 0);
    return hipGetLastError();
}

// Planes of op(A) and op(B) in one launch.
hipError_t planes2(const double* a, int64_t lda, int m, bool kca, const int* sig_a, int8_t* pa, int64_t plane_a,
                   const double* b, int64_t ldb, int n, bool kcb, const int* sig_b, int8_t* pb, int64_t plane_b,
                   int k, int kp, int nplanes, bool crt, const CrtConsts* cc, hipStream_t s, bool bf16) {
    PlaneArgs pa_;
    memset(&pa_, 0, sizeof(pa_));
    auto op = [&](const double* x, int64_t ld, int rows, bool kc, const int* sig, int8_t* out, int64_t stride) {
        PlaneOp o;
        o.x = x; o.ld = ld; o.rows = rows; o.kcontig = kc ? 1 : 0; o.sigma = sig; o.out = out;
        o.plane_stride = stride;
        o.blocks = (rows + ZG_PLANE_ROWS - 1) / ZG_PLANE_ROWS;
        return o;
    };
    pa_.op[0] = op(a, lda, m, kca, sig_a, pa, plane_a);
    pa_.op[1] = op(b, ldb, n, kcb, sig_b, pb, plane_b);
    pa_.k = k; pa_.kp = kp; pa_.nplanes = nplanes; pa_.crt = crt ? 1 : 0; pa_.bf16 = bf16 ? 1 : 0;
    if (crt) {
        for (int l = 0; l < nplanes; ++l) {
            const int md = cc->m[l];
            pa_.m[l] = md;
            pa_.inv_m[l] = cc->inv_m[l];
            const uint64_t c32 = (1ull << 32) % (uint64_t)md;
            pa_.c32[l] = (double)c32;
            pa_.c64[l] = (double)((c32 * c32) % (uint64_t)md);
        }
    }
    // |a| <= 2^(alpha+1) fits two parts below 2^63
    const bool three = crt && cc->alpha + 1 >= 63;
    if (pa_.op[0].blocks + pa_.op[1].blocks == 0) return hipSuccess;
    dim3 grid(pa_.op[0].blocks + pa_.op[1].blocks, kp / kPlaneTK);
    if (three) plane_kernel<3><<<grid, 256, 0, s>>>(pa_);
    else plane_kernel<2><<<grid, 256, 0, s>>>(pa_);
    return hipGetLastError();
}
}  // namespace

template <int BM, int BN, int BK, int WM, int WN, int OCC, class MF = I8Mfma<16>, bool SWZ = false>
static hipError_t i8_launch_cfg(I8Args p, int nout, hipStream_t s) {
    p.tiles_m = (p.m + BM - 1) / BM;
    p.tiles_n = (p.n + BN - 1) / BN;
    if (p.kp % BK != 0) return hipErrorInvalidValue;
    const size_t lds = 2 * (size_t)(BM + BN) * (SWZ ? BK : BK + ZG_I8_LDS_PAD);
    auto kern = i8_gemm_kernel<BM, BN, BK, WM, WN, OCC, MF, SWZ>;
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(kern), hipFuncAttributeMaxDynamicSharedMemorySize,
                                       (int)lds);
    if (e != hipSuccess) return e;
    dim3 grid(p.tiles_m * p.tiles_n, 1, nout);
    kern<<<grid, 64 * WM * WN, lds, s>>>(p);
    return hipGetLastError();
}

template <int BM, int BN, int WM, int WN, int NS, int OCC, class MF>
static hipError_t i8_dma_launch_cfg(I8Args p, int nout, hipStream_t s) {
    p.tiles_m = (p.m + BM - 1) / BM;
    p.tiles_n = (p.n + BN - 1) / BN;
    if (p.kp % 64 != 0 || p.m < 1 || p.n < 1) return hipErrorInvalidValue;
    const size_t lds = (size_t)NS * (BM + BN) * 64;
    auto kern = i8_dma_kernel<BM, BN, WM, WN, NS, OCC, MF>;
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(kern), hipFuncAttributeMaxDynamicSharedMemorySize,
                                       (int)lds);
    if (e != hipSuccess) return e;
    dim3 grid(p.tiles_m * p.tiles_n, 1, nout);
    kern<<<grid, 64 * WM * WN, lds, s>>>(p);
    return hipGetLastError();
}

// INT8 tile configurations (BM x BN x BK, wave grid); kp must be a multiple of BK.
hipError_t i8_gemm_launch(const I8Args& args, int nout, hipStream_t s, int tile) {
    switch (tile) {
        case 0: return i8_launch_cfg<256, 256, 128, 2, 2, 1>(args, nout, s);
        case 1: return i8_launch_cfg<256, 128, 64, 2, 2, 2>(args, nout, s);
        case 2: return i8_launch_cfg<128, 128, 128, 2, 2, 2>(args, nout, s);
        case 3: return i8_launch_cfg<256, 256, 64, 2, 2, 1>(args, nout, s);
        case 4: return i8_launch_cfg<256, 256, 128, 4, 2, 2>(args, nout, s);
        case 5: return i8_launch_cfg<256, 256, 64, 4, 2, 2>(args, nout, s);
        case 6: return i8_launch_cfg<256, 256, 128, 4, 2, 2, I8Mfma<32>>(args, nout, s);
        case 7: return i8_launch_cfg<256, 256, 64, 4, 2, 2, I8Mfma<32>>(args, nout, s);
        case 8: return i8_launch_cfg<256, 128, 64, 2, 2, 2, I8Mfma<32>>(args, nout, s);
        case 9: return i8_launch_cfg<128, 128, 128, 2, 2, 2, I8Mfma<32>>(args, nout, s);
        // direct-to-LDS ring (BK 64 per stage): stages, 32x32x32 MFMA
        case 10: return i8_dma_launch_cfg<256, 256, 4, 2, 4, 2, I8Mfma<32>>(args, nout, s);
        case 11: return i8_dma_launch_cfg<256, 256, 4, 2, 3, 2, I8Mfma<32>>(args, nout, s);
        case 12: return i8_dma_launch_cfg<256, 128, 4, 2, 4, 2, I8Mfma<32>>(args, nout, s);
        case 13: return i8_dma_launch_cfg<256, 256, 2, 2, 4, 1, I8Mfma<32>>(args, nout, s);
        // unpadded LDS rows with the XOR chunk swizzle
        case 14: return i8_launch_cfg<256, 256, 128, 4, 2, 2, I8Mfma<16>, true>(args, nout, s);
        case 15: return i8_launch_cfg<256, 256, 128, 2, 2, 1, I8Mfma<16>, true>(args, nout, s);
        case 16: return i8_launch_cfg<256, 256, 128, 4, 2, 2, I8Mfma<32>, true>(args, nout, s);
        case 17: return i8_launch_cfg<128, 128, 128, 2, 2, 2, I8Mfma<16>, true>(args, nout, s);
        // 16 waves of 64 x 64 (4 per SIMD: twice the latency hiding of the 8-wave tiles)
        case 18: return i8_launch_cfg<256, 256, 64, 4, 4, 4, I8Mfma<16>>(args, nout, s);
        case 19: return i8_launch_cfg<256, 256, 128, 4, 4, 4, I8Mfma<16>, true>(args, nout, s);
        case 20: return i8_launch_cfg<256, 256, 64, 4, 4, 4, I8Mfma<32>>(args, nout, s);
        case 21: return i8_launch_cfg<256, 256, 128, 4, 4, 4, I8Mfma<32>, true>(args, nout, s);
        // 8 waves of 64 x 64, two workgroups per CU (independent barriers)
        case 22: return i8_launch_cfg<128, 256, 64, 2, 4, 4, I8Mfma<16>>(args, nout, s);
        case 23: return i8_launch_cfg<256, 128, 64, 4, 2, 4, I8Mfma<16>>(args, nout, s);
        case 24: return i8_launch_cfg<128, 256, 128, 2, 4, 4, I8Mfma<16>, true>(args, nout, s);
        case 25: return i8_launch_cfg<256, 256, 128, 4, 4, 4, I8Mfma<16>>(args, nout, s);
        default: return hipErrorInvalidValue;
    }
}

// BF16 tile configurations over bf16 planes (BK in bytes: 64 bf16 per step)
hipError_t bf16_gemm_launch(const I8Args& args, int nout, hipStream_t s, int tile) {
    switch (tile) {
        case 0: return i8_launch_cfg<128, 128, 128, 4, 2, 2, Bf16Mfma<16>>(args, nout, s);
        case 1: return i8_launch_cfg<128, 128, 128, 4, 2, 2, Bf16Mfma<32>>(args, nout, s);
        case 2: return i8_launch_cfg<256, 128, 128, 4, 2, 1, Bf16Mfma<32>>(args, nout, s);
        case 3: return i8_dma_launch_cfg<256, 128, 4, 2, 4, 1, Bf16Mfma<32>>(args, nout, s);
        default: return hipErrorInvalidValue;
    }
}

namespace {
struct I8TuneRow {
    long long min_tiles256;
    int cfg;
};
#include "i8_tune.inc"
}  // namespace

// kp is a multiple of 128, which every configuration accepts
int i8_pick_tile(int m, int n, int kp) {
    (void)kp;
    const long long tiles = (long long)((m + 255) / 256) * ((n + 255) / 256);
    for (const I8TuneRow& r : kI8TuneTable)
        if (tiles >= r.min_tiles256) return r.cfg;
    return 1;
}

int ozaki_dgemm(bool ta, bool tb, int m, int n, int k, double alpha, const double* a, int64_t lda, const double* b,
                int64_t ldb, double beta, double* c, int64_t ldc, hipStream_t s, int mode, int mode_param, int tile) {
    const bool crt = (mode == ZYNUM_GPU_MODE_OZAKI_CRT);
    const bool bf16 = (mode == ZYNUM_GPU_MODE_OZAKI_BF16);  // slices as bf16, FP32 MFMA accumulation
    const int eb = bf16 ? 2 : 1;
    const int planes_n = crt ? (mode_param ? mode_param : 14) : (mode_param ? mode_param : 7);
    if (crt && (planes_n < 8 || planes_n > kMaxModuli)) return ZYNUM_GPU_DECLINE_UNSUPPORTED;
    if (!crt && (planes_n < 2 || planes_n > 8)) return ZYNUM_GPU_DECLINE_UNSUPPORTED;
    // INT32 headroom: CRT residues are in [-128, 127], so K products of at most 2^14 stay below 2^31 for
    // K < 2^17; a slice level sums up to s pairs of products of at most 2^12, so s K < 2^19.  The BF16 form
    // flushes its FP32 sums into FP64 and has no such limit (2^18 bounds its workspace).
    const int64_t kmax = crt ? (1 << 17) - 1 : bf16 ? (1 << 18) : ((1 << 19) - 1) / planes_n;
    if ((int64_t)k > kmax) return ZYNUM_GPU_DECLINE_UNSUPPORTED;
    CrtConsts cc;
    if (crt && !crt_setup(planes_n, k, cc)) return ZYNUM_GPU_DECLINE_UNSUPPORTED;

    const int kp = (k + 127) / 128 * 128;
    const int ldu = (m + kCrtRows - 1) / kCrtRows * kCrtRows;  // residue plane rows (the GEMM writes 4 at a time)
    const size_t plane_a = (size_t)m * kp * eb, plane_b = (size_t)n * kp * eb;  // bytes
    size_t bytes = 0;
    const int nchunks = (k + kStatChunk - 1) / kStatChunk;
    bytes += ((size_t)m + (size_t)n) * 8 * (1 + (size_t)nchunks) + 4096;  // max bits + chunk sums of squares
    bytes += ((size_t)m + (size_t)n) * 4 + 1024;                 // sigma
    bytes += (plane_a + plane_b) * planes_n + 1024;              // planes
    bytes += crt ? (size_t)ldu * n * planes_n + 256 : (size_t)m * n * 8 + 256;  // u bytes / FP64 level sums
    bytes += 4096;
    void* wsp = workspace_acquire(bytes, s);
    if (!wsp) return ZYNUM_GPU_DECLINE_ALLOC;
    Workspace ws{static_cast<char*>(wsp), 0};
    uint64_t* max_a = ws.take<uint64_t>(m);
    uint64_t* max_b = ws.take<uint64_t>(n);
    double* ss_a = ws.take<double>((size_t)m * nchunks);
    double* ss_b = ws.take<double>((size_t)n * nchunks);
    int* sig_a = ws.take<int>(m);
    int* sig_b = ws.take<int>(n);
    int8_t* pa = ws.take<int8_t>(plane_a * planes_n);
    int8_t* pb = ws.take<int8_t>(plane_b * planes_n);

    // op(A) rows are contiguous in k when transposed; op(B) "rows" (its
    // columns) when not.  Each stage is one launch over both operands.
    //
    // Large CRT products run in ZG_CRT_PARTS parts of N on two streams: s
    // prepares op(A) and op(B)'s first part and starts that part's INT8
    // products (MFMA-bound) while the side stream prepares the rest of op(B)
    // (memory-bound); after that, the side stream reconstructs each part while
    // s runs the next part's products.  Every kernel computes what it
    // computes unsplit, so the results are the same bits.
    const int parts = (crt && n >= kCrtSplitMinN) ? kCrtParts : 1;
    const int nfirst = (parts > 1) ? (n / parts + 255) / 256 * 256 : n;
    const double* b_rest = tb ? b + nfirst : b + (int64_t)nfirst * ldb;  // op(B) columns nfirst..n
    hipError_t e = stats_zero(max_a, m, max_b, n, s);
    if (e == hipSuccess)
        e = stats_pass(a, lda, m, ta, max_a, ss_a, sig_a, b, ldb, nfirst, !tb, max_b, ss_b, sig_b, k, nchunks,
                       crt ? cc.alpha : 0, crt, s);
    if (e == hipSuccess)
        e = planes2(a, lda, m, ta, sig_a, pa, plane_a, b, ldb, nfirst, !tb, sig_b, pb, plane_b, k, kp, planes_n, crt,
                    &cc, s, bf16);
    hipStream_t side = nullptr;
    const bool split = (e == hipSuccess) && nfirst < n && aux_fork(s, &side);
    const hipStream_t sr = split ? side : s;  // where the rest of op(B) is prepared
    if (e == hipSuccess && nfirst < n) {
        double* ss_rest = ss_b + (size_t)nfirst * nchunks;
        e = stats_pass(nullptr, 0, 0, false, nullptr, nullptr, nullptr, b_rest, ldb, n - nfirst, !tb, max_b + nfirst,
                       ss_rest, sig_b + nfirst, k, nchunks, crt ? cc.alpha : 0, crt, sr);
        if (e == hipSuccess)
            e = planes2(nullptr, 0, 0, false, nullptr, nullptr, 0, b_rest, ldb, n - nfirst, !tb, sig_b + nfirst,
                        pb + (int64_t)nfirst * kp * eb, plane_b, k, kp, planes_n, crt, &cc, sr, bf16);
    }

    I8Args ia;
    memset(&ia, 0, sizeof(ia));
    ia.m = m; ia.n = n; ia.kp = kp * eb;
    ia.a = pa; ia.a_plane = plane_a; ia.b = pb; ia.b_plane = plane_b;
    if (crt && e == hipSuccess) {
        uint8_t* u = ws.take<uint8_t>((size_t)ldu * n * planes_n);
        ia.mode = 0;
        ia.u = u; ia.u_plane = (int64_t)ldu * n; ia.ldu = ldu;
        for (int l = 0; l < planes_n; ++l) {
            ia.outs[l].npairs = 1;
            ia.outs[l].pairs[0] = {l, l};
            ia.outs[l].modulus = cc.m[l];
            ia.outs[l].y = cc.y[l];
            ia.outs[l].inv_m = cc.inv_m[l];
        }
        // products and reconstruction of output columns [c0, c0 + nc)
        auto products = [&](int c0, int nc) {
            I8Args h = ia;
            h.n = nc;
            h.b = pb + (int64_t)c0 * kp * eb;
            h.u = u + (int64_t)c0 * ldu;
            return i8_gemm_launch(h, planes_n, s, tile >= 0 ? tile : i8_pick_tile(m, nc, kp));
        };
        auto reconstruct = [&](int c0, int nc, hipStream_t st) {
            CrtArgs ca;
            ca.m = m; ca.n = nc; ca.u = u + (int64_t)c0 * ldu; ca.u_plane = ia.u_plane; ca.ldu = ldu;
            ca.sig_a = sig_a; ca.sig_b = sig_b + c0; ca.alpha = alpha; ca.beta = beta; ca.c = c + (int64_t)c0 * ldc;
            ca.ldc = ldc;
            ca.cc = cc;
            dim3 grid((m + kCrtThreads * kCrtRows - 1) / (kCrtThreads * kCrtRows), nc);
            switch (cc.ndig) {
                case 2: crt_reconstruct_kernel<2><<<grid, kCrtThreads, 0, st>>>(ca); break;
                case 3: crt_reconstruct_kernel<3><<<grid, kCrtThreads, 0, st>>>(ca); break;
                case 4: crt_reconstruct_kernel<4><<<grid, kCrtThreads, 0, st>>>(ca); break;
                default: crt_reconstruct_kernel<5><<<grid, kCrtThreads, 0, st>>>(ca); break;
            }
            return hipGetLastError();
        };
        if (!split) {
            e = products(0, n);
            if (e == hipSuccess) e = reconstruct(0, n, s);
        } else {
            // part p covers columns [p * nfirst, min(n, (p + 1) * nfirst))
            for (int c0 = 0; c0 < n && e == hipSuccess; c0 += nfirst) {
                const int nc = std::min(nfirst, n - c0);
                e = products(c0, nc);
                if (e != hipSuccess) break;
                if (c0 == 0) e = aux_join(s, side);  // the later parts' planes are ready
                if (e != hipSuccess) break;
                if (c0 + nc < n) {
                    hipStream_t sd = nullptr;  // the side stream, after this part's products
                    if (!aux_fork(s, &sd)) e = hipErrorUnknown;
                    if (e == hipSuccess) e = reconstruct(c0, nc, sd);
                } else {
                    e = reconstruct(c0, nc, s);
                }
            }
        }
    }
    if (split) {
        const hipError_t ej = aux_join(s, side);  // always joined: the workspace is released after s drains
        if (e == hipSuccess) e = ej;
    }
    if (!crt && e == hipSuccess) {
        double* w = ws.take<double>((size_t)m * n);
        ia.mode = 1;
        ia.w = w; ia.ldw = m;
        // level t (2..s+1): pairs (i, t - i) of 1-based slices, scale 2^-(12 + 7 (t - 2))
        for (int t = 2; t <= planes_n + 1 && e == hipSuccess; ++t) {
            I8Args lv = ia;
            int np = 0;
            for (int i = 1; i < t; ++i) {
                const int j = t - i;
                if (i > planes_n || j > planes_n) continue;
                lv.outs[0].pairs[np++] = {i - 1, j - 1};
            }
            lv.outs[0].npairs = np;
            lv.outs[0].scale = ldexp(1.0, -(12 + 7 * (t - 2)));
            lv.accumulate = (t == 2) ? 0 : 1;
            e = bf16 ? bf16_gemm_launch(lv, 1, s, tile >= 0 ? tile : 2)  // measured best at 8192^3
                     : i8_gemm_launch(lv, 1, s, tile >= 0 ? tile : i8_pick_tile(m, n, kp));
        }
        if (e == hipSuccess) {
            dim3 grid((m + 255) / 256, n);
            slices_finish_kernel<<<grid, 256, 0, s>>>(m, n, w, m, sig_a, sig_b, alpha, beta, c, ldc);
            e = hipGetLastError();
        }
    }
    int st = status_from_hip(e);
    if (split && st != ZYNUM_GPU_OK) (void)hipStreamSynchronize(side);
    workspace_release(s);  // the side stream's work is joined into s
    return st;
}

}  // namespace zg
