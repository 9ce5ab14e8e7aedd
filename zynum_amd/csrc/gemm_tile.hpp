// The gfx950 MFMA GEMM tile routine shared by the plain kernel and the
// persistent grouped kernel.
#pragma once

#include "common.hpp"

namespace zg {

template <typename T, int N>
using vec_t = T __attribute__((ext_vector_type(N)));

template <typename T>
struct GemmArgs {
    int m, n, k;
    int64_t lda, ldb, ldc;
    T alpha, beta;
    const T* __restrict__ a;
    const T* __restrict__ b;
    T* __restrict__ c;
    int tiles_m, tiles_n, group_m;
};

// Moves one R x BK slice of an operand (R = tile rows of op(X)) from global
// memory to registers and from registers to a k-major LDS panel [BK][R + PAD].
// KCONTIG: element (r, kk) is at x[(r0 + r) * ldx + k0 + kk] (contiguous in k),
// otherwise at x[(k0 + kk) * ldx + r0 + r] (contiguous in r).
template <typename T, int R, int BK, bool KCONTIG, int VEC, int NT, int PAD>
struct Stage {
    static constexpr int E = R * BK / NT;  // elements per thread
    static constexpr int NV = E / VEC;     // vector loads per thread
    static_assert(E % VEC == 0, "tile not divisible");
    static constexpr int PER = KCONTIG ? BK / VEC : R / VEC;  // vectors along the contiguous dim

    T v[E];

    __device__ __forceinline__ void load(const T* __restrict__ xp, int64_t ldx, int r0, int k0, int rlim,
                                         int klim, bool full, int tid) {
        const gptr<const T> x = to_global(xp);  // global loads also for a pointer read from memory
#pragma unroll
        for (int i = 0; i < NV; ++i) {
            const int idx = tid + NT * i;
            const int c = (idx % PER) * VEC;  // position along the contiguous dim
            const int o = idx / PER;          // position along the other dim
            const int r = KCONTIG ? o : c;
            const int kk = KCONTIG ? c : o;
            const int64_t off = KCONTIG ? (int64_t)(r0 + r) * ldx + (k0 + kk) : (int64_t)(k0 + kk) * ldx + (r0 + r);
            if (full) {
                if constexpr (VEC == 1) {
                    v[i] = x[off];
                } else {
                    vec_t<T, VEC> t = *(gptr<const vec_t<T, VEC>>)(x + off);
#pragma unroll
                    for (int e = 0; e < VEC; ++e) v[i * VEC + e] = t[e];
                }
            } else {
#pragma unroll
                for (int e = 0; e < VEC; ++e) {
                    const int rr = r0 + r + (KCONTIG ? 0 : e);
                    const int kv = k0 + kk + (KCONTIG ? e : 0);
                    v[i * VEC + e] = (rr < rlim && kv < klim) ? x[off + e] : T(0);
                }
            }
        }
    }

    __device__ __forceinline__ void store(T* __restrict__ lds, int tid) const {
#pragma unroll
        for (int i = 0; i < NV; ++i) {
            const int idx = tid + NT * i;
            const int c = (idx % PER) * VEC;
            const int o = idx / PER;
            if constexpr (KCONTIG) {
                // r = o, kk = c..c+VEC-1: transpose into the k-major panel
#pragma unroll
                for (int e = 0; e < VEC; ++e) lds[(c + e) * (R + PAD) + o] = v[i * VEC + e];
            } else {
                // kk = o, r = c..c+VEC-1: contiguous in the panel row
                if constexpr (VEC == 1) {
                    lds[o * (R + PAD) + c] = v[i];
                } else {
                    vec_t<T, VEC> t;
#pragma unroll
                    for (int e = 0; e < VEC; ++e) t[e] = v[i * VEC + e];
                    *reinterpret_cast<vec_t<T, VEC>*>(lds + o * (R + PAD) + c) = t;
                }
            }
        }
    }
};

// Tile configuration: BM x BN tile per workgroup, BK-deep K slices, WM x WN
// wave grid (each wave owns (BM/WM) x (BN/WN)), OCC = waves per SIMD the
// register budget is sized for (0 = compiler's choice).
template <int BM_, int BN_, int BK_, int WM_, int WN_, int OCC_>
struct TileCfg {
    static constexpr int BM = BM_, BN = BN_, BK = BK_, WM = WM_, WN = WN_, OCC = OCC_;
    static constexpr int NT = 64 * WM_ * WN_;
};

// Padding (bytes) of each k-row of an LDS operand panel; a build knob for tuning.
#ifndef ZG_LDS_PAD_BYTES
#define ZG_LDS_PAD_BYTES 16
#endif
template <typename T>
constexpr int lds_pad() { return ZG_LDS_PAD_BYTES / (int)sizeof(T); }

// LDS elements one tile configuration needs for its two operand stages.
template <typename T, class CF>
constexpr int tile_lds_a() { return 2 * CF::BK * (CF::BM + lds_pad<T>()); }
template <typename T, class CF>
constexpr int tile_lds_b() { return 2 * CF::BK * (CF::BN + lds_pad<T>()); }

// C[m0:m0+BM, n0:n0+BN] = alpha * op(A) op(B) + beta * C for one output tile,
// run by all CF::NT threads of the workgroup.  The K loop, the per-element
// MFMA accumulation order (16x16x4 steps in increasing k) and the epilogue are
// the same for every tile configuration, so any configuration gives the same
// bits for the same element when the whole K range is walked by one tile.
// (The split-K form sums K ranges separately and then adds the ranges, a
// different order: bitwise only on inputs where every order is exact.)
// TRI selects the stored part of C: 0 all of it, 1 the lower triangle
// (row >= col), 2 the upper triangle (row <= col); other elements are not
// touched (SYRK on the GEMM core).
template <typename T, class CF, bool TA, bool TB, int VEC, int TRI = 0>
__device__ __forceinline__ void gemm_tile(const GemmArgs<T>& p, int m0, int n0, T* __restrict__ lds_a,
                                          T* __restrict__ lds_b) {
    constexpr int BM = CF::BM, BN = CF::BN, BK = CF::BK, NT = CF::NT;
    constexpr int PAD = lds_pad<T>();
    constexpr int WTM = BM / CF::WM, WTN = BN / CF::WN;
    constexpr int FM = WTM / 16, FN = WTN / 16;
    static_assert(FM >= 1 && FN >= 1, "wave tile smaller than one MFMA block");
    using M = Mfma<T>;
    using Acc = typename M::acc_t;

    // lds_a: 2 stages of BK * (BM + PAD), lds_b: 2 stages of BK * (BN + PAD)

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int wm0 = (wave % CF::WM) * WTM;
    const int wn0 = (wave / CF::WM) * WTN;

    // op(A) tile rows are M (A: M-contiguous unless TA); op(B) rows are N
    // (B^T: N-contiguous when TB, otherwise K-contiguous).
    using StA = Stage<T, BM, BK, TA, VEC, NT, PAD>;
    using StB = Stage<T, BN, BK, !TB, VEC, NT, PAD>;

    Acc acc[FM][FN];
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j) acc[i][j] = M::zero();

    const int nk = (p.k + BK - 1) / BK;
    const bool mn_full = (m0 + BM <= p.m) && (n0 + BN <= p.n);

    const int lr = lane & 15;  // column within a 16x16 block (B, D)
    const int lk = lane >> 4;  // k within a 4-deep step
    const int ar = M::a_row(lr);  // A row fed by this lane

    auto load = [&](StA& sa, StB& sb, int t) {
        const int k0 = t * BK;
        const bool full = mn_full && (k0 + BK <= p.k);
        sa.load(p.a, p.lda, m0, k0, p.m, p.k, full, tid);
        sb.load(p.b, p.ldb, n0, k0, p.n, p.k, full, tid);
    };
    auto store = [&](const StA& sa, const StB& sb, int buf) {
        sa.store(lds_a + buf * (BK * (BM + PAD)), tid);
        sb.store(lds_b + buf * (BK * (BN + PAD)), tid);
    };
    auto multiply = [&](int buf) {
        const T* __restrict__ la = lds_a + buf * (BK * (BM + PAD));
        const T* __restrict__ lb = lds_b + buf * (BK * (BN + PAD));
#pragma unroll
        for (int kk = 0; kk < BK; kk += 4) {
            T af[FM], bf[FN];
#pragma unroll
            for (int i = 0; i < FM; ++i) af[i] = la[(kk + lk) * (BM + PAD) + wm0 + i * 16 + ar];
#pragma unroll
            for (int j = 0; j < FN; ++j) bf[j] = lb[(kk + lk) * (BN + PAD) + wn0 + j * 16 + lr];
#pragma unroll
            for (int i = 0; i < FM; ++i)
#pragma unroll
                for (int j = 0; j < FN; ++j) acc[i][j] = M::mma(af[i], bf[j], acc[i][j]);
        }
    };

    StA sa;
    StB sb;
    load(sa, sb, 0);
    store(sa, sb, 0);
    __syncthreads();
    for (int t = 0; t < nk; ++t) {
        const bool more = (t + 1) < nk;
        if (more) load(sa, sb, t + 1);
        multiply(t & 1);
        if (more) store(sa, sb, (t + 1) & 1);
        __syncthreads();
    }

    // Epilogue: lane holds rows 4*(lane/16)+r, column lane%16 of each block,
    // i.e. 4 consecutive rows of one C column.
    const T alpha = p.alpha, beta = p.beta;
#pragma unroll
    for (int i = 0; i < FM; ++i) {
#pragma unroll
        for (int j = 0; j < FN; ++j) {
            const int col = n0 + wn0 + j * 16 + lr;
            const int row = m0 + wm0 + i * 16 + 4 * lk;
            if (col < p.n) {
                T* cp = p.c + (int64_t)col * p.ldc + row;
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    if (TRI == 1 && row + r < col) continue;
                    if (TRI == 2 && row + r > col) continue;
                    if (row + r < p.m) cp[r] = epilogue<T>(alpha, beta, acc[i][j][r], cp + r);
                }
            }
        }
    }
}


}  // namespace zg
