// The FP64 128x128 GEMM tile fed by direct-to-LDS loads, shared by the plain
// GEMM kernel and the SYRK triangle kernel.
#pragma once

#include "common.hpp"
#include "gemm_tile.hpp"

namespace zg {

// FP64 128x128 tile fed by direct-to-LDS loads (global_load_lds_dwordx4), for
// full tiles: m and n multiples of 128, k of 8, A and B 16-byte aligned with
// even leading dimensions.  No registers stage the operands: stages of 8 k
// sit in a 4-stage LDS ring (66 KB, two workgroups per CU), loaded 3 stages
// ahead as 16 direct loads per stage (8 for op(A), 8 for op(B)), 16 / NW per
// wave.  NW waves (4, 8 or 16) in a 4 x NW/4 grid each multiply a 32 x 512/NW
// sub-tile in 16x16x4 MFMA steps in increasing k, exactly as gemm_tile does,
// so results are bitwise those of the register-staged kernel; the freed
// registers hold the fragments of the next 4-k half, read while the current
// half's MFMAs run.  Wider sub-tiles read fewer fragment bytes per MFMA;
// fewer waves leave less to cover the waits: at 8192^3 4 waves are fastest
// for NN (72.8 TFLOP/s), 8 for the other pairs (71.1-72.6), 16 slower
// (69.6-70.9).
//
// A direct load cannot transpose, so each operand keeps the layout it has in
// memory.  An operand contiguous along m / n (A, B^T) is stored as 8 k-rows
// of 128 + 2 doubles (the padding of the register-staged tile), one direct
// load per k-row.  An operand contiguous along k (A^T, B) is stored as 128
// rows of 8 k, 16 rows per direct load, the four 16-byte k pairs of row r
// rotated by (r >> 2) & 3, so the 16 rows of one fragment read at one k fall
// on distinct LDS banks.
constexpr int kDmaStages = 4;
// The tile at (m0, n0), computed by a workgroup of 64 NW threads.  TRI as in
// gemm_tile: 1 stores only entries on or below the diagonal, 2 on or above.
template <bool TA, bool TB, int NW, int TRI = 0>
__device__ __forceinline__ void gemm_tile_dma(const GemmArgs<double>& p, int m0, int n0) {
    constexpr int BM = 128, BN = 128, BK = 8, NS = kDmaStages;
    constexpr bool KRA = !TA, KRB = TB;  // operand stored as k-rows (contiguous along m / n)
    constexpr int RK = BM + 2;            // k-row stride
    constexpr int SA = KRA ? BK * RK : BM * BK;
    constexpr int SB = KRB ? BK * RK : BN * BK;
    constexpr int SS = SA + SB;
    __shared__ __attribute__((aligned(16))) double lds[NS * SS];

    static_assert(NW == 4 || NW == 8 || NW == 16, "4, 8 or 16 waves");
    constexpr int WN = NW / 4, WTN = 128 / WN, FN = WTN / 16, LPW = 16 / NW;
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int wm0 = (wave % 4) * 32, wn0 = (wave / 4) * WTN;

    // direct load idx = wave + NW l (0..15) fills part q = idx % 8 of op(A)
    // (idx < 8) or op(B): k-row q (k-row layout) or rows 16q..16q+15 (row
    // layout; lane: slot lane % 4 of row lane / 4, holding k pair
    // (lane % 4) ^ ((row >> 2) & 3)); load t of it reads src + t * kstep
    const double* src[LPW];
    int64_t kstep[LPW];
    int dst[LPW];
#pragma unroll
    for (int l = 0; l < LPW; ++l) {
        const int idx = wave + NW * l;
        const bool is_a = idx < 8;
        const int q = idx % 8;
        const bool kr = is_a ? KRA : KRB;
        const double* x = is_a ? p.a : p.b;
        const int64_t ldx = is_a ? p.lda : p.ldb;
        const int x0 = is_a ? m0 : n0;
        if (kr) {
            src[l] = x + (int64_t)q * ldx + x0 + 2 * lane;
            kstep[l] = (int64_t)BK * ldx;
            dst[l] = (is_a ? 0 : SA) + q * RK;
        } else {
            const int rl = q * 16 + (lane >> 2);
            const int pair = (lane & 3) ^ ((rl >> 2) & 3);
            src[l] = x + (int64_t)(x0 + rl) * ldx + 2 * pair;
            kstep[l] = BK;
            dst[l] = (is_a ? 0 : SA) + q * 128;
        }
    }
    auto issue = [&](int t) {
#pragma unroll
        for (int l = 0; l < LPW; ++l)
            __builtin_amdgcn_global_load_lds(
                (const void __attribute__((address_space(1)))*)(src[l] + (int64_t)t * kstep[l]),
                (void __attribute__((address_space(3)))*)(lds + (t % NS) * SS + dst[l]), 16, 0, 0);
    };

    const int lr = lane & 15, lk = lane >> 4;
    const int ar = Mfma<double>::a_row(lr);
    auto at = [](bool krl, int base, int r, int k) {  // element (row r, k) of an operand part
        return krl ? base + k * RK + r : base + r * BK + ((((k >> 1) ^ ((r >> 2) & 3))) << 1) + (k & 1);
    };
    // a k-row operand's second half is a constant 4 k-rows further (folded
    // into the read's offset); a row-layout operand keeps an offset per half
    int aoff[KRA ? 1 : 2][2], boff[KRB ? 1 : 2][FN];  // [half][block]
#pragma unroll
    for (int h = 0; h < (KRA ? 1 : 2); ++h)
#pragma unroll
        for (int i = 0; i < 2; ++i) aoff[h][i] = at(KRA, 0, wm0 + i * 16 + ar, 4 * h + lk);
#pragma unroll
    for (int h = 0; h < (KRB ? 1 : 2); ++h)
#pragma unroll
        for (int i = 0; i < FN; ++i) boff[h][i] = at(KRB, SA, wn0 + i * 16 + lr, 4 * h + lk);
    v4d acc[2][FN];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j) acc[i][j] = Mfma<double>::zero();
    double fa[2][2], fb[2][FN];  // [half][block]
    auto read = [&](int h, int t) {
        const double* base = lds + (t % NS) * SS;
#pragma unroll
        for (int i = 0; i < 2; ++i) fa[h][i] = KRA ? base[aoff[0][i] + 4 * h * RK] : base[aoff[KRA ? 0 : h][i]];
#pragma unroll
        for (int i = 0; i < FN; ++i) fb[h][i] = KRB ? base[boff[0][i] + 4 * h * RK] : base[boff[KRB ? 0 : h][i]];
    };
    auto mma_row = [&](int h, int i) {
#pragma unroll
        for (int j = 0; j < FN; ++j) acc[i][j] = Mfma<double>::mma(fa[h][i], fb[h][j], acc[i][j]);
    };

    const int nk = p.k / BK;
#pragma unroll
    for (int t = 0; t < NS - 1; ++t)
        if (t < nk) issue(t);
    if (nk >= NS - 1) __builtin_amdgcn_s_waitcnt(0x0F70 | (2 * LPW));  // vmcnt(2 LPW): this wave's stage 0 has landed
    else __builtin_amdgcn_s_waitcnt(0x0F70);               // vmcnt(0)
    __builtin_amdgcn_s_barrier();
    read(0, 0);
    for (int s = 0; s < nk; ++s) {
        // this wave's stage s+1 has landed (stages up to s + NS - 2 are issued)
        if (s + NS - 2 <= nk - 1) __builtin_amdgcn_s_waitcnt(0x0F70 | LPW);  // vmcnt(LPW)
        else __builtin_amdgcn_s_waitcnt(0x0F70);                       // vmcnt(0)
        __builtin_amdgcn_sched_barrier(0);
        __builtin_amdgcn_s_barrier();  // every wave's stage s+1 is in LDS, and every wave has read stage s-1
        __builtin_amdgcn_sched_barrier(0);
        if (s + NS - 1 < nk) issue(s + NS - 1);  // into stage s-1's buffer
        // Each half's fragment reads are issued two MFMAs before the wait
        // for them (the compiler waits for every outstanding LDS read here).
        mma_row(0, 0);
        __builtin_amdgcn_sched_barrier(0);
        read(1, s);
        __builtin_amdgcn_sched_barrier(0);
        mma_row(0, 1);
        __builtin_amdgcn_sched_barrier(0);
        mma_row(1, 0);
        __builtin_amdgcn_sched_barrier(0);
        read(0, s + 1);  // also on the last step (a valid buffer; the values are unused)
        __builtin_amdgcn_sched_barrier(0);
        mma_row(1, 1);
    }

    const double alpha = p.alpha, beta = p.beta;
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j) {
            const int col = n0 + wn0 + j * 16 + lr;
            const int row = m0 + wm0 + i * 16 + 4 * lk;
            double* cp = p.c + (int64_t)col * p.ldc + row;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                if (TRI == 1 && row + r < col) continue;
                if (TRI == 2 && row + r > col) continue;
                cp[r] = epilogue<double>(alpha, beta, acc[i][j][r], cp + r);
            }
        }
}


// Whether the direct-load tile takes an FP64 product of this shape: full
// 128x128 tiles, k a multiple of 8, 16-byte aligned operands with even
// leading dimensions.
inline bool dma_shape_ok(int m, int n, int k, const void* a, int64_t lda, const void* b, int64_t ldb) {
    const uintptr_t mask = 15;
    return m % 128 == 0 && n % 128 == 0 && k >= 8 && k % 8 == 0 && (reinterpret_cast<uintptr_t>(a) & mask) == 0 &&
           (reinterpret_cast<uintptr_t>(b) & mask) == 0 && lda % 2 == 0 && ldb % 2 == 0;
}

}  // namespace zg
