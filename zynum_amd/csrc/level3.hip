// DTRSM and DSYRK on the GEMM core.
//
// TRSM (triangular.zig:155-249 solves one column at a time with trsv and
// streams the whole triangle per right-hand side) is recast recursively, as
// in the reference's blocked template (structured_blocked.zig:315-374):
// split the triangle in two halves, solve the first, update the second with
// one GEMM (B2 -= op(A)21 X1, plain FP64 MFMA or split precision), recurse.
// Triangles of at most 256 unknowns are one launch: 64-row blocks solved in
// FP64 with four lanes per vector (column of B for side L, row for side R)
// and the vector's unknowns in registers, in the reference's per-step order
// x_k /= a_kk; x_i = fma(a_ik, -x_k, x_i) (matrix_vector/triangular.zig:
// 375-385), their coupling subtracted with MFMAs in between.
//
// SYRK (symmetric.zig:180-279, one strided dot per element) runs as one launch
// over the lower (or upper) triangle of output tiles with the GEMM tile routine;
// diagonal tiles write only their stored triangle, so the other triangle is
// never touched (structured_blocked.zig:158-201 commits the same way).
#include <math.h>

#include <algorithm>
#include <utility>

#include "common.hpp"
#include "gemm.hpp"
#include "gemm_dma.hpp"
#include "gemm_tile.hpp"
#include "ops.hpp"
#include "runtime.hpp"

namespace zg {

// ---------------------------------------------------------------------------
// B <- alpha * B (a plain multiply, as scaleDenseColumns, triangular.zig:52-59)
// ---------------------------------------------------------------------------
template <typename T>
__global__ void mul_kernel(int m, int n, T alpha, T* __restrict__ b, int64_t ldb) {
    const int64_t total = (int64_t)m * n;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t j = e / m, i = e % m;
        b[j * ldb + i] = alpha * b[j * ldb + i];
    }
}

// ---------------------------------------------------------------------------
// leaf solve: nb <= 64 unknowns per vector
// ---------------------------------------------------------------------------
constexpr int kLeaf = 64;

template <typename T>
struct LeafArgs {
    const T* a;
    int64_t lda;
    T* b;
    int64_t ldb;
    int left;      // side L: vectors are columns of B, unknowns are rows i0..i0+nb
    int lower_eff; // op(A) is lower triangular
    int trans, unit;
    int i0, nb;    // unknowns [i0, i0 + nb) along the triangle
    int nvec;      // number of vectors (n for side L, m for side R)
};

template <typename T>
__device__ __forceinline__ T op_a(const LeafArgs<T>& p, int r, int c) {
    return p.trans ? p.a[(int64_t)r * p.lda + c] : p.a[(int64_t)c * p.lda + r];
}

// Leaf kernel: a quad of lanes per vector, 16 vectors per workgroup (one
// wave).  The unknowns are numbered in solve order (u = i forward, u = 63 - i
// backward, so a backward solve is a forward one in u); lane q of a quad holds
// unknowns u = 4j + q in x[j].  Step k takes the owner lane's (q = k % 4)
// final x_k to the whole quad with one DPP broadcast, divides it by the
// diagonal in every lane and eliminates it, one FMA per later unknown of the
// lane, in the reference's per-column order x_k /= a_kk;
// x_i = fma(a_ik, -x_k, x_i) (matrix_vector/triangular.zig:375-385).  The
// coefficients of a step are laid out in LDS per lane position and negated,
// Cs[k][q][j] = -a_{4j+q, k} (so no negation of x_k sits between the
// division and the FMAs), and read two steps ahead of their use; the
// diagonals are read as broadcasts.  A lane does a quarter of the FMAs and
// reads a quarter of the coefficients of the one-lane-per-vector form, and
// the grid has four times the waves: with one wave per SIMD the solve is
// bound by instruction issue, and the per-step division, which every lane
// executes, is then the largest share.  Unknowns past nb are padding:
// diagonal 1 and zero coefficients, so they stay 0 and eliminate nothing.
constexpr int kQuad = 4;                  // lanes per vector
constexpr int kLeafU = kLeaf / kQuad;     // unknowns per lane
constexpr int kLeafThreads = 64;
constexpr int kLeafVecs = kLeafThreads / kQuad;  // vectors per workgroup
constexpr int kLeafAhead = 2;             // steps whose coefficients are read ahead
constexpr int kLdsQ = kLeafU + 2;         // per-lane-position row (odd in 16-byte units: no bank conflict)
constexpr int kLdsK = kQuad * kLdsQ;      // per-step row

template <typename T>
__device__ __forceinline__ T quad_bcast(T v, int src);
template <>
__device__ __forceinline__ double quad_bcast<double>(double v, int src) {
    // DPP quad_perm [src, src, src, src] on both halves; src is a constant at every call
    const uint64_t u = __double_as_longlong(v);
    int lo = (int)(uint32_t)u, hi = (int)(uint32_t)(u >> 32);
    switch (src) {
        case 0: lo = __builtin_amdgcn_mov_dpp(lo, 0x00, 0xf, 0xf, false); hi = __builtin_amdgcn_mov_dpp(hi, 0x00, 0xf, 0xf, false); break;
        case 1: lo = __builtin_amdgcn_mov_dpp(lo, 0x55, 0xf, 0xf, false); hi = __builtin_amdgcn_mov_dpp(hi, 0x55, 0xf, 0xf, false); break;
        case 2: lo = __builtin_amdgcn_mov_dpp(lo, 0xaa, 0xf, 0xf, false); hi = __builtin_amdgcn_mov_dpp(hi, 0xaa, 0xf, 0xf, false); break;
        default: lo = __builtin_amdgcn_mov_dpp(lo, 0xff, 0xf, 0xf, false); hi = __builtin_amdgcn_mov_dpp(hi, 0xff, 0xf, 0xf, false); break;
    }
    return __longlong_as_double((long long)(((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo));
}
template <>
__device__ __forceinline__ float quad_bcast<float>(float v, int src) {
    int x = __float_as_int(v);
    switch (src) {
        case 0: x = __builtin_amdgcn_mov_dpp(x, 0x00, 0xf, 0xf, false); break;
        case 1: x = __builtin_amdgcn_mov_dpp(x, 0x55, 0xf, 0xf, false); break;
        case 2: x = __builtin_amdgcn_mov_dpp(x, 0xaa, 0xf, 0xf, false); break;
        default: x = __builtin_amdgcn_mov_dpp(x, 0xff, 0xf, 0xf, false); break;
    }
    return __int_as_float(x);
}

// coefficients (and diagonal) of step K into ring slot K % (kLeafAhead + 1);
// cq points at this lane's Cs[0][q][0]
template <typename T, bool UNIT, int K>
__device__ __forceinline__ void leaf_fetch(T (&c)[kLeafAhead + 1][kLeafU], T (&d)[kLeafAhead + 1], const T* cq,
                                           const T* Dg) {
    if constexpr (K < kLeaf) {
        constexpr int slot = K % (kLeafAhead + 1), j0 = K / kQuad;
#pragma unroll
        for (int j = j0; j < kLeafU; ++j) c[slot][j] = cq[K * kLdsK + j];
        if constexpr (!UNIT) d[slot] = Dg[K];
    }
}

template <typename T, bool UNIT, int K>
__device__ __forceinline__ void leaf_step(T (&x)[kLeafU], T (&c)[kLeafAhead + 1][kLeafU], T (&d)[kLeafAhead + 1],
                                          const T* cq, const T* Dg, int q, T& upd) {
    leaf_fetch<T, UNIT, K + kLeafAhead>(c, d, cq, Dg);
    constexpr int slot = K % (kLeafAhead + 1), jo = K / kQuad, o = K % kQuad;
    // x_k from its owner lane: for o > 0 the previous step's update of slot
    // jo, which is what lane o kept there (q > o - 1), taken before the selects
    T xk = quad_bcast<T>(o == 0 ? x[jo] : upd, o);
    if constexpr (!UNIT) xk = xk / d[slot];
    // the coefficients are stored negated: fma(-a_ik, x_k, x_i) is
    // fma(a_ik, -x_k, x_i) bit for bit
#pragma unroll
    for (int j = jo + 1; j < kLeafU; ++j) x[j] = fma_t(c[slot][j], xk, x[j]);
    // slot jo: unknown k in the owner, earlier (solved) unknowns below it,
    // later ones above
    // (unpredictable: selects, not branches around the FMA and the division)
    upd = fma_t(c[slot][jo], xk, x[jo]);
    const T keep = __builtin_unpredictable(q > o) ? upd : x[jo];
    x[jo] = __builtin_unpredictable(q == o) ? xk : keep;
    // one scheduling region per step: the reads stay ahead, register pressure
    // and compile time stay bounded
    __builtin_amdgcn_sched_barrier(0);
}

template <typename T, bool UNIT, int... K>
__device__ __forceinline__ void leaf_solve(T (&x)[kLeafU], const T* cq, const T* Dg, int q,
                                           std::integer_sequence<int, K...>) {
    T c[kLeafAhead + 1][kLeafU];
    T d[kLeafAhead + 1];
#pragma unroll
    for (int s = 0; s < kLeafAhead; ++s) {
#pragma unroll
        for (int j = s / kQuad; j < kLeafU; ++j) c[s][j] = cq[s * kLdsK + j];
        if constexpr (!UNIT) d[s] = Dg[s];
    }
    T upd = T(0);
    (leaf_step<T, UNIT, K>(x, c, d, cq, Dg, q, upd), ...);
}

// Coefficients of the 64-block whose first stored element is ablk: thread t
// takes stored row t, av[c] = stored column c (the wave reads one contiguous
// column per load), and dv = the diagonal of equation t.  Every load is
// issued before the first is used.  A full block (nb = 64) needs no clamping;
// otherwise rows and columns past nb repeat the last one.
template <typename T, typename P>
__device__ __forceinline__ void leaf_coeffs_load(T (&av)[kLeaf], T& dv, P ablk, int64_t lda, int nb, int t) {
    const int tc = min(t, nb - 1);
    if (nb == kLeaf) {
        P col = ablk + t;
#pragma unroll
        for (int c = 0; c < kLeaf; ++c, col += lda) av[c] = *col;
    } else {
#pragma unroll
        for (int c = 0; c < kLeaf; ++c) av[c] = ablk[(int64_t)min(c, nb - 1) * lda + tc];
    }
    dv = ablk[(int64_t)tc * lda + tc];
}

// The block's coefficients into LDS in solve order, negated,
// Cs[k][q][j] = -M_{4j+q, k} (M the equation matrix of the block), and its
// diagonal into Dg.  Side L
// solves op(A) x = b (M = op(A) block), side R solves x op(A) = b, i.e.
// op(A)^T x^T = b^T (M = op(A)^T block), so the stored element (row r,
// column c) is M[r][c] for side L without transpose and for side R with it,
// and M[c][r] otherwise (swap).  Every (equation, unknown) pair has its own
// slot.  A full block needs no zeroing: every slot a step reads for a later
// unknown holds its coefficient, and the other slots only feed updates that
// the owner selection discards.  A partial block zeroes the slots of padding
// and of the other triangle; padding equations get diagonal 1.
template <typename T, bool FWD, bool UNIT>
__device__ __forceinline__ void leaf_coeffs_store(T* Cs, T* Dg, const T (&av)[kLeaf], T dv, int nb, int t,
                                                  bool swap) {
    if constexpr (!UNIT) Dg[FWD ? t : kLeaf - 1 - t] = (t < nb) ? dv : T(1);
    auto slot = [](int i, int k) {  // slot of M[i][k]
        const int ui = FWD ? i : kLeaf - 1 - i, uk = FWD ? k : kLeaf - 1 - k;
        return uk * kLdsK + (ui % kQuad) * kLdsQ + ui / kQuad;
    };
    auto keep = [&](int i, int k) {  // a coefficient some step reads
        const int ui = FWD ? i : kLeaf - 1 - i, uk = FWD ? k : kLeaf - 1 - k;
        return i < nb && k < nb && ui > uk;
    };
    if (nb == kLeaf) {
        if (swap) {
#pragma unroll
            for (int c = 0; c < kLeaf; ++c) Cs[slot(c, t)] = -av[c];
        } else {
#pragma unroll
            for (int c = 0; c < kLeaf; ++c) Cs[slot(t, c)] = -av[c];
        }
    } else {
        if (swap) {
#pragma unroll
            for (int c = 0; c < kLeaf; ++c) Cs[slot(c, t)] = -(keep(c, t) ? av[c] : T(0));
        } else {
#pragma unroll
            for (int c = 0; c < kLeaf; ++c) Cs[slot(t, c)] = -(keep(t, c) ? av[c] : T(0));
        }
    }
}

template <typename T, bool FWD, bool UNIT>
__global__ __launch_bounds__(kLeafThreads) void trsm_leaf_kernel(LeafArgs<T> p) {
    __shared__ __attribute__((aligned(16))) T Cs[kLeaf * kLdsK];  // Cs[k][q][j] = -a_{4j+q, k} (solve order)
    __shared__ T Dg[kLeaf];                      // diagonal in solve order (1 for padding)
    __shared__ T Xs[kLeafVecs * (kLeaf + 1)];    // side L: vectors staged for the transposition
    const int t = threadIdx.x, q = t & (kQuad - 1), vv = t / kQuad;
    const int v0 = blockIdx.x * kLeafVecs, v = v0 + vv;  // this quad's vector
    const int nb = p.nb;
    const int tc = min(t, nb - 1);
    auto idx = [](int u) { return FWD ? u : kLeaf - 1 - u; };  // solve order -> index in the block
    // Every load of a phase is issued before the first is used (one memory
    // round trip per phase).
    T av[kLeaf], dv;
    leaf_coeffs_load(av, dv, p.a + (int64_t)p.i0 * p.lda + p.i0, p.lda, nb, t);
    T x[kLeafU];
    if (p.left) {
        // vector v is column v of B, unknown i row i0 + i: the wave reads
        // columns (contiguous rows), transposed through LDS below
        T bv[kLeafVecs];
        const T* bblk = p.b + p.i0 + tc;
        if (v0 + kLeafVecs <= p.nvec) {
            const T* col = bblk + (int64_t)v0 * p.ldb;
#pragma unroll
            for (int w = 0; w < kLeafVecs; ++w, col += p.ldb) bv[w] = *col;
        } else {
#pragma unroll
            for (int w = 0; w < kLeafVecs; ++w) bv[w] = bblk[(int64_t)min(v0 + w, p.nvec - 1) * p.ldb];
        }
#pragma unroll
        for (int w = 0; w < kLeafVecs; ++w) Xs[w * (kLeaf + 1) + t] = bv[w];
    }
    // vector v is row v of B for side R, unknown i column i0 + i
    const T* bvec = p.b + (int64_t)p.i0 * p.ldb + min(v, p.nvec - 1);
    auto load_x = [&](T (&xx)[kLeafU]) {
        if (p.left) {
#pragma unroll
            for (int j = 0; j < kLeafU; ++j) xx[j] = Xs[vv * (kLeaf + 1) + idx(kQuad * j + q)];
        } else {
#pragma unroll
            for (int j = 0; j < kLeafU; ++j) xx[j] = bvec[(int64_t)min(idx(kQuad * j + q), nb - 1) * p.ldb];
        }
#pragma unroll
        for (int j = 0; j < kLeafU; ++j)
            if (idx(kQuad * j + q) >= nb) xx[j] = T(0);
    };
    if (!p.left) load_x(x);
    leaf_coeffs_store<T, FWD, UNIT>(Cs, Dg, av, dv, nb, t, (p.left != 0) == (p.trans != 0));
    __syncthreads();
    if (p.left) load_x(x);
    leaf_solve<T, UNIT>(x, Cs + q * kLdsQ, Dg, q, std::make_integer_sequence<int, kLeaf>{});
    if (p.left) {
        __syncthreads();
#pragma unroll
        for (int j = 0; j < kLeafU; ++j) Xs[vv * (kLeaf + 1) + idx(kQuad * j + q)] = x[j];
        __syncthreads();
        T* bblk = p.b + p.i0 + t;
        if (t < nb) {
            if (v0 + kLeafVecs <= p.nvec) {
                T* col = bblk + (int64_t)v0 * p.ldb;
#pragma unroll
                for (int w = 0; w < kLeafVecs; ++w, col += p.ldb) *col = Xs[w * (kLeaf + 1) + t];
            } else {
#pragma unroll
                for (int w = 0; w < kLeafVecs; ++w)
                    if (v0 + w < p.nvec) bblk[(int64_t)(v0 + w) * p.ldb] = Xs[w * (kLeaf + 1) + t];
            }
        }
    } else if (v < p.nvec) {
#pragma unroll
        for (int j = 0; j < kLeafU; ++j)
            if (idx(kQuad * j + q) < nb) p.b[(int64_t)(p.i0 + idx(kQuad * j + q)) * p.ldb + v] = x[j];
    }
}

// ---------------------------------------------------------------------------
// block solve: up to 256 unknowns per vector in one launch
// ---------------------------------------------------------------------------
// Below 256 unknowns the recursion's launches are latency-bound: a 64-row
// leaf over 8192 vectors and the 64- and 128-deep updates between leaves
// each take 10-15 us on MI355X for a few microseconds of work.  The block
// kernel keeps a workgroup's 16 vectors (all 256 unknowns) in LDS and runs
// the same recursion inside: each 64-block is solved by the leaf's quad
// solve, and its coupling to every later 64-block of the 256 is subtracted
// with 16 x 16 x 4 MFMAs, the workgroup's 16 vectors being the MFMA's 16
// columns (B_t -= M_ts X_s, as the recursion's GEMM updates).
//
// Two waves: wave 0 solves, wave 1 updates.  Step k (64-blocks in solve
// order) has two phases.  A(k): wave 1 subtracts the coupling of block k to
// block k - 1 (the near update, whose coefficients it read during B(k - 1)),
// while wave 0 writes block k's coefficients to LDS (read at the end of
// B(k - 1)).  B(k): wave 0 solves block k; wave 1 reads the
// coefficients of the near update A(k + 1) and subtracts the coupling of
// blocks k + 1 and later to block k - 1 (the far updates), which the solve
// of block k does not need.  Only the solves and the near updates are on
// the critical path.
constexpr int kBlock = 256;
constexpr int kBlockThreads = 2 * kLeafThreads;
constexpr int kXsV = kBlock + 4;  // LDS row per vector (4 mod 32 doubles: the MFMA reads spread over the banks)
constexpr int kTileK = kLeaf / 4;  // MFMA k-steps per 64-block

// row of item r of lane l in a 16 x 16 MFMA result (column l % 16)
template <typename T>
__device__ __forceinline__ int mfma_d_row(int r, int l);
template <>
__device__ __forceinline__ int mfma_d_row<double>(int r, int l) { return 4 * r + (l >> 4); }
template <>
__device__ __forceinline__ int mfma_d_row<float>(int r, int l) { return 4 * (l >> 4) + r; }

template <typename T, bool FWD, bool UNIT>
__global__ __launch_bounds__(kBlockThreads) void trsm_block_kernel(LeafArgs<T> p) {
    __shared__ __attribute__((aligned(16))) T Cs[kLeaf * kLdsK];  // current 64-block, as in the leaf
    __shared__ T Dg[kLeaf];
    __shared__ T Xs[kLeafVecs * kXsV];  // Xs[w][i]: unknown i of vector v0 + w (0 past len)
    const int tid = threadIdx.x;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int t = tid & (kLeafThreads - 1), q = t & (kQuad - 1), vv = t / kQuad;
    const int lr = t & 15, lk = t >> 4;  // MFMA operand lane coordinates
    const int v0 = blockIdx.x * kLeafVecs;
    const int len = p.nb, nsub = (len + kLeaf - 1) / kLeaf;
    const bool swap = (p.left != 0) == (p.trans != 0);
    const gptr<const T> a0 = to_global(p.a) + (int64_t)p.i0 * p.lda + p.i0;
    const gptr<T> b = to_global(p.b);
    auto sub = [&](int k) { return FWD ? k : nsub - 1 - k; };  // k-th 64-block in solve order
    auto idx = [](int u) { return FWD ? u : kLeaf - 1 - u; };

    // wave 0: the 64-blocks' own coefficients
    T av[kLeaf], dv;
    auto load_block = [&](int k) {  // coefficients of the k-th 64-block into av, dv
        const int s0 = sub(k) * kLeaf;
        int64_t ld = p.lda;
        asm volatile("" : "+s"(ld));  // the 64 column offsets are recomputed per block, not kept in registers
        leaf_coeffs_load(av, dv, a0 + (int64_t)s0 * ld + s0, ld, min(kLeaf, len - s0), t);
    };
    auto store_block = [&](int k) {
        leaf_coeffs_store<T, FWD, UNIT>(Cs, Dg, av, dv, min(kLeaf, len - sub(k) * kLeaf), t, swap);
    };

    // wave 1: coupling tiles.  Tile (kt, rt) of source ks: rows sub(kt) * 64 +
    // 16 rt + [0, 16) of M against the source block's 64 unknowns; A operand:
    // lane l holds M[row0 + l % 16][s0 + 4 kk + l / 16].  Indices are clamped
    // into [0, len): a row past len only feeds a padding unknown (solved
    // last, coupled to nothing, never stored); a column past len (a partial
    // first block of a backward solve) is masked to zero before use.
    const int64_t si = swap ? p.lda : 1, sk = swap ? 1 : p.lda;
    auto load_tile = [&](int kt, int rt, int ks, T (&am)[kTileK]) {
        const int s0 = sub(ks) * kLeaf;
        const gptr<const T> ai = a0 + (int64_t)min(sub(kt) * kLeaf + 16 * rt + lr, len - 1) * si;
#pragma unroll
        for (int kk = 0; kk < kTileK; ++kk) am[kk] = ai[(int64_t)min(s0 + 4 * kk + lk, len - 1) * sk];
        __builtin_amdgcn_sched_barrier(0);  // issued here, consumed later
    };
    // B_t -= M_ts X_s as B_t + M_ts (-X_s); xb is -X_s as the B operand:
    // lane l holds -X[s0 + 4 kk + l / 16][vector l % 16]
    auto mma_tile = [&](int kt, int rt, int ks, T (&am)[kTileK], const T (&xb)[kTileK]) {
        const int s0 = sub(ks) * kLeaf;
        if (len - s0 < kLeaf) {
#pragma unroll
            for (int kk = 0; kk < kTileK; ++kk)
                if (s0 + 4 * kk + lk >= len) am[kk] = T(0);
        }
        const int i0 = sub(kt) * kLeaf + 16 * rt;
        typename Mfma<T>::acc_t acc;
#pragma unroll
        for (int r = 0; r < 4; ++r) acc[r] = Xs[lr * kXsV + i0 + mfma_d_row<T>(r, t)];
#pragma unroll
        for (int kk = 0; kk < kTileK; ++kk) acc = Mfma<T>::mma(am[kk], xb[kk], acc);
#pragma unroll
        for (int r = 0; r < 4; ++r) Xs[lr * kXsV + i0 + mfma_d_row<T>(r, t)] = acc[r];
    };
    // The vectors' unknowns of one 64-block, rows [r0, r0 + 64), between B
    // and Xs, by NTH threads (this one is th): every load (or LDS read) is
    // issued before the first store.  Rows past len are 0 in Xs and not
    // written back; vectors past nvec are not written back.
    auto rows_in = [&](int r0, auto nth_c, int th) {
        constexpr int NTH = decltype(nth_c)::value, E = kLeaf * kLeafVecs / NTH;
        T xv[E];
        if (p.left) {
            // consecutive threads read consecutive rows of one column
#pragma unroll
            for (int e = 0; e < E; ++e) {
                const int f = th + NTH * e, w = f / kLeaf, i = r0 + f % kLeaf;
                xv[e] = b[(int64_t)min(v0 + w, p.nvec - 1) * p.ldb + p.i0 + min(i, len - 1)];
            }
#pragma unroll
            for (int e = 0; e < E; ++e) {
                const int f = th + NTH * e, w = f / kLeaf, i = r0 + f % kLeaf;
                Xs[w * kXsV + i] = (i < len) ? xv[e] : T(0);
            }
        } else {
            // consecutive threads read consecutive vectors of one row
            const int w = th % kLeafVecs;
            const gptr<const T> col = b + (int64_t)p.i0 * p.ldb + min(v0 + w, p.nvec - 1);
#pragma unroll
            for (int e = 0; e < E; ++e) {
                const int i = r0 + (th + NTH * e) / kLeafVecs;
                xv[e] = col[(int64_t)min(i, len - 1) * p.ldb];
            }
#pragma unroll
            for (int e = 0; e < E; ++e) {
                const int i = r0 + (th + NTH * e) / kLeafVecs;
                Xs[w * kXsV + i] = (i < len) ? xv[e] : T(0);
            }
        }
    };
    auto rows_out = [&](int r0, auto nth_c, int th) {
        constexpr int NTH = decltype(nth_c)::value, E = kLeaf * kLeafVecs / NTH;
        T xv[E];
        if (p.left) {
#pragma unroll
            for (int e = 0; e < E; ++e) {
                const int f = th + NTH * e;
                xv[e] = Xs[(f / kLeaf) * kXsV + r0 + f % kLeaf];
            }
#pragma unroll
            for (int e = 0; e < E; ++e) {
                const int f = th + NTH * e, w = f / kLeaf, i = r0 + f % kLeaf;
                if (v0 + w < p.nvec && i < len) b[(int64_t)(v0 + w) * p.ldb + p.i0 + i] = xv[e];
            }
        } else {
            const int w = th % kLeafVecs;
#pragma unroll
            for (int e = 0; e < E; ++e) xv[e] = Xs[w * kXsV + r0 + (th + NTH * e) / kLeafVecs];
            if (v0 + w < p.nvec) {
                const gptr<T> col = b + (int64_t)p.i0 * p.ldb + v0 + w;
#pragma unroll
                for (int e = 0; e < E; ++e) {
                    const int i = r0 + (th + NTH * e) / kLeafVecs;
                    if (i < len) col[(int64_t)i * p.ldb] = xv[e];
                }
            }
        }
    };
    using Both = std::integral_constant<int, kBlockThreads>;
    using One = std::integral_constant<int, kLeafThreads>;

    if (wave == 0) load_block(0);
    // the first-solved block's unknowns (both waves); wave 1 reads the others
    // during its solve, and writes each block back during the next one's
    rows_in(sub(0) * kLeaf, Both{}, tid);
    // Each wave runs its own loop (its registers are allocated for its own
    // work, not for both); the two loops meet at the same barriers, two per
    // 64-block, which end phases A(k) and B(k).
    if (wave == 0) {
        for (int k = 0; k < nsub; ++k) {
            store_block(k);
            __syncthreads();  // end of A(k)
            const int r0 = sub(k) * kLeaf;
            T x[kLeafU];
#pragma unroll
            for (int j = 0; j < kLeafU; ++j) x[j] = Xs[vv * kXsV + r0 + idx(kQuad * j + q)];
            leaf_solve<T, UNIT>(x, Cs + q * kLdsQ, Dg, q, std::make_integer_sequence<int, kLeaf>{});
#pragma unroll
            for (int j = 0; j < kLeafU; ++j) Xs[vv * kXsV + r0 + idx(kQuad * j + q)] = x[j];
            // the next block's coefficients: in flight during the barrier and
            // the near update (read during the solve, they cost it issue slots)
            if (k + 1 < nsub) load_block(k + 1);
            __syncthreads();  // end of B(k)
        }
    } else {
        T xb[kTileK];     // -X of the last solved block
        T nm[4][kTileK];  // the near update's coefficients
        for (int k = 0; k < nsub; ++k) {
            if (k > 0) {
#pragma unroll
                for (int kk = 0; kk < kTileK; ++kk) xb[kk] = -Xs[lr * kXsV + sub(k - 1) * kLeaf + 4 * kk + lk];
#pragma unroll
                for (int rt = 0; rt < 4; ++rt) mma_tile(k, rt, k - 1, nm[rt], xb);
            }
            __syncthreads();  // end of A(k)
            if (k + 1 < nsub) {
#pragma unroll
                for (int rt = 0; rt < 4; ++rt) load_tile(k + 1, rt, k, nm[rt]);
            }
            if (k == 0) {
                // the other blocks' unknowns, before the first update reads them
#pragma unroll 1
                for (int k2 = 1; k2 < nsub; ++k2) rows_in(sub(k2) * kLeaf, One{}, t);
            } else {
                rows_out(sub(k - 1) * kLeaf, One{}, t);  // final since B(k - 1)
            }
            // far updates: blocks k + 1 .. nsub - 1 from block k - 1, two
            // operand sets (a tile's loads in flight during the previous tile's MFMAs)
            const int ntiles = k > 0 ? (nsub - 1 - k) * 4 : 0;
            if (ntiles > 0) {
                T am0[kTileK], am1[kTileK];
                load_tile(k + 1, 0, k - 1, am0);
                for (int j = 0; j < ntiles; j += 2) {
                    load_tile(k + 1 + (j + 1) / 4, (j + 1) % 4, k - 1, am1);
                    mma_tile(k + 1 + j / 4, j % 4, k - 1, am0, xb);
                    const int jn = min(j + 2, ntiles - 1);
                    load_tile(k + 1 + jn / 4, jn % 4, k - 1, am0);
                    mma_tile(k + 1 + (j + 1) / 4, (j + 1) % 4, k - 1, am1, xb);
                }
            }
            __syncthreads();  // end of B(k)
        }
    }
    // the last-solved block back to B (both waves)
    rows_out(sub(nsub - 1) * kLeaf, Both{}, tid);
}

// ---------------------------------------------------------------------------
// recursion
// ---------------------------------------------------------------------------
namespace {
template <typename T>
struct TrsmCtx {
    bool left, lower_eff, trans, unit;
    int m, n;
    const T* a;
    int64_t lda;
    T* b;
    int64_t ldb;
    hipStream_t s;
    int mode, mode_param;
    // lookahead (plain updates only): the second stream and the call's own
    // fork / join events; null when not used
    hipStream_t side = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
};

// op(A)[r0.., c0..] as a GEMM operand (pointer, transpose flag)
template <typename T>
const T* sub_opa(const TrsmCtx<T>& c, int r0, int c0) {
    return c.trans ? c.a + r0 * c.lda + c0 : c.a + c0 * c.lda + r0;
}

// smallest m, n and k of an off-diagonal update that runs in split precision
#ifndef ZG_SPLIT_UPDATE_MIN
#define ZG_SPLIT_UPDATE_MIN 1024
#endif
constexpr int kSplitUpdateMin = ZG_SPLIT_UPDATE_MIN;

template <typename T>
int gemm_update(const TrsmCtx<T>& c, bool ta, bool tb, int m, int n, int k, const T* a, int64_t lda, const T* b,
                int64_t ldb, T* cc, int64_t ldc) {
    if (m <= 0 || n <= 0 || k <= 0) return ZYNUM_GPU_OK;
    if constexpr (sizeof(T) == 8) {
        // Split precision pays only on large updates: its operand preparation
        // and reconstruction are fixed costs that the small updates at the
        // bottom of the recursion cannot amortise, so those run on FP64 MFMA
        // (same or smaller error).
        if (c.mode != ZYNUM_GPU_MODE_PLAIN && std::min(std::min(m, n), k) >= kSplitUpdateMin)
            return ozaki_dgemm(ta, tb, m, n, k, -1.0, a, lda, b, ldb, 1.0, cc, ldc, c.s, c.mode, c.mode_param);
    }
    GemmDesc<T> d;
    d.ta = ta; d.tb = tb; d.m = m; d.n = n; d.k = k;
    d.lda = lda; d.ldb = ldb; d.ldc = ldc;
    d.alpha = T(-1); d.beta = T(1);
    d.a = a; d.b = b; d.c = cc;
    // The thin updates at the bottom of the recursion (at most 256 rows, 256
    // deep) are latency-bound: the 64 x 64 tile with BK 32 keeps more
    // workgroups in flight than the tuned large-shape choice (MI355X, 256 x
    // 8192 x 256: 34 us against 53 us).
    const int tile = (std::min(m, n) <= 256 && k <= 256) ? kTile64x64k32 : kTileAuto;
    return status_from_hip(gemm_launch<T>(d, c.s, tile));
}

template <typename T>
int leaf(const TrsmCtx<T>& c, int i0, int nb) {
    LeafArgs<T> p;
    p.a = c.a; p.lda = c.lda; p.b = c.b; p.ldb = c.ldb;
    p.left = c.left; p.lower_eff = c.lower_eff; p.trans = c.trans; p.unit = c.unit;
    p.i0 = i0; p.nb = nb;
    p.nvec = c.left ? c.n : c.m;
    const int grid = (p.nvec + kLeafVecs - 1) / kLeafVecs;
    // forward substitution when the equations are lower triangular
    const bool fwd = c.left ? c.lower_eff : !c.lower_eff;
    if (fwd) {
        if (c.unit) trsm_leaf_kernel<T, true, true><<<grid, kLeafThreads, 0, c.s>>>(p);
        else trsm_leaf_kernel<T, true, false><<<grid, kLeafThreads, 0, c.s>>>(p);
    } else {
        if (c.unit) trsm_leaf_kernel<T, false, true><<<grid, kLeafThreads, 0, c.s>>>(p);
        else trsm_leaf_kernel<T, false, false><<<grid, kLeafThreads, 0, c.s>>>(p);
    }
    return status_from_hip(hipGetLastError());
}

// Solves the unknowns [i0, i0 + len), 64 < len <= kBlock, in one launch.
template <typename T>
int block(const TrsmCtx<T>& c, int i0, int len) {
    LeafArgs<T> p;
    p.a = c.a; p.lda = c.lda; p.b = c.b; p.ldb = c.ldb;
    p.left = c.left; p.lower_eff = c.lower_eff; p.trans = c.trans; p.unit = c.unit;
    p.i0 = i0; p.nb = len;
    p.nvec = c.left ? c.n : c.m;
    const int grid = (p.nvec + kLeafVecs - 1) / kLeafVecs;
    const bool fwd = c.left ? c.lower_eff : !c.lower_eff;
    if (fwd) {
        if (c.unit) trsm_block_kernel<T, true, true><<<grid, kBlockThreads, 0, c.s>>>(p);
        else trsm_block_kernel<T, true, false><<<grid, kBlockThreads, 0, c.s>>>(p);
    } else {
        if (c.unit) trsm_block_kernel<T, false, true><<<grid, kBlockThreads, 0, c.s>>>(p);
        else trsm_block_kernel<T, false, false><<<grid, kBlockThreads, 0, c.s>>>(p);
    }
    return status_from_hip(hipGetLastError());
}

// A range of unknowns (rows of X for side L, columns for side R).
struct Range {
    int lo, len;
};

// The two halves of the unknowns [i0, i0 + len) in solve order: f is solved
// first.  Forward when the equations are lower triangular (side L with op(A)
// lower, side R with op(A) upper).
template <typename T>
void halves(const TrsmCtx<T>& c, int i0, int len, Range& f, Range& g) {
    int h = (len / 2 + kLeaf - 1) / kLeaf * kLeaf;
    if (h >= len) h = len - kLeaf;
    const bool forward = c.left ? c.lower_eff : !c.lower_eff;
    if (forward) {
        f = {i0, h};
        g = {i0 + h, len - h};
    } else {
        f = {i0 + h, len - h};
        g = {i0, h};
    }
}

// Unknowns r -= coupling to the solved unknowns s, on stream st:
// side L, B[r, :] -= op(A)[r, s] X[s, :]; side R, B[:, r] -= X[:, s] op(A)[s, r].
template <typename T>
int update_range(const TrsmCtx<T>& c, Range r, Range s, hipStream_t st) {
    TrsmCtx<T> cs = c;
    cs.s = st;
    if (c.left)
        return gemm_update(cs, c.trans, false, r.len, c.n, s.len, sub_opa(c, r.lo, s.lo), c.lda, c.b + s.lo, c.ldb,
                           c.b + r.lo, c.ldb);
    return gemm_update(cs, false, c.trans, c.m, r.len, s.len, c.b + (int64_t)s.lo * c.ldb, c.ldb,
                       sub_opa(c, s.lo, r.lo), c.lda, c.b + (int64_t)r.lo * c.ldb, c.ldb);
}

// Smallest second half that is solved with lookahead.
constexpr int kTrsmLookaheadMin = 2048;

// Solves the unknowns [i0, i0 + len): the first half, the update of the
// second half, the second half (structured_blocked.zig:315-374 blocks the
// same way).  With lookahead (a second stream, plain updates, a second half
// of at least kTrsmLookaheadMin), the second half's update is split at the
// second half's own first half g1: g1's part runs on the main stream and is
// solved there while the rest of the update (g2's part, the larger GEMM)
// runs on the second stream; the main stream waits for it only before g2's
// next update.  The leaf solves and thin updates inside g1 are latency-bound
// and leave most of the chip to that GEMM.  Every update writes what it
// writes without lookahead.  A lookahead is not nested inside g1's solve
// (allow_la), so one update at a time is on the second stream.
template <typename T>
int solve(const TrsmCtx<T>& c, int i0, int len, bool allow_la = true) {
    if (len <= kLeaf) return leaf(c, i0, len);
    if (len <= kBlock) return block(c, i0, len);
    Range f, g;
    halves(c, i0, len, f, g);
    int st;
    if ((st = solve(c, f.lo, f.len, allow_la))) return st;
    if (!(allow_la && c.side != nullptr && g.len >= kTrsmLookaheadMin)) {
        if ((st = update_range(c, g, f, c.s))) return st;
        return solve(c, g.lo, g.len, allow_la);
    }
    Range g1, g2;
    halves(c, g.lo, g.len, g1, g2);
    // g2's part of the update on the second stream, after everything queued so far
    if ((st = status_from_hip(hipEventRecord(c.ev_fork, c.s)))) return st;
    if ((st = status_from_hip(hipStreamWaitEvent(c.side, c.ev_fork, 0)))) return st;
    if ((st = update_range(c, g2, f, c.side))) return st;
    if ((st = status_from_hip(hipEventRecord(c.ev_join, c.side)))) return st;
    // g1's part and its solve on the main stream, beside it
    if ((st = update_range(c, g1, f, c.s))) return st;
    if ((st = solve(c, g1.lo, g1.len, false))) return st;
    if ((st = status_from_hip(hipStreamWaitEvent(c.s, c.ev_join, 0)))) return st;
    if ((st = update_range(c, g2, g1, c.s))) return st;
    return solve(c, g2.lo, g2.len, allow_la);
}
}  // namespace

template <typename T>
int trsm_device(char side, char uplo, bool trans, char diag, int m, int n, T alpha, const T* a, int64_t lda, T* b,
                int64_t ldb, hipStream_t s, int mode, int mode_param) {
    if (m <= 0 || n <= 0) return ZYNUM_GPU_OK;
    if (alpha != T(1)) {
        const int64_t total = (int64_t)m * n;
        const int blocks = (int)std::min<int64_t>(65536, (total + 255) / 256);
        mul_kernel<T><<<blocks, 256, 0, s>>>(m, n, alpha, b, ldb);
        const int st = status_from_hip(hipGetLastError());
        if (st) return st;
    }
    TrsmCtx<T> c;
    c.left = (side == 'L');
    c.lower_eff = (uplo == 'L') != trans;
    c.trans = trans;
    c.unit = (diag == 'U');
    c.m = m; c.n = n; c.a = a; c.lda = lda; c.b = b; c.ldb = ldb; c.s = s;
    c.mode = mode; c.mode_param = mode_param;
    const int len = c.left ? m : n;
    // lookahead needs a second half of at least kTrsmLookaheadMin below the top
    if (mode == ZYNUM_GPU_MODE_PLAIN && len >= 2 * kTrsmLookaheadMin) {
        hipStream_t side = aux_stream();
        if (side != nullptr && hipEventCreateWithFlags(&c.ev_fork, hipEventDisableTiming) == hipSuccess &&
            hipEventCreateWithFlags(&c.ev_join, hipEventDisableTiming) == hipSuccess) {
            c.side = side;
        }
        (void)hipGetLastError();
    }
    const int st = solve(c, 0, len);
    // (destroying an event the device has yet to reach is deferred by the runtime)
    if (c.ev_fork) (void)hipEventDestroy(c.ev_fork);
    if (c.ev_join) (void)hipEventDestroy(c.ev_join);
    return st;
}

// ---------------------------------------------------------------------------
// SYRK
// ---------------------------------------------------------------------------
// The large-shape GEMM tile: 16 waves of 32x32, two workgroups per CU; BK 16
// for FP64, 32 for FP32 (the same bytes per slice).
template <typename T>
using SyrkCfg = TileCfg<128, 128, (sizeof(T) == 8) ? 16 : 32, 4, 4, 4>;

// Triangle tiles in a cache-friendly order: the lower triangle of the tn x tn
// tile grid is cut into 8 x 8 super-tiles, walked in triangular order, with
// the tiles of each super-tile in row order; the workgroup index is first
// remapped so each XCD takes a contiguous run of that order.  The workgroups
// running together on one XCD then share a few A and B panels, as the grouped
// order of the plain GEMM does.  TRI 2 (upper) takes the transposed tile.
constexpr int kSyrkGroup = 8;

template <int TRI>
__device__ __forceinline__ void tri_tile(int t, int tn, int& ti, int& tj) {
    const int G = kSyrkGroup, ns = (tn + G - 1) / G;
    int I = 0, J = 0;
    for (I = 0; I < ns; ++I) {
        const int r0 = I * G, r1 = min(tn, r0 + G);
        bool found = false;
        for (J = 0; J <= I; ++J) {
            const int c0 = J * G, c1 = min(tn, c0 + G);
            int cnt = 0;  // tiles (i, j) of the super-tile with j <= i
            if (J < I) cnt = (r1 - r0) * (c1 - c0);
            else for (int i = r0; i < r1; ++i) cnt += min(c1, i + 1) - c0;
            if (t < cnt) { found = true; break; }
            t -= cnt;
        }
        if (found) break;
    }
    const int r0 = I * G, c0 = J * G, c1 = min(tn, c0 + G);
    int i = r0;
    for (;; ++i) {
        const int w = (J < I) ? (c1 - c0) : (min(c1, i + 1) - c0);
        if (t < w) break;
        t -= w;
    }
    const int j = c0 + t;
    ti = (TRI == 1) ? i : j;
    tj = (TRI == 1) ? j : i;
}

template <typename T, bool TA, int TRI, int VEC>
__global__ __launch_bounds__(SyrkCfg<T>::NT) __attribute__((amdgpu_waves_per_eu(SyrkCfg<T>::OCC)))
void syrk_kernel(GemmArgs<T> p, int tiles) {
    using CF = SyrkCfg<T>;
    __shared__ __attribute__((aligned(16))) T lds_a[tile_lds_a<T, CF>()];
    __shared__ __attribute__((aligned(16))) T lds_b[tile_lds_b<T, CF>()];
    int ti, tj;
    tri_tile<TRI>(xcd_remap(blockIdx.x, tiles), p.tiles_m, ti, tj);
    // C = alpha * op(A) op(A)^T + beta * C: the B operand is A with the other transpose
    gemm_tile<T, CF, TA, !TA, VEC, TRI>(p, ti * CF::BM, tj * CF::BN, lds_a, lds_b);
}

// FP64 triangle tiles on the direct-load tile (full tiles, aligned operands).
// 8 waves of 32x64 (the products are NT or TN, where 8 waves are fastest).
template <bool TA, int TRI>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(4))) void syrk_dma_kernel(GemmArgs<double> p,
                                                                                                int tiles) {
    int ti, tj;
    tri_tile<TRI>(xcd_remap(blockIdx.x, tiles), p.tiles_m, ti, tj);
    gemm_tile_dma<TA, !TA, 8, TRI>(p, ti * 128, tj * 128);
}

// Triangle-tile launch: the direct-load tile for FP64 full tiles, else the
// register-staged one with 16-byte operand loads when both operands allow them.
template <typename T>
hipError_t syrk_launch(const GemmArgs<T>& p, int tiles, bool lower, bool trans, hipStream_t s) {
    if constexpr (sizeof(T) == 8) {
        if (dma_shape_ok(p.m, p.n, p.k, p.a, p.lda, p.b, p.ldb)) {
            if (lower) {
                if (trans) syrk_dma_kernel<true, 1><<<tiles, 512, 0, s>>>(p, tiles);
                else syrk_dma_kernel<false, 1><<<tiles, 512, 0, s>>>(p, tiles);
            } else {
                if (trans) syrk_dma_kernel<true, 2><<<tiles, 512, 0, s>>>(p, tiles);
                else syrk_dma_kernel<false, 2><<<tiles, 512, 0, s>>>(p, tiles);
            }
            return hipGetLastError();
        }
    }
    constexpr int V = 16 / sizeof(T);
    const bool vec = ((reinterpret_cast<uintptr_t>(p.a) & 15) == 0) && ((reinterpret_cast<uintptr_t>(p.b) & 15) == 0) &&
                     (p.lda % V == 0) && (p.ldb % V == 0);
#define ZG_SYRK(TA, TRI, VV) syrk_kernel<T, TA, TRI, VV><<<tiles, SyrkCfg<T>::NT, 0, s>>>(p, tiles)
    if (vec) {
        if (lower) { if (trans) ZG_SYRK(true, 1, V); else ZG_SYRK(false, 1, V); }
        else { if (trans) ZG_SYRK(true, 2, V); else ZG_SYRK(false, 2, V); }
    } else {
        if (lower) { if (trans) ZG_SYRK(true, 1, 1); else ZG_SYRK(false, 1, 1); }
        else { if (trans) ZG_SYRK(true, 2, 1); else ZG_SYRK(false, 2, 1); }
    }
#undef ZG_SYRK
    return hipGetLastError();
}

// stored triangle of C <- W + beta * C (W = alpha * op(A) op(A)^T computed in
// full), or C <- beta * C when w is null (k == 0 or alpha == 0, exact 0 for
// beta == 0 as symmetric.zig:196)
template <typename T>
__global__ void syrk_commit_kernel(int n, int lower, const T* __restrict__ w, int64_t ldw, T beta, T* __restrict__ c,
                                   int64_t ldc) {
    const int i = blockIdx.x * 256 + threadIdx.x, j = blockIdx.y;
    if (i >= n || (lower ? i < j : i > j)) return;
    T* cp = c + (int64_t)j * ldc + i;
    const T s = (beta == T(0)) ? T(0) : beta * *cp;
    *cp = (w != nullptr) ? w[(int64_t)j * ldw + i] + s : s;
}

template <typename T>
int syrk_device(char uplo, bool trans, int n, int k, T alpha, const T* a, int64_t lda, T beta, T* c, int64_t ldc,
                hipStream_t s, int mode, int mode_param) {
    if (n <= 0) return ZYNUM_GPU_OK;
    const bool lower = (uplo == 'L');
    if (k <= 0 || alpha == T(0)) {
        // C = beta * C on the stored triangle only
        dim3 grid((n + 255) / 256, n);
        syrk_commit_kernel<T><<<grid, 256, 0, s>>>(n, lower, nullptr, 0, beta, c, ldc);
        return status_from_hip(hipGetLastError());
    }
    if (sizeof(T) == 8 && mode != ZYNUM_GPU_MODE_PLAIN) {
        if constexpr (sizeof(T) == 8) {
            T* w = nullptr;
            if (hipMalloc(&w, (size_t)n * n * sizeof(T)) != hipSuccess) {
                (void)hipGetLastError();
                return ZYNUM_GPU_DECLINE_ALLOC;
            }
            int st = ozaki_dgemm(trans, !trans, n, n, k, alpha, a, lda, a, lda, 0.0, w, n, s, mode, mode_param);
            if (st == ZYNUM_GPU_OK) {
                dim3 grid((n + 255) / 256, n);
                syrk_commit_kernel<T><<<grid, 256, 0, s>>>(n, lower, w, n, beta, c, ldc);
                st = status_from_hip(hipGetLastError());
                if (st == ZYNUM_GPU_OK) st = status_from_hip(hipStreamSynchronize(s));
            }
            (void)hipFree(w);
            return st;
        }
    }
    GemmArgs<T> p;
    p.m = n; p.n = n; p.k = k;
    p.lda = lda; p.ldb = lda; p.ldc = ldc;
    p.alpha = alpha; p.beta = beta;
    p.a = a; p.b = a; p.c = c;
    const int tn = (n + SyrkCfg<T>::BM - 1) / SyrkCfg<T>::BM;
    p.tiles_m = p.tiles_n = tn;
    p.group_m = 1;
    const int tiles = tn * (tn + 1) / 2;
    return status_from_hip(syrk_launch<T>(p, tiles, lower, trans, s));
}

// ---------------------------------------------------------------------------
// SYMM, TRMM, SYR2K on the GEMM core
// ---------------------------------------------------------------------------
inline int64_t ld_round(int64_t rows) { return (rows + 7) / 8 * 8; }
inline size_t bytes_round(size_t b) { return (b + 255) / 256 * 256; }

// S (n x n) <- the full symmetric matrix whose stored triangle is A's
// (symValue, access.zig:31-42).  One 32 x 32 tile per workgroup: the mirrored
// tile of A is read along its contiguous dimension and transposed in LDS.
template <typename T>
__global__ __launch_bounds__(256) void symm_fill_kernel(int n, int lower, const T* __restrict__ a, int64_t lda,
                                                        T* __restrict__ s, int64_t lds, int tiles) {
    __shared__ T t[32][33];
    const int bi = (int)(blockIdx.x % (unsigned)tiles), bj = (int)(blockIdx.x / (unsigned)tiles);
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
#pragma unroll
    for (int c = ty; c < 32; c += 8) {  // t[c][r] = A(bj*32 + r, bi*32 + c)
        const int r = bj * 32 + tx, cc = bi * 32 + c;
        if (r < n && cc < n) t[c][tx] = a[(int64_t)cc * lda + r];
    }
    __syncthreads();
#pragma unroll
    for (int c = ty; c < 32; c += 8) {  // S(i, j), i = bi*32 + tx, j = bj*32 + c
        const int i = bi * 32 + tx, j = bj * 32 + c;
        if (i < n && j < n) {
            const bool stored = lower ? (i >= j) : (i <= j);
            s[(int64_t)j * lds + i] = stored ? a[(int64_t)j * lda + i] : t[tx][c];
        }
    }
}

// T (n x n) <- the triangle of A, zeros elsewhere, ones on a unit diagonal
// (triValue, access.zig:54-63, without the transpose).
template <typename T>
__global__ void trmm_fill_kernel(int n, int lower, int unit, const T* __restrict__ a, int64_t lda, T* __restrict__ t,
                                 int64_t ldt) {
    const int64_t total = (int64_t)n * n;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t j = e / n, i = e % n;
        T v = T(0);
        if (i == j) v = unit ? T(1) : a[j * lda + i];
        else if (lower ? (i > j) : (i < j)) v = a[j * lda + i];
        t[j * ldt + i] = v;
    }
}

// C = alpha * S * B + beta * C (side L, S m x m) or alpha * B * S + beta * C
// (side R, S n x n), S symmetric from the stored triangle of A.  The reference
// (symmetric.zig:56-70) forms each element as one sum over the full row of S;
// here S is expanded once into the workspace and the tuned GEMM runs on it.
// The reference has no alpha == 0 quick return (alpha * sum is formed, so a
// NaN or an infinity in S or B reaches C): neither has this product.
template <typename T>
int symm_device(char side, char uplo, int m, int n, T alpha, const T* a, int64_t lda, const T* b, int64_t ldb, T beta,
                T* c, int64_t ldc, hipStream_t s) {
    if (m <= 0 || n <= 0) return ZYNUM_GPU_OK;
    const bool left = (side == 'L');
    GemmDesc<T> d;
    d.m = m; d.n = n;
    d.alpha = alpha; d.beta = beta;
    d.c = c; d.ldc = ldc;
    const int na = left ? m : n;
    const int64_t lds = ld_round(na);
    T* sym = static_cast<T*>(workspace_acquire(bytes_round((size_t)lds * na * sizeof(T)), s));
    if (!sym) return ZYNUM_GPU_DECLINE_ALLOC;
    const int tiles = (na + 31) / 32;
    const int64_t blocks = (int64_t)tiles * tiles;
    hipError_t e = hipSuccess;
    if (blocks > 0x7fffffff) e = hipErrorInvalidValue;
    if (e == hipSuccess) {
        symm_fill_kernel<T><<<(unsigned)blocks, 256, 0, s>>>(na, uplo == 'L', a, lda, sym, lds, tiles);
        e = hipGetLastError();
    }
    if (e == hipSuccess) {
        d.ta = false; d.tb = false;
        if (left) {
            d.k = m; d.a = sym; d.lda = lds; d.b = b; d.ldb = ldb;
        } else {
            d.k = n; d.a = b; d.lda = ldb; d.b = sym; d.ldb = lds;
        }
        e = gemm_launch_product<T>(d, s);
    }
    const int st = status_from_hip(e);
    workspace_release(s);
    return st;
}

// B = alpha * op(A) * B (side L) or alpha * B * op(A) (side R), A triangular.
// The triangle is expanded into a zero-filled square matrix and B copied
// aside, then the product runs as GEMMs over row (L) or column (R) panels of
// the result, each with only the K range where op(A) is not zero, so the work
// is about half a full GEMM.  triangular.zig:202-231 multiplies column by
// column (trmv) or row by row.
template <typename T>
int trmm_device(char side, char uplo, bool trans, char diag, int m, int n, T alpha, const T* a, int64_t lda, T* b,
                int64_t ldb, hipStream_t s) {
    if (m <= 0 || n <= 0) return ZYNUM_GPU_OK;
    if (alpha == T(0)) return status_from_hip(scale_launch<T>(m, n, T(0), b, ldb, s));
    const bool left = (side == 'L');
    const int na = left ? m : n;
    const int64_t ldt = ld_round(na), ldw = ld_round(m);
    const size_t t_bytes = bytes_round((size_t)ldt * na * sizeof(T));
    char* ws = static_cast<char*>(workspace_acquire(t_bytes + (size_t)ldw * n * sizeof(T), s));
    if (!ws) return ZYNUM_GPU_DECLINE_ALLOC;
    T* tri = reinterpret_cast<T*>(ws);
    T* b0 = reinterpret_cast<T*>(ws + t_bytes);
    const int64_t total = (int64_t)na * na;
    const int blocks = (int)std::min<int64_t>(65536, (total + 255) / 256);
    trmm_fill_kernel<T><<<blocks, 256, 0, s>>>(na, uplo == 'L', diag == 'U', a, lda, tri, ldt);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess)
        e = hipMemcpy2DAsync(b0, (size_t)ldw * sizeof(T), b, (size_t)ldb * sizeof(T), (size_t)m * sizeof(T), (size_t)n,
                             hipMemcpyDeviceToDevice, s);
    const bool lower_eff = (uplo == 'L') != trans;  // op(A) lower triangular
    const int extent = left ? m : n;               // panels along this dimension
    int panel = (extent + 7) / 8;
    panel = std::max(256, (panel + 127) / 128 * 128);
    for (int x0 = 0; e == hipSuccess && x0 < extent; x0 += panel) {
        const int x1 = std::min(extent, x0 + panel);
        GemmDesc<T> d;
        d.alpha = alpha;
        d.beta = T(0);
        if (left) {  // rows x0..x1 of op(A) are nonzero in columns [0, x1) (lower) or [x0, m) (upper)
            const int k0 = lower_eff ? 0 : x0, k1 = lower_eff ? x1 : m;
            d.m = x1 - x0; d.n = n; d.k = k1 - k0;
            d.ta = trans;
            d.a = trans ? tri + k0 + (int64_t)x0 * ldt : tri + x0 + (int64_t)k0 * ldt;
            d.lda = ldt;
            d.tb = false; d.b = b0 + k0; d.ldb = ldw;
            d.c = b + x0; d.ldc = ldb;
        } else {  // columns x0..x1 of op(A) are nonzero in rows [x0, n) (lower) or [0, x1) (upper)
            const int k0 = lower_eff ? x0 : 0, k1 = lower_eff ? n : x1;
            d.m = m; d.n = x1 - x0; d.k = k1 - k0;
            d.ta = false; d.a = b0 + (int64_t)k0 * ldw; d.lda = ldw;
            d.tb = trans;
            d.b = trans ? tri + x0 + (int64_t)k0 * ldt : tri + k0 + (int64_t)x0 * ldt;
            d.ldb = ldt;
            d.c = b + (int64_t)x0 * ldb; d.ldc = ldb;
        }
        e = gemm_launch<T>(d, s);
    }
    const int st = status_from_hip(e);
    workspace_release(s);
    return st;
}

// C_tri = alpha * (op(A) op(B)^T + op(B) op(A)^T) + beta * C_tri on the stored
// triangle (symmetric.zig:301-329): two launches of the triangle-tile kernel,
// the second adding alpha * op(B) op(A)^T to the first's result.
template <typename T>
int syr2k_device(char uplo, bool trans, int n, int k, T alpha, const T* a, int64_t lda, const T* b, int64_t ldb,
                 T beta, T* c, int64_t ldc, hipStream_t s) {
    if (n <= 0) return ZYNUM_GPU_OK;
    const bool lower = (uplo == 'L');
    if (k <= 0 || alpha == T(0)) {
        dim3 grid((n + 255) / 256, n);
        syrk_commit_kernel<T><<<grid, 256, 0, s>>>(n, lower, nullptr, 0, beta, c, ldc);
        return status_from_hip(hipGetLastError());
    }
    GemmArgs<T> p;
    p.m = n; p.n = n; p.k = k;
    p.ldc = ldc;
    p.alpha = alpha;
    p.c = c;
    const int tn = (n + SyrkCfg<T>::BM - 1) / SyrkCfg<T>::BM;
    p.tiles_m = p.tiles_n = tn;
    p.group_m = 1;
    const int tiles = tn * (tn + 1) / 2;
    hipError_t e = hipSuccess;
    for (int pass = 0; pass < 2 && e == hipSuccess; ++pass) {
        p.a = pass ? b : a;
        p.lda = pass ? ldb : lda;
        p.b = pass ? a : b;
        p.ldb = pass ? lda : ldb;
        p.beta = pass ? T(1) : beta;
        e = syrk_launch<T>(p, tiles, lower, trans, s);
    }
    return status_from_hip(e);
}

template int symm_device<double>(char, char, int, int, double, const double*, int64_t, const double*, int64_t, double,
                                 double*, int64_t, hipStream_t);
template int symm_device<float>(char, char, int, int, float, const float*, int64_t, const float*, int64_t, float,
                                float*, int64_t, hipStream_t);
template int trmm_device<double>(char, char, bool, char, int, int, double, const double*, int64_t, double*, int64_t,
                                 hipStream_t);
template int trmm_device<float>(char, char, bool, char, int, int, float, const float*, int64_t, float*, int64_t,
                                hipStream_t);
template int syr2k_device<double>(char, bool, int, int, double, const double*, int64_t, const double*, int64_t,
                                  double, double*, int64_t, hipStream_t);
template int syr2k_device<float>(char, bool, int, int, float, const float*, int64_t, const float*, int64_t, float,
                                 float*, int64_t, hipStream_t);
template int trsm_device<double>(char, char, bool, char, int, int, double, const double*, int64_t, double*, int64_t,
                                 hipStream_t, int, int);
template int trsm_device<float>(char, char, bool, char, int, int, float, const float*, int64_t, float*, int64_t,
                                hipStream_t, int, int);
template int syrk_device<double>(char, bool, int, int, double, const double*, int64_t, double, double*, int64_t,
                                 hipStream_t, int, int);
template int syrk_device<float>(char, bool, int, int, float, const float*, int64_t, float, float*, int64_t,
                                hipStream_t, int, int);

}  // namespace zg
