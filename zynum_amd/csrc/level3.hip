// DTRSM and DSYRK on the GEMM core.
//
// TRSM (triangular.zig:155-249 solves one column at a time with trsv and
// streams the whole triangle per right-hand side) is recast recursively, as
// in the reference's blocked template (structured_blocked.zig:315-374):
// split the triangle in two halves, solve the first, update the second with
// one GEMM (B2 -= op(A)21 X1, plain FP64 MFMA or split precision), recurse.
// Leaves of at most 64 rows are solved in FP64 from LDS: the 64x64 triangle is
// staged once per workgroup and each wave64 solves four vectors (columns of B
// for side L, rows for side R) as interleaved chains with one lane per
// unknown, in the reference's per-step order x_k /= a_kk;
// x_i = fma(a_ik, -x_k, x_i) (matrix_vector/triangular.zig:375-385).
//
// SYRK (symmetric.zig:180-279, one strided dot per element) runs as one launch
// over the lower (or upper) triangle of output tiles with the GEMM tile routine;
// diagonal tiles write only their stored triangle, so the other triangle is
// never touched (structured_blocked.zig:158-201 commits the same way).
#include <math.h>

#include <algorithm>

#include "common.hpp"
#include "gemm.hpp"
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

// grid.x: groups of kLeafVec vectors; 256 threads = 4 waves, each wave solves
// kLeafVec / 4 vectors as independent interleaved chains (one lane per
// unknown), so the per-step shuffle and FMA latencies of one vector overlap
// with the others'.
constexpr int kLeafVec = 16;

// Value of v in lane `lane` (wave-uniform) of this wavefront.
__device__ __forceinline__ double read_lane(double v, int lane) {
    const uint64_t u = __double_as_longlong(v);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)u, lane);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(u >> 32), lane);
    return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}
__device__ __forceinline__ float read_lane(float v, int lane) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}
constexpr int kChains = kLeafVec / 4;

template <typename T>
__global__ __launch_bounds__(256) void trsm_leaf_kernel(LeafArgs<T> p) {
    __shared__ T M[kLeaf][kLeaf + 1];     // M[i][k]: coefficient of unknown k in equation i
    __shared__ T V[kLeafVec][kLeaf + 1];  // V[v][i]: vector v, unknown i
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int v0 = blockIdx.x * kLeafVec;
    const int nb = p.nb;
    // Equation matrix: side L solves op(A) x = b  (M = op(A) block),
    // side R solves x op(A) = b, i.e. op(A)^T x^T = b^T (M = op(A)^T block).
    // M is lower iff (side L and op(A) lower) or (side R and op(A) upper).
    const bool m_lower = p.left ? p.lower_eff : !p.lower_eff;
    // All loads of a thread are issued before any is used: indices are
    // clamped into the block (always in bounds) and the unused values are
    // replaced afterwards, so no load waits behind another one's branch.
    constexpr int kMLoads = kLeaf * kLeaf / 256, kVLoads = kLeafVec * kLeaf / 256;
    T mv[kMLoads];
#pragma unroll
    for (int t = 0; t < kMLoads; ++t) {
        const int e = tid + 256 * t;
        const int i = min(e % kLeaf, nb - 1), k = min(e / kLeaf, nb - 1);
        mv[t] = p.left ? op_a(p, p.i0 + i, p.i0 + k) : op_a(p, p.i0 + k, p.i0 + i);
    }
    T bv[kVLoads];
#pragma unroll
    for (int t = 0; t < kVLoads; ++t) {
        const int e = tid + 256 * t;
        int v, i;
        if (p.left) { i = e % kLeaf; v = e / kLeaf; } else { v = e % kLeafVec; i = e / kLeafVec; }
        i = min(i, nb - 1);
        v = min(v0 + v, p.nvec - 1);
        bv[t] = p.left ? p.b[(int64_t)v * p.ldb + p.i0 + i] : p.b[(int64_t)(p.i0 + i) * p.ldb + v];
    }
#pragma unroll
    for (int t = 0; t < kMLoads; ++t) {
        const int e = tid + 256 * t;
        const int i = e % kLeaf, k = e / kLeaf;
        T v = T(0);
        if (i < nb && k < nb) {
            if (i == k) v = p.unit ? T(1) : mv[t];
            else if (m_lower ? (i > k) : (i < k)) v = mv[t];
        }
        M[i][k] = v;
    }
    // vectors: side L, vector v = column v0 + v of B, unknown i = row i0 + i;
    // side R, vector v = row v0 + v, unknown i = column i0 + i
#pragma unroll
    for (int t = 0; t < kVLoads; ++t) {
        const int e = tid + 256 * t;
        int v, i;
        if (p.left) { i = e % kLeaf; v = e / kLeaf; } else { v = e % kLeafVec; i = e / kLeafVec; }
        V[v][i] = (i < nb && v0 + v < p.nvec) ? bv[t] : T(0);
    }
    __syncthreads();
    T x[kChains];
#pragma unroll
    for (int c = 0; c < kChains; ++c) x[c] = V[wave * kChains + c][lane];
    // Step k: unknown k of every chain is read from lane k into scalar
    // registers (k is wave-uniform), divided by the diagonal (lane c of one
    // vector division holds chain c's numerator, so all chains share one
    // correctly rounded division), kept by lane k and eliminated from the
    // equations of the other lanes.  The chains are updated with selects, not
    // branches.
    const bool unit = p.unit;
    for (int step = 0; step < nb; ++step) {
        const int k = m_lower ? step : nb - 1 - step;
        const T mk = M[lane][k];
        const T dk = M[k][k];
        const bool below = m_lower ? (lane > k && lane < nb) : (lane < k);
        T xk[kChains];
#pragma unroll
        for (int c = 0; c < kChains; ++c) xk[c] = read_lane(x[c], k);
        if (!unit) {
            // one division for all chains: lane c divides chain c's value
            T num = xk[0];
#pragma unroll
            for (int c = 1; c < kChains; ++c) num = (lane == c) ? xk[c] : num;
            const T quo = num / dk;
#pragma unroll
            for (int c = 0; c < kChains; ++c) xk[c] = read_lane(quo, c);
        }
#pragma unroll
        for (int c = 0; c < kChains; ++c) {
            const T upd = fma_t(mk, -xk[c], x[c]);
            x[c] = (lane == k) ? xk[c] : (below ? upd : x[c]);
        }
    }
#pragma unroll
    for (int c = 0; c < kChains; ++c) V[wave * kChains + c][lane] = x[c];
    __syncthreads();
    for (int e = tid; e < kLeafVec * kLeaf; e += 256) {
        int v, i;
        if (p.left) { i = e % kLeaf; v = e / kLeaf; } else { v = e % kLeafVec; i = e / kLeafVec; }
        if (i < nb && v0 + v < p.nvec) {
            if (p.left) p.b[(int64_t)(v0 + v) * p.ldb + p.i0 + i] = V[v][i];
            else p.b[(int64_t)(p.i0 + i) * p.ldb + v0 + v] = V[v][i];
        }
    }
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
};

// op(A)[r0.., c0..] as a GEMM operand (pointer, transpose flag)
template <typename T>
const T* sub_opa(const TrsmCtx<T>& c, int r0, int c0) {
    return c.trans ? c.a + r0 * c.lda + c0 : c.a + c0 * c.lda + r0;
}

template <typename T>
int gemm_update(const TrsmCtx<T>& c, bool ta, bool tb, int m, int n, int k, const T* a, int64_t lda, const T* b,
                int64_t ldb, T* cc, int64_t ldc) {
    if (m <= 0 || n <= 0 || k <= 0) return ZYNUM_GPU_OK;
    if constexpr (sizeof(T) == 8) {
        if (c.mode != ZYNUM_GPU_MODE_PLAIN)
            return ozaki_dgemm(ta, tb, m, n, k, -1.0, a, lda, b, ldb, 1.0, cc, ldc, c.s, c.mode, c.mode_param);
    }
    GemmDesc<T> d;
    d.ta = ta; d.tb = tb; d.m = m; d.n = n; d.k = k;
    d.lda = lda; d.ldb = ldb; d.ldc = ldc;
    d.alpha = T(-1); d.beta = T(1);
    d.a = a; d.b = b; d.c = cc;
    return status_from_hip(gemm_launch<T>(d, c.s));
}

template <typename T>
int leaf(const TrsmCtx<T>& c, int i0, int nb) {
    LeafArgs<T> p;
    p.a = c.a; p.lda = c.lda; p.b = c.b; p.ldb = c.ldb;
    p.left = c.left; p.lower_eff = c.lower_eff; p.trans = c.trans; p.unit = c.unit;
    p.i0 = i0; p.nb = nb;
    p.nvec = c.left ? c.n : c.m;
    const int grid = (p.nvec + kLeafVec - 1) / kLeafVec;
    trsm_leaf_kernel<T><<<grid, 256, 0, c.s>>>(p);
    return status_from_hip(hipGetLastError());
}

// solves the unknowns [i0, i0 + len) (rows of X for side L, columns for side R)
template <typename T>
int solve(const TrsmCtx<T>& c, int i0, int len) {
    if (len <= kLeaf) return leaf(c, i0, len);
    int h = (len / 2 + kLeaf - 1) / kLeaf * kLeaf;
    if (h >= len) h = len - kLeaf;
    const int i1 = i0 + h, len2 = len - h;
    // forward when the equations are lower triangular: side L with op(A) lower,
    // side R with op(A) upper
    const bool forward = c.left ? c.lower_eff : !c.lower_eff;
    int st;
    if (c.left) {
        if (forward) {
            if ((st = solve(c, i0, h))) return st;
            // B[i1:, :] -= op(A)[i1:, i0:i1] X[i0:i1, :]
            st = gemm_update(c, c.trans, false, len2, c.n, h, sub_opa(c, i1, i0), c.lda, c.b + i0, c.ldb, c.b + i1,
                             c.ldb);
            if (st) return st;
            return solve(c, i1, len2);
        }
        if ((st = solve(c, i1, len2))) return st;
        // B[i0:i1, :] -= op(A)[i0:i1, i1:] X[i1:, :]
        st = gemm_update(c, c.trans, false, h, c.n, len2, sub_opa(c, i0, i1), c.lda, c.b + i1, c.ldb, c.b + i0,
                         c.ldb);
        if (st) return st;
        return solve(c, i0, h);
    }
    if (forward) {
        if ((st = solve(c, i0, h))) return st;
        // B[:, i1:] -= X[:, i0:i1] op(A)[i0:i1, i1:]
        st = gemm_update(c, false, c.trans, c.m, len2, h, c.b + (int64_t)i0 * c.ldb, c.ldb, sub_opa(c, i0, i1),
                         c.lda, c.b + (int64_t)i1 * c.ldb, c.ldb);
        if (st) return st;
        return solve(c, i1, len2);
    }
    if ((st = solve(c, i1, len2))) return st;
    // B[:, i0:i1] -= X[:, i1:] op(A)[i1:, i0:i1]
    st = gemm_update(c, false, c.trans, c.m, h, len2, c.b + (int64_t)i1 * c.ldb, c.ldb, sub_opa(c, i1, i0), c.lda,
                     c.b + (int64_t)i0 * c.ldb, c.ldb);
    if (st) return st;
    return solve(c, i0, h);
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
    return solve(c, 0, c.left ? m : n);
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

// Triangle-tile launch: 16-byte operand loads when both operands allow them.
template <typename T>
hipError_t syrk_launch(const GemmArgs<T>& p, int tiles, bool lower, bool trans, hipStream_t s) {
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
template <typename T>
int symm_device(char side, char uplo, int m, int n, T alpha, const T* a, int64_t lda, const T* b, int64_t ldb, T beta,
                T* c, int64_t ldc, hipStream_t s) {
    if (m <= 0 || n <= 0) return ZYNUM_GPU_OK;
    const bool left = (side == 'L');
    GemmDesc<T> d;
    d.m = m; d.n = n;
    d.alpha = alpha; d.beta = beta;
    d.c = c; d.ldc = ldc;
    if (alpha == T(0)) {  // C = beta * C (exact 0 for beta == 0)
        d.k = 0;
        return status_from_hip(gemm_launch<T>(d, s));
    }
    const int na = left ? m : n;
    const int64_t lds = ld_round(na);
    T* sym = static_cast<T*>(workspace_acquire(bytes_round((size_t)lds * na * sizeof(T))));
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
        e = gemm_launch<T>(d, s);
    }
    int st = status_from_hip(e);
    if (st == ZYNUM_GPU_OK) st = status_from_hip(hipStreamSynchronize(s));  // workspace reuse is host-ordered
    workspace_release();
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
    char* ws = static_cast<char*>(workspace_acquire(t_bytes + (size_t)ldw * n * sizeof(T)));
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
    int st = status_from_hip(e);
    if (st == ZYNUM_GPU_OK) st = status_from_hip(hipStreamSynchronize(s));
    workspace_release();
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
