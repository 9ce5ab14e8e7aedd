// Plain FP64 / FP32 GEMM on gfx950 matrix cores (v_mfma_f64_16x16x4_f64,
// v_mfma_f32_16x16x4_f32).
//
// Replaces Zynum's packed-B CPU micro-kernel (packed_simd.zig:140-199), its
// B-panel / A-transpose packing (packing.zig:45-80,131-139) and the alpha/beta
// epilogue (epilogue.zig:10-39).  Design:
//   * one workgroup of 256 threads (4 wave64s in a 2x2 grid) owns a BM x BN
//     tile of C; every wave owns (BM/2) x (BN/2) accumulated in 16x16 MFMA
//     blocks held in registers for the whole K loop;
//   * op(A) and op(B) K-slices (BK deep) are staged in LDS as k-major
//     [BK][BM] / [BK][BN] panels, double buffered: global loads for slice t+1
//     are issued into registers before the MFMAs of slice t and written to the
//     other LDS buffer afterwards, one barrier per slice;
//   * column-major operands are read with 16-byte coalesced loads along the
//     contiguous dimension whatever the transpose (M/N-contiguous for A / B^T,
//     K-contiguous for A^T / B); transposition happens on the LDS write;
//   * edge tiles (M, N, K not multiples of the tile) use guarded scalar loads
//     with zero fill; interior tiles run unguarded;
//   * tile ids are remapped so each of the 8 XCDs gets a contiguous block of
//     tiles in grouped (super-row) order, keeping shared A/B panels in its L2.
#include "common.hpp"
#include "gemm.hpp"

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

    __device__ __forceinline__ void load(const T* __restrict__ x, int64_t ldx, int r0, int k0, int rlim,
                                         int klim, bool full, int tid) {
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
                    vec_t<T, VEC> t = *reinterpret_cast<const vec_t<T, VEC>*>(x + off);
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

template <typename T, class CF, bool TA, bool TB, int VEC>
__global__ __launch_bounds__(CF::NT) __attribute__((amdgpu_waves_per_eu(CF::OCC > 0 ? CF::OCC : 1)))
void gemm_mfma_kernel(GemmArgs<T> p) {
    constexpr int BM = CF::BM, BN = CF::BN, BK = CF::BK, NT = CF::NT;
    constexpr int PAD = 16 / sizeof(T);
    constexpr int WTM = BM / CF::WM, WTN = BN / CF::WN;
    constexpr int FM = WTM / 16, FN = WTN / 16;
    static_assert(FM >= 1 && FN >= 1, "wave tile smaller than one MFMA block");
    using M = Mfma<T>;
    using Acc = typename M::acc_t;

    __shared__ __attribute__((aligned(16))) T lds_a[2][BK * (BM + PAD)];
    __shared__ __attribute__((aligned(16))) T lds_b[2][BK * (BN + PAD)];

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int wm0 = (wave % CF::WM) * WTM;
    const int wn0 = (wave / CF::WM) * WTN;

    const int total = p.tiles_m * p.tiles_n;
    int tm, tn;
    grouped_tile(xcd_remap(blockIdx.x, total), p.tiles_m, p.tiles_n, p.group_m, tm, tn);
    const int m0 = tm * BM, n0 = tn * BN;

    // op(A) tile rows are M (A: M-contiguous unless TA); op(B) rows are N
    // (B^T: N-contiguous when TB, otherwise K-contiguous).
    Stage<T, BM, BK, TA, VEC, NT, PAD> sa;
    Stage<T, BN, BK, !TB, VEC, NT, PAD> sb;

    Acc acc[FM][FN];
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j) acc[i][j] = M::zero();

    const int nk = (p.k + BK - 1) / BK;
    const bool mn_full = (m0 + BM <= p.m) && (n0 + BN <= p.n);

    sa.load(p.a, p.lda, m0, 0, p.m, p.k, mn_full && BK <= p.k, tid);
    sb.load(p.b, p.ldb, n0, 0, p.n, p.k, mn_full && BK <= p.k, tid);
    sa.store(lds_a[0], tid);
    sb.store(lds_b[0], tid);
    __syncthreads();

    const int lr = lane & 15;  // column within a 16x16 block (B, D)
    const int lk = lane >> 4;  // k within a 4-deep step
    const int ar = M::a_row(lr);  // A row fed by this lane

    for (int t = 0; t < nk; ++t) {
        const int cur = t & 1;
        const bool more = (t + 1) < nk;
        if (more) {
            const int k0 = (t + 1) * BK;
            const bool full = mn_full && (k0 + BK <= p.k);
            sa.load(p.a, p.lda, m0, k0, p.m, p.k, full, tid);
            sb.load(p.b, p.ldb, n0, k0, p.n, p.k, full, tid);
        }
        const T* __restrict__ la = lds_a[cur];
        const T* __restrict__ lb = lds_b[cur];
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
        if (more) {
            sa.store(lds_a[cur ^ 1], tid);
            sb.store(lds_b[cur ^ 1], tid);
        }
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
                    if (row + r < p.m) cp[r] = epilogue<T>(alpha, beta, acc[i][j][r], cp + r);
                }
            }
        }
    }
}

template <typename T>
__global__ void scale_kernel(int m, int n, T beta, T* __restrict__ c, int64_t ldc) {
    const int64_t total = (int64_t)m * n;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t j = e / m, i = e % m;
        T* cp = c + j * ldc + i;
        *cp = (beta == T(0)) ? T(0) : beta * *cp;
    }
}

template <typename T>
hipError_t scale_launch(int m, int n, T beta, T* c, int64_t ldc, hipStream_t stream) {
    if (m <= 0 || n <= 0) return hipSuccess;
    const int64_t total = (int64_t)m * n;
    int blocks = (int)((total + 255) / 256);
    if (blocks > 65536) blocks = 65536;
    scale_kernel<T><<<blocks, 256, 0, stream>>>(m, n, beta, c, ldc);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// dispatch
// ---------------------------------------------------------------------------
template <typename T, class CF, int VEC>
static hipError_t launch_cfg(const GemmDesc<T>& d, hipStream_t stream) {
    GemmArgs<T> p;
    p.m = d.m; p.n = d.n; p.k = d.k;
    p.lda = d.lda; p.ldb = d.ldb; p.ldc = d.ldc;
    p.alpha = d.alpha; p.beta = d.beta;
    p.a = d.a; p.b = d.b; p.c = d.c;
    p.tiles_m = (d.m + CF::BM - 1) / CF::BM;
    p.tiles_n = (d.n + CF::BN - 1) / CF::BN;
    p.group_m = 8;
    const int grid = p.tiles_m * p.tiles_n;
    if (!d.ta && !d.tb) gemm_mfma_kernel<T, CF, false, false, VEC><<<grid, CF::NT, 0, stream>>>(p);
    else if (!d.ta && d.tb) gemm_mfma_kernel<T, CF, false, true, VEC><<<grid, CF::NT, 0, stream>>>(p);
    else if (d.ta && !d.tb) gemm_mfma_kernel<T, CF, true, false, VEC><<<grid, CF::NT, 0, stream>>>(p);
    else gemm_mfma_kernel<T, CF, true, true, VEC><<<grid, CF::NT, 0, stream>>>(p);
    return hipGetLastError();
}

template <typename T, class CF>
static hipError_t launch_vec(const GemmDesc<T>& d, hipStream_t stream) {
    constexpr int V = 16 / sizeof(T);
    const uintptr_t mask = 15;
    const bool aligned = ((reinterpret_cast<uintptr_t>(d.a) & mask) == 0) &&
                         ((reinterpret_cast<uintptr_t>(d.b) & mask) == 0) && (d.lda % V == 0) &&
                         (d.ldb % V == 0);
    if (aligned) return launch_cfg<T, CF, V>(d, stream);
    return launch_cfg<T, CF, 1>(d, stream);
}

// BK is given for FP64; FP32 uses twice the depth (same bytes per slice).
template <typename T, int BM, int BN, int BK64, int WM, int WN, int OCC>
static hipError_t launch_shape(const GemmDesc<T>& d, hipStream_t stream) {
    constexpr int BK = (sizeof(T) == 8) ? BK64 : 2 * BK64;
    return launch_vec<T, TileCfg<BM, BN, BK, WM, WN, OCC>>(d, stream);
}

template <typename T>
hipError_t gemm_launch(const GemmDesc<T>& d, hipStream_t stream, int tile) {
    if (d.m <= 0 || d.n <= 0) return hipSuccess;
    if (d.k <= 0 || d.alpha == T(0)) return scale_launch<T>(d.m, d.n, d.beta, d.c, d.ldc, stream);
    if (tile == kTileAuto) tile = pick_tile(d.m, d.n, d.k, (int)sizeof(T));
    switch (tile) {
        case kTile64x64: return launch_shape<T, 64, 64, 16, 2, 2, 0>(d, stream);
        case kTile128x128: return launch_shape<T, 128, 128, 16, 2, 2, 1>(d, stream);
        case kTile128x128o2: return launch_shape<T, 128, 128, 16, 2, 2, 2>(d, stream);
        case kTile128x64: return launch_shape<T, 128, 64, 16, 2, 2, 2>(d, stream);
        case kTile64x128: return launch_shape<T, 64, 128, 16, 2, 2, 2>(d, stream);
        case kTile128x128k32: return launch_shape<T, 128, 128, 32, 2, 2, 1>(d, stream);
        case kTile128x128w8: return launch_shape<T, 128, 128, 16, 4, 2, 2>(d, stream);
        case kTile64x64k32: return launch_shape<T, 64, 64, 32, 2, 2, 0>(d, stream);
        default: return hipErrorInvalidValue;
    }
}

template hipError_t gemm_launch<double>(const GemmDesc<double>&, hipStream_t, int);
template hipError_t gemm_launch<float>(const GemmDesc<float>&, hipStream_t, int);
template hipError_t scale_launch<double>(int, int, double, double*, int64_t, hipStream_t);
template hipError_t scale_launch<float>(int, int, float, float*, int64_t, hipStream_t);

}  // namespace zg
