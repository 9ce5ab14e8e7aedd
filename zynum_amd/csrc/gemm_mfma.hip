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
//     tiles in grouped (super-row) order, keeping shared A/B panels in its L2;
//   * FP64 full tiles of the 16-wave 128x128 configuration skip the
//     register staging: direct-to-LDS loads fill a 4-stage ring, with the
//     same MFMA order, so the same results; automatic tile choice sends FP64
//     shapes of at least 128 such tiles there whatever the transpose pair.
#include <algorithm>

#include "common.hpp"
#include "gemm.hpp"
#include "gemm_tile.hpp"

// Tile rows per super-row of the grouped tile order; a build knob for tuning.
#ifndef ZG_GROUP_M
#define ZG_GROUP_M 8
#endif

namespace zg {

template <typename T, class CF, bool TA, bool TB, int VEC>
__global__ __launch_bounds__(CF::NT) __attribute__((amdgpu_waves_per_eu(CF::OCC > 0 ? CF::OCC : 1)))
void gemm_mfma_kernel(GemmArgs<T> p) {
    __shared__ __attribute__((aligned(16))) T lds_a[tile_lds_a<T, CF>()];
    __shared__ __attribute__((aligned(16))) T lds_b[tile_lds_b<T, CF>()];
    const int total = p.tiles_m * p.tiles_n;
    int tm, tn;
    grouped_tile(xcd_remap(blockIdx.x, total), p.tiles_m, p.tiles_n, p.group_m, tm, tn);
    gemm_tile<T, CF, TA, TB, VEC>(p, tm * CF::BM, tn * CF::BN, lds_a, lds_b);
}

// FP64 form of the 16-wave 128x128 tile fed by direct-to-LDS loads
// (global_load_lds_dwordx4), for full tiles: m and n multiples of 128, k of
// 8, A and B 16-byte aligned with even leading dimensions.  No registers
// stage the operands: stages of 8 k sit in a 4-stage LDS ring (66 KB, two
// workgroups per CU as with the register-staged tile), loaded 3 stages
// ahead, one direct load per wave per stage (waves 0..7 op(A), waves 8..15
// op(B)).  Each wave multiplies its 32x32 sub-tile in 16x16x4 MFMA steps in
// increasing k, exactly as gemm_tile does, so results are bitwise those of
// the register-staged kernel; the freed registers hold the fragments of the
// next 4-k half, read while the current half's MFMAs run.
//
// A direct load cannot transpose, so each operand keeps the layout it has in
// memory.  An operand contiguous along m / n (A, B^T) is stored as 8 k-rows
// of 128 + 2 doubles (the padding of the register-staged tile), one direct
// load per k-row.  An operand contiguous along k (A^T, B) is stored as 128
// rows of 8 k, 16 rows per direct load, the four 16-byte k pairs of row r
// rotated by (r >> 2) & 3, so the 16 rows of one fragment read at one k fall
// on distinct LDS banks.
constexpr int kDmaStages = 4;
template <bool TA, bool TB>
__global__ __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu(8))) void gemm_f64_dma_kernel(
    GemmArgs<double> p) {
    constexpr int BM = 128, BN = 128, BK = 8, NS = kDmaStages;
    constexpr bool KRA = !TA, KRB = TB;  // operand stored as k-rows (contiguous along m / n)
    constexpr int RK = BM + 2;            // k-row stride
    constexpr int SA = KRA ? BK * RK : BM * BK;
    constexpr int SB = KRB ? BK * RK : BN * BK;
    constexpr int SS = SA + SB;
    __shared__ __attribute__((aligned(16))) double lds[NS * SS];

    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int wm0 = (wave % 4) * 32, wn0 = (wave / 4) * 32;
    const int total = p.tiles_m * p.tiles_n;
    int tm, tn;
    grouped_tile(xcd_remap(blockIdx.x, total), p.tiles_m, p.tiles_n, p.group_m, tm, tn);
    const int m0 = tm * BM, n0 = tn * BN;

    // waves 0..7 load op(A), waves 8..15 op(B): a k-row of 128 (k-row layout)
    // or 16 rows of 8 k (row layout, lane: slot lane % 4 of row lane / 4,
    // holding k pair (lane % 4) ^ ((row >> 2) & 3))
    const bool is_a = wave < 8;
    const int q = is_a ? wave : wave - 8;
    const bool kr = is_a ? KRA : KRB;
    const double* x = is_a ? p.a : p.b;
    const int64_t ldx = is_a ? p.lda : p.ldb;
    const int x0 = is_a ? m0 : n0;
    const double* src;
    int64_t kstep;
    int dst;
    if (kr) {
        src = x + (int64_t)q * ldx + x0 + 2 * lane;
        kstep = (int64_t)BK * ldx;
        dst = (is_a ? 0 : SA) + q * RK;
    } else {
        const int rl = q * 16 + (lane >> 2);
        const int pair = (lane & 3) ^ ((rl >> 2) & 3);
        src = x + (int64_t)(x0 + rl) * ldx + 2 * pair;
        kstep = BK;
        dst = (is_a ? 0 : SA) + q * 128;
    }
    auto issue = [&](int t) {
        __builtin_amdgcn_global_load_lds((const void __attribute__((address_space(1)))*)(src + (int64_t)t * kstep),
                                         (void __attribute__((address_space(3)))*)(lds + (t % NS) * SS + dst), 16, 0,
                                         0);
    };

    const int lr = lane & 15, lk = lane >> 4;
    const int ar = Mfma<double>::a_row(lr);
    auto at = [](bool krl, int base, int r, int k) {  // element (row r, k) of an operand part
        return krl ? base + k * RK + r : base + r * BK + ((((k >> 1) ^ ((r >> 2) & 3))) << 1) + (k & 1);
    };
    // a k-row operand's second half is a constant 4 k-rows further (folded
    // into the read's offset); a row-layout operand keeps an offset per half
    int aoff[KRA ? 1 : 2][2], boff[KRB ? 1 : 2][2];  // [half][block]
#pragma unroll
    for (int h = 0; h < (KRA ? 1 : 2); ++h)
#pragma unroll
        for (int i = 0; i < 2; ++i) aoff[h][i] = at(KRA, 0, wm0 + i * 16 + ar, 4 * h + lk);
#pragma unroll
    for (int h = 0; h < (KRB ? 1 : 2); ++h)
#pragma unroll
        for (int i = 0; i < 2; ++i) boff[h][i] = at(KRB, SA, wn0 + i * 16 + lr, 4 * h + lk);
    v4d acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = Mfma<double>::zero();
    double fa[2][2], fb[2][2];  // [half][block]
    auto read = [&](int h, int t) {
        const double* base = lds + (t % NS) * SS;
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            fa[h][i] = KRA ? base[aoff[0][i] + 4 * h * RK] : base[aoff[KRA ? 0 : h][i]];
            fb[h][i] = KRB ? base[boff[0][i] + 4 * h * RK] : base[boff[KRB ? 0 : h][i]];
        }
    };
    auto mma_row = [&](int h, int i) {
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = Mfma<double>::mma(fa[h][i], fb[h][j], acc[i][j]);
    };

    const int nk = p.k / BK;
#pragma unroll
    for (int t = 0; t < NS - 1; ++t)
        if (t < nk) issue(t);
    if (nk >= NS - 1) __builtin_amdgcn_s_waitcnt(0x0F72);  // vmcnt(2): this wave's stage 0 has landed
    else __builtin_amdgcn_s_waitcnt(0x0F70);               // vmcnt(0)
    __builtin_amdgcn_s_barrier();
    read(0, 0);
    for (int s = 0; s < nk; ++s) {
        // this wave's stage s+1 has landed (stages up to s + NS - 2 are issued)
        if (s + NS - 2 <= nk - 1) __builtin_amdgcn_s_waitcnt(0x0F71);  // vmcnt(1)
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
        for (int j = 0; j < 2; ++j) {
            const int col = n0 + wn0 + j * 16 + lr;
            const int row = m0 + wm0 + i * 16 + 4 * lk;
            double* cp = p.c + (int64_t)col * p.ldc + row;
#pragma unroll
            for (int r = 0; r < 4; ++r) cp[r] = epilogue<double>(alpha, beta, acc[i][j][r], cp + r);
        }
}

// Whether gemm_f64_dma_kernel takes the shape.
template <typename T>
static bool dma_eligible(const GemmDesc<T>& d) {
    if constexpr (sizeof(T) != 8) {
        return false;
    } else {
        const uintptr_t mask = 15;
        return d.m % 128 == 0 && d.n % 128 == 0 && d.k >= 8 && d.k % 8 == 0 &&
               (reinterpret_cast<uintptr_t>(d.a) & mask) == 0 && (reinterpret_cast<uintptr_t>(d.b) & mask) == 0 &&
               d.lda % 2 == 0 && d.ldb % 2 == 0;
    }
}

// Split-K form for shapes with few output tiles and a long K: blockIdx.y
// selects a K range [k0, k0 + kchunk) of the product, written unscaled
// (alpha 1, beta 0: the accumulator itself) to its own slice of the partial
// buffer; splitk_reduce_kernel sums the slices in increasing order and applies
// the alpha/beta epilogue.
template <typename T, class CF, bool TA, bool TB, int VEC>
__global__ __launch_bounds__(CF::NT) __attribute__((amdgpu_waves_per_eu(CF::OCC > 0 ? CF::OCC : 1)))
void gemm_splitk_kernel(GemmArgs<T> p, int kchunk, int64_t part_stride) {
    __shared__ __attribute__((aligned(16))) T lds_a[tile_lds_a<T, CF>()];
    __shared__ __attribute__((aligned(16))) T lds_b[tile_lds_b<T, CF>()];
    const int k0 = (int)blockIdx.y * kchunk;
    GemmArgs<T> q = p;
    q.k = min(kchunk, p.k - k0);
    q.a = TA ? p.a + k0 : p.a + (int64_t)k0 * p.lda;
    q.b = TB ? p.b + (int64_t)k0 * p.ldb : p.b + k0;
    q.c = p.c + (int64_t)blockIdx.y * part_stride;
    const int total = p.tiles_m * p.tiles_n;
    int tm, tn;
    grouped_tile(xcd_remap(blockIdx.x, total), p.tiles_m, p.tiles_n, p.group_m, tm, tn);
    gemm_tile<T, CF, TA, TB, VEC>(q, tm * CF::BM, tn * CF::BN, lds_a, lds_b);
}

template <typename T>
__global__ void splitk_reduce_kernel(int m, int n, int parts, const T* __restrict__ part, int64_t part_stride,
                                     T alpha, T beta, T* __restrict__ c, int64_t ldc) {
    const int64_t total = (int64_t)m * n;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
        T acc = part[e];
        for (int s = 1; s < parts; ++s) acc += part[s * part_stride + e];
        T* cp = c + (e / m) * ldc + e % m;
        *cp = epilogue(alpha, beta, acc, cp);
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
    p.group_m = ZG_GROUP_M;
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
    return gemm_launch_product<T>(d, stream, tile);
}

template <typename T>
hipError_t gemm_launch_product(const GemmDesc<T>& d, hipStream_t stream, int tile) {
    if (d.m <= 0 || d.n <= 0) return hipSuccess;
    if (d.k <= 0) return scale_launch<T>(d.m, d.n, d.beta, d.c, d.ldc, stream);
    if (tile == kTileAuto) {
        tile = pick_tile(d.m, d.n, d.k, (int)sizeof(T), d.ta, d.tb);
        // the direct-load tile beats the register-staged ones the table picks
        // on FP64 shapes it takes from 128 output tiles up, for all transpose
        // pairs (8192^3 TN: 68.2 against 66.5 TFLOP/s; 2048^3: 54-62 against
        // 51-58; 1536 x 2048 x 4096 NN / TT: 47.7 / 48.3 against 45.5 / 45.0)
        if (dma_eligible(d) && (int64_t)(d.m / 128) * (d.n / 128) >= 128) tile = kTile128x128w16;
    }
    switch (tile) {
        case kTile64x64: return launch_shape<T, 64, 64, 16, 2, 2, 0>(d, stream);
        case kTile128x128: return launch_shape<T, 128, 128, 16, 2, 2, 1>(d, stream);
        case kTile128x128o2: return launch_shape<T, 128, 128, 16, 2, 2, 2>(d, stream);
        case kTile128x64: return launch_shape<T, 128, 64, 16, 2, 2, 2>(d, stream);
        case kTile64x128: return launch_shape<T, 64, 128, 16, 2, 2, 2>(d, stream);
        case kTile128x128k32: return launch_shape<T, 128, 128, 32, 2, 2, 1>(d, stream);
        case kTile128x128w8: return launch_shape<T, 128, 128, 16, 4, 2, 2>(d, stream);
        case kTile64x64k32: return launch_shape<T, 64, 64, 32, 2, 2, 0>(d, stream);
        case kTile256x128w8: return launch_shape<T, 256, 128, 16, 4, 2, 1>(d, stream);
        case kTile128x256w8: return launch_shape<T, 128, 256, 16, 2, 4, 1>(d, stream);
        case kTile128x128w8k8: return launch_shape<T, 128, 128, 8, 4, 2, 2>(d, stream);
        case kTile128x128w16:
            if (dma_eligible(d)) {
                GemmArgs<double> p;
                p.m = d.m; p.n = d.n; p.k = d.k;
                p.lda = d.lda; p.ldb = d.ldb; p.ldc = d.ldc;
                p.alpha = d.alpha; p.beta = d.beta;
                p.a = reinterpret_cast<const double*>(d.a);
                p.b = reinterpret_cast<const double*>(d.b);
                p.c = reinterpret_cast<double*>(d.c);
                p.tiles_m = d.m / 128;
                p.tiles_n = d.n / 128;
                p.group_m = ZG_GROUP_M;
                const int g = p.tiles_m * p.tiles_n;
                if (!d.ta && !d.tb) gemm_f64_dma_kernel<false, false><<<g, 1024, 0, stream>>>(p);
                else if (!d.ta && d.tb) gemm_f64_dma_kernel<false, true><<<g, 1024, 0, stream>>>(p);
                else if (d.ta && !d.tb) gemm_f64_dma_kernel<true, false><<<g, 1024, 0, stream>>>(p);
                else gemm_f64_dma_kernel<true, true><<<g, 1024, 0, stream>>>(p);
                return hipGetLastError();
            }
            return launch_shape<T, 128, 128, 16, 4, 4, 4>(d, stream);
        default: return hipErrorInvalidValue;
    }
}

// Split-K plan on the 64x64 tile: split when fewer than 128 tiles would run
// and K >= 1024, into enough K ranges (each a multiple of 32 and at least
// 256 deep, at most 64 of them) to give about 512 workgroups.
int gemm_splitk_plan(int m, int n, int k, int* kchunk) {
    const int64_t tiles = (int64_t)((m + 63) / 64) * ((n + 63) / 64);
    if (tiles >= 128 || k < 1024) return 1;
    int64_t parts = std::min<int64_t>({(512 + tiles - 1) / tiles, (int64_t)k / 256, 64});
    if (parts < 2) return 1;
    int64_t kc = ((k + parts - 1) / parts + 31) / 32 * 32;
    parts = (k + kc - 1) / kc;
    if (parts < 2) return 1;
    *kchunk = (int)kc;
    return (int)parts;
}

template <typename T>
hipError_t gemm_splitk_launch(const GemmDesc<T>& d, hipStream_t stream, T* part, int parts, int kchunk) {
    using CF = TileCfg<64, 64, (sizeof(T) == 8) ? 16 : 32, 2, 2, 0>;
    GemmArgs<T> p;
    p.m = d.m; p.n = d.n; p.k = d.k;
    p.lda = d.lda; p.ldb = d.ldb; p.ldc = d.m;
    p.alpha = T(1); p.beta = T(0);
    p.a = d.a; p.b = d.b; p.c = part;
    p.tiles_m = (d.m + CF::BM - 1) / CF::BM;
    p.tiles_n = (d.n + CF::BN - 1) / CF::BN;
    p.group_m = ZG_GROUP_M;
    const int64_t stride = (int64_t)d.m * d.n;
    constexpr int V = 16 / sizeof(T);
    const bool vec = ((reinterpret_cast<uintptr_t>(d.a) & 15) == 0) && ((reinterpret_cast<uintptr_t>(d.b) & 15) == 0) &&
                     (d.lda % V == 0) && (d.ldb % V == 0) && (kchunk % V == 0);
    const dim3 grid(p.tiles_m * p.tiles_n, parts);
#define ZG_SPLITK(TA, TB, VV) gemm_splitk_kernel<T, CF, TA, TB, VV><<<grid, CF::NT, 0, stream>>>(p, kchunk, stride)
    if (vec) {
        if (!d.ta && !d.tb) ZG_SPLITK(false, false, V);
        else if (!d.ta && d.tb) ZG_SPLITK(false, true, V);
        else if (d.ta && !d.tb) ZG_SPLITK(true, false, V);
        else ZG_SPLITK(true, true, V);
    } else {
        if (!d.ta && !d.tb) ZG_SPLITK(false, false, 1);
        else if (!d.ta && d.tb) ZG_SPLITK(false, true, 1);
        else if (d.ta && !d.tb) ZG_SPLITK(true, false, 1);
        else ZG_SPLITK(true, true, 1);
    }
#undef ZG_SPLITK
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    int blocks = (int)std::min<int64_t>(4096, (stride + 255) / 256);
    splitk_reduce_kernel<T><<<blocks, 256, 0, stream>>>(d.m, d.n, parts, part, stride, d.alpha, d.beta, d.c, d.ldc);
    return hipGetLastError();
}

template hipError_t gemm_splitk_launch<double>(const GemmDesc<double>&, hipStream_t, double*, int, int);
template hipError_t gemm_splitk_launch<float>(const GemmDesc<float>&, hipStream_t, float*, int, int);
template hipError_t gemm_launch<double>(const GemmDesc<double>&, hipStream_t, int);
template hipError_t gemm_launch<float>(const GemmDesc<float>&, hipStream_t, int);
template hipError_t gemm_launch_product<double>(const GemmDesc<double>&, hipStream_t, int);
template hipError_t gemm_launch_product<float>(const GemmDesc<float>&, hipStream_t, int);
template hipError_t scale_launch<double>(int, int, double, double*, int64_t, hipStream_t);
template hipError_t scale_launch<float>(int, int, float, float*, int64_t, hipStream_t);

}  // namespace zg
