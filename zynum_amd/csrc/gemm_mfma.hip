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
#include "gemm_dma.hpp"
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

// FP64 direct-load form of config 11 (gemm_dma.hpp), NW waves.
template <bool TA, bool TB, int NW>
__global__ __launch_bounds__(64 * NW) __attribute__((amdgpu_waves_per_eu(NW / 2))) void gemm_f64_dma_kernel(
    GemmArgs<double> p) {
    int tm, tn;
    grouped_tile(xcd_remap(blockIdx.x, p.tiles_m * p.tiles_n), p.tiles_m, p.tiles_n, p.group_m, tm, tn);
    gemm_tile_dma<TA, TB, NW>(p, tm * 128, tn * 128);
}

// Whether gemm_f64_dma_kernel takes the shape.
template <typename T>
static bool dma_eligible(const GemmDesc<T>& d) {
    if constexpr (sizeof(T) != 8) {
        return false;
    } else {
        return dma_shape_ok(d.m, d.n, d.k, d.a, d.lda, d.b, d.ldb);
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
                // 4 waves of 32x128 for NN, 8 of 32x64 for the other pairs (fastest of 4 / 8 / 16)
                if (!d.ta && !d.tb) gemm_f64_dma_kernel<false, false, 4><<<g, 256, 0, stream>>>(p);
                else if (!d.ta && d.tb) gemm_f64_dma_kernel<false, true, 8><<<g, 512, 0, stream>>>(p);
                else if (d.ta && !d.tb) gemm_f64_dma_kernel<true, false, 8><<<g, 512, 0, stream>>>(p);
                else gemm_f64_dma_kernel<true, true, 8><<<g, 512, 0, stream>>>(p);
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
