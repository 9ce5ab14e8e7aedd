// Host-side entry points of the gfx950 GEMM core (device pointers only).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace zg {

// One column-major GEMM: C = alpha * op(A) * op(B) + beta * C.
template <typename T>
struct GemmDesc {
    bool ta = false, tb = false;
    int m = 0, n = 0, k = 0;
    int64_t lda = 1, ldb = 1, ldc = 1;
    T alpha = T(1), beta = T(0);
    const T* a = nullptr;
    const T* b = nullptr;
    T* c = nullptr;
};

// Tile configurations the launcher can pick (see gemm_mfma.hip).
enum TileConfig : int {
    kTileAuto = -1,
    kTile64x64 = 0,       // 256 threads, 2x2 waves of 32x32, BK 16
    kTile128x128 = 1,     // 256 threads, 2x2 waves of 64x64, BK 16
    kTile128x128o2 = 2,   // as 1, registers capped for 2 waves per SIMD
    kTile128x64 = 3,      // 2x2 waves of 64x32, 2 waves per SIMD
    kTile64x128 = 4,      // 2x2 waves of 32x64, 2 waves per SIMD
    kTile128x128k32 = 5,  // 2x2 waves of 64x64, BK 32
    kTile128x128w8 = 6,   // 512 threads, 4x2 waves of 32x64
    kTile64x64k32 = 7,    // as 0 with BK 32
    kTile256x128w8 = 8,   // 512 threads, 4x2 waves of 64x64, one workgroup per CU
    kTile128x256w8 = 9,   // 512 threads, 2x4 waves of 64x64, one workgroup per CU
    kTile128x128w8k8 = 10,  // as 6 with BK 8 (half the LDS: more workgroups per CU)
    kTile128x128w16 = 11,   // 1024 threads, 4x4 waves of 32x32, BK 16, 8 waves per SIMD
    kTileConfigs = 12
};

// Tile choice for one GEMM shape from the measured tuning table
// (tune_table.inc, written by zynum_amd/autotune.py).
int pick_tile(int m, int n, int k, int elem_bytes, bool ta, bool tb);

// Launch the plain MFMA GEMM (FP64 or FP32). Returns a hipError_t.
// k == 0 or alpha == 0 takes the quick return C = beta * C (gemm.zig:170-180).
template <typename T>
hipError_t gemm_launch(const GemmDesc<T>& d, hipStream_t stream, int tile = kTileAuto);
// The same product without the alpha == 0 quick return: alpha * sum is formed,
// so a NaN or an infinity in A or B still reaches C (for routines whose
// reference has no such quick return, symmetric.zig:56-70).
template <typename T>
hipError_t gemm_launch_product(const GemmDesc<T>& d, hipStream_t stream, int tile = kTileAuto);

// Split-K for few output tiles and a long K: the number of K ranges (1 = do
// not split) and, through kchunk, their depth.
int gemm_splitk_plan(int m, int n, int k, int* kchunk);
// The split-K GEMM: parts x m x n partial products in `part` (device memory
// the caller provides and keeps until the stream has passed the launch), then
// C = alpha * (sum of the parts in increasing order) + beta * C.
template <typename T>
hipError_t gemm_splitk_launch(const GemmDesc<T>& d, hipStream_t stream, T* part, int parts, int kchunk);

// C = beta * C (beta == 0 writes exact zeros), the k == 0 / alpha == 0 path.
template <typename T>
hipError_t scale_launch(int m, int n, T beta, T* c, int64_t ldc, hipStream_t stream);

}  // namespace zg
