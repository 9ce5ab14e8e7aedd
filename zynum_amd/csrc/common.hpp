// Shared device/host definitions for the gfx950 Level-3 kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace zg {

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));
typedef int v4i __attribute__((ext_vector_type(4)));

constexpr int kWave = 64;

// 16x16x4 MFMA for the two real scalar types.  Operand layout (both types):
//   A (16x4):  lane l holds A[l % 16][l / 16]
//   B (4x16):  lane l holds B[l / 16][l % 16]
// Result layout differs:
//   f32: lane l, item r holds D[4 * (l / 16) + r][l % 16]
//   f64: lane l, item r holds D[4 * r + l / 16][l % 16]
// a_row(i) is the A row a lane with l % 16 == i feeds.  For f64 it permutes
// the rows (i -> 4 (i % 4) + i / 4) so that the f64 result also lands as four
// consecutive rows per lane, like f32; the epilogue is then shared.
template <typename T> struct Mfma;
template <> struct Mfma<double> {
    typedef v4d acc_t;
    static __device__ __forceinline__ int a_row(int i) { return 4 * (i & 3) + (i >> 2); }
    static __device__ __forceinline__ acc_t zero() { return acc_t{0.0, 0.0, 0.0, 0.0}; }
    static __device__ __forceinline__ acc_t mma(double a, double b, acc_t c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
};
template <> struct Mfma<float> {
    typedef v4f acc_t;
    static __device__ __forceinline__ int a_row(int i) { return i; }
    static __device__ __forceinline__ acc_t zero() { return acc_t{0.f, 0.f, 0.f, 0.f}; }
    static __device__ __forceinline__ acc_t mma(float a, float b, acc_t c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
};

__device__ __forceinline__ float fma_t(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fma_t(double a, double b, double c) { return __builtin_fma(a, b, c); }

// Pointers into global memory by address space.  A pointer read from a
// descriptor in memory is generic, and accesses through it are flat
// instructions, which count against lgkmcnt as well as vmcnt: an LDS wait then
// also waits for them.  Accesses through a gptr are global instructions.
template <typename T>
using gptr = __attribute__((address_space(1))) T*;
template <typename T>
__device__ __forceinline__ gptr<T> to_global(T* p) {
    return (gptr<T>)p;
}

// Zynum's alpha/beta write-back (epilogue.zig:10-39):
//   alpha == 1 && beta == 0 -> acc
//   otherwise fma(alpha, acc, s), s = 0 (beta == 0), old (beta == 1), RN(beta * old)
// C is not read when beta == 0, so NaNs in C never propagate.
template <typename T>
__device__ __forceinline__ T epilogue(T alpha, T beta, T acc, const T* cptr) {
    if (alpha == T(1) && beta == T(0)) return acc;
    T s = T(0);
    if (beta != T(0)) {
        T old = *cptr;
        s = (beta == T(1)) ? old : beta * old;
    }
    return fma_t(alpha, acc, s);
}

// The same write-back through a global pointer.
template <typename T>
__device__ __forceinline__ T epilogue_g(T alpha, T beta, T acc, gptr<const T> cptr) {
    if (alpha == T(1) && beta == T(0)) return acc;
    T s = T(0);
    if (beta != T(0)) {
        T old = *cptr;
        s = (beta == T(1)) ? old : beta * old;
    }
    return fma_t(alpha, acc, s);
}

// Number of XCDs on MI355X; consecutive workgroup ids are dealt round-robin
// across them, so tile ids are remapped to give each XCD a contiguous block of
// tiles (shared A row panels / B column panels stay in that XCD's L2).
constexpr int kXcds = 8;

__device__ __forceinline__ int xcd_remap(int wg, int total) {
    if (total % kXcds != 0) return wg;
    int per = total / kXcds;
    return (wg % kXcds) * per + wg / kXcds;
}

// Grouped ("super-row") ordering of a tiles_m x tiles_n grid.
__device__ __forceinline__ void grouped_tile(int t, int tiles_m, int tiles_n, int group_m, int& tm, int& tn) {
    int per_group = group_m * tiles_n;
    int g = t / per_group;
    int first = g * group_m;
    int rows = min(tiles_m - first, group_m);
    int r = t % per_group;
    tm = first + r % rows;
    tn = r / rows;
}

}  // namespace zg
