// Complex GEMM (CGEMM / ZGEMM) on the real MFMA GEMM core.
//
// Replaces Zynum's complex GEMM routes (src/blas/core/matrix_matrix/gemm.zig:
// 1234-1246 portable loop, 1260-1280 route choice, 1039-1232 the real-buffer
// route) with one real GEMM per call:
//
//   * op(A) is expanded into the real 2m x 2k matrix A2 whose 2x2 block
//     (i, p) is [[vr, -vi], [vi, vr]] for v = alpha * op(A)(i, p): column 2p
//     of A2 is the interleaved complex column alpha * op(A)(:, p), column
//     2p + 1 is i * alpha * op(A)(:, p);
//   * a column-major interleaved complex matrix X (r x c, leading dimension
//     ldx) read as real numbers is the 2r x c real matrix with leading
//     dimension 2 ldx whose rows 2i, 2i + 1 are Re and Im of row i;
//   * so with B and C read that way, the real GEMM
//         C (2m x n) = A2 (2m x 2k) * B (2k x n) + beta * C
//     forms row 2i = sum_p vr br - vi bi = Re(v b) and row 2i + 1 = Im(v b),
//     which is C = alpha op(A) op(B) + beta C.
//
// B and C are used in place; only op(A) is packed, plus op(B) into a k x n
// complex copy when B is transposed or conjugated.  The real GEMM does 8mnk
// flops, the complex operation count, so no arithmetic is wasted, and it runs
// on the tuned FP64 / FP32 tile configurations.  A complex beta scales C in a
// separate pass first (the real epilogue then adds the product); a real beta
// goes to the real epilogue directly, and beta == 0 never reads C.
#include <stdint.h>

#include "common.hpp"
#include "gemm.hpp"
#include "gemm_tile.hpp"
#include "ops.hpp"
#include "runtime.hpp"

namespace zg {

namespace {

template <typename T>
using cplx = vec_t<T, 2>;  // (re, im)

// Zynum's complex multiply (src/blas/core/shared/scalar.zig:77-82): four
// products, no fused operations (the library is built without contraction).
template <typename T>
__device__ __forceinline__ cplx<T> cmul(cplx<T> a, cplx<T> b) {
    cplx<T> r;
    r.x = a.x * b.x - a.y * b.y;
    r.y = a.x * b.y + a.y * b.x;
    return r;
}

// Caller matrices only guarantee the alignment of one real scalar, so their
// elements are read and written as two scalars.
template <typename T>
__device__ __forceinline__ cplx<T> cload(const T* p) {
    cplx<T> r;
    r.x = p[0];
    r.y = p[1];
    return r;
}

constexpr int kPackTile = 32;  // complex elements per side of a pack tile
constexpr int kPackThreads = 256;

// A2 (rows 0 .. 2 mc, columns 0 .. 2k) from rows i0 .. i0 + mc of op(A).
// TR: 0 op(A) = A, 1 A^T, 2 A^H.  ONE: alpha == 1 (the copy is exact).
template <typename T, int TR, bool ONE>
__global__ __launch_bounds__(kPackThreads) void cpack_a_kernel(int mc, int k, int i0, T alpha_re, T alpha_im,
                                                               const T* __restrict__ a, int64_t lda,
                                                               T* __restrict__ a2, int64_t ld2, int tiles_i) {
    __shared__ cplx<T> s[kPackTile][kPackTile + 1];  // [p][i]
    const int ti = (int)(blockIdx.x % (unsigned)tiles_i);
    const int tp = (int)(blockIdx.x / (unsigned)tiles_i);
    const int ib = ti * kPackTile, pb = tp * kPackTile;
    const int tx = threadIdx.x & (kPackTile - 1), ty = threadIdx.x / kPackTile;
    constexpr int kRowsPerPass = kPackThreads / kPackTile;
    // the source's contiguous dimension runs along tx
#pragma unroll
    for (int r = ty; r < kPackTile; r += kRowsPerPass) {
        if constexpr (TR == 0) {  // op(A)(i, p) = a[i + p lda]
            const int i = ib + tx, p = pb + r;
            if (i < mc && p < k) s[r][tx] = cload(a + 2 * ((int64_t)(i0 + i) + (int64_t)p * lda));
        } else {  // op(A)(i, p) = a[p + i lda] (conjugated for A^H)
            const int i = ib + r, p = pb + tx;
            if (i < mc && p < k) s[tx][r] = cload(a + 2 * ((int64_t)p + (int64_t)(i0 + i) * lda));
        }
    }
    __syncthreads();
    cplx<T> al;
    al.x = alpha_re;
    al.y = alpha_im;
#pragma unroll
    for (int r = ty; r < kPackTile; r += kRowsPerPass) {
        const int i = ib + tx, p = pb + r;
        if (i < mc && p < k) {
            cplx<T> x = s[r][tx];
            if constexpr (TR == 2) x.y = -x.y;
            const cplx<T> v = ONE ? x : cmul(al, x);
            cplx<T> w;
            w.x = -v.y;
            w.y = v.x;
            T* col = a2 + (int64_t)(2 * p) * ld2 + 2 * i;  // A2 columns are 2-scalar aligned
            *reinterpret_cast<cplx<T>*>(col) = v;
            *reinterpret_cast<cplx<T>*>(col + ld2) = w;
        }
    }
}

// bt (k x n, leading dimension ldt) = op(B) for op(B) = B^T (CONJ false) or B^H.
template <typename T, bool CONJ>
__global__ __launch_bounds__(kPackThreads) void cpack_bt_kernel(int k, int n, const T* __restrict__ b, int64_t ldb,
                                                                T* __restrict__ bt, int64_t ldt, int tiles_p) {
    __shared__ cplx<T> s[kPackTile][kPackTile + 1];  // [p][j]
    const int tp = (int)(blockIdx.x % (unsigned)tiles_p);
    const int tj = (int)(blockIdx.x / (unsigned)tiles_p);
    const int pb = tp * kPackTile, jb = tj * kPackTile;
    const int tx = threadIdx.x & (kPackTile - 1), ty = threadIdx.x / kPackTile;
    constexpr int kRowsPerPass = kPackThreads / kPackTile;
#pragma unroll
    for (int r = ty; r < kPackTile; r += kRowsPerPass) {  // op(B)(p, j) = b[j + p ldb]: contiguous in j = tx
        const int j = jb + tx, p = pb + r;
        if (j < n && p < k) s[r][tx] = cload(b + 2 * ((int64_t)j + (int64_t)p * ldb));
    }
    __syncthreads();
#pragma unroll
    for (int r = ty; r < kPackTile; r += kRowsPerPass) {  // bt column j is contiguous in p = tx
        const int p = pb + tx, j = jb + r;
        if (j < n && p < k) {
            cplx<T> x = s[tx][r];
            if constexpr (CONJ) x.y = -x.y;
            *reinterpret_cast<cplx<T>*>(bt + 2 * ((int64_t)p + (int64_t)j * ldt)) = x;
        }
    }
}

// C = beta * C (complex), beta == 0 writes zeros without reading C.
template <typename T>
__global__ void cscale_kernel(int m, int n, T beta_re, T beta_im, T* __restrict__ c, int64_t ldc) {
    const int64_t total = (int64_t)m * n;
    const bool zero = beta_re == T(0) && beta_im == T(0);
    cplx<T> be;
    be.x = beta_re;
    be.y = beta_im;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t j = e / m, i = e % m;
        T* cp = c + 2 * (j * ldc + i);
        cplx<T> v;
        if (zero) {
            v.x = T(0);
            v.y = T(0);
        } else {
            v = cmul(be, cload(cp));
        }
        cp[0] = v.x;
        cp[1] = v.y;
    }
}

template <typename T>
hipError_t cscale_launch(int m, int n, T beta_re, T beta_im, T* c, int64_t ldc, hipStream_t s) {
    const int64_t total = (int64_t)m * n;
    int blocks = (int)((total + 255) / 256);
    if (blocks > 65536) blocks = 65536;
    cscale_kernel<T><<<blocks, 256, 0, s>>>(m, n, beta_re, beta_im, c, ldc);
    return hipGetLastError();
}

template <typename T, int TR>
hipError_t cpack_a_tr(int mc, int k, int i0, T are, T aim, const T* a, int64_t lda, T* a2, int64_t ld2,
                      hipStream_t s) {
    const int tiles_i = (mc + kPackTile - 1) / kPackTile;
    const int64_t blocks = (int64_t)tiles_i * ((k + kPackTile - 1) / kPackTile);
    if (are == T(1) && aim == T(0))
        cpack_a_kernel<T, TR, true><<<(unsigned)blocks, kPackThreads, 0, s>>>(mc, k, i0, are, aim, a, lda, a2, ld2,
                                                                              tiles_i);
    else
        cpack_a_kernel<T, TR, false><<<(unsigned)blocks, kPackThreads, 0, s>>>(mc, k, i0, are, aim, a, lda, a2, ld2,
                                                                               tiles_i);
    return hipGetLastError();
}

template <typename T>
hipError_t cpack_a(char ta, int mc, int k, int i0, T are, T aim, const T* a, int64_t lda, T* a2, int64_t ld2,
                   hipStream_t s) {
    if (ta == 'N') return cpack_a_tr<T, 0>(mc, k, i0, are, aim, a, lda, a2, ld2, s);
    if (ta == 'T') return cpack_a_tr<T, 1>(mc, k, i0, are, aim, a, lda, a2, ld2, s);
    return cpack_a_tr<T, 2>(mc, k, i0, are, aim, a, lda, a2, ld2, s);
}

template <typename T>
hipError_t cpack_bt(bool conj, int k, int n, const T* b, int64_t ldb, T* bt, int64_t ldt, hipStream_t s) {
    const int tiles_p = (k + kPackTile - 1) / kPackTile;
    const int64_t blocks = (int64_t)tiles_p * ((n + kPackTile - 1) / kPackTile);
    if (conj)
        cpack_bt_kernel<T, true><<<(unsigned)blocks, kPackThreads, 0, s>>>(k, n, b, ldb, bt, ldt, tiles_p);
    else
        cpack_bt_kernel<T, false><<<(unsigned)blocks, kPackThreads, 0, s>>>(k, n, b, ldb, bt, ldt, tiles_p);
    return hipGetLastError();
}

inline int64_t round_up(int64_t x, int64_t q) { return (x + q - 1) / q * q; }

// Leading dimension (real scalars) of a packed operand of `rows` real rows:
// a multiple of 8 scalars plus 8, so columns stay 64-byte aligned without a
// power-of-two stride.
inline int64_t packed_ld(int64_t rows) { return round_up(rows, 8) + 8; }

}  // namespace

// Workspace budget of the expanded A2 per GEMM launch; larger problems are
// packed and multiplied in row chunks (no extra traffic on C: each chunk owns
// its rows of C).
constexpr size_t kComplexPackBudget = size_t(4) << 30;

template <typename T>
int cgemm_device(char ta, char tb, int m, int n, int k, T alpha_re, T alpha_im, const T* a, int64_t lda, const T* b,
                 int64_t ldb, T beta_re, T beta_im, T* c, int64_t ldc, hipStream_t s, int tile) {
    if (m <= 0 || n <= 0) return ZYNUM_GPU_OK;
    const bool alpha_zero = alpha_re == T(0) && alpha_im == T(0);
    if (k <= 0 || alpha_zero) return status_from_hip(cscale_launch<T>(m, n, beta_re, beta_im, c, ldc, s));
    // the real GEMM has 2m rows and depth 2k
    if (m > (1 << 29) || k > (1 << 29)) return ZYNUM_GPU_DECLINE_UNSUPPORTED;

    // row chunk of A2 within the budget (all of op(A) when it fits)
    const size_t col_bytes_per_row = (size_t)2 * (size_t)k * 2 * sizeof(T);  // two real rows, 2k columns
    int64_t mc = m;
    if ((size_t)packed_ld(2 * (int64_t)m) * 2 * k * sizeof(T) > kComplexPackBudget) {
        mc = (int64_t)(kComplexPackBudget / col_bytes_per_row) / 64 * 64;
        if (mc < 64) mc = 64;
        if (mc > m) mc = m;
    }
    const bool pack_b = tb != 'N';
    // one pack workgroup per 32 x 32 complex tile must fit a 1-D grid
    const int64_t max_grid = 0x7fffffff;
    const int64_t pt = kPackTile;
    if (((mc + pt - 1) / pt) * ((k + pt - 1) / pt) > max_grid) return ZYNUM_GPU_DECLINE_UNSUPPORTED;
    if (pack_b && ((k + pt - 1) / pt) * (((int64_t)n + pt - 1) / pt) > max_grid) return ZYNUM_GPU_DECLINE_UNSUPPORTED;
    const int64_t ld2 = packed_ld(2 * mc);
    const size_t a2_bytes = round_up((int64_t)((size_t)ld2 * 2 * k * sizeof(T)), 256);
    const int64_t ldt = pack_b ? packed_ld(2 * (int64_t)k) / 2 : 0;  // complex elements
    const size_t bt_bytes = pack_b ? (size_t)ldt * n * 2 * sizeof(T) : 0;

    char* ws = static_cast<char*>(workspace_acquire(a2_bytes + bt_bytes + 256, s));
    if (!ws) return ZYNUM_GPU_DECLINE_ALLOC;
    T* a2 = reinterpret_cast<T*>(ws);
    T* bt = reinterpret_cast<T*>(ws + a2_bytes);

    hipError_t e = hipSuccess;
    T real_beta = beta_re;
    if (beta_im != T(0)) {  // C = beta C first, then C += A2 B
        e = cscale_launch<T>(m, n, beta_re, beta_im, c, ldc, s);
        real_beta = T(1);
    }
    const T* breal = b;
    int64_t ldbr = 2 * ldb;
    if (e == hipSuccess && pack_b) {
        e = cpack_bt<T>(tb == 'C', k, n, b, ldb, bt, ldt, s);
        breal = bt;
        ldbr = 2 * ldt;
    }
    for (int64_t i0 = 0; e == hipSuccess && i0 < m; i0 += mc) {
        const int rows = (int)((m - i0 < mc) ? (m - i0) : mc);
        e = cpack_a<T>(ta, rows, k, (int)i0, alpha_re, alpha_im, a, lda, a2, ld2, s);
        if (e != hipSuccess) break;
        GemmDesc<T> d;
        d.ta = false;
        d.tb = false;
        d.m = 2 * rows;
        d.n = n;
        d.k = 2 * k;
        d.lda = ld2;
        d.ldb = ldbr;
        d.ldc = 2 * ldc;
        d.alpha = T(1);
        d.beta = real_beta;
        d.a = a2;
        d.b = breal;
        d.c = c + 2 * i0;
        int kchunk = 0;  // few output tiles and a long K: split K (partials from the scratch cache)
        const int parts = tile == kTileAuto ? gemm_splitk_plan(d.m, d.n, d.k, &kchunk) : 1;
        size_t got = 0;
        void* part = parts > 1 ? scratch_acquire((size_t)parts * d.m * d.n * sizeof(T), &got) : nullptr;
        if (part != nullptr) {
            e = gemm_splitk_launch<T>(d, s, static_cast<T*>(part), parts, kchunk);
            const hipError_t e2 = hipStreamSynchronize(s);
            if (e == hipSuccess) e = e2;
            if (e == hipSuccess) scratch_release(part, got);
        } else {
            e = gemm_launch<T>(d, s, tile);
        }
    }
    const int st = status_from_hip(e);
    workspace_release(s);
    return st;
}

template int cgemm_device<double>(char, char, int, int, int, double, double, const double*, int64_t, const double*,
                                  int64_t, double, double, double*, int64_t, hipStream_t, int);
template int cgemm_device<float>(char, char, int, int, int, float, float, const float*, int64_t, const float*,
                                 int64_t, float, float, float*, int64_t, hipStream_t, int);

}  // namespace zg
