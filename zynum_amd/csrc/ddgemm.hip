// Double-double GEMM on FP64 VALU: the on-device accuracy reference the
// split-precision forms are measured against.  Each product is formed exactly
// with TwoProd (fma), accumulated with TwoSum (Knuth, TAOCP vol. 2; Ogita,
// Rump and Oishi 2005), then alpha and beta*C are applied in double-double.
#include "common.hpp"
#include "ops.hpp"
#include "runtime.hpp"

namespace zg {

struct dd {
    double hi, lo;
};

__device__ __forceinline__ dd two_sum(double a, double b) {
    double s = a + b;
    double bb = s - a;
    return {s, (a - (s - bb)) + (b - bb)};
}
__device__ __forceinline__ dd fast_two_sum(double a, double b) {
    double s = a + b;
    return {s, b - (s - a)};
}
__device__ __forceinline__ dd dd_add(dd x, dd y) {
    dd s = two_sum(x.hi, y.hi);
    dd t = two_sum(x.lo, y.lo);
    s.lo += t.hi;
    s = fast_two_sum(s.hi, s.lo);
    s.lo += t.lo;
    return fast_two_sum(s.hi, s.lo);
}
__device__ __forceinline__ dd two_prod(double a, double b) {
    double p = a * b;
    return {p, __builtin_fma(a, b, -p)};
}
__device__ __forceinline__ dd dd_mul_d(dd x, double y) {
    double p = x.hi * y;
    double e = __builtin_fma(x.hi, y, -p);
    e = __builtin_fma(x.lo, y, e);
    return fast_two_sum(p, e);
}

constexpr int DT = 16;  // 16 x 16 outputs per workgroup, 16-deep K slices in LDS

template <bool TA, bool TB>
__global__ __launch_bounds__(256) void dd_gemm_kernel(int m, int n, int k, double alpha, const double* __restrict__ a,
                                                       int64_t lda, const double* __restrict__ b, int64_t ldb,
                                                       double beta, const double* __restrict__ c0, int64_t ldc0,
                                                       double* __restrict__ hi, double* __restrict__ lo,
                                                       int64_t ldo) {
    __shared__ double sa[DT][DT + 1];  // sa[kk][i]
    __shared__ double sb[DT][DT + 1];  // sb[kk][j]
    const int tx = threadIdx.x % DT, ty = threadIdx.x / DT;
    const int i0 = blockIdx.x * DT, j0 = blockIdx.y * DT;
    const int i = i0 + tx, j = j0 + ty;
    dd s = {0.0, 0.0};
    for (int k0 = 0; k0 < k; k0 += DT) {
        {  // load op(A)(i0 + ty', k0 + tx') and op(B)(k0 + ., j0 + .)
            const int r = i0 + (TA ? ty : tx), kk = k0 + (TA ? tx : ty);
            double v = 0.0;
            if (r < m && kk < k) v = TA ? a[(int64_t)r * lda + kk] : a[(int64_t)kk * lda + r];
            if (TA) sa[tx][ty] = v; else sa[ty][tx] = v;
            const int cidx = j0 + (TB ? tx : ty), kb = k0 + (TB ? ty : tx);
            double w = 0.0;
            if (cidx < n && kb < k) w = TB ? b[(int64_t)kb * ldb + cidx] : b[(int64_t)cidx * ldb + kb];
            if (TB) sb[ty][tx] = w; else sb[tx][ty] = w;
        }
        __syncthreads();
        const int kend = min(DT, k - k0);
        for (int kk = 0; kk < kend; ++kk) s = dd_add(s, two_prod(sa[kk][tx], sb[kk][ty]));
        __syncthreads();
    }
    if (i < m && j < n) {
        s = dd_mul_d(s, alpha);
        if (beta != 0.0) s = dd_add(s, two_prod(beta, c0[(int64_t)j * ldc0 + i]));
        hi[(int64_t)j * ldo + i] = s.hi;
        if (lo) lo[(int64_t)j * ldo + i] = s.lo;
    }
}

hipError_t dd_gemm_launch(bool ta, bool tb, int m, int n, int k, double alpha, const double* a, int64_t lda,
                          const double* b, int64_t ldb, double beta, const double* c0, int64_t ldc0, double* hi,
                          double* lo, int64_t ldo, hipStream_t s) {
    dim3 grid((m + DT - 1) / DT, (n + DT - 1) / DT);
    if (!ta && !tb) dd_gemm_kernel<false, false><<<grid, 256, 0, s>>>(m, n, k, alpha, a, lda, b, ldb, beta, c0, ldc0, hi, lo, ldo);
    else if (!ta && tb) dd_gemm_kernel<false, true><<<grid, 256, 0, s>>>(m, n, k, alpha, a, lda, b, ldb, beta, c0, ldc0, hi, lo, ldo);
    else if (ta && !tb) dd_gemm_kernel<true, false><<<grid, 256, 0, s>>>(m, n, k, alpha, a, lda, b, ldb, beta, c0, ldc0, hi, lo, ldo);
    else dd_gemm_kernel<true, true><<<grid, 256, 0, s>>>(m, n, k, alpha, a, lda, b, ldb, beta, c0, ldc0, hi, lo, ldo);
    return hipGetLastError();
}

// C <- RN(dd(alpha*op(A)op(B) + beta*C)); C is read before it is written by the
// same thread, so in-place use is safe.
int dd_dgemm(bool ta, bool tb, int m, int n, int k, double alpha, const double* a, int64_t lda, const double* b,
             int64_t ldb, double beta, double* c, int64_t ldc, hipStream_t s) {
    return status_from_hip(dd_gemm_launch(ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, c, nullptr, ldc, s));
}

}  // namespace zg

extern "C" int zynum_gpu_dd_gemm(char transa, char transb, int m, int n, int k, double alpha, const double* a,
                                 int lda, const double* b, int ldb, double beta, const double* c0, int ldc0,
                                 double* hi, double* lo, int ldo, void* stream) {
    if (m < 0 || n < 0 || k < 0 || ldo < (m > 1 ? m : 1)) return ZYNUM_GPU_DECLINE_ARGUMENT;
    if (!zg::device_ready()) return ZYNUM_GPU_DECLINE_NO_DEVICE;
    if (m == 0 || n == 0) return ZYNUM_GPU_OK;
    const bool ta = transa != 'N' && transa != 'n', tb = transb != 'N' && transb != 'n';
    hipStream_t s = stream ? reinterpret_cast<hipStream_t>(stream) : zg::blas_stream();
    int st = zg::status_from_hip(zg::dd_gemm_launch(ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c0, ldc0, hi, lo, ldo, s));
    if (st != ZYNUM_GPU_OK) return st;
    return zg::status_from_hip(hipStreamSynchronize(s));
}
