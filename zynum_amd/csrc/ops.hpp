// Checked, residency-aware operations behind the C ABI.  Each returns a
// ZYNUM_GPU_* code; nonzero codes other than ZYNUM_GPU_ERR_LAUNCH are
// returned before any caller-visible memory is written.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace zg {

template <typename T>
int gemm_op(bool ta, bool tb, int m, int n, int k, T alpha, const T* a, int64_t lda, const T* b, int64_t ldb,
            T beta, T* c, int64_t ldc, hipStream_t s, int mode, int mode_param, int flags, int tile = -1);

template <typename T>
int trsm_op(char side, char uplo, bool trans, char diag, int m, int n, T alpha, const T* a, int64_t lda, T* b,
            int64_t ldb, hipStream_t s, int mode, int mode_param, int flags);

template <typename T>
int syrk_op(char uplo, bool trans, int n, int k, T alpha, const T* a, int64_t lda, T beta, T* c, int64_t ldc,
            hipStream_t s, int mode, int mode_param, int flags);

template <typename T>
int symm_op(char side, char uplo, int m, int n, T alpha, const T* a, int64_t lda, const T* b, int64_t ldb, T beta,
            T* c, int64_t ldc, hipStream_t s, int flags);

template <typename T>
int trmm_op(char side, char uplo, bool trans, char diag, int m, int n, T alpha, const T* a, int64_t lda, T* b,
            int64_t ldb, hipStream_t s, int flags);

template <typename T>
int syr2k_op(char uplo, bool trans, int n, int k, T alpha, const T* a, int64_t lda, const T* b, int64_t ldb, T beta,
             T* c, int64_t ldc, hipStream_t s, int flags);

// Complex GEMM; a, b, c point at interleaved (re, im) storage, leading
// dimensions and alpha/beta pointers are in complex elements.  ta/tb: 'N', 'T', 'C'.
template <typename T>
int cgemm_op(char ta, char tb, int m, int n, int k, const T* alpha, const T* a, int64_t lda, const T* b,
             int64_t ldb, const T* beta, T* c, int64_t ldc, hipStream_t s, int flags, int tile = -1);

// Device-pointer implementations (no staging), stream ordered.
template <typename T>
int cgemm_device(char ta, char tb, int m, int n, int k, T alpha_re, T alpha_im, const T* a, int64_t lda, const T* b,
                 int64_t ldb, T beta_re, T beta_im, T* c, int64_t ldc, hipStream_t s, int tile);
int ozaki_dgemm(bool ta, bool tb, int m, int n, int k, double alpha, const double* a, int64_t lda,
                const double* b, int64_t ldb, double beta, double* c, int64_t ldc, hipStream_t s, int mode,
                int mode_param, int tile = -1);
int dd_dgemm(bool ta, bool tb, int m, int n, int k, double alpha, const double* a, int64_t lda, const double* b,
             int64_t ldb, double beta, double* c, int64_t ldc, hipStream_t s);
template <typename T>
int trsm_device(char side, char uplo, bool trans, char diag, int m, int n, T alpha, const T* a, int64_t lda, T* b,
                int64_t ldb, hipStream_t s, int mode, int mode_param);
template <typename T>
int syrk_device(char uplo, bool trans, int n, int k, T alpha, const T* a, int64_t lda, T beta, T* c, int64_t ldc,
                hipStream_t s, int mode, int mode_param);
template <typename T>
int symm_device(char side, char uplo, int m, int n, T alpha, const T* a, int64_t lda, const T* b, int64_t ldb, T beta,
                T* c, int64_t ldc, hipStream_t s);
template <typename T>
int trmm_device(char side, char uplo, bool trans, char diag, int m, int n, T alpha, const T* a, int64_t lda, T* b,
                int64_t ldb, hipStream_t s);
template <typename T>
int syr2k_device(char uplo, bool trans, int n, int k, T alpha, const T* a, int64_t lda, const T* b, int64_t ldb,
                 T beta, T* c, int64_t ldc, hipStream_t s);

}  // namespace zg
