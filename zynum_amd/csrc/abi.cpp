// Exported BLAS symbols and the private GPU-tier request ABI.
//
// Argument checking, xerbla reporting and CBLAS row-major normalisation keep
// Zynum's semantics (src/blas/abi/fortran.zig:16-47,88-117,239-282,760-876;
// src/blas/abi/cblas.zig:16-116,1415-1489,1550-1559,1605-1627,1677-1753).
// Execution is always on the GPU: device operands are used in place and host
// operands are staged through device memory.  There is no CPU compute path.
#include <ctype.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>

#include "abi_common.hpp"
#include "gemm.hpp"
#include "ops.hpp"
#include "ozaki.hpp"
#include "runtime.hpp"

using zg::gemm_error;
using zg::scalar_from_bits;

namespace {

thread_local int t_last_status = ZYNUM_GPU_OK;

inline char up(const char* p) { return (char)toupper((unsigned char)p[0]); }
inline bool valid_trans(char c) { return c == 'N' || c == 'T' || c == 'C'; }
inline bool valid_uplo(char c) { return c == 'U' || c == 'L'; }
inline bool valid_diag(char c) { return c == 'N' || c == 'U'; }
inline bool valid_side(char c) { return c == 'L' || c == 'R'; }
inline int32_t max1(int32_t x) { return x > 1 ? x : 1; }

// fortran.zig:105-117
int triangular_error(char side, char uplo, char trans, char diag, int32_t m, int32_t n, int32_t lda, int32_t ldb) {
    if (!valid_side(side)) return 1;
    if (!valid_uplo(uplo)) return 2;
    if (!valid_trans(trans)) return 3;
    if (!valid_diag(diag)) return 4;
    if (m < 0) return 5;
    if (n < 0) return 6;
    const int32_t nrowa = (side == 'L') ? m : n;
    if (lda < max1(nrowa)) return 9;
    if (ldb < max1(m)) return 11;
    return 0;
}

// fortran.zig:252-262
int rank_k_error(char uplo, char trans, int32_t n, int32_t k, int32_t lda, int32_t ldc) {
    if (!valid_uplo(uplo)) return 1;
    if (!valid_trans(trans)) return 2;
    if (n < 0) return 3;
    if (k < 0) return 4;
    const int32_t nrowa = (trans == 'N') ? n : k;
    if (lda < max1(nrowa)) return 7;
    if (ldc < max1(n)) return 10;
    return 0;
}

// fortran.zig:239-250
int symmetric_error(char side, char uplo, int32_t m, int32_t n, int32_t lda, int32_t ldb, int32_t ldc) {
    if (!valid_side(side)) return 1;
    if (!valid_uplo(uplo)) return 2;
    if (m < 0) return 3;
    if (n < 0) return 4;
    const int32_t nrowa = (side == 'L') ? m : n;
    if (lda < max1(nrowa)) return 7;
    if (ldb < max1(m)) return 9;
    if (ldc < max1(m)) return 12;
    return 0;
}

// fortran.zig:264-275
int rank_2k_error(char uplo, char trans, int32_t n, int32_t k, int32_t lda, int32_t ldb, int32_t ldc) {
    if (!valid_uplo(uplo)) return 1;
    if (!valid_trans(trans)) return 2;
    if (n < 0) return 3;
    if (k < 0) return 4;
    const int32_t nrow = (trans == 'N') ? n : k;
    if (lda < max1(nrow)) return 7;
    if (ldb < max1(nrow)) return 9;
    if (ldc < max1(n)) return 12;
    return 0;
}

// fortran.zig:277-282: report through xerbla_ with the padded routine name
bool report_error(const char* name, int info) {
    if (info == 0) return false;
    zynum_blas_int i = info;
    xerbla_(name, &i);
    return true;
}

void note_status(const char* routine, int st) {
    t_last_status = st;
    if (st != ZYNUM_GPU_OK && st != ZYNUM_GPU_DECLINE_FORCED) {
        fprintf(stderr, " ** %s: GPU tier did not run (%s); outputs %s\n", routine, zynum_gpu_status_string(st),
                st == ZYNUM_GPU_ERR_LAUNCH ? "undefined" : "unchanged");
    }
}

}  // namespace

namespace zg {

template <typename T>
int gemm_op(bool ta, bool tb, int m, int n, int k, T alpha, const T* a, int64_t lda, const T* b, int64_t ldb,
            T beta, T* c, int64_t ldc, hipStream_t s, int mode, int mode_param, int flags, int tile) {
    if (m <= 0 || n <= 0) return ZYNUM_GPU_OK;  // gemm.zig:1393
    if (take_forced_decline()) return ZYNUM_GPU_DECLINE_FORCED;
    if (!device_ready()) return ZYNUM_GPU_DECLINE_NO_DEVICE;
    if (mode != ZYNUM_GPU_MODE_PLAIN && sizeof(T) != 8) return ZYNUM_GPU_DECLINE_UNSUPPORTED;
    if (mode < ZYNUM_GPU_MODE_PLAIN || mode > ZYNUM_GPU_MODE_DD) return ZYNUM_GPU_DECLINE_UNSUPPORTED;

    const bool reads_ab = (k > 0) && (alpha != T(0));
    Operand oa, ob, oc;
    int st = ZYNUM_GPU_OK;
    if (reads_ab) {
        st = oa.bind_in(a, ta ? k : m, ta ? m : k, lda, sizeof(T), s);
        if (st == ZYNUM_GPU_OK) st = ob.bind_in(b, tb ? n : k, tb ? k : n, ldb, sizeof(T), s);
    }
    if (st == ZYNUM_GPU_OK) st = oc.bind_out(c, m, n, ldc, sizeof(T), beta != T(0), s);
    if (st != ZYNUM_GPU_OK) return st;

    const T* da = static_cast<const T*>(oa.dev);
    const T* db = static_cast<const T*>(ob.dev);
    T* dc = static_cast<T*>(oc.dev);
    if (!reads_ab || mode == ZYNUM_GPU_MODE_PLAIN) {
        GemmDesc<T> d;
        d.ta = ta; d.tb = tb; d.m = m; d.n = n; d.k = k;
        d.lda = lda; d.ldb = ldb; d.ldc = ldc;
        d.alpha = alpha; d.beta = beta;
        d.a = da; d.b = db; d.c = dc;
        int kchunk = 0;
        const int parts = (reads_ab && tile == kTileAuto) ? gemm_splitk_plan(m, n, k, &kchunk) : 1;
        size_t got = 0;
        void* part = parts > 1 ? scratch_acquire((size_t)parts * m * n * sizeof(T), &got) : nullptr;
        if (part != nullptr) {
            hipError_t e = gemm_splitk_launch<T>(d, s, static_cast<T*>(part), parts, kchunk);
            const hipError_t e2 = hipStreamSynchronize(s);  // the partial buffer goes back to the cache
            st = status_from_hip(e != hipSuccess ? e : e2);
            if (e == hipSuccess && e2 == hipSuccess) scratch_release(part, got);  // else leaked, never reused in flight
        } else {
            st = status_from_hip(gemm_launch<T>(d, s, tile));
        }
    } else if (mode == ZYNUM_GPU_MODE_DD) {
        if constexpr (sizeof(T) == 8) st = dd_dgemm(ta, tb, m, n, k, alpha, da, lda, db, ldb, beta, dc, ldc, s);
    } else {
        if constexpr (sizeof(T) == 8)
            st = ozaki_dgemm(ta, tb, m, n, k, alpha, da, lda, db, ldb, beta, dc, ldc, s, mode, mode_param, tile);
    }
    if (st != ZYNUM_GPU_OK) return st;
    if (oc.staged) return oc.finish(s);
    if (!(flags & ZYNUM_GPU_FLAG_ASYNC) || oa.staged || ob.staged) return status_from_hip(hipStreamSynchronize(s));
    return ZYNUM_GPU_OK;
}

template <typename T>
int cgemm_op(char ta, char tb, int m, int n, int k, const T* alpha, const T* a, int64_t lda, const T* b,
             int64_t ldb, const T* beta, T* c, int64_t ldc, hipStream_t s, int flags, int tile) {
    if (m <= 0 || n <= 0) return ZYNUM_GPU_OK;  // gemm.zig:1393
    if (take_forced_decline()) return ZYNUM_GPU_DECLINE_FORCED;
    if (!device_ready()) return ZYNUM_GPU_DECLINE_NO_DEVICE;
    const size_t es = 2 * sizeof(T);
    const bool reads_ab = (k > 0) && !(alpha[0] == T(0) && alpha[1] == T(0));
    const bool beta_zero = beta[0] == T(0) && beta[1] == T(0);
    Operand oa, ob, oc;
    int st = ZYNUM_GPU_OK;
    if (reads_ab) {
        st = oa.bind_in(a, ta != 'N' ? k : m, ta != 'N' ? m : k, lda, es, s);
        if (st == ZYNUM_GPU_OK) st = ob.bind_in(b, tb != 'N' ? n : k, tb != 'N' ? k : n, ldb, es, s);
    }
    if (st == ZYNUM_GPU_OK) st = oc.bind_out(c, m, n, ldc, es, !beta_zero, s);
    if (st != ZYNUM_GPU_OK) return st;
    st = cgemm_device<T>(ta, tb, m, n, reads_ab ? k : 0, alpha[0], alpha[1], static_cast<const T*>(oa.dev), lda,
                         static_cast<const T*>(ob.dev), ldb, beta[0], beta[1], static_cast<T*>(oc.dev), ldc, s, tile);
    if (st != ZYNUM_GPU_OK) return st;
    if (oc.staged) return oc.finish(s);
    if (!(flags & ZYNUM_GPU_FLAG_ASYNC) || oa.staged || ob.staged) return status_from_hip(hipStreamSynchronize(s));
    return ZYNUM_GPU_OK;
}

template <typename T>
int trsm_op(char side, char uplo, bool trans, char diag, int m, int n, T alpha, const T* a, int64_t lda, T* b,
            int64_t ldb, hipStream_t s, int mode, int mode_param, int flags) {
    if (m <= 0 || n <= 0) return ZYNUM_GPU_OK;  // triangular.zig:234
    if (take_forced_decline()) return ZYNUM_GPU_DECLINE_FORCED;
    if (!device_ready()) return ZYNUM_GPU_DECLINE_NO_DEVICE;
    if (mode != ZYNUM_GPU_MODE_PLAIN && sizeof(T) != 8) return ZYNUM_GPU_DECLINE_UNSUPPORTED;
    const int na = (side == 'L') ? m : n;
    Operand oa, ob;
    int st = oa.bind_in(a, na, na, lda, sizeof(T), s);
    if (st == ZYNUM_GPU_OK) st = ob.bind_out(b, m, n, ldb, sizeof(T), true, s);
    if (st != ZYNUM_GPU_OK) return st;
    st = trsm_device<T>(side, uplo, trans, diag, m, n, alpha, static_cast<const T*>(oa.dev), lda,
                        static_cast<T*>(ob.dev), ldb, s, mode, mode_param);
    if (st != ZYNUM_GPU_OK) return st;
    if (ob.staged) return ob.finish(s);
    if (!(flags & ZYNUM_GPU_FLAG_ASYNC) || oa.staged) return status_from_hip(hipStreamSynchronize(s));
    return ZYNUM_GPU_OK;
}

template <typename T>
int syrk_op(char uplo, bool trans, int n, int k, T alpha, const T* a, int64_t lda, T beta, T* c, int64_t ldc,
            hipStream_t s, int mode, int mode_param, int flags) {
    if (n <= 0) return ZYNUM_GPU_OK;  // symmetric.zig:238
    if (take_forced_decline()) return ZYNUM_GPU_DECLINE_FORCED;
    if (!device_ready()) return ZYNUM_GPU_DECLINE_NO_DEVICE;
    if (mode != ZYNUM_GPU_MODE_PLAIN && sizeof(T) != 8) return ZYNUM_GPU_DECLINE_UNSUPPORTED;
    Operand oa, oc;
    int st = ZYNUM_GPU_OK;
    if (k > 0) st = oa.bind_in(a, trans ? k : n, trans ? n : k, lda, sizeof(T), s);
    // the unstored triangle must come back unchanged, so C is always copied in
    if (st == ZYNUM_GPU_OK) st = oc.bind_out(c, n, n, ldc, sizeof(T), true, s);
    if (st != ZYNUM_GPU_OK) return st;
    st = syrk_device<T>(uplo, trans, n, k, alpha, static_cast<const T*>(oa.dev), lda, beta, static_cast<T*>(oc.dev),
                        ldc, s, mode, mode_param);
    if (st != ZYNUM_GPU_OK) return st;
    if (oc.staged) return oc.finish(s);
    if (!(flags & ZYNUM_GPU_FLAG_ASYNC) || oa.staged) return status_from_hip(hipStreamSynchronize(s));
    return ZYNUM_GPU_OK;
}

template <typename T>
int symm_op(char side, char uplo, int m, int n, T alpha, const T* a, int64_t lda, const T* b, int64_t ldb, T beta,
            T* c, int64_t ldc, hipStream_t s, int flags) {
    if (m <= 0 || n <= 0) return ZYNUM_GPU_OK;  // symmetric.zig:108
    if (take_forced_decline()) return ZYNUM_GPU_DECLINE_FORCED;
    if (!device_ready()) return ZYNUM_GPU_DECLINE_NO_DEVICE;
    const int na = (side == 'L') ? m : n;
    Operand oa, ob, oc;
    // A and B are read whatever alpha is (symmetric.zig:56-70 has no alpha == 0 quick return)
    int st = oa.bind_in(a, na, na, lda, sizeof(T), s);
    if (st == ZYNUM_GPU_OK) st = ob.bind_in(b, m, n, ldb, sizeof(T), s);
    if (st == ZYNUM_GPU_OK) st = oc.bind_out(c, m, n, ldc, sizeof(T), beta != T(0), s);
    if (st != ZYNUM_GPU_OK) return st;
    st = symm_device<T>(side, uplo, m, n, alpha, static_cast<const T*>(oa.dev), lda, static_cast<const T*>(ob.dev), ldb,
                        beta, static_cast<T*>(oc.dev), ldc, s);
    if (st != ZYNUM_GPU_OK) return st;
    if (oc.staged) return oc.finish(s);
    if (!(flags & ZYNUM_GPU_FLAG_ASYNC) || oa.staged || ob.staged) return status_from_hip(hipStreamSynchronize(s));
    return ZYNUM_GPU_OK;
}

template <typename T>
int trmm_op(char side, char uplo, bool trans, char diag, int m, int n, T alpha, const T* a, int64_t lda, T* b,
            int64_t ldb, hipStream_t s, int flags) {
    if (m <= 0 || n <= 0) return ZYNUM_GPU_OK;  // triangular.zig:203
    if (take_forced_decline()) return ZYNUM_GPU_DECLINE_FORCED;
    if (!device_ready()) return ZYNUM_GPU_DECLINE_NO_DEVICE;
    const int na = (side == 'L') ? m : n;
    Operand oa, ob;
    int st = ZYNUM_GPU_OK;
    if (alpha != T(0)) st = oa.bind_in(a, na, na, lda, sizeof(T), s);
    if (st == ZYNUM_GPU_OK) st = ob.bind_out(b, m, n, ldb, sizeof(T), alpha != T(0), s);
    if (st != ZYNUM_GPU_OK) return st;
    st = trmm_device<T>(side, uplo, trans, diag, m, n, alpha, static_cast<const T*>(oa.dev), lda,
                        static_cast<T*>(ob.dev), ldb, s);
    if (st != ZYNUM_GPU_OK) return st;
    if (ob.staged) return ob.finish(s);
    if (!(flags & ZYNUM_GPU_FLAG_ASYNC) || oa.staged) return status_from_hip(hipStreamSynchronize(s));
    return ZYNUM_GPU_OK;
}

template <typename T>
int syr2k_op(char uplo, bool trans, int n, int k, T alpha, const T* a, int64_t lda, const T* b, int64_t ldb, T beta,
             T* c, int64_t ldc, hipStream_t s, int flags) {
    if (n <= 0) return ZYNUM_GPU_OK;  // symmetric.zig:367
    if (take_forced_decline()) return ZYNUM_GPU_DECLINE_FORCED;
    if (!device_ready()) return ZYNUM_GPU_DECLINE_NO_DEVICE;
    Operand oa, ob, oc;
    int st = ZYNUM_GPU_OK;
    if (k > 0 && alpha != T(0)) {
        st = oa.bind_in(a, trans ? k : n, trans ? n : k, lda, sizeof(T), s);
        if (st == ZYNUM_GPU_OK) st = ob.bind_in(b, trans ? k : n, trans ? n : k, ldb, sizeof(T), s);
    }
    // the unstored triangle must come back unchanged, so C is always copied in
    if (st == ZYNUM_GPU_OK) st = oc.bind_out(c, n, n, ldc, sizeof(T), true, s);
    if (st != ZYNUM_GPU_OK) return st;
    st = syr2k_device<T>(uplo, trans, n, k, alpha, static_cast<const T*>(oa.dev), lda, static_cast<const T*>(ob.dev),
                         ldb, beta, static_cast<T*>(oc.dev), ldc, s);
    if (st != ZYNUM_GPU_OK) return st;
    if (oc.staged) return oc.finish(s);
    if (!(flags & ZYNUM_GPU_FLAG_ASYNC) || oa.staged || ob.staged) return status_from_hip(hipStreamSynchronize(s));
    return ZYNUM_GPU_OK;
}

template int gemm_op<double>(bool, bool, int, int, int, double, const double*, int64_t, const double*, int64_t,
                             double, double*, int64_t, hipStream_t, int, int, int, int);
template int gemm_op<float>(bool, bool, int, int, int, float, const float*, int64_t, const float*, int64_t, float,
                            float*, int64_t, hipStream_t, int, int, int, int);
template int cgemm_op<double>(char, char, int, int, int, const double*, const double*, int64_t, const double*,
                              int64_t, const double*, double*, int64_t, hipStream_t, int, int);
template int cgemm_op<float>(char, char, int, int, int, const float*, const float*, int64_t, const float*, int64_t,
                             const float*, float*, int64_t, hipStream_t, int, int);

}  // namespace zg

// ---------------------------------------------------------------------------
// Fortran ABI (fortran.zig)
// ---------------------------------------------------------------------------
extern "C" {

int lsame_(const char* ca, const char* cb) { return toupper((unsigned char)ca[0]) == toupper((unsigned char)cb[0]); }

// fortran.zig:20-39: the routine name is read up to 6 characters or a NUL and
// trimmed of spaces.
void xerbla_(const char* srname, const zynum_blas_int* info) {
    char name[8];
    size_t len = 0;
    while (len < 6 && srname[len] != 0) {
        name[len] = srname[len];
        ++len;
    }
    size_t start = 0;
    while (start < len && name[start] == ' ') ++start;
    while (len > start && name[len - 1] == ' ') --len;
    name[len] = 0;
    const char* routine = (len > start) ? name + start : "UNKNOWN";
    fprintf(stderr, " ** On entry to %s parameter number %d had an illegal value\n", routine, (int)*info);
}

void zynum_blas_shutdown(void) { zg::workspace_free_all(); }
void zynum_blas_shutdown_(void) { zg::workspace_free_all(); }

void sgemm_(const char* ta, const char* tb, const zynum_blas_int* m, const zynum_blas_int* n, const zynum_blas_int* k,
            const float* alpha, const float* a, const zynum_blas_int* lda, const float* b, const zynum_blas_int* ldb,
            const float* beta, float* c, const zynum_blas_int* ldc) {
    const char A = up(ta), B = up(tb);
    if (report_error("SGEMM ", gemm_error(A, B, *m, *n, *k, *lda, *ldb, *ldc))) return;
    note_status("SGEMM", zg::gemm_op<float>(A != 'N', B != 'N', *m, *n, *k, *alpha, a, *lda, b, *ldb, *beta, c, *ldc,
                                            zg::blas_stream(), ZYNUM_GPU_MODE_PLAIN, 0, 0));
}

void dgemm_(const char* ta, const char* tb, const zynum_blas_int* m, const zynum_blas_int* n, const zynum_blas_int* k,
            const double* alpha, const double* a, const zynum_blas_int* lda, const double* b,
            const zynum_blas_int* ldb, const double* beta, double* c, const zynum_blas_int* ldc) {
    const char A = up(ta), B = up(tb);
    if (report_error("DGEMM ", gemm_error(A, B, *m, *n, *k, *lda, *ldb, *ldc))) return;
    note_status("DGEMM", zg::gemm_op<double>(A != 'N', B != 'N', *m, *n, *k, *alpha, a, *lda, b, *ldb, *beta, c,
                                             *ldc, zg::blas_stream(), ZYNUM_GPU_MODE_PLAIN, 0, 0));
}

// fortran.zig:768-775: 'C' is the conjugate transpose for the complex forms.
void cgemm_(const char* ta, const char* tb, const zynum_blas_int* m, const zynum_blas_int* n, const zynum_blas_int* k,
            const zynum_blas_complex_float* alpha, const zynum_blas_complex_float* a, const zynum_blas_int* lda,
            const zynum_blas_complex_float* b, const zynum_blas_int* ldb, const zynum_blas_complex_float* beta,
            zynum_blas_complex_float* c, const zynum_blas_int* ldc) {
    const char A = up(ta), B = up(tb);
    if (report_error("CGEMM ", gemm_error(A, B, *m, *n, *k, *lda, *ldb, *ldc))) return;
    note_status("CGEMM", zg::cgemm_op<float>(A, B, *m, *n, *k, &alpha->real, reinterpret_cast<const float*>(a), *lda,
                                             reinterpret_cast<const float*>(b), *ldb, &beta->real,
                                             reinterpret_cast<float*>(c), *ldc, zg::blas_stream(), 0));
}

void zgemm_(const char* ta, const char* tb, const zynum_blas_int* m, const zynum_blas_int* n, const zynum_blas_int* k,
            const zynum_blas_complex_double* alpha, const zynum_blas_complex_double* a, const zynum_blas_int* lda,
            const zynum_blas_complex_double* b, const zynum_blas_int* ldb, const zynum_blas_complex_double* beta,
            zynum_blas_complex_double* c, const zynum_blas_int* ldc) {
    const char A = up(ta), B = up(tb);
    if (report_error("ZGEMM ", gemm_error(A, B, *m, *n, *k, *lda, *ldb, *ldc))) return;
    note_status("ZGEMM", zg::cgemm_op<double>(A, B, *m, *n, *k, &alpha->real, reinterpret_cast<const double*>(a), *lda,
                                              reinterpret_cast<const double*>(b), *ldb, &beta->real,
                                              reinterpret_cast<double*>(c), *ldc, zg::blas_stream(), 0));
}

void ssyrk_(const char* u, const char* t, const zynum_blas_int* n, const zynum_blas_int* k, const float* alpha,
            const float* a, const zynum_blas_int* lda, const float* beta, float* c, const zynum_blas_int* ldc) {
    const char U = up(u), T = up(t);
    if (report_error("SSYRK ", rank_k_error(U, T, *n, *k, *lda, *ldc))) return;
    note_status("SSYRK", zg::syrk_op<float>(U, T != 'N', *n, *k, *alpha, a, *lda, *beta, c, *ldc, zg::blas_stream(),
                                            ZYNUM_GPU_MODE_PLAIN, 0, 0));
}

void dsyrk_(const char* u, const char* t, const zynum_blas_int* n, const zynum_blas_int* k, const double* alpha,
            const double* a, const zynum_blas_int* lda, const double* beta, double* c, const zynum_blas_int* ldc) {
    const char U = up(u), T = up(t);
    if (report_error("DSYRK ", rank_k_error(U, T, *n, *k, *lda, *ldc))) return;
    note_status("DSYRK", zg::syrk_op<double>(U, T != 'N', *n, *k, *alpha, a, *lda, *beta, c, *ldc,
                                             zg::blas_stream(), ZYNUM_GPU_MODE_PLAIN, 0, 0));
}

// fortran.zig:777-784
void ssymm_(const char* side, const char* u, const zynum_blas_int* m, const zynum_blas_int* n, const float* alpha,
            const float* a, const zynum_blas_int* lda, const float* b, const zynum_blas_int* ldb, const float* beta,
            float* c, const zynum_blas_int* ldc) {
    const char S = up(side), U = up(u);
    if (report_error("SSYMM ", symmetric_error(S, U, *m, *n, *lda, *ldb, *ldc))) return;
    note_status("SSYMM", zg::symm_op<float>(S, U, *m, *n, *alpha, a, *lda, b, *ldb, *beta, c, *ldc, zg::blas_stream(), 0));
}

void dsymm_(const char* side, const char* u, const zynum_blas_int* m, const zynum_blas_int* n, const double* alpha,
            const double* a, const zynum_blas_int* lda, const double* b, const zynum_blas_int* ldb, const double* beta,
            double* c, const zynum_blas_int* ldc) {
    const char S = up(side), U = up(u);
    if (report_error("DSYMM ", symmetric_error(S, U, *m, *n, *lda, *ldb, *ldc))) return;
    note_status("DSYMM",
                zg::symm_op<double>(S, U, *m, *n, *alpha, a, *lda, b, *ldb, *beta, c, *ldc, zg::blas_stream(), 0));
}

// fortran.zig:827-834
void ssyr2k_(const char* u, const char* t, const zynum_blas_int* n, const zynum_blas_int* k, const float* alpha,
             const float* a, const zynum_blas_int* lda, const float* b, const zynum_blas_int* ldb, const float* beta,
             float* c, const zynum_blas_int* ldc) {
    const char U = up(u), T = up(t);
    if (report_error("SSYR2K", rank_2k_error(U, T, *n, *k, *lda, *ldb, *ldc))) return;
    note_status("SSYR2K", zg::syr2k_op<float>(U, T != 'N', *n, *k, *alpha, a, *lda, b, *ldb, *beta, c, *ldc,
                                              zg::blas_stream(), 0));
}

void dsyr2k_(const char* u, const char* t, const zynum_blas_int* n, const zynum_blas_int* k, const double* alpha,
             const double* a, const zynum_blas_int* lda, const double* b, const zynum_blas_int* ldb,
             const double* beta, double* c, const zynum_blas_int* ldc) {
    const char U = up(u), T = up(t);
    if (report_error("DSYR2K", rank_2k_error(U, T, *n, *k, *lda, *ldb, *ldc))) return;
    note_status("DSYR2K", zg::syr2k_op<double>(U, T != 'N', *n, *k, *alpha, a, *lda, b, *ldb, *beta, c, *ldc,
                                               zg::blas_stream(), 0));
}

// fortran.zig:852-859
void strmm_(const char* side, const char* u, const char* t, const char* d, const zynum_blas_int* m,
            const zynum_blas_int* n, const float* alpha, const float* a, const zynum_blas_int* lda, float* b,
            const zynum_blas_int* ldb) {
    const char S = up(side), U = up(u), T = up(t), D = up(d);
    if (report_error("STRMM ", triangular_error(S, U, T, D, *m, *n, *lda, *ldb))) return;
    note_status("STRMM", zg::trmm_op<float>(S, U, T != 'N', D, *m, *n, *alpha, a, *lda, b, *ldb, zg::blas_stream(), 0));
}

void dtrmm_(const char* side, const char* u, const char* t, const char* d, const zynum_blas_int* m,
            const zynum_blas_int* n, const double* alpha, const double* a, const zynum_blas_int* lda, double* b,
            const zynum_blas_int* ldb) {
    const char S = up(side), U = up(u), T = up(t), D = up(d);
    if (report_error("DTRMM ", triangular_error(S, U, T, D, *m, *n, *lda, *ldb))) return;
    note_status("DTRMM",
                zg::trmm_op<double>(S, U, T != 'N', D, *m, *n, *alpha, a, *lda, b, *ldb, zg::blas_stream(), 0));
}

void strsm_(const char* side, const char* u, const char* t, const char* d, const zynum_blas_int* m,
            const zynum_blas_int* n, const float* alpha, const float* a, const zynum_blas_int* lda, float* b,
            const zynum_blas_int* ldb) {
    const char S = up(side), U = up(u), T = up(t), D = up(d);
    if (report_error("STRSM ", triangular_error(S, U, T, D, *m, *n, *lda, *ldb))) return;
    note_status("STRSM", zg::trsm_op<float>(S, U, T != 'N', D, *m, *n, *alpha, a, *lda, b, *ldb, zg::blas_stream(),
                                            ZYNUM_GPU_MODE_PLAIN, 0, 0));
}

void dtrsm_(const char* side, const char* u, const char* t, const char* d, const zynum_blas_int* m,
            const zynum_blas_int* n, const double* alpha, const double* a, const zynum_blas_int* lda, double* b,
            const zynum_blas_int* ldb) {
    const char S = up(side), U = up(u), T = up(t), D = up(d);
    if (report_error("DTRSM ", triangular_error(S, U, T, D, *m, *n, *lda, *ldb))) return;
    note_status("DTRSM", zg::trsm_op<double>(S, U, T != 'N', D, *m, *n, *alpha, a, *lda, b, *ldb, zg::blas_stream(),
                                             ZYNUM_GPU_MODE_PLAIN, 0, 0));
}

// ---------------------------------------------------------------------------
// CBLAS ABI (cblas.zig): invalid enums return silently with outputs
// unchanged; numeric checks are the Fortran ones after the row-major swap.
// ---------------------------------------------------------------------------
static inline bool cb_layout(int l) { return l == CblasRowMajor || l == CblasColMajor; }
static inline bool cb_trans(int t) { return t == CblasNoTrans || t == CblasTrans || t == CblasConjTrans; }
static inline char cb_trans_char(int t) {
    return t == CblasNoTrans ? 'N' : t == CblasTrans ? 'T' : t == CblasConjTrans ? 'C' : 'X';
}
static inline char cb_uplo_char(int u) { return u == CblasUpper ? 'U' : u == CblasLower ? 'L' : 'X'; }
static inline char cb_row_uplo_char(int u) { return u == CblasUpper ? 'L' : u == CblasLower ? 'U' : 'X'; }
static inline char cb_side_char(int s) { return s == CblasLeft ? 'L' : s == CblasRight ? 'R' : 'X'; }
static inline char cb_row_side_char(int s) { return s == CblasLeft ? 'R' : s == CblasRight ? 'L' : 'X'; }
static inline char cb_diag_char(int d) { return d == CblasNonUnit ? 'N' : d == CblasUnit ? 'U' : 'X'; }
static inline char cb_row_trans_char(int t) {
    return t == CblasNoTrans ? 'T' : t == CblasTrans ? 'N' : t == CblasConjTrans ? 'C' : 'X';
}

void cblas_sgemm(int layout, int transa, int transb, int m, int n, int k, float alpha, const float* a, int lda,
                 const float* b, int ldb, float beta, float* c, int ldc) {
    if (!cb_layout(layout) || !cb_trans(transa) || !cb_trans(transb)) return;
    const bool row = layout == CblasRowMajor;
    char ta = cb_trans_char(row ? transb : transa), tb = cb_trans_char(row ? transa : transb);
    zynum_blas_int mm = row ? n : m, nn = row ? m : n, kk = k, la = row ? ldb : lda, lb = row ? lda : ldb, lc = ldc;
    sgemm_(&ta, &tb, &mm, &nn, &kk, &alpha, row ? b : a, &la, row ? a : b, &lb, &beta, c, &lc);
}

void cblas_dgemm(int layout, int transa, int transb, int m, int n, int k, double alpha, const double* a, int lda,
                 const double* b, int ldb, double beta, double* c, int ldc) {
    if (!cb_layout(layout) || !cb_trans(transa) || !cb_trans(transb)) return;
    const bool row = layout == CblasRowMajor;
    char ta = cb_trans_char(row ? transb : transa), tb = cb_trans_char(row ? transa : transb);
    zynum_blas_int mm = row ? n : m, nn = row ? m : n, kk = k, la = row ? ldb : lda, lb = row ? lda : ldb, lc = ldc;
    dgemm_(&ta, &tb, &mm, &nn, &kk, &alpha, row ? b : a, &la, row ? a : b, &lb, &beta, c, &lc);
}

// cblas.zig:1441-1465: row-major swaps the operands like the real forms; the
// transpose codes keep their meaning (a conjugate transpose stays one).
void cblas_cgemm(int layout, int transa, int transb, int m, int n, int k, const void* alpha, const void* a, int lda,
                 const void* b, int ldb, const void* beta, void* c, int ldc) {
    if (!cb_layout(layout) || !cb_trans(transa) || !cb_trans(transb)) return;
    const bool row = layout == CblasRowMajor;
    char ta = cb_trans_char(row ? transb : transa), tb = cb_trans_char(row ? transa : transb);
    zynum_blas_int mm = row ? n : m, nn = row ? m : n, kk = k, la = row ? ldb : lda, lb = row ? lda : ldb, lc = ldc;
    typedef zynum_blas_complex_float cf;
    cgemm_(&ta, &tb, &mm, &nn, &kk, static_cast<const cf*>(alpha), static_cast<const cf*>(row ? b : a), &la,
           static_cast<const cf*>(row ? a : b), &lb, static_cast<const cf*>(beta), static_cast<cf*>(c), &lc);
}

void cblas_zgemm(int layout, int transa, int transb, int m, int n, int k, const void* alpha, const void* a, int lda,
                 const void* b, int ldb, const void* beta, void* c, int ldc) {
    if (!cb_layout(layout) || !cb_trans(transa) || !cb_trans(transb)) return;
    const bool row = layout == CblasRowMajor;
    char ta = cb_trans_char(row ? transb : transa), tb = cb_trans_char(row ? transa : transb);
    zynum_blas_int mm = row ? n : m, nn = row ? m : n, kk = k, la = row ? ldb : lda, lb = row ? lda : ldb, lc = ldc;
    typedef zynum_blas_complex_double cd;
    zgemm_(&ta, &tb, &mm, &nn, &kk, static_cast<const cd*>(alpha), static_cast<const cd*>(row ? b : a), &la,
           static_cast<const cd*>(row ? a : b), &lb, static_cast<const cd*>(beta), static_cast<cd*>(c), &lc);
}

// cblas.zig:1467-1489: only the layout is checked here; side and uplo map to
// 'X' when invalid and the Fortran check reports them.
void cblas_ssymm(int layout, int side, int uplo, int m, int n, float alpha, const float* a, int lda, const float* b,
                 int ldb, float beta, float* c, int ldc) {
    if (!cb_layout(layout)) return;
    const bool row = layout == CblasRowMajor;
    char ss = row ? cb_row_side_char(side) : cb_side_char(side);
    char uu = row ? cb_row_uplo_char(uplo) : cb_uplo_char(uplo);
    zynum_blas_int mm = row ? n : m, nn = row ? m : n, la = lda, lb = ldb, lc = ldc;
    ssymm_(&ss, &uu, &mm, &nn, &alpha, a, &la, b, &lb, &beta, c, &lc);
}

void cblas_dsymm(int layout, int side, int uplo, int m, int n, double alpha, const double* a, int lda,
                 const double* b, int ldb, double beta, double* c, int ldc) {
    if (!cb_layout(layout)) return;
    const bool row = layout == CblasRowMajor;
    char ss = row ? cb_row_side_char(side) : cb_side_char(side);
    char uu = row ? cb_row_uplo_char(uplo) : cb_uplo_char(uplo);
    zynum_blas_int mm = row ? n : m, nn = row ? m : n, la = lda, lb = ldb, lc = ldc;
    dsymm_(&ss, &uu, &mm, &nn, &alpha, a, &la, b, &lb, &beta, c, &lc);
}

// cblas.zig:1605-1627
void cblas_ssyr2k(int layout, int uplo, int trans, int n, int k, float alpha, const float* a, int lda, const float* b,
                  int ldb, float beta, float* c, int ldc) {
    if (!cb_layout(layout)) return;
    const bool row = layout == CblasRowMajor;
    char uu = row ? cb_row_uplo_char(uplo) : cb_uplo_char(uplo);
    char tt = row ? cb_row_trans_char(trans) : cb_trans_char(trans);
    zynum_blas_int nn = n, kk = k, la = lda, lb = ldb, lc = ldc;
    ssyr2k_(&uu, &tt, &nn, &kk, &alpha, a, &la, b, &lb, &beta, c, &lc);
}

void cblas_dsyr2k(int layout, int uplo, int trans, int n, int k, double alpha, const double* a, int lda,
                  const double* b, int ldb, double beta, double* c, int ldc) {
    if (!cb_layout(layout)) return;
    const bool row = layout == CblasRowMajor;
    char uu = row ? cb_row_uplo_char(uplo) : cb_uplo_char(uplo);
    char tt = row ? cb_row_trans_char(trans) : cb_trans_char(trans);
    zynum_blas_int nn = n, kk = k, la = lda, lb = ldb, lc = ldc;
    dsyr2k_(&uu, &tt, &nn, &kk, &alpha, a, &la, b, &lb, &beta, c, &lc);
}

// cblas.zig:1677-1700: every enum is checked (invalid returns silently)
void cblas_strmm(int layout, int side, int uplo, int transa, int diag, int m, int n, float alpha, const float* a,
                 int lda, float* b, int ldb) {
    if (!cb_layout(layout) || !(side == CblasLeft || side == CblasRight) ||
        !(uplo == CblasUpper || uplo == CblasLower) || !cb_trans(transa) ||
        !(diag == CblasNonUnit || diag == CblasUnit))
        return;
    const bool row = layout == CblasRowMajor;
    char ss = row ? cb_row_side_char(side) : cb_side_char(side);
    char uu = row ? cb_row_uplo_char(uplo) : cb_uplo_char(uplo);
    char tt = cb_trans_char(transa), dd = cb_diag_char(diag);
    zynum_blas_int mm = row ? n : m, nn = row ? m : n, la = lda, lb = ldb;
    strmm_(&ss, &uu, &tt, &dd, &mm, &nn, &alpha, a, &la, b, &lb);
}

void cblas_dtrmm(int layout, int side, int uplo, int transa, int diag, int m, int n, double alpha, const double* a,
                 int lda, double* b, int ldb) {
    if (!cb_layout(layout) || !(side == CblasLeft || side == CblasRight) ||
        !(uplo == CblasUpper || uplo == CblasLower) || !cb_trans(transa) ||
        !(diag == CblasNonUnit || diag == CblasUnit))
        return;
    const bool row = layout == CblasRowMajor;
    char ss = row ? cb_row_side_char(side) : cb_side_char(side);
    char uu = row ? cb_row_uplo_char(uplo) : cb_uplo_char(uplo);
    char tt = cb_trans_char(transa), dd = cb_diag_char(diag);
    zynum_blas_int mm = row ? n : m, nn = row ? m : n, la = lda, lb = ldb;
    dtrmm_(&ss, &uu, &tt, &dd, &mm, &nn, &alpha, a, &la, b, &lb);
}

void cblas_dsyrk(int layout, int uplo, int trans, int n, int k, double alpha, const double* a, int lda, double beta,
                 double* c, int ldc) {
    if (!cb_layout(layout)) return;
    const bool row = layout == CblasRowMajor;
    char uu = row ? cb_row_uplo_char(uplo) : cb_uplo_char(uplo);
    char tt = row ? cb_row_trans_char(trans) : cb_trans_char(trans);
    zynum_blas_int nn = n, kk = k, la = lda, lc = ldc;
    dsyrk_(&uu, &tt, &nn, &kk, &alpha, a, &la, &beta, c, &lc);
}

void cblas_dtrsm(int layout, int side, int uplo, int transa, int diag, int m, int n, double alpha, const double* a,
                 int lda, double* b, int ldb) {
    if (!cb_layout(layout) || !(side == CblasLeft || side == CblasRight) ||
        !(uplo == CblasUpper || uplo == CblasLower) || !cb_trans(transa) ||
        !(diag == CblasNonUnit || diag == CblasUnit))
        return;
    const bool row = layout == CblasRowMajor;
    char ss = row ? cb_row_side_char(side) : cb_side_char(side);
    char uu = row ? cb_row_uplo_char(uplo) : cb_uplo_char(uplo);
    char tt = cb_trans_char(transa), dd = cb_diag_char(diag);
    zynum_blas_int mm = row ? n : m, nn = row ? m : n, la = lda, lb = ldb;
    dtrsm_(&ss, &uu, &tt, &dd, &mm, &nn, &alpha, a, &la, b, &lb);
}

// ---------------------------------------------------------------------------
// GPU-tier private ABI
// ---------------------------------------------------------------------------
int zynum_gpu_gemm(const zynum_gpu_gemm_request* r) {
    if (r == nullptr) return ZYNUM_GPU_DECLINE_ARGUMENT;
    const bool cplx = r->scalar == ZYNUM_GPU_C32 || r->scalar == ZYNUM_GPU_C64;
    auto code = [cplx](uint8_t t) { return zg::request_trans_char(t, cplx); };
    const char ta = code(r->transa), tb = code(r->transb);
    if (gemm_error(ta, tb, r->m, r->n, r->k, r->lda, r->ldb, r->ldc) != 0) return ZYNUM_GPU_DECLINE_ARGUMENT;
    const int tiles = r->mode == ZYNUM_GPU_MODE_PLAIN       ? zg::kTileConfigs
                      : r->mode == ZYNUM_GPU_MODE_OZAKI_BF16 ? zg::kBf16Tiles
                                                             : zg::kI8Tiles;
    if (r->tile > tiles) return ZYNUM_GPU_DECLINE_UNSUPPORTED;
    hipStream_t s = r->stream ? reinterpret_cast<hipStream_t>(r->stream) : zg::blas_stream();
    if (cplx) {
        if (r->mode != ZYNUM_GPU_MODE_PLAIN) return ZYNUM_GPU_DECLINE_UNSUPPORTED;
        if (r->scalar == ZYNUM_GPU_C64) {
            const double alpha[2] = {scalar_from_bits<double>(r->alpha_re), scalar_from_bits<double>(r->alpha_im)};
            const double beta[2] = {scalar_from_bits<double>(r->beta_re), scalar_from_bits<double>(r->beta_im)};
            return zg::cgemm_op<double>(ta, tb, r->m, r->n, r->k, alpha, static_cast<const double*>(r->a), r->lda,
                                        static_cast<const double*>(r->b), r->ldb, beta, static_cast<double*>(r->c),
                                        r->ldc, s, r->flags, (int)r->tile - 1);
        }
        const float alpha[2] = {scalar_from_bits<float>(r->alpha_re), scalar_from_bits<float>(r->alpha_im)};
        const float beta[2] = {scalar_from_bits<float>(r->beta_re), scalar_from_bits<float>(r->beta_im)};
        return zg::cgemm_op<float>(ta, tb, r->m, r->n, r->k, alpha, static_cast<const float*>(r->a), r->lda,
                                   static_cast<const float*>(r->b), r->ldb, beta, static_cast<float*>(r->c), r->ldc,
                                   s, r->flags, (int)r->tile - 1);
    }
    if (r->alpha_im != 0 || r->beta_im != 0) return ZYNUM_GPU_DECLINE_UNSUPPORTED;
    if (r->scalar == ZYNUM_GPU_F64) {
        return zg::gemm_op<double>(ta == 'T', tb == 'T', r->m, r->n, r->k, scalar_from_bits<double>(r->alpha_re),
                                   static_cast<const double*>(r->a), r->lda, static_cast<const double*>(r->b), r->ldb,
                                   scalar_from_bits<double>(r->beta_re), static_cast<double*>(r->c), r->ldc, s,
                                   r->mode, r->mode_param, r->flags, (int)r->tile - 1);
    }
    if (r->scalar == ZYNUM_GPU_F32) {
        return zg::gemm_op<float>(ta == 'T', tb == 'T', r->m, r->n, r->k, scalar_from_bits<float>(r->alpha_re),
                                  static_cast<const float*>(r->a), r->lda, static_cast<const float*>(r->b), r->ldb,
                                  scalar_from_bits<float>(r->beta_re), static_cast<float*>(r->c), r->ldc, s, r->mode,
                                  r->mode_param, r->flags, (int)r->tile - 1);
    }
    return ZYNUM_GPU_DECLINE_UNSUPPORTED;
}

int zynum_gpu_dtrsm(char side, char uplo, char transa, char diag, int m, int n, double alpha, const double* a,
                    int lda, double* b, int ldb, void* stream, int mode, int mode_param, int flags) {
    const char S = (char)toupper(side), U = (char)toupper(uplo), T = (char)toupper(transa), D = (char)toupper(diag);
    if (triangular_error(S, U, T, D, m, n, lda, ldb) != 0) return ZYNUM_GPU_DECLINE_ARGUMENT;
    hipStream_t s = stream ? reinterpret_cast<hipStream_t>(stream) : zg::blas_stream();
    return zg::trsm_op<double>(S, U, T != 'N', D, m, n, alpha, a, lda, b, ldb, s, mode, mode_param, flags);
}

int zynum_gpu_dsyrk(char uplo, char trans, int n, int k, double alpha, const double* a, int lda, double beta,
                    double* c, int ldc, void* stream, int mode, int mode_param, int flags) {
    const char U = (char)toupper(uplo), T = (char)toupper(trans);
    if (rank_k_error(U, T, n, k, lda, ldc) != 0) return ZYNUM_GPU_DECLINE_ARGUMENT;
    hipStream_t s = stream ? reinterpret_cast<hipStream_t>(stream) : zg::blas_stream();
    return zg::syrk_op<double>(U, T != 'N', n, k, alpha, a, lda, beta, c, ldc, s, mode, mode_param, flags);
}

int zynum_gpu_dsymm(char side, char uplo, int m, int n, double alpha, const double* a, int lda, const double* b,
                    int ldb, double beta, double* c, int ldc, void* stream, int flags) {
    const char S = (char)toupper(side), U = (char)toupper(uplo);
    if (symmetric_error(S, U, m, n, lda, ldb, ldc) != 0) return ZYNUM_GPU_DECLINE_ARGUMENT;
    hipStream_t s = stream ? reinterpret_cast<hipStream_t>(stream) : zg::blas_stream();
    return zg::symm_op<double>(S, U, m, n, alpha, a, lda, b, ldb, beta, c, ldc, s, flags);
}

int zynum_gpu_dtrmm(char side, char uplo, char transa, char diag, int m, int n, double alpha, const double* a,
                    int lda, double* b, int ldb, void* stream, int flags) {
    const char S = (char)toupper(side), U = (char)toupper(uplo), T = (char)toupper(transa), D = (char)toupper(diag);
    if (triangular_error(S, U, T, D, m, n, lda, ldb) != 0) return ZYNUM_GPU_DECLINE_ARGUMENT;
    hipStream_t s = stream ? reinterpret_cast<hipStream_t>(stream) : zg::blas_stream();
    return zg::trmm_op<double>(S, U, T != 'N', D, m, n, alpha, a, lda, b, ldb, s, flags);
}

int zynum_gpu_dsyr2k(char uplo, char trans, int n, int k, double alpha, const double* a, int lda, const double* b,
                     int ldb, double beta, double* c, int ldc, void* stream, int flags) {
    const char U = (char)toupper(uplo), T = (char)toupper(trans);
    if (rank_2k_error(U, T, n, k, lda, ldb, ldc) != 0) return ZYNUM_GPU_DECLINE_ARGUMENT;
    hipStream_t s = stream ? reinterpret_cast<hipStream_t>(stream) : zg::blas_stream();
    return zg::syr2k_op<double>(U, T != 'N', n, k, alpha, a, lda, b, ldb, beta, c, ldc, s, flags);
}

int zynum_gpu_last_status(void) { return t_last_status; }

const char* zynum_gpu_status_string(int status) {
    switch (status) {
        case ZYNUM_GPU_OK: return "ok";
        case ZYNUM_GPU_DECLINE_NO_DEVICE: return "no gfx950 device";
        case ZYNUM_GPU_DECLINE_ALLOC: return "device allocation failed";
        case ZYNUM_GPU_DECLINE_UNSUPPORTED: return "unsupported form";
        case ZYNUM_GPU_DECLINE_ARGUMENT: return "invalid argument";
        case ZYNUM_GPU_DECLINE_FORCED: return "forced decline";
        case ZYNUM_GPU_ERR_LAUNCH: return "HIP launch error";
        case ZYNUM_GPU_DECLINE_SCRATCH: return "needs more scratch than the fixed no-allocation buffer holds";
        default: return "unknown status";
    }
}

const char* zynum_gpu_version(void) { return "zynum-gpu 0.1 gfx950"; }

}  // extern "C"
