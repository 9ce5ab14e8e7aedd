/*
 * TEST INFRASTRUCTURE ONLY -- the oracle.
 *
 * A plain-C restatement of Zynum's portable CPU Level-3 path (GEMM, left/right
 * TRSM, SYRK, SYMM, SYR2K, TRMM, complex GEMM) plus double-double GEMMs used
 * to state errors.  Only the test
 * suite, the smoke check and the benchmark's `cpu_baseline` leg load this
 * library, and only as the checker or as the timed CPU baseline.  Nothing in
 * the zynum_amd product links or calls it.
 *
 * The arithmetic order follows the reference (paths relative to the Zynum
 * source tree):
 *   - routing:   src/blas/core/matrix_matrix/gemm.zig:168-247,1392-1430 and
 *                src/blas/core/matrix_matrix/planner.zig:327-383
 *   - per-element FMA chain over p ascending:
 *                src/blas/kernels/shared/matrix_matrix/packed_simd.zig:140-199
 *   - K-major B panel (k x 8) and A-transpose pack:
 *                src/blas/kernels/shared/matrix_matrix/packing.zig:45-80,131-139
 *   - alpha/beta epilogue:
 *                src/blas/kernels/shared/matrix_matrix/epilogue.zig:10-39
 *   - non-fused portable loop: gemm.zig:1422-1429
 *   - TRSM: core/matrix_matrix/triangular.zig:52-63,72-84,103-118,233-249 and
 *           core/matrix_vector/triangular.zig:44-57,362-412
 *   - SYRK: core/matrix_matrix/symmetric.zig:180-200,237-279
 *   - SYMM, SYR2K: symmetric.zig:56-160,301-412; TRMM: triangular.zig:202-231
 *           with matrix_vector/triangular.zig:44-133 (see below)
 *   - complex GEMM (CGEMM / ZGEMM): gemm.zig:1234-1246,1404-1418 with the
 *           complex scalar rules of core/shared/scalar.zig (see below)
 *
 * Pinning (see tests/test_oracle.py): known-answer vectors from the
 * reference's own tests, bitwise equality against exact integer arithmetic on
 * integer-valued inputs (transposes, ld, tails, epilogue), exact-rational
 * tolerance checks, and double-double against exact fractions.  Shapes the
 * reference routes to GEMV edges (gemm.zig:181-224, m==1 or n==1 special
 * cases) are restated with the same FMA chain; their exact rounding is
 * "parity unpinned" (the reference's GEMV reduction order is not observable
 * without running it).
 */
#define _GNU_SOURCE
#include <math.h>
#include <pthread.h>
#include <sched.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include <immintrin.h>

/* ------------------------------------------------------------------------ */
/* splitmix64 synthetic input generator (counter form).                       */
/* value_i = (double)(z_i >> 11) * 2^-52 - 1, z_i = mix(seed + (i+1)*golden)  */
/* ------------------------------------------------------------------------ */
static inline uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

void oracle_fill_uniform(uint64_t seed, uint64_t offset, double *out, int64_t count) {
    for (int64_t i = 0; i < count; ++i) {
        uint64_t z = mix64(seed + (offset + (uint64_t)i + 1) * 0x9e3779b97f4a7c15ULL);
        out[i] = (double)(z >> 11) * 0x1p-52 - 1.0;
    }
}

/* ------------------------------------------------------------------------ */
/* helpers                                                                   */
/* ------------------------------------------------------------------------ */
static inline int is_no_trans(char t) { return t == 'N' || t == 'n'; }
static inline size_t mi(int ld, size_t r, size_t c) { return r + c * (size_t)ld; }

/* epilogue.zig:10-39: out = acc if alpha==1 && beta==0, else
   fma(alpha, acc, s) with s = 0 (beta==0), old (beta==1), RN(beta*old). */
static inline double epi_d(double alpha, double beta, double acc, const double *cold) {
    if (alpha == 1.0 && beta == 0.0) return acc;
    double s = (beta == 0.0) ? 0.0 : (beta == 1.0 ? *cold : beta * *cold);
    return fma(alpha, acc, s);
}
static inline float epi_s(float alpha, float beta, float acc, const float *cold) {
    if (alpha == 1.0f && beta == 0.0f) return acc;
    float s = (beta == 0.0f) ? 0.0f : (beta == 1.0f ? *cold : beta * *cold);
    return fmaf(alpha, acc, s);
}

/* ------------------------------------------------------------------------ */
/* threading: column-panel tasks over a small pthread pool                   */
/* planner.zig:50-69 (aligned column split), :296-325 (run tasks)            */
/* ------------------------------------------------------------------------ */
static int g_threads = 0; /* 0 -> min(affinity cores, 64) */

int oracle_set_threads(int t) { g_threads = t; return t; }

int oracle_threads(void) {
    if (g_threads > 0) return g_threads;
    cpu_set_t set;
    CPU_ZERO(&set);
    int n = 1;
    if (sched_getaffinity(0, sizeof(set), &set) == 0) n = CPU_COUNT(&set);
    if (n < 1) n = 1;
    if (n > 64) n = 64; /* thread_pool.zig:8 max_tasks */
    return n;
}

typedef struct {
    void (*fn)(void *, size_t, size_t);
    void *ctx;
    size_t j0, j1;
} panel_task;

static void *panel_thread(void *p) {
    panel_task *t = (panel_task *)p;
    t->fn(t->ctx, t->j0, t->j1);
    return NULL;
}

static void run_column_tasks(void (*fn)(void *, size_t, size_t), void *ctx, size_t n, size_t align) {
    int threads = oracle_threads();
    size_t panels = (n + align - 1) / align;
    size_t tasks = (size_t)threads < panels ? (size_t)threads : panels;
    if (tasks <= 1) { fn(ctx, 0, n); return; }
    panel_task tk[64];
    pthread_t th[64];
    size_t base = panels / tasks, extra = panels % tasks, p0 = 0;
    for (size_t t = 0; t < tasks; ++t) {
        size_t cnt = base + (t < extra);
        tk[t].fn = fn; tk[t].ctx = ctx;
        tk[t].j0 = p0 * align;
        tk[t].j1 = (p0 + cnt) * align < n ? (p0 + cnt) * align : n;
        p0 += cnt;
    }
    for (size_t t = 1; t < tasks; ++t) pthread_create(&th[t], NULL, panel_thread, &tk[t]);
    panel_thread(&tk[0]); /* task 0 on the caller, as thread_pool.zig:227-245 */
    for (size_t t = 1; t < tasks; ++t) pthread_join(th[t], NULL);
}

/* ------------------------------------------------------------------------ */
/* packed path: per element acc = fma chain over p ascending                 */
/* ------------------------------------------------------------------------ */
typedef struct {
    size_t m, k;
    double alpha, beta;
    const double *a; int lda;       /* op(A) as m x k column-major (packed if TN/TT) */
    const double *b; int ldb; int b_trans;
    double *c; int ldc;
} dtask;

typedef struct {
    size_t m, k;
    float alpha, beta;
    const float *a; int lda;
    const float *b; int ldb; int b_trans;
    float *c; int ldc;
} stask;

/* packing.zig:131-139: b_pack[p*8 + c] = op(B)(p, j+c) */
static void pack_b_d(const dtask *t, size_t j, size_t w, double *bp) {
    for (size_t p = 0; p < t->k; ++p)
        for (size_t c = 0; c < w; ++c)
            bp[p * 8 + c] = t->b_trans ? t->b[mi(t->ldb, j + c, p)] : t->b[mi(t->ldb, p, j + c)];
}
static void pack_b_s(const stask *t, size_t j, size_t w, float *bp) {
    for (size_t p = 0; p < t->k; ++p)
        for (size_t c = 0; c < w; ++c)
            bp[p * 8 + c] = t->b_trans ? t->b[mi(t->ldb, j + c, p)] : t->b[mi(t->ldb, p, j + c)];
}

/* scalar rows of one panel (tails and non-SIMD hosts) */
static void rows_scalar_d(const dtask *t, const double *bp, size_t i0, size_t i1, size_t j, size_t w) {
    for (size_t i = i0; i < i1; ++i) {
        double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        for (size_t p = 0; p < t->k; ++p) {
            double av = t->a[mi(t->lda, i, p)];
            for (size_t c = 0; c < w; ++c) acc[c] = fma(av, bp[p * 8 + c], acc[c]);
        }
        for (size_t c = 0; c < w; ++c) {
            double *cp = &t->c[mi(t->ldc, i, j + c)];
            *cp = epi_d(t->alpha, t->beta, acc[c], cp);
        }
    }
}
static void rows_scalar_s(const stask *t, const float *bp, size_t i0, size_t i1, size_t j, size_t w) {
    for (size_t i = i0; i < i1; ++i) {
        float acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        for (size_t p = 0; p < t->k; ++p) {
            float av = t->a[mi(t->lda, i, p)];
            for (size_t c = 0; c < w; ++c) acc[c] = fmaf(av, bp[p * 8 + c], acc[c]);
        }
        for (size_t c = 0; c < w; ++c) {
            float *cp = &t->c[mi(t->ldc, i, j + c)];
            *cp = epi_s(t->alpha, t->beta, acc[c], cp);
        }
    }
}

/* 16 x 8 register tile on AVX-512 (packed_params.zig:45-67: 8 lanes x 2 row
   groups x tile_n 8), same per-element FMA order as the scalar rows. */
__attribute__((target("avx512f,fma")))
static size_t rows_avx512_d(const dtask *t, const double *bp, size_t j) {
    size_t i = 0;
    for (; i + 16 <= t->m; i += 16) {
        __m512d acc0[8], acc1[8];
        for (int c = 0; c < 8; ++c) { acc0[c] = _mm512_setzero_pd(); acc1[c] = _mm512_setzero_pd(); }
        for (size_t p = 0; p < t->k; ++p) {
            const double *ap = &t->a[mi(t->lda, i, p)];
            __m512d a0 = _mm512_loadu_pd(ap), a1 = _mm512_loadu_pd(ap + 8);
            for (int c = 0; c < 8; ++c) {
                __m512d bv = _mm512_set1_pd(bp[p * 8 + c]);
                acc0[c] = _mm512_fmadd_pd(a0, bv, acc0[c]);
                acc1[c] = _mm512_fmadd_pd(a1, bv, acc1[c]);
            }
        }
        for (int c = 0; c < 8; ++c) {
            double tmp[16];
            _mm512_storeu_pd(tmp, acc0[c]);
            _mm512_storeu_pd(tmp + 8, acc1[c]);
            for (int r = 0; r < 16; ++r) {
                double *cp = &t->c[mi(t->ldc, i + r, j + c)];
                *cp = epi_d(t->alpha, t->beta, tmp[r], cp);
            }
        }
    }
    return i;
}
__attribute__((target("avx512f,fma")))
static size_t rows_avx512_s(const stask *t, const float *bp, size_t j) {
    size_t i = 0;
    for (; i + 32 <= t->m; i += 32) {
        __m512 acc0[8], acc1[8];
        for (int c = 0; c < 8; ++c) { acc0[c] = _mm512_setzero_ps(); acc1[c] = _mm512_setzero_ps(); }
        for (size_t p = 0; p < t->k; ++p) {
            const float *ap = &t->a[mi(t->lda, i, p)];
            __m512 a0 = _mm512_loadu_ps(ap), a1 = _mm512_loadu_ps(ap + 16);
            for (int c = 0; c < 8; ++c) {
                __m512 bv = _mm512_set1_ps(bp[p * 8 + c]);
                acc0[c] = _mm512_fmadd_ps(a0, bv, acc0[c]);
                acc1[c] = _mm512_fmadd_ps(a1, bv, acc1[c]);
            }
        }
        for (int c = 0; c < 8; ++c) {
            float tmp[32];
            _mm512_storeu_ps(tmp, acc0[c]);
            _mm512_storeu_ps(tmp + 16, acc1[c]);
            for (int r = 0; r < 32; ++r) {
                float *cp = &t->c[mi(t->ldc, i + r, j + c)];
                *cp = epi_s(t->alpha, t->beta, tmp[r], cp);
            }
        }
    }
    return i;
}
/* 4 x 8 tile on AVX2 (packed_params.zig:58-66) */
__attribute__((target("avx2,fma")))
static size_t rows_avx2_d(const dtask *t, const double *bp, size_t j) {
    size_t i = 0;
    for (; i + 4 <= t->m; i += 4) {
        __m256d acc[8];
        for (int c = 0; c < 8; ++c) acc[c] = _mm256_setzero_pd();
        for (size_t p = 0; p < t->k; ++p) {
            __m256d a0 = _mm256_loadu_pd(&t->a[mi(t->lda, i, p)]);
            for (int c = 0; c < 8; ++c) acc[c] = _mm256_fmadd_pd(a0, _mm256_set1_pd(bp[p * 8 + c]), acc[c]);
        }
        for (int c = 0; c < 8; ++c) {
            double tmp[4];
            _mm256_storeu_pd(tmp, acc[c]);
            for (int r = 0; r < 4; ++r) {
                double *cp = &t->c[mi(t->ldc, i + r, j + c)];
                *cp = epi_d(t->alpha, t->beta, tmp[r], cp);
            }
        }
    }
    return i;
}
__attribute__((target("avx2,fma")))
static size_t rows_avx2_s(const stask *t, const float *bp, size_t j) {
    size_t i = 0;
    for (; i + 8 <= t->m; i += 8) {
        __m256 acc[8];
        for (int c = 0; c < 8; ++c) acc[c] = _mm256_setzero_ps();
        for (size_t p = 0; p < t->k; ++p) {
            __m256 a0 = _mm256_loadu_ps(&t->a[mi(t->lda, i, p)]);
            for (int c = 0; c < 8; ++c) acc[c] = _mm256_fmadd_ps(a0, _mm256_set1_ps(bp[p * 8 + c]), acc[c]);
        }
        for (int c = 0; c < 8; ++c) {
            float tmp[8];
            _mm256_storeu_ps(tmp, acc[c]);
            for (int r = 0; r < 8; ++r) {
                float *cp = &t->c[mi(t->ldc, i + r, j + c)];
                *cp = epi_s(t->alpha, t->beta, tmp[r], cp);
            }
        }
    }
    return i;
}

static int g_isa = -1; /* 2 avx512, 1 avx2, 0 scalar */
int oracle_isa(void) {
    if (g_isa < 0) {
        __builtin_cpu_init();
        g_isa = __builtin_cpu_supports("avx512f") ? 2 : (__builtin_cpu_supports("avx2") && __builtin_cpu_supports("fma")) ? 1 : 0;
    }
    return g_isa;
}
int oracle_force_isa(int isa) { g_isa = isa; return isa; }

/* packed_simd.zig:277-290: for each 8-column panel, pack B, run row tiles */
static void packed_panels_d(void *ctx, size_t j0, size_t j1) {
    const dtask *t = (const dtask *)ctx;
    double *bp = (double *)malloc(sizeof(double) * (t->k * 8 + 8));
    int isa = oracle_isa();
    for (size_t j = j0; j < j1; j += 8) {
        size_t w = (j1 - j) < 8 ? (j1 - j) : 8;
        pack_b_d(t, j, w, bp);
        size_t i = 0;
        if (w == 8 && isa == 2) i = rows_avx512_d(t, bp, j);
        else if (w == 8 && isa == 1) i = rows_avx2_d(t, bp, j);
        rows_scalar_d(t, bp, i, t->m, j, w);
    }
    free(bp);
}
static void packed_panels_s(void *ctx, size_t j0, size_t j1) {
    const stask *t = (const stask *)ctx;
    float *bp = (float *)malloc(sizeof(float) * (t->k * 8 + 8));
    int isa = oracle_isa();
    for (size_t j = j0; j < j1; j += 8) {
        size_t w = (j1 - j) < 8 ? (j1 - j) : 8;
        pack_b_s(t, j, w, bp);
        size_t i = 0;
        if (w == 8 && isa == 2) i = rows_avx512_s(t, bp, j);
        else if (w == 8 && isa == 1) i = rows_avx2_s(t, bp, j);
        rows_scalar_s(t, bp, i, t->m, j, w);
    }
    free(bp);
}

/* ------------------------------------------------------------------------ */
/* routing: gemm.zig:1392-1430                                               */
/* ------------------------------------------------------------------------ */
#define DEFINE_GEMM(NAME, T, TASK, PANELS, FMAF)                                               \
    void NAME(char ta, char tb, int m, int n, int k, T alpha, const T *a, int lda,             \
              const T *b, int ldb, T beta, T *c, int ldc) {                                    \
        if (m <= 0 || n <= 0) return;                                                          \
        int nota = is_no_trans(ta), notb = is_no_trans(tb);                                   \
        int packed = 0;                                                                        \
        if (k > 0 && alpha != (T)0) {                                                          \
            if (nota) packed = 1; /* NN and NT take the packed planner */                      \
            else if (!(m == 1 || n == 1) && !(m <= 33 && n <= 33 && k <= 33)) packed = 2;      \
        }                                                                                      \
        if (nota && !packed) { /* gemm.zig:170-180 / planner.zig:333-339 */                   \
            for (int j = 0; j < n; ++j)                                                        \
                for (int i = 0; i < m; ++i) {                                                  \
                    T *cp = &c[mi(ldc, i, j)];                                                 \
                    *cp = (beta == (T)0) ? (T)0 : beta * *cp;                                  \
                }                                                                              \
            return;                                                                            \
        }                                                                                      \
        if (packed) {                                                                          \
            TASK t;                                                                            \
            T *apack = NULL;                                                                   \
            t.m = (size_t)m; t.k = (size_t)k; t.alpha = alpha; t.beta = beta;                  \
            t.b = b; t.ldb = ldb; t.b_trans = !notb; t.c = c; t.ldc = ldc;                     \
            if (nota) { t.a = a; t.lda = lda; }                                                \
            else { /* planner.zig:376-381: a_pack[p*m + i] = A[p + i*lda] */                   \
                apack = (T *)malloc(sizeof(T) * (size_t)m * (size_t)k);                        \
                for (size_t p = 0; p < (size_t)k; ++p)                                         \
                    for (size_t i = 0; i < (size_t)m; ++i) apack[p * m + i] = a[mi(lda, p, i)]; \
                t.a = apack; t.lda = m;                                                        \
            }                                                                                  \
            run_column_tasks(PANELS, &t, (size_t)n, 8);                                        \
            free(apack);                                                                       \
            return;                                                                            \
        }                                                                                      \
        /* portable non-fused loop, gemm.zig:1412-1429 */                                      \
        for (int j = 0; j < n; ++j)                                                            \
            for (int i = 0; i < m; ++i) {                                                      \
                T *cp = &c[mi(ldc, i, j)];                                                     \
                if (k <= 0 || alpha == (T)0) { *cp = (beta == (T)0) ? (T)0 : beta * *cp; continue; } \
                T sum = 0;                                                                     \
                for (int p = 0; p < k; ++p) {                                                  \
                    T av = nota ? a[mi(lda, i, p)] : a[mi(lda, p, i)];                         \
                    T bv = notb ? b[mi(ldb, p, j)] : b[mi(ldb, j, p)];                         \
                    T prod = av * bv;                                                          \
                    sum = sum + prod;                                                          \
                }                                                                              \
                T as = alpha * sum;                                                            \
                T bs = (beta == (T)0) ? (T)0 : beta * *cp;                                     \
                *cp = as + bs;                                                                 \
            }                                                                                  \
    }

DEFINE_GEMM(oracle_dgemm, double, dtask, packed_panels_d, fma)
DEFINE_GEMM(oracle_sgemm, float, stask, packed_panels_s, fmaf)

/* ------------------------------------------------------------------------ */
/* TRSM: triangular.zig:233-249                                              */
/* ------------------------------------------------------------------------ */
static inline double tri_val(char uplo, char diag, int trans, const double *a, int lda, size_t r, size_t cidx) {
    /* value of op(A)(r, c) honoring uplo and unit diagonal */
    size_t rr = trans ? cidx : r, cc = trans ? r : cidx;
    int upper = (uplo == 'U' || uplo == 'u');
    if (rr == cc) return (diag == 'U' || diag == 'u') ? 1.0 : a[mi(lda, rr, cc)];
    if (upper ? (rr > cc) : (rr < cc)) return 0.0;
    return a[mi(lda, rr, cc)];
}

/* matrix_vector/triangular.zig:362-412 (trsvUnit), unit stride */
static void trsv_unit(char uplo, int trans, char diag, size_t n, const double *a, int lda, double *x) {
    int upper = (uplo == 'U' || uplo == 'u');
    int nonunit = !(diag == 'U' || diag == 'u');
    if (!trans) {
        if (upper) {
            for (size_t j = n; j-- > 0;) {
                const double *col = a + j * (size_t)lda;
                double xj = x[j];
                if (nonunit) xj = xj / col[j];
                x[j] = xj;
                if (xj != 0.0) for (size_t i = 0; i < j; ++i) x[i] = fma(col[i], -xj, x[i]);
            }
        } else {
            for (size_t j = 0; j < n; ++j) {
                const double *col = a + j * (size_t)lda;
                double xj = x[j];
                if (nonunit) xj = xj / col[j];
                x[j] = xj;
                if (xj != 0.0) for (size_t i = j + 1; i < n; ++i) x[i] = fma(col[i], -xj, x[i]);
            }
        }
        return;
    }
    if (upper) {
        for (size_t j = 0; j < n; ++j) {
            const double *col = a + j * (size_t)lda;
            double dot = 0.0;
            for (size_t i = 0; i < j; ++i) dot = fma(col[i], x[i], dot);
            double xj = x[j] - dot;
            if (nonunit) xj = xj / col[j];
            x[j] = xj;
        }
    } else {
        for (size_t j = n; j-- > 0;) {
            const double *col = a + j * (size_t)lda;
            double dot = 0.0;
            for (size_t i = j + 1; i < n; ++i) dot = fma(col[i], x[i], dot);
            double xj = x[j] - dot;
            if (nonunit) xj = xj / col[j];
            x[j] = xj;
        }
    }
}

typedef struct { char uplo, diag; int trans; size_t m; double alpha; const double *a; int lda; double *b; int ldb; } trsm_left_task;

static void trsm_left_cols(void *ctx, size_t j0, size_t j1) {
    const trsm_left_task *t = (const trsm_left_task *)ctx;
    for (size_t j = j0; j < j1; ++j) {
        double *x = t->b + j * (size_t)t->ldb;
        if (t->alpha != 1.0) for (size_t i = 0; i < t->m; ++i) x[i] = t->alpha * x[i];
        trsv_unit(t->uplo, t->trans, t->diag, t->m, t->a, t->lda, x);
    }
}

void oracle_dtrsm(char side, char uplo, char transa, char diag, int m, int n, double alpha,
                  const double *a, int lda, double *b, int ldb) {
    if (m <= 0 || n <= 0) return;
    int trans = !is_no_trans(transa);
    if (side == 'L' || side == 'l') {
        trsm_left_task t = {uplo, diag, trans, (size_t)m, alpha, a, lda, b, ldb};
        run_column_tasks(trsm_left_cols, &t, (size_t)n, 1);
        return;
    }
    /* right side: B <- alpha*B, then per row a strided solve with the
       transposed operation (triangular.zig:72-84; trsvFallback non-fused) */
    if (alpha != 1.0)
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < m; ++i) b[mi(ldb, i, j)] = alpha * b[mi(ldb, i, j)];
    int vtrans = !trans; /* row x solves op(A)^T x = row */
    int nonunit = !(diag == 'U' || diag == 'u');
    int eff_upper = ((uplo == 'U' || uplo == 'u') && !vtrans) || (!(uplo == 'U' || uplo == 'u') && vtrans);
    for (int i = 0; i < m; ++i) {
        double *x = b + i;
        size_t inc = (size_t)ldb;
        size_t nn = (size_t)n;
        if (eff_upper) {
            for (size_t rr = nn; rr-- > 0;) {
                double tv = x[rr * inc];
                for (size_t j = rr + 1; j < nn; ++j) { double pr = tri_val(uplo, diag, vtrans, a, lda, rr, j) * x[j * inc]; tv = tv - pr; }
                if (nonunit) tv = tv / tri_val(uplo, diag, vtrans, a, lda, rr, rr);
                x[rr * inc] = tv;
            }
        } else {
            for (size_t r = 0; r < nn; ++r) {
                double tv = x[r * inc];
                for (size_t j = 0; j < r; ++j) { double pr = tri_val(uplo, diag, vtrans, a, lda, r, j) * x[j * inc]; tv = tv - pr; }
                if (nonunit) tv = tv / tri_val(uplo, diag, vtrans, a, lda, r, r);
                x[r * inc] = tv;
            }
        }
    }
}

/* ------------------------------------------------------------------------ */
/* SYRK: symmetric.zig:180-200 (non-fused dot, stored triangle only)         */
/* ------------------------------------------------------------------------ */
typedef struct { int upper, trans; size_t n, k; double alpha, beta; const double *a; int lda; double *c; int ldc; } syrk_task;

static void syrk_cols(void *ctx, size_t j0, size_t j1) {
    const syrk_task *t = (const syrk_task *)ctx;
    for (size_t j = j0; j < j1; ++j) {
        size_t r0 = t->upper ? 0 : j, r1 = t->upper ? j + 1 : t->n;
        for (size_t i = r0; i < r1; ++i) {
            double sum = 0.0;
            for (size_t p = 0; p < t->k; ++p) {
                double ai = t->trans ? t->a[mi(t->lda, p, i)] : t->a[mi(t->lda, i, p)];
                double aj = t->trans ? t->a[mi(t->lda, p, j)] : t->a[mi(t->lda, j, p)];
                double pr = ai * aj;
                sum = sum + pr;
            }
            double *cp = &t->c[mi(t->ldc, i, j)];
            double as = t->alpha * sum;
            double bs = (t->beta == 0.0) ? 0.0 : t->beta * *cp;
            *cp = as + bs;
        }
    }
}

void oracle_dsyrk(char uplo, char trans, int n, int k, double alpha, const double *a, int lda,
                  double beta, double *c, int ldc) {
    if (n <= 0) return;
    syrk_task t = {(uplo == 'U' || uplo == 'u'), !is_no_trans(trans), (size_t)n, (size_t)(k < 0 ? 0 : k),
                   alpha, beta, a, lda, c, ldc};
    run_column_tasks(syrk_cols, &t, (size_t)n, 1);
}

/* ------------------------------------------------------------------------ */
/* SYMM: symmetric.zig:56-70,107-160 with symValue (access.zig:31-42):       */
/* per element sum = sum + S(i,p) B(p,j) (left) or B(i,p) S(p,j) (right),    */
/* p ascending, non-fused; then alpha*sum + (beta == 0 ? 0 : beta*C).        */
/* There is no alpha == 0 quick return in the reference.                      */
/* ------------------------------------------------------------------------ */
static inline double sym_val(char uplo, const double *a, int lda, size_t r, size_t c) {
    int upper = (uplo == 'U' || uplo == 'u');
    if ((upper && r <= c) || (!upper && r >= c)) return a[mi(lda, r, c)];
    return a[mi(lda, c, r)];
}

void oracle_dsymm(char side, char uplo, int m, int n, double alpha, const double *a, int lda,
                  const double *b, int ldb, double beta, double *c, int ldc) {
    if (m <= 0 || n <= 0) return;
    int left = (side == 'L' || side == 'l');
    for (size_t j = 0; j < (size_t)n; ++j)
        for (size_t i = 0; i < (size_t)m; ++i) {
            double sum = 0.0;
            if (left)
                for (size_t p = 0; p < (size_t)m; ++p) sum = sum + sym_val(uplo, a, lda, i, p) * b[mi(ldb, p, j)];
            else
                for (size_t p = 0; p < (size_t)n; ++p) sum = sum + b[mi(ldb, i, p)] * sym_val(uplo, a, lda, p, j);
            double *cp = &c[mi(ldc, i, j)];
            double s = (beta == 0.0) ? 0.0 : beta * *cp;
            *cp = alpha * sum + s;
        }
}

/* ------------------------------------------------------------------------ */
/* SYR2K: symmetric.zig:301-329,366-412: for the stored triangle,             */
/* sum = sum + (a_ip b_jp + b_ip a_jp) over p ascending (op(X)(i,p) read by  */
/* matrixValue), non-fused; then alpha*sum + (beta == 0 ? 0 : beta*C).       */
/* ------------------------------------------------------------------------ */
void oracle_dsyr2k(char uplo, char trans, int n, int k, double alpha, const double *a, int lda,
                   const double *b, int ldb, double beta, double *c, int ldc) {
    if (n <= 0) return;
    int upper = (uplo == 'U' || uplo == 'u');
    int tr = !is_no_trans(trans);
    size_t kk = (size_t)(k < 0 ? 0 : k);
    for (size_t j = 0; j < (size_t)n; ++j) {
        size_t r0 = upper ? 0 : j, r1 = upper ? j + 1 : (size_t)n;
        for (size_t i = r0; i < r1; ++i) {
            double sum = 0.0;
            for (size_t p = 0; p < kk; ++p) {
                double ai = tr ? a[mi(lda, p, i)] : a[mi(lda, i, p)];
                double bi = tr ? b[mi(ldb, p, i)] : b[mi(ldb, i, p)];
                double aj = tr ? a[mi(lda, p, j)] : a[mi(lda, j, p)];
                double bj = tr ? b[mi(ldb, p, j)] : b[mi(ldb, j, p)];
                double t1 = ai * bj;
                double t2 = bi * aj;
                sum = sum + (t1 + t2);
            }
            double *cp = &c[mi(ldc, i, j)];
            double s = (beta == 0.0) ? 0.0 : beta * *cp;
            *cp = alpha * sum + s;
        }
    }
}

/* ------------------------------------------------------------------------ */
/* TRMM: triangular.zig:202-231.                                              */
/*  - left: per column, trmv (the scalar trmvUnit of                          */
/*    matrix_vector/triangular.zig:44-57,59-78,80-133: axpy/dot steps,        */
/*    non-fused), then B(:,j) = alpha*B(:,j) unless alpha == 1                */
/*    (scaleDenseColumns, triangular.zig:52-59).  The reference may run       */
/*    tuned SIMD axpy/dot bodies here; their rounding is parity unpinned.     */
/*  - right: per row, sum over all p of B(i,p) op(A)(p,j) (triValue gives 0   */
/*    outside the triangle and 1 on a unit diagonal), B(i,j) = alpha*sum,     */
/*    columns descending when op(A) is upper so the row is read before it    */
/*    is overwritten.                                                         */
/* ------------------------------------------------------------------------ */
static void trmv_unit(char uplo, int trans, char diag, size_t n, const double *a, int lda, double *x) {
    int upper = (uplo == 'U' || uplo == 'u');
    int nonunit = !(diag == 'U' || diag == 'u');
    if (!trans) {
        if (upper) {
            for (size_t j = 0; j < n; ++j) {
                const double *col = a + j * (size_t)lda;
                double xj = x[j];
                if (xj != 0.0) for (size_t i = 0; i < j; ++i) x[i] = x[i] + xj * col[i];
                if (nonunit) x[j] = xj * col[j];
            }
        } else {
            for (size_t j = n; j-- > 0;) {
                const double *col = a + j * (size_t)lda;
                double xj = x[j];
                if (nonunit) x[j] = xj * col[j];
                if (xj != 0.0) for (size_t i = j + 1; i < n; ++i) x[i] = x[i] + xj * col[i];
            }
        }
        return;
    }
    if (upper) {
        for (size_t j = n; j-- > 0;) {
            const double *col = a + j * (size_t)lda;
            double sum = 0.0;
            for (size_t i = 0; i < j; ++i) sum = sum + col[i] * x[i];
            double d = nonunit ? col[j] : 1.0;
            sum = sum + d * x[j];
            x[j] = sum;
        }
    } else {
        for (size_t j = 0; j < n; ++j) {
            const double *col = a + j * (size_t)lda;
            double d = nonunit ? col[j] : 1.0;
            double sum = d * x[j];
            double dot = 0.0;
            for (size_t i = j + 1; i < n; ++i) dot = dot + col[i] * x[i];
            sum = sum + dot;
            x[j] = sum;
        }
    }
}

void oracle_dtrmm(char side, char uplo, char transa, char diag, int m, int n, double alpha,
                  const double *a, int lda, double *b, int ldb) {
    if (m <= 0 || n <= 0) return;
    int trans = !is_no_trans(transa);
    int upper = (uplo == 'U' || uplo == 'u');
    if (side == 'L' || side == 'l') {
        for (size_t j = 0; j < (size_t)n; ++j) {
            double *x = b + j * (size_t)ldb;
            trmv_unit(uplo, trans, diag, (size_t)m, a, lda, x);
            if (alpha != 1.0)
                for (size_t i = 0; i < (size_t)m; ++i) x[i] = alpha * x[i];
        }
        return;
    }
    int op_upper = (!trans && upper) || (trans && !upper);
    for (size_t i = 0; i < (size_t)m; ++i) {
        for (size_t t = 0; t < (size_t)n; ++t) {
            size_t jj = op_upper ? (size_t)n - 1 - t : t;
            double sum = 0.0;
            for (size_t p = 0; p < (size_t)n; ++p) sum = sum + b[mi(ldb, i, p)] * tri_val(uplo, diag, trans, a, lda, p, jj);
            b[mi(ldb, i, jj)] = alpha * sum;
        }
    }
}

/* ------------------------------------------------------------------------ */
/* double-double GEMM (error oracle)                                         */
/* TwoSum (Knuth, TAOCP vol. 2), TwoProd via FMA (Ogita-Rump-Oishi 2005)     */
/* ------------------------------------------------------------------------ */
typedef struct { double hi, lo; } dd;

static inline dd two_sum(double a, double b) {
    double s = a + b;
    double bb = s - a;
    double e = (a - (s - bb)) + (b - bb);
    dd r = {s, e};
    return r;
}
static inline dd fast_two_sum(double a, double b) {
    double s = a + b;
    dd r = {s, b - (s - a)};
    return r;
}
static inline dd dd_add(dd x, dd y) {
    dd s = two_sum(x.hi, y.hi);
    dd t = two_sum(x.lo, y.lo);
    s.lo += t.hi;
    s = fast_two_sum(s.hi, s.lo);
    s.lo += t.lo;
    return fast_two_sum(s.hi, s.lo);
}
static inline dd dd_mul_d(dd x, double y) {
    double p = x.hi * y;
    double e = fma(x.hi, y, -p);
    e = fma(x.lo, y, e);
    return fast_two_sum(p, e);
}
static inline dd two_prod(double a, double b) {
    double p = a * b;
    dd r = {p, fma(a, b, -p)};
    return r;
}

/* dd value of alpha*sum_p op(A)(i,p) op(B)(p,j) + beta*C0(i,j) for a list of
   (row, col) entries; c0 may be NULL when beta == 0. */
void oracle_dd_gemm_entries(char ta, char tb, int k, double alpha, const double *a, int lda,
                            const double *b, int ldb, double beta, const double *c0, int ldc,
                            const int64_t *rows, const int64_t *cols, int64_t count,
                            double *out_hi, double *out_lo) {
    int nota = is_no_trans(ta), notb = is_no_trans(tb);
    for (int64_t e = 0; e < count; ++e) {
        size_t i = (size_t)rows[e], j = (size_t)cols[e];
        dd s = {0.0, 0.0};
        for (int p = 0; p < k; ++p) {
            double av = nota ? a[mi(lda, i, p)] : a[mi(lda, p, i)];
            double bv = notb ? b[mi(ldb, p, j)] : b[mi(ldb, j, p)];
            s = dd_add(s, two_prod(av, bv));
        }
        s = dd_mul_d(s, alpha);
        if (beta != 0.0) s = dd_add(s, two_prod(beta, c0[mi(ldc, i, j)]));
        out_hi[e] = s.hi;
        out_lo[e] = s.lo;
    }
}

/* ------------------------------------------------------------------------ */
/* Complex GEMM.  Storage is interleaved (re, im); leading dimensions count    */
/* complex elements.                                                          */
/*   - scalar add / mul / conj: core/shared/scalar.zig:56-60,77-82,91-95       */
/*     (four separately rounded products, no fused steps)                     */
/*   - element access with conjugation: core/matrix_vector/access.zig:23-29    */
/*   - k == 0 or alpha == 0 (both parts): gemm.zig:1407-1414                  */
/*   - portable loop: gemm.zig:1234-1246, sum over p ascending, then           */
/*     alpha*sum + (beta == 0 ? 0 : beta*c)                                   */
/* The reference's tuned complex routes (expanded real, 3M, compact,          */
/* gemm.zig:1260-1280) regroup the same sum; their exact roundings are        */
/* "parity unpinned" here and the tests compare within a stated bound.        */
/* ------------------------------------------------------------------------ */
#define DEFINE_CGEMM(NAME, ENTRIES, T)                                                         \
    static inline void NAME##_mul(T ar, T ai, T br, T bi, T *rr, T *ri) {                    \
        *rr = ar * br - ai * bi;                                                               \
        *ri = ar * bi + ai * br;                                                               \
    }                                                                                          \
    static inline void NAME##_at(char t, const T *x, int ld, size_t r, size_t c, T *re, T *im) { \
        const T *p = is_no_trans(t) ? &x[2 * mi(ld, r, c)] : &x[2 * mi(ld, c, r)];             \
        *re = p[0];                                                                            \
        *im = (t == 'C' || t == 'c') ? -p[1] : p[1];                                           \
    }                                                                                          \
    static inline void NAME##_one(char ta, char tb, int k, const T *alpha, const T *a, int lda, \
                                  const T *b, int ldb, const T *beta, const T *cold, size_t i,  \
                                  size_t j, T *out) {                                          \
        int bzero = beta[0] == (T)0 && beta[1] == (T)0;                                        \
        T br_ = 0, bi_ = 0;                                                                    \
        if (!bzero) NAME##_mul(beta[0], beta[1], cold[0], cold[1], &br_, &bi_);                \
        if (k <= 0 || (alpha[0] == (T)0 && alpha[1] == (T)0)) {                                \
            out[0] = br_;                                                                      \
            out[1] = bi_;                                                                      \
            return;                                                                            \
        }                                                                                      \
        T sr = 0, si = 0;                                                                      \
        for (int p = 0; p < k; ++p) {                                                          \
            T xr, xi, yr, yi, pr, pi;                                                          \
            NAME##_at(ta, a, lda, i, (size_t)p, &xr, &xi);                                     \
            NAME##_at(tb, b, ldb, (size_t)p, j, &yr, &yi);                                     \
            NAME##_mul(xr, xi, yr, yi, &pr, &pi);                                              \
            sr = sr + pr;                                                                      \
            si = si + pi;                                                                      \
        }                                                                                      \
        T ar_, ai_;                                                                            \
        NAME##_mul(alpha[0], alpha[1], sr, si, &ar_, &ai_);                                    \
        out[0] = ar_ + br_;                                                                    \
        out[1] = ai_ + bi_;                                                                    \
    }                                                                                          \
    void NAME(char ta, char tb, int m, int n, int k, const T *alpha, const T *a, int lda,      \
              const T *b, int ldb, const T *beta, T *c, int ldc) {                             \
        if (m <= 0 || n <= 0) return;                                                          \
        for (int j = 0; j < n; ++j)                                                            \
            for (int i = 0; i < m; ++i) {                                                      \
                T *cp = &c[2 * mi(ldc, i, j)];                                                 \
                T out[2];                                                                      \
                NAME##_one(ta, tb, k, alpha, a, lda, b, ldb, beta, cp, i, j, out);             \
                cp[0] = out[0];                                                                \
                cp[1] = out[1];                                                                \
            }                                                                                  \
    }                                                                                          \
    /* the same value for a list of (row, col) entries of C; c0 may be NULL   */               \
    /* when beta == 0; out holds interleaved (re, im) per entry               */               \
    void ENTRIES(char ta, char tb, int k, const T *alpha, const T *a, int lda, const T *b,     \
                 int ldb, const T *beta, const T *c0, int ldc, const int64_t *rows,            \
                 const int64_t *cols, int64_t count, T *out) {                                 \
        static const T zero2[2] = {0, 0};                                                      \
        for (int64_t e = 0; e < count; ++e) {                                                  \
            size_t i = (size_t)rows[e], j = (size_t)cols[e];                                   \
            const T *cold = c0 ? &c0[2 * mi(ldc, i, j)] : zero2;                               \
            NAME##_one(ta, tb, k, alpha, a, lda, b, ldb, beta, cold, i, j, &out[2 * e]);       \
        }                                                                                      \
    }

DEFINE_CGEMM(oracle_zgemm, oracle_zgemm_entries, double)
DEFINE_CGEMM(oracle_cgemm, oracle_cgemm_entries, float)

/* dd value of alpha*sum_p op(A)(i,p) op(B)(p,j) + beta*C0(i,j), complex, for a
   list of entries: out_hi / out_lo hold interleaved (re, im). */
void oracle_zdd_gemm_entries(char ta, char tb, int k, const double *alpha, const double *a, int lda,
                             const double *b, int ldb, const double *beta, const double *c0, int ldc,
                             const int64_t *rows, const int64_t *cols, int64_t count, double *out_hi,
                             double *out_lo) {
    for (int64_t e = 0; e < count; ++e) {
        size_t i = (size_t)rows[e], j = (size_t)cols[e];
        dd sr = {0.0, 0.0}, si = {0.0, 0.0};
        for (int p = 0; p < k; ++p) {
            double xr, xi, yr, yi;
            oracle_zgemm_at(ta, a, lda, i, (size_t)p, &xr, &xi);
            oracle_zgemm_at(tb, b, ldb, (size_t)p, j, &yr, &yi);
            sr = dd_add(sr, two_prod(xr, yr));
            sr = dd_add(sr, two_prod(-xi, yi));
            si = dd_add(si, two_prod(xr, yi));
            si = dd_add(si, two_prod(xi, yr));
        }
        dd rr = dd_add(dd_mul_d(sr, alpha[0]), dd_mul_d(si, -alpha[1]));
        dd ri = dd_add(dd_mul_d(si, alpha[0]), dd_mul_d(sr, alpha[1]));
        if (beta[0] != 0.0 || beta[1] != 0.0) {
            const double *cp = &c0[2 * mi(ldc, i, j)];
            rr = dd_add(rr, dd_add(two_prod(beta[0], cp[0]), two_prod(-beta[1], cp[1])));
            ri = dd_add(ri, dd_add(two_prod(beta[0], cp[1]), two_prod(beta[1], cp[0])));
        }
        out_hi[2 * e] = rr.hi;
        out_lo[2 * e] = rr.lo;
        out_hi[2 * e + 1] = ri.hi;
        out_lo[2 * e + 1] = ri.lo;
    }
}

/* wall clock for the CPU-baseline leg */
double oracle_now(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}
