/*
 * zynum_gpu.h -- C ABI of the MI355X (gfx950) Level-3 tier.
 *
 * Part 1 keeps Zynum's public BLAS symbols byte-for-byte (32-bit LP64 BLAS
 * integers, Fortran arguments by pointer with no hidden string lengths).  A
 * call whose operands live in device memory runs on the GPU; host operands are
 * staged through device memory by the same kernels.  Part 2 is the private
 * request ABI the host planner (or any caller holding a HIP stream) uses to
 * reach the GPU tier directly, including the split-precision and grouped
 * forms.  All signatures use plain pointers and sizes only.
 *
 * Reference interfaces (paths relative to the Zynum source tree):
 *   cblas_sgemm / cblas_dgemm  include/zynum/blas/cblas.h:1336,1353,
 *                              src/blas/abi/cblas.zig:1415-1439
 *   sgemm_ / dgemm_            include/zynum/blas/blas.h:161-162,
 *                              src/blas/abi/fortran.zig:760-767
 *   cgemm_ / zgemm_            include/zynum/blas/blas.h:163-164,
 *                              src/blas/abi/fortran.zig:768-775
 *   cblas_cgemm / cblas_zgemm  include/zynum/blas/cblas.h:1370-1402,
 *                              src/blas/abi/cblas.zig:1441-1465
 *   zynum_blas_complex_*       include/zynum/blas/blas.h:15-28
 *   dsyrk_ / cblas_dsyrk       src/blas/abi/fortran.zig:806-809,
 *                              src/blas/abi/cblas.zig:1550-1559, cblas.h:1514
 *   dtrsm_ / cblas_dtrsm       src/blas/abi/fortran.zig:873-876,
 *                              src/blas/abi/cblas.zig:1742-1753, cblas.h:1755
 *   ssyrk_ / strsm_            src/blas/abi/fortran.zig:802-805,869-872
 *   ssymm_ / dsymm_            src/blas/abi/fortran.zig:777-784 (checks :239-250)
 *   ssyr2k_ / dsyr2k_          src/blas/abi/fortran.zig:827-834 (checks :264-275)
 *   strmm_ / dtrmm_            src/blas/abi/fortran.zig:852-859 (checks :105-117)
 *   cblas_{s,d}symm, cblas_{s,d}syr2k, cblas_{s,d}trmm
 *                              src/blas/abi/cblas.zig:1467-1489,1605-1627,1677-1700
 *   xerbla_ / lsame_           src/blas/abi/fortran.zig:16-22
 *   zynum_blas_shutdown(_)     src/blas/abi/fortran.zig:41-47
 *   zynum_gpu_gemm_request     modeled on GemmRequest,
 *                              src/blas/kernels/isolated/x86_64_structured_abi.zig:50-70
 *                              (first 88 bytes identical; GPU fields appended)
 */
#ifndef ZYNUM_GPU_H
#define ZYNUM_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* every symbol declared here is exported from libzynum_gpu.so */
#pragma GCC visibility push(default)

typedef int32_t zynum_blas_int;
#ifndef ZYNUM_BLAS_TYPES_DEFINED
#define ZYNUM_BLAS_TYPES_DEFINED
typedef struct {
    float real;
    float imag;
} zynum_blas_complex_float;
typedef struct {
    double real;
    double imag;
} zynum_blas_complex_double;
typedef zynum_blas_complex_float zynum_blas_complexF32;
typedef zynum_blas_complex_double zynum_blas_complexF64;
#endif

/* ---------------------------------------------------------------------- */
/* Part 1: Zynum public BLAS symbols                                       */
/* ---------------------------------------------------------------------- */

enum {
    CblasRowMajor = 101, CblasColMajor = 102,
    CblasNoTrans = 111, CblasTrans = 112, CblasConjTrans = 113,
    CblasUpper = 121, CblasLower = 122,
    CblasNonUnit = 131, CblasUnit = 132,
    CblasLeft = 141, CblasRight = 142
};

void cblas_sgemm(int layout, int transa, int transb, int m, int n, int k, float alpha,
                 const float *a, int lda, const float *b, int ldb, float beta, float *c, int ldc);
void cblas_dgemm(int layout, int transa, int transb, int m, int n, int k, double alpha,
                 const double *a, int lda, const double *b, int ldb, double beta, double *c, int ldc);
/* complex forms: alpha and beta point at one complex scalar; a, b, c at interleaved (re, im) storage */
void cblas_cgemm(int layout, int transa, int transb, int m, int n, int k, const void *alpha,
                 const void *a, int lda, const void *b, int ldb, const void *beta, void *c, int ldc);
void cblas_zgemm(int layout, int transa, int transb, int m, int n, int k, const void *alpha,
                 const void *a, int lda, const void *b, int ldb, const void *beta, void *c, int ldc);
void cblas_ssymm(int layout, int side, int uplo, int m, int n, float alpha, const float *a, int lda,
                 const float *b, int ldb, float beta, float *c, int ldc);
void cblas_dsymm(int layout, int side, int uplo, int m, int n, double alpha, const double *a, int lda,
                 const double *b, int ldb, double beta, double *c, int ldc);
void cblas_ssyr2k(int layout, int uplo, int trans, int n, int k, float alpha, const float *a, int lda,
                  const float *b, int ldb, float beta, float *c, int ldc);
void cblas_dsyr2k(int layout, int uplo, int trans, int n, int k, double alpha, const double *a, int lda,
                  const double *b, int ldb, double beta, double *c, int ldc);
void cblas_strmm(int layout, int side, int uplo, int transa, int diag, int m, int n, float alpha,
                 const float *a, int lda, float *b, int ldb);
void cblas_dtrmm(int layout, int side, int uplo, int transa, int diag, int m, int n, double alpha,
                 const double *a, int lda, double *b, int ldb);
void cblas_dsyrk(int layout, int uplo, int trans, int n, int k, double alpha, const double *a,
                 int lda, double beta, double *c, int ldc);
void cblas_dtrsm(int layout, int side, int uplo, int transa, int diag, int m, int n, double alpha,
                 const double *a, int lda, double *b, int ldb);

void sgemm_(const char *transa, const char *transb, const zynum_blas_int *m, const zynum_blas_int *n,
            const zynum_blas_int *k, const float *alpha, const float *a, const zynum_blas_int *lda,
            const float *b, const zynum_blas_int *ldb, const float *beta, float *c,
            const zynum_blas_int *ldc);
void dgemm_(const char *transa, const char *transb, const zynum_blas_int *m, const zynum_blas_int *n,
            const zynum_blas_int *k, const double *alpha, const double *a, const zynum_blas_int *lda,
            const double *b, const zynum_blas_int *ldb, const double *beta, double *c,
            const zynum_blas_int *ldc);
void cgemm_(const char *transa, const char *transb, const zynum_blas_int *m, const zynum_blas_int *n,
            const zynum_blas_int *k, const zynum_blas_complex_float *alpha,
            const zynum_blas_complex_float *a, const zynum_blas_int *lda,
            const zynum_blas_complex_float *b, const zynum_blas_int *ldb,
            const zynum_blas_complex_float *beta, zynum_blas_complex_float *c,
            const zynum_blas_int *ldc);
void zgemm_(const char *transa, const char *transb, const zynum_blas_int *m, const zynum_blas_int *n,
            const zynum_blas_int *k, const zynum_blas_complex_double *alpha,
            const zynum_blas_complex_double *a, const zynum_blas_int *lda,
            const zynum_blas_complex_double *b, const zynum_blas_int *ldb,
            const zynum_blas_complex_double *beta, zynum_blas_complex_double *c,
            const zynum_blas_int *ldc);
void ssyrk_(const char *uplo, const char *trans, const zynum_blas_int *n, const zynum_blas_int *k,
            const float *alpha, const float *a, const zynum_blas_int *lda, const float *beta,
            float *c, const zynum_blas_int *ldc);
void dsyrk_(const char *uplo, const char *trans, const zynum_blas_int *n, const zynum_blas_int *k,
            const double *alpha, const double *a, const zynum_blas_int *lda, const double *beta,
            double *c, const zynum_blas_int *ldc);
void ssymm_(const char *side, const char *uplo, const zynum_blas_int *m, const zynum_blas_int *n,
            const float *alpha, const float *a, const zynum_blas_int *lda, const float *b,
            const zynum_blas_int *ldb, const float *beta, float *c, const zynum_blas_int *ldc);
void dsymm_(const char *side, const char *uplo, const zynum_blas_int *m, const zynum_blas_int *n,
            const double *alpha, const double *a, const zynum_blas_int *lda, const double *b,
            const zynum_blas_int *ldb, const double *beta, double *c, const zynum_blas_int *ldc);
void ssyr2k_(const char *uplo, const char *trans, const zynum_blas_int *n, const zynum_blas_int *k,
             const float *alpha, const float *a, const zynum_blas_int *lda, const float *b,
             const zynum_blas_int *ldb, const float *beta, float *c, const zynum_blas_int *ldc);
void dsyr2k_(const char *uplo, const char *trans, const zynum_blas_int *n, const zynum_blas_int *k,
             const double *alpha, const double *a, const zynum_blas_int *lda, const double *b,
             const zynum_blas_int *ldb, const double *beta, double *c, const zynum_blas_int *ldc);
void strmm_(const char *side, const char *uplo, const char *transa, const char *diag,
            const zynum_blas_int *m, const zynum_blas_int *n, const float *alpha, const float *a,
            const zynum_blas_int *lda, float *b, const zynum_blas_int *ldb);
void dtrmm_(const char *side, const char *uplo, const char *transa, const char *diag,
            const zynum_blas_int *m, const zynum_blas_int *n, const double *alpha, const double *a,
            const zynum_blas_int *lda, double *b, const zynum_blas_int *ldb);
void strsm_(const char *side, const char *uplo, const char *transa, const char *diag,
            const zynum_blas_int *m, const zynum_blas_int *n, const float *alpha, const float *a,
            const zynum_blas_int *lda, float *b, const zynum_blas_int *ldb);
void dtrsm_(const char *side, const char *uplo, const char *transa, const char *diag,
            const zynum_blas_int *m, const zynum_blas_int *n, const double *alpha, const double *a,
            const zynum_blas_int *lda, double *b, const zynum_blas_int *ldb);

int lsame_(const char *ca, const char *cb);
void xerbla_(const char *srname, const zynum_blas_int *info);
void zynum_blas_shutdown(void);
void zynum_blas_shutdown_(void);

/* ---------------------------------------------------------------------- */
/* Part 2: GPU-tier private ABI (no counterpart in the reference beyond    */
/* the GemmRequest layout it extends)                                      */
/* ---------------------------------------------------------------------- */

/* Return codes.  Every nonzero code except ZYNUM_GPU_ERR_LAUNCH is returned
   before any caller-visible memory is written (decline-before-mutation). */
enum {
    ZYNUM_GPU_OK = 0,
    ZYNUM_GPU_DECLINE_NO_DEVICE = 1,   /* no usable gfx950 device */
    ZYNUM_GPU_DECLINE_ALLOC = 2,       /* staging or workspace allocation failed */
    ZYNUM_GPU_DECLINE_UNSUPPORTED = 3, /* form/mode/scalar not provided */
    ZYNUM_GPU_DECLINE_ARGUMENT = 4,    /* invalid argument (xerbla-style check failed) */
    ZYNUM_GPU_DECLINE_FORCED = 5,      /* decline injected by zynum_gpu_debug_force_decline */
    ZYNUM_GPU_ERR_LAUNCH = 6,          /* HIP error after launch; outputs undefined */
    ZYNUM_GPU_DECLINE_SCRATCH = 7      /* zynum_gpu_gemm_noalloc: the call needs more than its fixed scratch */
};

/* scalar codes of the reference's structured ABI (x86_64_structured_abi.zig:14-19) */
enum { ZYNUM_GPU_F32 = 0, ZYNUM_GPU_F64 = 1, ZYNUM_GPU_C32 = 2, ZYNUM_GPU_C64 = 3 };
enum { ZYNUM_GPU_NO_TRANS = 0, ZYNUM_GPU_TRANS = 1, ZYNUM_GPU_CONJ_TRANS = 2 };

/* Precision modes of a DGEMM request. */
enum {
    ZYNUM_GPU_MODE_PLAIN = 0,      /* FP64 (or FP32) MFMA */
    ZYNUM_GPU_MODE_OZAKI_INT8 = 1, /* split precision: mode_param INT8 slices (2..8), INT32 accumulation */
    ZYNUM_GPU_MODE_OZAKI_BF16 = 2, /* split precision: mode_param BF16 slices (2..8), FP32 accumulation */
    ZYNUM_GPU_MODE_OZAKI_CRT = 3,  /* split precision: mode_param INT8 residue GEMMs (8..20) reconstructed
                                      by the Chinese remainder theorem (Ozaki scheme II) */
    ZYNUM_GPU_MODE_DD = 4          /* double-double FP64 VALU reference (accuracy oracle on the device) */
};
/* Split-precision modes: K < 2^17 (CRT) and slices * K < 2^19 (INT8 slices) keep the INT32 sums
   exact; larger K declines with ZYNUM_GPU_DECLINE_UNSUPPORTED before writing.  A NaN or an
   infinity in a row of op(A) or a column of op(B) makes that row / column of the result NaN. */

/* request flags */
enum {
    ZYNUM_GPU_FLAG_ASYNC = 1 /* do not synchronize the stream before returning */
};

typedef struct zynum_gpu_gemm_request {
    uint8_t transa;      /* ZYNUM_GPU_NO_TRANS / TRANS / CONJ_TRANS */
    uint8_t transb;
    uint8_t scalar;      /* ZYNUM_GPU_F32 / F64 / C32 / C64 */
    uint8_t reserved0[5];
    int32_t m, n, k, lda, ldb, ldc;
    uint64_t alpha_re;   /* bit pattern of alpha (double, or float in the low 32 bits) */
    uint64_t alpha_im;   /* imaginary parts: complex scalars only, must be 0 for F32 / F64 */
    uint64_t beta_re;
    uint64_t beta_im;
    const void *a;
    const void *b;
    void *c;
    /* GPU-tier fields (offset 88) */
    void *stream;        /* hipStream_t; NULL uses the stream set by zynum_gpu_set_stream */
    uint8_t mode;        /* ZYNUM_GPU_MODE_* */
    uint8_t mode_param;  /* slices or moduli count; 0 picks the default */
    uint8_t flags;       /* ZYNUM_GPU_FLAG_* */
    uint8_t tile;        /* tile configuration of the GEMM core: 0 = tuned choice, i + 1 = configuration i
                            (plain: FP64/FP32 MFMA tiles 0..11; INT8 modes: 0..25; BF16: 0..3) */
    uint8_t reserved1[4];
} zynum_gpu_gemm_request; /* 104 bytes */

/* Run one GEMM request: C = alpha*op(A)*op(B) + beta*C (column-major).
   Arguments follow dgemm_ rules (gemmError, fortran.zig:88-103).  Complex
   requests (C32 / C64, interleaved storage, leading dimensions in complex
   elements) take CONJ_TRANS as the conjugate transpose and mode PLAIN only; they
   run one real MFMA GEMM on op(A) expanded into 2x2 real blocks. */
int zynum_gpu_gemm(const zynum_gpu_gemm_request *request);

/* zynum_gpu_gemm for callers that cannot allocate device memory (code under
   HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
   No counterpart in the reference.  Plain F32 / F64 requests on device (or
   managed) operands only; other modes, complex scalars and host operands
   decline with ZYNUM_GPU_DECLINE_UNSUPPORTED.  Results are bitwise those of
   zynum_gpu_gemm for the same request.  The split-K partials of shapes with
   few output tiles and a long K live in a fixed 24 MiB device array per GPU
   (held until the process ends) instead of allocated memory; a call that
   would need more declines with ZYNUM_GPU_DECLINE_SCRATCH (the split-K plan
   needs at most 20.9 MB).  Calls no hipMalloc, hipFree or allocating routine
   and creates no stream: it runs on the request's stream, else the stream
   set by zynum_gpu_set_stream, else the null stream.  Every call on a GPU
   shares that GPU's array: calls that overlap on one GPU race. */
int zynum_gpu_gemm_noalloc(const zynum_gpu_gemm_request *request);

/* One problem of a grouped GEMM batch (device pointers, column-major). */
typedef struct zynum_gpu_gemm_problem {
    uint8_t transa, transb, reserved0[6];
    int32_t m, n, k, lda, ldb, ldc;
    double alpha, beta;  /* used as float for an FP32 batch */
    const void *a;
    const void *b;
    void *c;
} zynum_gpu_gemm_problem; /* 72 bytes */

/* Opaque grouped-GEMM plan: the problem list is validated, bucketed by size,
   assigned a tile shape per bucket by the autotuner table (tile = 0), or a
   forced tile configuration (tile = i + 1: 0 small 32x64 / 64x32 tiles, 1 64x64,
   2 128x128), expanded
   into output tiles in longest-first order and uploaded once.
   zynum_gpu_grouped_run issues one persistent launch per batch; it allocates
   nothing and may be captured into a HIP graph.  Problem pointers are
   captured by the plan: rerun it after rewriting the same buffers. */
typedef struct zynum_gpu_grouped_plan zynum_gpu_grouped_plan;

int zynum_gpu_grouped_plan_create(int scalar, const zynum_gpu_gemm_problem *problems, int count, int tile,
                                  zynum_gpu_grouped_plan **plan_out);
int zynum_gpu_grouped_run(zynum_gpu_grouped_plan *plan, void *stream);
/* Capture the persistent launch into a HIP graph owned by the plan and replay it. */
int zynum_gpu_grouped_graph_run(zynum_gpu_grouped_plan *plan, void *stream);
int zynum_gpu_grouped_plan_info(const zynum_gpu_grouped_plan *plan, int64_t *tiles_out,
                                int64_t *workgroups_out, double *flops_out);
void zynum_gpu_grouped_plan_destroy(zynum_gpu_grouped_plan *plan);

/* Reference loop for the same batch: one plain GEMM launch per problem. */
int zynum_gpu_grouped_loop(int scalar, const zynum_gpu_gemm_problem *problems, int count, void *stream);

/* Level-3 on the GEMM core, device pointers, stream-ordered.  Arguments follow
   dtrsm_/dsyrk_ (triangularMatrixError, rankKMatrixError).  `mode` is
   ZYNUM_GPU_MODE_PLAIN or a split-precision mode used for the off-diagonal
   GEMM updates (diagonal blocks are always solved in FP64). */
int zynum_gpu_dtrsm(char side, char uplo, char transa, char diag, int m, int n, double alpha,
                    const double *a, int lda, double *b, int ldb, void *stream, int mode,
                    int mode_param, int flags);
int zynum_gpu_dsyrk(char uplo, char trans, int n, int k, double alpha, const double *a, int lda,
                    double beta, double *c, int ldc, void *stream, int mode, int mode_param,
                    int flags);

/* DSYMM, DTRMM and DSYR2K on the GEMM core with a caller stream (arguments follow
   dsymm_ / dtrmm_ / dsyr2k_, symmetricMatrixError / triangularMatrixError /
   rank2KMatrixError).  SYMM expands the symmetric operand and TRMM the
   triangle into the library workspace (synchronous whatever the flags);
   SYR2K allocates nothing. */
int zynum_gpu_dsymm(char side, char uplo, int m, int n, double alpha, const double *a, int lda,
                    const double *b, int ldb, double beta, double *c, int ldc, void *stream, int flags);
int zynum_gpu_dtrmm(char side, char uplo, char transa, char diag, int m, int n, double alpha,
                    const double *a, int lda, double *b, int ldb, void *stream, int flags);
int zynum_gpu_dsyr2k(char uplo, char trans, int n, int k, double alpha, const double *a, int lda,
                     const double *b, int ldb, double beta, double *c, int ldc, void *stream, int flags);

/* Stream used by the Part 1 symbols and by private calls that pass no stream.
   NULL (the default) selects the library's own stream on the current device:
   non-blocking, default priority, and ordered after the work already queued
   on the device's null stream when each call starts.  Synchronous calls wait
   for their stream before returning. */
void zynum_gpu_set_stream(void *stream);
/* Host threads that copy host operands into the pinned staging buffers
   (1..64; 0 restores the default, min(8, hardware threads / 2)).  Takes
   effect from the next staged call.  Returns ZYNUM_GPU_OK, or
   ZYNUM_GPU_DECLINE_ARGUMENT for a count out of range. */
int zynum_gpu_set_copy_threads(int threads);
/* Last return code seen by a Part 1 symbol on this thread. */
int zynum_gpu_last_status(void);
/* Test hook: the next `count` GPU-tier calls decline with ZYNUM_GPU_DECLINE_FORCED. */
void zynum_gpu_debug_force_decline(int count);
/* Human-readable text for a return code. */
const char *zynum_gpu_status_string(int status);
/* Library version string. */
const char *zynum_gpu_version(void);

/* Host-pointer offload crossover.  Replaces the work thresholds of the
   reference's planner (src/blas/core/matrix_matrix/planner.zig:13,88-154,
   which pick a CPU thread count from m*n*k) with a GPU-vs-CPU decision for a
   planner that has this GPU tier: host operands are staged through cached
   pinned bounce buffers, and the choice weighs that staging against the CPU
   tier's time.  The model (measured device call times and staging rates,
   written by zynum_amd/autotune.py) is host arithmetic only: these functions
   touch no device and are safe to call on machines without a GPU.
   host_mask: bit 0 A, bit 1 B, bit 2 C in host memory.  beta_nonzero: C is
   read (staged in as well as out).  cpu_gflops: the CPU tier's rate for this
   call, in GFLOP/s counted as 2mnk (8mnk complex). */
/* Predicted wall nanoseconds of one synchronous GPU-tier call; 0 for m or
   n = 0; negative for an invalid scalar, dimension or mask. */
double zynum_gpu_offload_predict_ns(int scalar, int64_t m, int64_t n, int64_t k, int host_mask,
                                    int beta_nonzero);
/* 1: the GPU tier is predicted faster; 0: keep the call on the CPU (also for
   m, n or k = 0); -ZYNUM_GPU_DECLINE_ARGUMENT for invalid arguments. */
int zynum_gpu_offload_choose(int scalar, int64_t m, int64_t n, int64_t k, int host_mask, int beta_nonzero,
                             double cpu_gflops);
/* Smallest square edge e (m = n = k = e) the GPU tier takes at this CPU rate,
   as a planner's precomputed threshold; -1 if none up to 65536;
   -ZYNUM_GPU_DECLINE_ARGUMENT for invalid arguments. */
int64_t zynum_gpu_offload_crossover(int scalar, int host_mask, int beta_nonzero, double cpu_gflops);

/* Double-double reference product on the device (FP64 VALU, TwoProd/TwoSum):
   hi + lo ~= alpha*op(A)*op(B) + beta*C0 to about 2^-104 relative.  C0 may be
   NULL when beta == 0; lo may be NULL.  Synchronous.  Used to state the
   error of the split-precision forms. */
int zynum_gpu_dd_gemm(char transa, char transb, int m, int n, int k, double alpha, const double *a,
                      int lda, const double *b, int ldb, double beta, const double *c0, int ldc0,
                      double *hi, double *lo, int ldo, void *stream);

#pragma GCC visibility pop

#ifdef __cplusplus
}
#endif

#endif /* ZYNUM_GPU_H */
