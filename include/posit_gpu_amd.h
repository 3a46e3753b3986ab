/*
 * posit_gpu_amd.h -- C ABI of the MI355X (gfx950) posit32 dense linear algebra engine.
 *
 * Plain C types only: a posit32 value is its 32-bit pattern (uint32_t), matrices are
 * column-major with leading dimensions (as in the reference's libRpack), and device
 * memory / streams are opaque pointers.  Every function returns a pga_status; an
 * illegal argument gives PGA_ERROR_ILLEGAL_ARGUMENT and pga_illegal_argument() then
 * reports the 1-based parameter number the reference's Mxerbla would have printed.
 *
 * The reference interface each entry point replaces is cited beside it (paths relative
 * to the reference tree).  The C++ drop-in of mpblas_posit.h / mplapack_posit.h
 * (include/mpblas_posit.h, include/mplapack_posit.h) is built on these functions.
 */
#ifndef POSIT_GPU_AMD_H
#define POSIT_GPU_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    PGA_SUCCESS = 0,
    PGA_ERROR_ILLEGAL_ARGUMENT = 1, /* see pga_illegal_argument() */
    PGA_ERROR_NOT_INITIALIZED = 2,
    PGA_ERROR_HIP = 3,              /* a HIP runtime call failed */
    PGA_ERROR_OUT_OF_MEMORY = 4,
    PGA_ERROR_NO_DEVICE = 5,
    PGA_ERROR_UNSUPPORTED = 6,
    PGA_ERROR_EXCEEDS_SCRATCH = 7   /* more than a no-allocation entry point's fixed scratch holds */
} pga_status;

typedef struct {
    char name[64];
    char arch[32];
    int compute_units;        /* hipDeviceProp_t.multiProcessorCount */
    int clock_khz;            /* hipDeviceProp_t.clockRate */
    int simd_per_cu;          /* 4 on CDNA */
    int lanes_per_simd;       /* 16 on CDNA: a wave64 VALU op takes 4 cycles */
    double int32_valu_peak;   /* CU * SIMD * lanes * clock, integer lane-ops per second */
    size_t hbm_bytes;
} pga_device_info;

/* ---- runtime ------------------------------------------------------------------- */

/* Select and initialise a device.  Replaces initOpenCL (libRpack/Rgemm_opencl.cpp:59-70)
 * and its one-time set-up (init_gpu, :22-57; OclHelper.hpp:43-285).  Idempotent.
 * device < 0 picks PGA_DEVICE from the environment, else device 0. */
pga_status pga_init(int device);
pga_status pga_finalize(void);
int pga_is_initialized(void);
const char *pga_status_string(pga_status s);
/* 1-based parameter number of the last illegal argument on this thread (0 if none). */
int pga_illegal_argument(void);
pga_status pga_get_device_info(pga_device_info *info);

/* GEMM implementation: 0 = tiled kernel (default), 1 = exact one-element-per-thread kernel
 * (validation).  Also set by PGA_GEMM_IMPL=tiled|exact.  PGA_GEMM_TILE=16|32|64 (read at
 * initialisation; 64 selects the 64 x 32 tile) fixes the tiled kernel's tile instead of the
 * per-shape choice (the role of the reference's OPENCL_GEMM_BLOCKSIZE,
 * Rgemm_opencl.cpp:62-67); the bits do not depend on it.  The 64 x 32 tile of GEMMs with
 * m*n*k >= 2^30 (PGA_GEMM_PANEL_MIN) converts A and alpha*B to FP64 operand panels once,
 * in workspace from a library memory pool, instead of in every tile; PGA_GEMM_PANELS=0
 * turns that off (same bits either way). */
pga_status pga_set_gemm_impl(int impl);

/* ---- device memory and streams (plumbing for resident data) ------------------- */
pga_status pga_malloc(void **dptr, size_t bytes);
pga_status pga_free(void *dptr);
pga_status pga_memcpy_h2d(void *dst, const void *src, size_t bytes);
pga_status pga_memcpy_d2h(void *dst, const void *src, size_t bytes);
pga_status pga_memcpy_d2d(void *dst, const void *src, size_t bytes);
pga_status pga_memset(void *dptr, int value, size_t bytes);
pga_status pga_stream_create(void **stream);
pga_status pga_stream_destroy(void *stream);
pga_status pga_stream_synchronize(void *stream); /* NULL: the library's default stream */
pga_status pga_device_synchronize(void);
pga_status pga_event_create(void **event);
pga_status pga_event_destroy(void *event);
pga_status pga_event_record(void *event, void *stream);
pga_status pga_event_synchronize(void *event);
pga_status pga_event_elapsed_ms(float *ms, void *start, void *stop);

/* ---- posit32 GEMM ----------------------------------------------------------------
 * C = alpha*op(A)*op(B) + beta*C with the per-element rounding order of the CPU Rgemm
 * (libRpack/Rgemm.cpp:62-251): NN/NT accumulate (alpha*b)*a into beta*C in ascending
 * k; TN/TT sum a*b from zero, then alpha*t + beta*C.
 *
 * pga_rgemm: host pointers, synchronous.  Replaces RgemmG (libRpack/Rgemm_opencl.cpp:
 * 199-250) and Rgemm_internal (:72-197); copies only the logical sub-matrices.  transa /
 * transb follow Rgemm's argument check ('N','T','C' in either case; 'C' means 'T').
 * pga_rgemm_device: the same on device pointers, asynchronous on `stream`. */
/* Autotuner of the plain GEMM: times the tile edges {64 (with FP64 operand panels where
 * m*n*k >= 2^30), 32, 16} -- the workgroup's block of C, which takes the place of
 * OPENCL_GEMM_BLOCKSIZE -- on scratch operands of this exact shape and records the fastest
 * for later calls of that shape (pga_rgemm, pga_rgemm_device, the drop-in RgemmG).  Every
 * tile gives the same bits.  *tile receives it (0 for an empty shape). */
pga_status pga_rgemm_autotune(char transa, char transb, int m, int n, int k, uint32_t alpha, uint32_t beta,
                              int *tile);
pga_status pga_rgemm(char transa, char transb, int m, int n, int k, uint32_t alpha,
                     const uint32_t *a, int lda, const uint32_t *b, int ldb, uint32_t beta,
                     uint32_t *c, int ldc);
pga_status pga_rgemm_device(char transa, char transb, int m, int n, int k, uint32_t alpha,
                            const uint32_t *a, int lda, const uint32_t *b, int ldb,
                            uint32_t beta, uint32_t *c, int ldc, void *stream);

/* ---- BLAS-2/3 (libRpack/mpblas_posit.h:37-45) -------------------------------------
 * Host-pointer forms are synchronous drop-ins for the CPU routines of the reference; the
 * *_device forms take device pointers and run asynchronously on `stream`.  Argument
 * checks and Mxerbla parameter numbers are those of the reference routine. */
/* Rtrsm (libRpack/Rtrsm.cpp:31-291): all side/uplo/transa/diag combinations */
pga_status pga_rtrsm(char side, char uplo, char transa, char diag, int m, int n, uint32_t alpha,
                     const uint32_t *a, int lda, uint32_t *b, int ldb);
pga_status pga_rtrsm_device(char side, char uplo, char transa, char diag, int m, int n,
                            uint32_t alpha, const uint32_t *a, int lda, uint32_t *b, int ldb,
                            void *stream);
/* Rsyrk (libRpack/Rsyrk.cpp:31-215) */
pga_status pga_rsyrk(char uplo, char trans, int n, int k, uint32_t alpha, const uint32_t *a,
                     int lda, uint32_t beta, uint32_t *c, int ldc);
pga_status pga_rsyrk_device(char uplo, char trans, int n, int k, uint32_t alpha, const uint32_t *a,
                            int lda, uint32_t beta, uint32_t *c, int ldc, void *stream);
/* Rgemv (libRpack/Rgemv.cpp:31-204) */
pga_status pga_rgemv(char trans, int m, int n, uint32_t alpha, const uint32_t *a, int lda,
                     const uint32_t *x, int incx, uint32_t beta, uint32_t *y, int incy);
pga_status pga_rgemv_device(char trans, int m, int n, uint32_t alpha, const uint32_t *a, int lda,
                            const uint32_t *x, int incx, uint32_t beta, uint32_t *y, int incy,
                            void *stream);
/* Rscal (libRpack/Rscal.cpp:62-119) */
pga_status pga_rscal(int n, uint32_t da, uint32_t *x, int incx);
pga_status pga_rscal_device(int n, uint32_t da, uint32_t *x, int incx, void *stream);
/* iRamax (libRpack/iRamax.cpp:31-88): *result = 1-based index (0 if n < 1 or incx <= 0);
 * the device form writes it to device memory */
pga_status pga_iramax(int n, const uint32_t *x, int incx, int *result);
pga_status pga_iramax_device(int n, const uint32_t *x, int incx, int *result, void *stream);
/* Rlaswp (libRpack/Rlaswp.cpp:32-111); ipiv is 1-based */
pga_status pga_rlaswp(int n, uint32_t *a, int lda, int k1, int k2, const int *ipiv, int incx);
pga_status pga_rlaswp_device(int n, uint32_t *a, int lda, int k1, int k2, const int *ipiv,
                             int incx, void *stream);

/* ---- LAPACK (libRpack/mplapack_posit.h:38-57) --------------------------------------
 * ipiv 1-based, info as in LAPACK: 0 ok, >0 first zero pivot (LU, factorization
 * continues) or order of the first non-positive leading minor (Cholesky, stops).
 * nb <= 0 selects iMlaenv's 64.  Device forms keep ipiv and info in device memory. */
/* Rgetrf_opencl (libRpack/Rgetrf_opencl.cpp:32-154, block size `nbb`) / Rgetrf */
pga_status pga_rgetrf(int m, int n, uint32_t *a, int lda, int *ipiv, int *info, int nb);
pga_status pga_rgetrf_device(int m, int n, uint32_t *a, int lda, int *ipiv, int *info, int nb,
                             void *stream);
/* Rgetrf2 / Rgetrf2_opencl (libRpack/Rgetrf2.cpp:32-185): recursive, unblocked */
pga_status pga_rgetrf2(int m, int n, uint32_t *a, int lda, int *ipiv, int *info);
pga_status pga_rgetrf2_device(int m, int n, uint32_t *a, int lda, int *ipiv, int *info,
                              void *stream);
/* Rgetrs (libRpack/Rgetrs.cpp:32-118) */
pga_status pga_rgetrs(char trans, int n, int nrhs, const uint32_t *a, int lda, const int *ipiv,
                      uint32_t *b, int ldb, int *info);
pga_status pga_rgetrs_device(char trans, int n, int nrhs, const uint32_t *a, int lda,
                             const int *ipiv, uint32_t *b, int ldb, void *stream);
/* Rpotrf_opencl / Rpotrf (libRpack/Rpotrf_opencl.cpp:32-166) */
pga_status pga_rpotrf(char uplo, int n, uint32_t *a, int lda, int *info, int nb);
pga_status pga_rpotrf_device(char uplo, int n, uint32_t *a, int lda, int *info, int nb,
                             void *stream);
/* Rpotrf2 (libRpack/Rpotrf2.cpp:32-151): recursive, unblocked */
pga_status pga_rpotrf2(char uplo, int n, uint32_t *a, int lda, int *info);
pga_status pga_rpotrf2_device(char uplo, int n, uint32_t *a, int lda, int *info, void *stream);
/* Rpotrs (libRpack/Rpotrs.cpp:32-110) */
pga_status pga_rpotrs(char uplo, int n, int nrhs, const uint32_t *a, int lda, uint32_t *b, int ldb,
                      int *info);
pga_status pga_rpotrs_device(char uplo, int n, int nrhs, const uint32_t *a, int lda, uint32_t *b,
                             int ldb, void *stream);

/* ---- batched GEMM (one persistent launch, hipGraph-capturable) -------------------
 * A plan describes `count` independent problems C_i = alpha_i*op(A_i)*op(B_i) +
 * beta_i*C_i (device pointers, one transa/transb for the batch) with the semantics of
 * pga_rgemm_device for each.  Executing it launches a one-lane counter reset and one
 * persistent kernel in which workgroups take tiles of all problems from a queue.  tile = 0
 * lets the autotuner time tile shapes {64,32,16} (the workgroup's block of C, which takes
 * the place of OPENCL_GEMM_BLOCKSIZE) x grid sizes {1,2,4,8} x CUs, then at the best of
 * them the problem order (largest first, the caller's, largest and smallest alternating),
 * on scratch outputs, and remember the winner per batch signature (also in the file named
 * by PGA_TUNE_CACHE); otherwise tile and grid (0 = 2 per CU) are used as given, largest
 * problems first.  Problems with alpha == 0 or k == 0 and beta != 1 are rejected
 * (PGA_ERROR_UNSUPPORTED).  The reference has no batched form: this extends RgemmG
 * (libRpack/Rgemm_opencl.cpp:199) to many problems per launch. */
typedef struct pga_batch_plan pga_batch_plan;
pga_status pga_batch_plan_create(pga_batch_plan **plan, char transa, char transb, int count,
                                 const int *m, const int *n, const int *k, const uint32_t *alpha,
                                 const uint32_t *const *a, const int *lda,
                                 const uint32_t *const *b, const int *ldb, const uint32_t *beta,
                                 uint32_t *const *c, const int *ldc, int tile, int grid);
pga_status pga_batch_plan_execute(pga_batch_plan *plan, void *stream);
pga_status pga_batch_plan_capture_graph(pga_batch_plan *plan, void *stream);
pga_status pga_batch_plan_launch_graph(pga_batch_plan *plan, void *stream);
pga_status pga_batch_plan_info(const pga_batch_plan *plan, int *tile, int *grid, int *total_tiles,
                               double *tune_ms);
/* The plan's problem order: 0 largest first, 1 the caller's, 2 largest and smallest alternating */
pga_status pga_batch_plan_schedule(const pga_batch_plan *plan, int *sched);
pga_status pga_batch_plan_destroy(pga_batch_plan *plan);

/* No-allocation batched GEMM.  No counterpart in the reference: it is for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  One call does what pga_batch_plan_create (with the
 * given tile, no autotuning) + pga_batch_plan_execute + pga_batch_plan_destroy do, with
 * the same kernel and the same bits, but calls no hipMalloc, hipFree or other allocating
 * routine: the problem descriptors, tile offsets and work counter of the batch live in a
 * fixed device array of about 240 KiB that the library holds on each device it runs on
 * until the process ends (set up by pga_init).  Arguments as in pga_batch_plan_create;
 * tile 0 picks the largest of 64, 32, 16 that gives every CU four tiles, grid 0 two
 * workgroups per CU.  At most PGA_NOALLOC_MAX_PROBLEMS problems (else
 * PGA_ERROR_EXCEEDS_SCRATCH).  Asynchronous on `stream`.  Every call on a device uses
 * that device's one array: calls on one stream run in order, but calls that overlap on
 * one GPU (different streams or threads without ordering between them) race and may
 * compute wrong results. */
#define PGA_NOALLOC_MAX_PROBLEMS 4096
pga_status pga_batch_gemm_noalloc(char transa, char transb, int count, const int *m, const int *n,
                                  const int *k, const uint32_t *alpha, const uint32_t *const *a,
                                  const int *lda, const uint32_t *const *b, const int *ldb,
                                  const uint32_t *beta, uint32_t *const *c, const int *ldc, int tile,
                                  int grid, void *stream);

/* ---- batched factorizations and solves (one persistent launch, hipGraph-capturable) --
 * Each task is one of the per-call routines below on its own device matrix, computed by
 * one workgroup with the same operations in the same order as the per-call driver, so
 * factors, pivots, info codes and solutions equal the per-call results bit for bit:
 *   GETRF  = pga_getrf_device  (Rgetrf_opencl.cpp:32-154, block size nb, 0 -> 64)
 *   POTRF  = pga_potrf_device  (Rpotrf_opencl.cpp:32-166, uplo, block size nb, 0 -> 64)
 *   GETRS  = pga_getrs_device 'N' on an earlier GETRF's factors (Rgetrs.cpp:86-100)
 *   POTRS  = pga_potrs_device  (Rpotrs.cpp:86-104)
 *   GESV / POSV = the factorization, then the solve when info == 0.
 * Tasks are square, n <= PGA_BLOCKED_TASK_MAX_N; a, ipiv, b, info are device pointers
 * (info: one int, written with the factorization's info, 0 for a solve).  Tasks with
 * n <= PGA_FACTOR_TASK_MAX_N run one workgroup each in one persistent launch, most
 * expensive first; larger tasks advance together one column block at a time (a few
 * launches per block for all of them).  Tasks must not share output memory.  The
 * reference has no batched form: this extends Rgetrf_opencl / Rpotrf_opencl / Rgetrs /
 * Rpotrs. */
#define PGA_FACTOR_TASK_MAX_N 512
#define PGA_BLOCKED_TASK_MAX_N 16384
enum {
  PGA_TASK_GETRF = 0,
  PGA_TASK_POTRF = 1,
  PGA_TASK_GETRS = 2,
  PGA_TASK_POTRS = 3,
  PGA_TASK_GESV = 4,
  PGA_TASK_POSV = 5
};
typedef struct {
  int kind;      /* PGA_TASK_* */
  char uplo;     /* 'L' or 'U' for the Cholesky tasks */
  int n;         /* order */
  int nrhs;      /* right-hand sides of the solve tasks */
  int nb;        /* block size of the mirrored per-call routine (0: iMlaenv's 64) */
  uint32_t *a;   /* n x n, column-major, leading dimension lda */
  int lda;
  int *ipiv;     /* n entries, 1-based (LU tasks) */
  uint32_t *b;   /* n x nrhs, leading dimension ldb (solve tasks) */
  int ldb;
  int *info;     /* may be NULL */
} pga_factor_task;
typedef struct pga_factor_plan pga_factor_plan;
/* grid: workgroups of the persistent launch (0: as many as fit on the device at once) */
pga_status pga_factor_plan_create(pga_factor_plan **plan, int count, const pga_factor_task *tasks,
                                  int grid);
pga_status pga_factor_plan_execute(pga_factor_plan *plan, void *stream);
pga_status pga_factor_plan_capture_graph(pga_factor_plan *plan, void *stream);
pga_status pga_factor_plan_launch_graph(pga_factor_plan *plan, void *stream);
pga_status pga_factor_plan_info(const pga_factor_plan *plan, int *count, int *grid);
/* The autotuner of a plan's persistent launch: times it on scratch copies of every task's
 * operands for each workgroup count per CU that fits (1 .. the occupancy limit) and each
 * task order (0 longest first, 1 the caller's, 2 long and short alternating), and keeps
 * the fastest.  Results are the same bits in every configuration.  A captured graph is
 * released (capture again to launch the tuned form).  grid, sched and tune_ms may be NULL. */
pga_status pga_factor_plan_autotune(pga_factor_plan *plan, int *grid, int *sched, double *tune_ms);
pga_status pga_factor_plan_destroy(pga_factor_plan *plan);

/* ---- serving: posit32 solve requests with continuous batching -----------------------
 * A server owns an HBM arena of hbm_budget_bytes.  Requests are admitted in arrival
 * order while their device memory (factor + right-hand sides) fits; factors kept for
 * later re-solves stay resident and are evicted least-recently-used when memory is
 * short.  In PGA_SERVER_BATCHED mode every admitted request with n <= batch_max_n joins
 * the next persistent task launch (pga_factor_plan semantics) on one of batch_slots
 * streams as soon as a slot is free, while larger requests are taken by `workers`
 * threads with a stream each: a worker runs every large request ready at that moment as
 * one blocked batch (pga_factor_plan semantics for orders above PGA_FACTOR_TASK_MAX_N;
 * orders above PGA_BLOCKED_TASK_MAX_N run alone through the per-call drivers); all of it
 * overlaps on the device.  In
 * PGA_SERVER_PLAIN mode one worker runs each request alone through the per-call drivers
 * (pga_rgetrf_device + pga_rgetrs_device, pga_rpotrf_device + pga_rpotrs_device), one
 * after another.  Both modes give identical bits: a factor-and-solve request computes
 * Rgetrf_opencl / Rpotrf_opencl (nb 64) and, if info == 0, Rgetrs 'N' / Rpotrs.
 * Host arrays a, b, x must stay valid until the request completes.  The reference has
 * no serving layer: this wraps the routines of mplapack_posit.h:38-55. */
enum { PGA_SERVER_BATCHED = 0, PGA_SERVER_PLAIN = 1 };
enum { PGA_REQ_LU_SOLVE = 0, PGA_REQ_CHOL_SOLVE = 1, PGA_REQ_RESOLVE = 2 };
/* request status in pga_request_result */
enum { PGA_REQ_OK = 0, PGA_REQ_NOT_RESIDENT = 1, PGA_REQ_TOO_LARGE = 2, PGA_REQ_FAILED = 3 };
typedef struct {
  size_t hbm_budget_bytes; /* device arena for inputs, factors and solutions */
  size_t staging_bytes;    /* pinned host staging of batched solutions (0: 256 MiB) */
  int mode;                /* PGA_SERVER_BATCHED or PGA_SERVER_PLAIN */
  int batch_max_n;         /* largest order that joins batched launches (<= PGA_FACTOR_TASK_MAX_N) */
  int max_batch;           /* requests per batched launch (0: 1024) */
  int batch_slots;         /* batched launches in flight at once (0: 2, at most 2) */
  int workers;             /* threads running larger requests (0: 4, at most 8) */
  char chol_uplo;          /* triangle of the Cholesky requests, 'U' or 'L' */
} pga_server_config;
typedef struct {
  int kind;                /* PGA_REQ_* */
  int n, nrhs;
  const uint32_t *a;       /* host n x n, column-major (factor-and-solve requests) */
  const uint32_t *b;       /* host n x nrhs right-hand sides */
  uint32_t *x;             /* host n x nrhs solution */
  int64_t factor;          /* PGA_REQ_RESOLVE: factor id (the id of the request that made it) */
  int keep_factor;         /* keep this request's factor resident for later re-solves */
} pga_request;
typedef struct {
  int64_t id, factor;
  int status;              /* PGA_REQ_OK, ... */
  int info;                /* of the factorization used */
  double t_submit, t_admit, t_done; /* seconds since the server was created */
} pga_request_result;
typedef struct {
  int64_t submitted, completed, solved, batches, batched_requests, evictions;
  int resident_factors;
  size_t bytes_in_use, bytes_peak;
} pga_server_stats;
typedef struct pga_server pga_server;
pga_status pga_server_create(pga_server **server, const pga_server_config *config);
pga_status pga_server_submit(pga_server *server, const pga_request *request, int64_t *id);
pga_status pga_server_wait(pga_server *server, int64_t id, pga_request_result *result);
pga_status pga_server_drain(pga_server *server);
pga_status pga_server_release_factor(pga_server *server, int64_t factor);
pga_status pga_server_get_stats(pga_server *server, pga_server_stats *stats);
pga_status pga_server_destroy(pga_server *server);

/* ---- conversions (device codecs) -------------------------------------------------
 * double <-> posit32 on device arrays: round to nearest, ties to the even pattern,
 * saturating at minpos/maxpos; NaN and infinities become NaR.  (The reference converts
 * with SoftPosit's posit32(double) constructor and toDouble(), e.g. main_gemm.cpp:99-104.) */
pga_status pga_p32_from_f64_device(const double *x, uint32_t *p, size_t n, void *stream);
pga_status pga_p32_to_f64_device(const uint32_t *p, double *x, size_t n, void *stream);

/* Synthetic inputs: p[i] = posit32(lo + (hi-lo)*u_i) with u_i = SplitMix64(seed, i)
 * mapped to [0,1) with 53 bits.  (The reference draws from std::random_device,
 * util.cpp:166-233; the configurations here fix seeded uniform inputs instead.) */
pga_status pga_fill_uniform_p32_device(uint32_t *p, size_t n, uint64_t seed, double lo, double hi,
                                       void *stream);

/* FP64 backward error of a solution x of A x = b (device arrays, n x n, one right-hand
 * side): ||A x - b||_inf / (||A||_inf ||x||_inf) with A, x, b decoded exactly and FP64
 * sums; synchronous, result on the host.  Reported beside LU / Cholesky results as the
 * reference's check programs report residuals (main_lu_check.cpp:234-257). */
pga_status pga_backward_error_device(int n, const uint32_t *a, int lda, const uint32_t *x,
                                     const uint32_t *b, double *result, void *stream);

/* ---- elementwise scalar arithmetic on device arrays (validation of the core) ---- */
/* op: 0 add, 1 sub, 2 mul, 3 div, 4 sqrt (b ignored) */
pga_status pga_p32_elementwise_device(int op, const uint32_t *a, const uint32_t *b,
                                      uint32_t *out, size_t n, void *stream);

/* ---- host scalar arithmetic for the posit32 value type ------------------------
 * Scalar operations for the drop-in posit32 type (include/softposit_cpp.h), which the
 * reference takes from the SoftPosit library (operators of posit32, e.g. used at
 * Rgetrf2.cpp:122 and main_gemm.cpp:99-104).  Same bits as the device arithmetic.
 * Matrix routines never use these: they run on the GPU. */
uint32_t pga_p32_add(uint32_t a, uint32_t b);
uint32_t pga_p32_sub(uint32_t a, uint32_t b);
uint32_t pga_p32_mul(uint32_t a, uint32_t b);
uint32_t pga_p32_div(uint32_t a, uint32_t b);
uint32_t pga_p32_sqrt(uint32_t a);
uint32_t pga_p32_from_double(double x);
double pga_p32_to_double(uint32_t a);

#ifdef __cplusplus
}
#endif
#endif /* POSIT_GPU_AMD_H */
