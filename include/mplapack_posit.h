/*
 * mplapack_posit.h -- posit32 LU and Cholesky of the libRpack drop-in (C++ linkage).
 *
 * Same prototypes as the reference's libRpack/mplapack_posit.h:38-57.  The
 * factorizations and solves run on the GPU (include/posit_gpu_amd.h) with the
 * reference's blocking, recursion and per-element rounding order, so factors, pivots
 * and info codes are bit-identical to the reference's.
 */
#ifndef PGA_MPLAPACK_POSIT_H
#define PGA_MPLAPACK_POSIT_H

#include "mpblas_posit.h"

typedef mplapackint INTEGER;
typedef posit32 REAL;

void Rpotrf(const char *uplo, INTEGER const n, REAL *a, INTEGER const lda, INTEGER &info);
void Rpotrf2(const char *uplo, INTEGER const n, REAL *a, INTEGER const lda, INTEGER &info);
bool Risnan(REAL const din);
void Rpotrs(const char *uplo, INTEGER const n, INTEGER const nrhs, REAL *a, INTEGER const lda,
            REAL *b, INTEGER const ldb, INTEGER &info);

void Rpotrf_opencl(const char *uplo, INTEGER const n, REAL *a, INTEGER const lda, INTEGER &info);
void Rpotrf2_opencl(const char *uplo, INTEGER const n, REAL *a, INTEGER const lda, INTEGER &info);

void Rgetrs(const char *trans, INTEGER const n, INTEGER const nrhs, REAL *a, INTEGER const lda,
            INTEGER *ipiv, REAL *b, INTEGER const ldb, INTEGER &info);

void Rgetrf(INTEGER const m, INTEGER const n, REAL *a, INTEGER const lda, INTEGER *ipiv,
            INTEGER &info);
void Rgetrf_opencl(INTEGER const m, INTEGER const n, REAL *a, INTEGER const lda, INTEGER *ipiv,
                   INTEGER &info, INTEGER);

void Rgetrf2(INTEGER const m, INTEGER const n, REAL *a, INTEGER const lda, INTEGER *ipiv,
             INTEGER &info);
void Rgetrf2_opencl(INTEGER const m, INTEGER const n, REAL *a, INTEGER const lda, INTEGER *ipiv,
                    INTEGER &info);

void Rlaswp(INTEGER const n, REAL *a, INTEGER const lda, INTEGER const k1, INTEGER const k2,
            INTEGER *ipiv, INTEGER const incx);
INTEGER iMlaenv(INTEGER const ispec, const char *name, const char *opts, INTEGER const n1,
                INTEGER const n2, INTEGER const n3, INTEGER const n4);

#endif
