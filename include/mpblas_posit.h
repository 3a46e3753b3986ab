/*
 * mpblas_posit.h -- posit32 BLAS of the libRpack drop-in (C++ linkage).
 *
 * Same prototypes as the reference's libRpack/mpblas_posit.h:37-55 (plus initOpenCL,
 * libRpack/def.h:63).  Every routine runs on the GPU through the C ABI of
 * include/posit_gpu_amd.h; host arrays in, host arrays out, synchronous.  Illegal
 * arguments call Mxerbla, which prints and exits as the reference's does
 * (libRpack/Mxerbla.cpp:70-73).
 */
#ifndef PGA_MPBLAS_POSIT_H
#define PGA_MPBLAS_POSIT_H

#include "softposit_cpp.h"

typedef int mplapackint;

void Rgemm(const char *transa, const char *transb, mplapackint const m, mplapackint const n,
           mplapackint const k, posit32 const alpha, posit32 *a, mplapackint const lda, posit32 *b,
           mplapackint const ldb, posit32 const beta, posit32 *c, mplapackint const ldc);
void Rtrsm(const char *side, const char *uplo, const char *transa, const char *diag,
           mplapackint const m, mplapackint const n, posit32 const alpha, posit32 *a,
           mplapackint const lda, posit32 *b, mplapackint const ldb);
void Rsyrk(const char *uplo, const char *trans, mplapackint const n, mplapackint const k,
           posit32 const alpha, posit32 *a, mplapackint const lda, posit32 const beta, posit32 *c,
           mplapackint const ldc);
void Rgemv(const char *trans, mplapackint const m, mplapackint const n, posit32 const alpha,
           posit32 *a, mplapackint const lda, posit32 *x, mplapackint const incx,
           posit32 const beta, posit32 *y, mplapackint const incy);

void Mxerbla(const char *srname, int info);
void Rscal(mplapackint const n, posit32 const da, posit32 *dx, mplapackint const incx);
mplapackint iRamax(mplapackint const n, posit32 *dx, mplapackint const incx);
bool Mlsame(const char *a, const char *b);

posit32 abs(posit32 x);

void RgemmG(const char *transa, const char *transb, mplapackint const m, mplapackint const n,
            mplapackint const k, posit32 const alpha, posit32 *A, mplapackint const lda,
            posit32 *B, mplapackint const ldb, posit32 const beta, posit32 *C,
            mplapackint const ldc);

void initOpenCL(void);

#endif
