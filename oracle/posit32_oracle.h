/*
 * posit32_oracle.h -- CPU oracle for the posit32 (es=2) dense linear algebra path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links, loads or calls
 * this code; only tests/, the smoke() check and the CPU-baseline leg of bench.py
 * use it, and only as the checker.
 *
 * The scalar routines restate the SoftPosit posit32 algorithms that the reference
 * carries in libRpack/kernel_entry.cl (add: 88-397, mul: 399-538).  The BLAS and
 * LAPACK routines restate the reference's libRpack C++ loops (file:line cited at
 * each function) with every posit32 operator replaced by the scalar routines here.
 *
 * All posit32 values are passed as raw 32-bit patterns (uint32_t).
 * Matrices are column-major with leading dimensions, as in the reference.
 */
#ifndef POSIT32_ORACLE_H
#define POSIT32_ORACLE_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_P32_ZERO 0x00000000u
#define ORC_P32_ONE 0x40000000u
#define ORC_P32_NAR 0x80000000u

/* ---- scalar posit32 arithmetic ---- */
uint32_t orc_p32_add(uint32_t a, uint32_t b);
uint32_t orc_p32_sub(uint32_t a, uint32_t b);
uint32_t orc_p32_mul(uint32_t a, uint32_t b);
uint32_t orc_p32_div(uint32_t a, uint32_t b);
uint32_t orc_p32_sqrt(uint32_t a);
uint32_t orc_p32_neg(uint32_t a);
uint32_t orc_p32_abs(uint32_t a);
int orc_p32_lt(uint32_t a, uint32_t b);
uint32_t orc_p32_from_double(double x);
double orc_p32_to_double(uint32_t a);

/* elementwise bulk forms (for tests that push many vectors through) */
void orc_p32_add_v(const uint32_t *a, const uint32_t *b, uint32_t *out, size_t n);
void orc_p32_mul_v(const uint32_t *a, const uint32_t *b, uint32_t *out, size_t n);
void orc_p32_sub_v(const uint32_t *a, const uint32_t *b, uint32_t *out, size_t n);
void orc_p32_div_v(const uint32_t *a, const uint32_t *b, uint32_t *out, size_t n);
void orc_p32_sqrt_v(const uint32_t *a, uint32_t *out, size_t n);
void orc_p32_from_double_v(const double *x, uint32_t *out, size_t n);
void orc_p32_to_double_v(const uint32_t *a, double *out, size_t n);

/* ---- BLAS (libRpack/Rgemm.cpp, Rtrsm.cpp, Rsyrk.cpp, Rgemv.cpp, Rscal.cpp, iRamax.cpp) ----
 * Return 0, or the Mxerbla parameter number (>0) the reference would report. */
int orc_rgemm(char transa, char transb, int m, int n, int k, uint32_t alpha,
              const uint32_t *a, int lda, const uint32_t *b, int ldb, uint32_t beta,
              uint32_t *c, int ldc, int nthreads);
int orc_rtrsm(char side, char uplo, char transa, char diag, int m, int n, uint32_t alpha,
              const uint32_t *a, int lda, uint32_t *b, int ldb);
int orc_rsyrk(char uplo, char trans, int n, int k, uint32_t alpha, const uint32_t *a, int lda,
              uint32_t beta, uint32_t *c, int ldc);
int orc_rgemv(char trans, int m, int n, uint32_t alpha, const uint32_t *a, int lda,
              const uint32_t *x, int incx, uint32_t beta, uint32_t *y, int incy);
void orc_rscal(int n, uint32_t da, uint32_t *dx, int incx);
int orc_iramax(int n, const uint32_t *dx, int incx);
void orc_rlaswp(int n, uint32_t *a, int lda, int k1, int k2, const int *ipiv, int incx);

/* ---- LAPACK (Rgetrf*.cpp, Rgetrs.cpp, Rpotrf*.cpp, Rpotrs.cpp) ----
 * Return value: 0, or minus the illegal-argument index (Mxerbla convention). *info gets
 * the LAPACK info code.  nb <= 0 selects the iMlaenv block size (64). */
int orc_rgetrf2(int m, int n, uint32_t *a, int lda, int *ipiv, int *info, int nthreads);
int orc_rgetrf(int m, int n, uint32_t *a, int lda, int *ipiv, int *info, int nb, int nthreads);
int orc_rgetrs(char trans, int n, int nrhs, const uint32_t *a, int lda, const int *ipiv,
               uint32_t *b, int ldb, int *info);
int orc_rpotrf2(char uplo, int n, uint32_t *a, int lda, int *info);
int orc_rpotrf(char uplo, int n, uint32_t *a, int lda, int *info, int nb, int nthreads);
int orc_rpotrs(char uplo, int n, int nrhs, const uint32_t *a, int lda, uint32_t *b, int ldb,
               int *info);

#ifdef __cplusplus
}
#endif
#endif
