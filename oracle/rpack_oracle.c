/*
 * rpack_oracle.c -- CPU oracle for the libRpack BLAS / LAPACK routines on posit32.
 *
 * TEST INFRASTRUCTURE ONLY (see posit32_oracle.h).
 *
 * Every routine restates the loop nest of the reference routine cited beside it, with
 * the same element order and the same rounding sequence: each posit32 operator of the
 * reference (`a + b`, `a - b`, `a * b`, `a / b`, `sqrt`, comparisons) becomes one call
 * to the scalar oracle.  Indices are 0-based here; the reference's are 1-based.
 * Illegal arguments are reported by return value instead of Mxerbla's exit().
 */
#include "posit32_oracle.h"

#include <ctype.h>
#include <stdlib.h>

#define ZERO ORC_P32_ZERO
#define ONE ORC_P32_ONE
#define MONE 0xC0000000u /* -1.0 */

#define ADD orc_p32_add
#define SUB orc_p32_sub
#define MUL orc_p32_mul
#define DIV orc_p32_div

static int same(char a, char b) { return toupper((unsigned char)a) == toupper((unsigned char)b); }
static int imax(int a, int b) { return a > b ? a : b; }
static int imin(int a, int b) { return a < b ? a : b; }
static int p_le(uint32_t a, uint32_t b) { return (int32_t)a <= (int32_t)b; }
static int p_gt(uint32_t a, uint32_t b) { return (int32_t)a > (int32_t)b; }

/* ------------------------------------------------------------------------- */
/* Rgemm.cpp:62-251                                                          */
/* ------------------------------------------------------------------------- */
int orc_rgemm(char transa, char transb, int m, int n, int k, uint32_t alpha, const uint32_t *a,
              int lda, const uint32_t *b, int ldb, uint32_t beta, uint32_t *c, int ldc, int nthreads)
{
    int nota = same(transa, 'N');
    int notb = same(transb, 'N');
    int nrowa = nota ? m : k;
    int nrowb = notb ? k : n;
    int info = 0;
    if (!nota && !same(transa, 'C') && !same(transa, 'T'))
        info = 1;
    else if (!notb && !same(transb, 'C') && !same(transb, 'T'))
        info = 2;
    else if (m < 0)
        info = 3;
    else if (n < 0)
        info = 4;
    else if (k < 0)
        info = 5;
    else if (lda < imax(1, nrowa))
        info = 8;
    else if (ldb < imax(1, nrowb))
        info = 10;
    else if (ldc < imax(1, m))
        info = 13;
    if (info)
        return info;

    if (m == 0 || n == 0 || ((alpha == ZERO || k == 0) && beta == ONE))
        return 0;
    if (nthreads < 1)
        nthreads = 1;

    if (alpha == ZERO) { /* Rgemm.cpp:142-157 */
        for (int j = 0; j < n; j++)
            for (int i = 0; i < m; i++)
                c[i + (size_t)j * ldc] = (beta == ZERO) ? ZERO : MUL(beta, c[i + (size_t)j * ldc]);
        return 0;
    }

    if (nota) {
        /* NN (Rgemm.cpp:163-185) and NT (:206-227): C = beta-op, then
         * c_ij += (alpha * b_lj) * a_il for ascending l */
#pragma omp parallel for num_threads(nthreads) schedule(static)
        for (int j = 0; j < n; j++) {
            uint32_t *cj = c + (size_t)j * ldc;
            if (beta == ZERO) {
                for (int i = 0; i < m; i++)
                    cj[i] = ZERO;
            } else if (beta != ONE) {
                for (int i = 0; i < m; i++)
                    cj[i] = MUL(beta, cj[i]);
            }
            for (int l = 0; l < k; l++) {
                uint32_t blj = notb ? b[l + (size_t)j * ldb] : b[j + (size_t)l * ldb];
                uint32_t temp = MUL(alpha, blj);
                const uint32_t *al = a + (size_t)l * lda;
                for (int i = 0; i < m; i++)
                    cj[i] = ADD(cj[i], MUL(temp, al[i]));
            }
        }
    } else {
        /* TN (Rgemm.cpp:186-204) and TT (:228-246): sum a*b from zero in ascending l,
         * then alpha*temp (+ beta*c) */
#pragma omp parallel for num_threads(nthreads) schedule(static)
        for (int j = 0; j < n; j++) {
            for (int i = 0; i < m; i++) {
                uint32_t temp = ZERO;
                const uint32_t *ai = a + (size_t)i * lda;
                for (int l = 0; l < k; l++) {
                    uint32_t blj = notb ? b[l + (size_t)j * ldb] : b[j + (size_t)l * ldb];
                    temp = ADD(temp, MUL(ai[l], blj));
                }
                uint32_t *cij = c + i + (size_t)j * ldc;
                if (beta == ZERO)
                    *cij = MUL(alpha, temp);
                else
                    *cij = ADD(MUL(alpha, temp), MUL(beta, *cij));
            }
        }
    }
    return 0;
}

/* ------------------------------------------------------------------------- */
/* Rtrsm.cpp:31-291                                                          */
/* ------------------------------------------------------------------------- */
int orc_rtrsm(char side, char uplo, char transa, char diag, int m, int n, uint32_t alpha,
              const uint32_t *a, int lda, uint32_t *b, int ldb)
{
    int lside = same(side, 'L');
    int nrowa = lside ? m : n;
    int nounit = same(diag, 'N');
    int upper = same(uplo, 'U');
    int info = 0;
    if (!lside && !same(side, 'R'))
        info = 1;
    else if (!upper && !same(uplo, 'L'))
        info = 2;
    else if (!same(transa, 'N') && !same(transa, 'T') && !same(transa, 'C'))
        info = 3;
    else if (!same(diag, 'U') && !same(diag, 'N'))
        info = 4;
    else if (m < 0)
        info = 5;
    else if (n < 0)
        info = 6;
    else if (lda < imax(1, nrowa))
        info = 9;
    else if (ldb < imax(1, m))
        info = 11;
    if (info)
        return info;
    if (m == 0 || n == 0)
        return 0;

#define A_(i, j) a[(i) + (size_t)(j) * lda]
#define B_(i, j) b[(i) + (size_t)(j) * ldb]
    if (alpha == ZERO) {
        for (int j = 0; j < n; j++)
            for (int i = 0; i < m; i++)
                B_(i, j) = ZERO;
        return 0;
    }
    if (lside) {
        if (same(transa, 'N')) {
            if (upper) { /* :120-137 */
                for (int j = 0; j < n; j++) {
                    if (alpha != ONE)
                        for (int i = 0; i < m; i++)
                            B_(i, j) = MUL(alpha, B_(i, j));
                    for (int kk = m - 1; kk >= 0; kk--) {
                        if (B_(kk, j) != ZERO) {
                            if (nounit)
                                B_(kk, j) = DIV(B_(kk, j), A_(kk, kk));
                            for (int i = 0; i < kk; i++)
                                B_(i, j) = SUB(B_(i, j), MUL(B_(kk, j), A_(i, kk)));
                        }
                    }
                }
            } else { /* :138-156 */
                for (int j = 0; j < n; j++) {
                    if (alpha != ONE)
                        for (int i = 0; i < m; i++)
                            B_(i, j) = MUL(alpha, B_(i, j));
                    for (int kk = 0; kk < m; kk++) {
                        if (B_(kk, j) != ZERO) {
                            if (nounit)
                                B_(kk, j) = DIV(B_(kk, j), A_(kk, kk));
                            for (int i = kk + 1; i < m; i++)
                                B_(i, j) = SUB(B_(i, j), MUL(B_(kk, j), A_(i, kk)));
                        }
                    }
                }
            }
        } else {
            if (upper) { /* :161-173 */
                for (int j = 0; j < n; j++) {
                    for (int i = 0; i < m; i++) {
                        uint32_t temp = MUL(alpha, B_(i, j));
                        for (int kk = 0; kk < i; kk++)
                            temp = SUB(temp, MUL(A_(kk, i), B_(kk, j)));
                        if (nounit)
                            temp = DIV(temp, A_(i, i));
                        B_(i, j) = temp;
                    }
                }
            } else { /* :174-186 */
                for (int j = 0; j < n; j++) {
                    for (int i = m - 1; i >= 0; i--) {
                        uint32_t temp = MUL(alpha, B_(i, j));
                        for (int kk = i + 1; kk < m; kk++)
                            temp = SUB(temp, MUL(A_(kk, i), B_(kk, j)));
                        if (nounit)
                            temp = DIV(temp, A_(i, i));
                        B_(i, j) = temp;
                    }
                }
            }
        }
    } else {
        if (same(transa, 'N')) {
            if (upper) { /* :194-214 */
                for (int j = 0; j < n; j++) {
                    if (alpha != ONE)
                        for (int i = 0; i < m; i++)
                            B_(i, j) = MUL(alpha, B_(i, j));
                    for (int kk = 0; kk < j; kk++) {
                        if (A_(kk, j) != ZERO)
                            for (int i = 0; i < m; i++)
                                B_(i, j) = SUB(B_(i, j), MUL(A_(kk, j), B_(i, kk)));
                    }
                    if (nounit) {
                        uint32_t temp = DIV(ONE, A_(j, j));
                        for (int i = 0; i < m; i++)
                            B_(i, j) = MUL(temp, B_(i, j));
                    }
                }
            } else { /* :215-236 */
                for (int j = n - 1; j >= 0; j--) {
                    if (alpha != ONE)
                        for (int i = 0; i < m; i++)
                            B_(i, j) = MUL(alpha, B_(i, j));
                    for (int kk = j + 1; kk < n; kk++) {
                        if (A_(kk, j) != ZERO)
                            for (int i = 0; i < m; i++)
                                B_(i, j) = SUB(B_(i, j), MUL(A_(kk, j), B_(i, kk)));
                    }
                    if (nounit) {
                        uint32_t temp = DIV(ONE, A_(j, j));
                        for (int i = 0; i < m; i++)
                            B_(i, j) = MUL(temp, B_(i, j));
                    }
                }
            }
        } else {
            if (upper) { /* :241-262 */
                for (int kk = n - 1; kk >= 0; kk--) {
                    if (nounit) {
                        uint32_t temp = DIV(ONE, A_(kk, kk));
                        for (int i = 0; i < m; i++)
                            B_(i, kk) = MUL(temp, B_(i, kk));
                    }
                    for (int j = 0; j < kk; j++) {
                        if (A_(j, kk) != ZERO) {
                            uint32_t temp = A_(j, kk);
                            for (int i = 0; i < m; i++)
                                B_(i, j) = SUB(B_(i, j), MUL(temp, B_(i, kk)));
                        }
                    }
                    if (alpha != ONE)
                        for (int i = 0; i < m; i++)
                            B_(i, kk) = MUL(alpha, B_(i, kk));
                }
            } else { /* :263-285 */
                for (int kk = 0; kk < n; kk++) {
                    if (nounit) {
                        uint32_t temp = DIV(ONE, A_(kk, kk));
                        for (int i = 0; i < m; i++)
                            B_(i, kk) = MUL(temp, B_(i, kk));
                    }
                    for (int j = kk + 1; j < n; j++) {
                        if (A_(j, kk) != ZERO) {
                            uint32_t temp = A_(j, kk);
                            for (int i = 0; i < m; i++)
                                B_(i, j) = SUB(B_(i, j), MUL(temp, B_(i, kk)));
                        }
                    }
                    if (alpha != ONE)
                        for (int i = 0; i < m; i++)
                            B_(i, kk) = MUL(alpha, B_(i, kk));
                }
            }
        }
    }
#undef A_
#undef B_
    return 0;
}

/* ------------------------------------------------------------------------- */
/* Rsyrk.cpp:31-215                                                          */
/* ------------------------------------------------------------------------- */
int orc_rsyrk(char uplo, char trans, int n, int k, uint32_t alpha, const uint32_t *a, int lda,
              uint32_t beta, uint32_t *c, int ldc)
{
    int nrowa = same(trans, 'N') ? n : k;
    int upper = same(uplo, 'U');
    int info = 0;
    if (!upper && !same(uplo, 'L'))
        info = 1;
    else if (!same(trans, 'N') && !same(trans, 'T') && !same(trans, 'C'))
        info = 2;
    else if (n < 0)
        info = 3;
    else if (k < 0)
        info = 4;
    else if (lda < imax(1, nrowa))
        info = 7;
    else if (ldc < imax(1, n))
        info = 10;
    if (info)
        return info;
    if (n == 0 || ((alpha == ZERO || k == 0) && beta == ONE))
        return 0;
#define A_(i, j) a[(i) + (size_t)(j) * lda]
#define C_(i, j) c[(i) + (size_t)(j) * ldc]
    if (alpha == ZERO) { /* :96-127 */
        for (int j = 0; j < n; j++) {
            int i0 = upper ? 0 : j, i1 = upper ? j + 1 : n;
            for (int i = i0; i < i1; i++)
                C_(i, j) = (beta == ZERO) ? ZERO : MUL(beta, C_(i, j));
        }
        return 0;
    }
    if (same(trans, 'N')) { /* :137-177 */
        for (int j = 0; j < n; j++) {
            int i0 = upper ? 0 : j, i1 = upper ? j + 1 : n;
            if (beta == ZERO) {
                for (int i = i0; i < i1; i++)
                    C_(i, j) = ZERO;
            } else if (beta != ONE) {
                for (int i = i0; i < i1; i++)
                    C_(i, j) = MUL(beta, C_(i, j));
            }
            for (int l = 0; l < k; l++) {
                if (A_(j, l) != ZERO) {
                    uint32_t temp = MUL(alpha, A_(j, l));
                    for (int i = i0; i < i1; i++)
                        C_(i, j) = ADD(C_(i, j), MUL(temp, A_(i, l)));
                }
            }
        }
    } else { /* :182-210 */
        for (int j = 0; j < n; j++) {
            int i0 = upper ? 0 : j, i1 = upper ? j + 1 : n;
            for (int i = i0; i < i1; i++) {
                uint32_t temp = ZERO;
                for (int l = 0; l < k; l++)
                    temp = ADD(temp, MUL(A_(l, i), A_(l, j)));
                if (beta == ZERO)
                    C_(i, j) = MUL(alpha, temp);
                else
                    C_(i, j) = ADD(MUL(alpha, temp), MUL(beta, C_(i, j)));
            }
        }
    }
#undef A_
#undef C_
    return 0;
}

/* ------------------------------------------------------------------------- */
/* Rgemv.cpp:31-204                                                          */
/* ------------------------------------------------------------------------- */
int orc_rgemv(char trans, int m, int n, uint32_t alpha, const uint32_t *a, int lda,
              const uint32_t *x, int incx, uint32_t beta, uint32_t *y, int incy)
{
    int info = 0;
    if (!same(trans, 'N') && !same(trans, 'T') && !same(trans, 'C'))
        info = 1;
    else if (m < 0)
        info = 2;
    else if (n < 0)
        info = 3;
    else if (lda < imax(1, m))
        info = 6;
    else if (incx == 0)
        info = 8;
    else if (incy == 0)
        info = 11;
    if (info)
        return info;
    if (m == 0 || n == 0 || (alpha == ZERO && beta == ONE))
        return 0;
    int notr = same(trans, 'N');
    int lenx = notr ? n : m;
    int leny = notr ? m : n;
    int kx = (incx > 0) ? 0 : -(lenx - 1) * incx;
    int ky = (incy > 0) ? 0 : -(leny - 1) * incy;
    if (beta != ONE) {
        int iy = ky;
        for (int i = 0; i < leny; i++, iy += incy)
            y[iy] = (beta == ZERO) ? ZERO : MUL(beta, y[iy]);
    }
    if (alpha == ZERO)
        return 0;
    if (notr) {
        int jx = kx;
        for (int j = 0; j < n; j++, jx += incx) {
            uint32_t temp = MUL(alpha, x[jx]);
            int iy = ky;
            for (int i = 0; i < m; i++, iy += incy)
                y[iy] = ADD(y[iy], MUL(temp, a[i + (size_t)j * lda]));
        }
    } else {
        int jy = ky;
        for (int j = 0; j < n; j++, jy += incy) {
            uint32_t temp = ZERO;
            int ix = kx;
            for (int i = 0; i < m; i++, ix += incx)
                temp = ADD(temp, MUL(a[i + (size_t)j * lda], x[ix]));
            y[jy] = ADD(y[jy], MUL(alpha, temp));
        }
    }
    return 0;
}

/* Rscal.cpp:62-119 (the 5-way unrolling does not change any element's result) */
void orc_rscal(int n, uint32_t da, uint32_t *dx, int incx)
{
    if (n <= 0 || incx <= 0)
        return;
    for (int i = 0; i < n; i++)
        dx[(size_t)i * incx] = MUL(da, dx[(size_t)i * incx]);
}

/* iRamax.cpp:31-88 with abs from util.cpp:12-17; strict '>' keeps the first maximum */
int orc_iramax(int n, const uint32_t *dx, int incx)
{
    if (n < 1 || incx <= 0)
        return 0;
    if (n == 1)
        return 1;
    int ret = 1;
    uint32_t dmax = orc_p32_abs(dx[0]);
    for (int i = 1; i < n; i++) {
        uint32_t v = orc_p32_abs(dx[(size_t)i * incx]);
        if (p_gt(v, dmax)) {
            ret = i + 1;
            dmax = v;
        }
    }
    return ret;
}

/* Rlaswp.cpp:32-111.  The reference walks columns in blocks of 32; row interchanges in
 * different columns are independent, so walking column by column gives the same array. */
void orc_rlaswp(int n, uint32_t *a, int lda, int k1, int k2, const int *ipiv, int incx)
{
    int ix0, i1, i2, inc;
    if (incx > 0) {
        ix0 = k1;
        i1 = k1;
        i2 = k2;
        inc = 1;
    } else if (incx < 0) {
        ix0 = k1 + (k1 - k2) * incx;
        i1 = k2;
        i2 = k1;
        inc = -1;
    } else {
        return;
    }
    for (int col = 0; col < n; col++) {
        uint32_t *ac = a + (size_t)col * lda;
        int ix = ix0;
        for (int i = i1; inc > 0 ? i <= i2 : i >= i2; i += inc) {
            int ip = ipiv[ix - 1];
            if (ip != i) {
                uint32_t t = ac[i - 1];
                ac[i - 1] = ac[ip - 1];
                ac[ip - 1] = t;
            }
            ix += incx;
        }
    }
}

/* ------------------------------------------------------------------------- */
/* Rgetrf2.cpp:32-185 (recursive panel LU)                                    */
/* ------------------------------------------------------------------------- */
int orc_rgetrf2(int m, int n, uint32_t *a, int lda, int *ipiv, int *info, int nthreads)
{
    *info = 0;
    if (m < 0)
        *info = -1;
    else if (n < 0)
        *info = -2;
    else if (lda < imax(1, m))
        *info = -4;
    if (*info)
        return *info;
    if (m == 0 || n == 0)
        return 0;
    if (m == 1) { /* :86-95 */
        ipiv[0] = 1;
        if (a[0] == ZERO)
            *info = 1;
    } else if (n == 1) { /* :96-132 */
        const uint32_t sfmin = orc_p32_from_double(1.17549435082228750797e-38); /* FLT_MIN */
        int i = orc_iramax(m, a, 1);
        ipiv[0] = i;
        if (a[i - 1] != ZERO) {
            if (i != 1) {
                uint32_t t = a[0];
                a[0] = a[i - 1];
                a[i - 1] = t;
            }
            if (!p_gt(sfmin, orc_p32_abs(a[0]))) { /* abs(a[0]) >= sfmin */
                orc_rscal(m - 1, DIV(ONE, a[0]), a + 1, 1);
            } else {
                for (int r = 1; r < m; r++)
                    a[r] = DIV(a[r], a[0]);
            }
        } else {
            *info = 1;
        }
    } else { /* :133-181 */
        int n1 = imin(m, n) / 2;
        int n2 = n - n1;
        int iinfo = 0;
        orc_rgetrf2(m, n1, a, lda, ipiv, &iinfo, nthreads);
        if (*info == 0 && iinfo > 0)
            *info = iinfo;
        orc_rlaswp(n2, a + (size_t)n1 * lda, lda, 1, n1, ipiv, 1);
        orc_rtrsm('L', 'L', 'N', 'U', n1, n2, ONE, a, lda, a + (size_t)n1 * lda, lda);
        orc_rgemm('N', 'N', m - n1, n2, n1, MONE, a + n1, lda, a + (size_t)n1 * lda, lda, ONE,
                  a + n1 + (size_t)n1 * lda, lda, nthreads);
        orc_rgetrf2(m - n1, n2, a + n1 + (size_t)n1 * lda, lda, ipiv + n1, &iinfo, nthreads);
        if (*info == 0 && iinfo > 0)
            *info = iinfo + n1;
        for (int i = n1; i < imin(m, n); i++)
            ipiv[i] += n1;
        orc_rlaswp(n1, a, lda, n1 + 1, imin(m, n), ipiv, 1);
    }
    return 0;
}

/* Rgetrf_opencl.cpp:32-154 / Rgetrf.cpp:32-137: blocked right-looking LU.  RgemmG
 * computes with Rgemm's per-element order, so both files compute the same bits. */
int orc_rgetrf(int m, int n, uint32_t *a, int lda, int *ipiv, int *info, int nb, int nthreads)
{
    *info = 0;
    if (m < 0)
        *info = -1;
    else if (n < 0)
        *info = -2;
    else if (lda < imax(1, m))
        *info = -4;
    if (*info)
        return *info;
    if (m == 0 || n == 0)
        return 0;
    if (nb <= 0)
        nb = 64; /* iMlaenv.cpp:145-151 */
    int mn = imin(m, n);
    if (nb <= 1 || nb >= mn)
        return orc_rgetrf2(m, n, a, lda, ipiv, info, nthreads);
    for (int j = 1; j <= mn; j += nb) {
        int jb = imin(mn - j + 1, nb);
        int iinfo = 0;
        uint32_t *ajj = a + (j - 1) + (size_t)(j - 1) * lda;
        orc_rgetrf2(m - j + 1, jb, ajj, lda, ipiv + j - 1, &iinfo, nthreads);
        if (*info == 0 && iinfo > 0)
            *info = iinfo + j - 1;
        for (int i = j; i <= imin(m, j + jb - 1); i++)
            ipiv[i - 1] += j - 1;
        orc_rlaswp(j - 1, a, lda, j, j + jb - 1, ipiv, 1);
        if (j + jb <= n) {
            orc_rlaswp(n - j - jb + 1, a + (size_t)(j + jb - 1) * lda, lda, j, j + jb - 1, ipiv, 1);
            orc_rtrsm('L', 'L', 'N', 'U', jb, n - j - jb + 1, ONE, ajj, lda,
                      a + (j - 1) + (size_t)(j + jb - 1) * lda, lda);
            if (j + jb <= m) {
                orc_rgemm('N', 'N', m - j - jb + 1, n - j - jb + 1, jb, MONE,
                          a + (j + jb - 1) + (size_t)(j - 1) * lda, lda,
                          a + (j - 1) + (size_t)(j + jb - 1) * lda, lda, ONE,
                          a + (j + jb - 1) + (size_t)(j + jb - 1) * lda, lda, nthreads);
            }
        }
    }
    return 0;
}

/* Rgetrs.cpp:32-118 */
int orc_rgetrs(char trans, int n, int nrhs, const uint32_t *a, int lda, const int *ipiv,
               uint32_t *b, int ldb, int *info)
{
    int notran = same(trans, 'N');
    *info = 0;
    if (!notran && !same(trans, 'T') && !same(trans, 'C'))
        *info = -1;
    else if (n < 0)
        *info = -2;
    else if (nrhs < 0)
        *info = -3;
    else if (lda < imax(1, n))
        *info = -5;
    else if (ldb < imax(1, n))
        *info = -8;
    if (*info)
        return *info;
    if (n == 0 || nrhs == 0)
        return 0;
    if (notran) {
        orc_rlaswp(nrhs, b, ldb, 1, n, ipiv, 1);
        orc_rtrsm('L', 'L', 'N', 'U', n, nrhs, ONE, a, lda, b, ldb);
        orc_rtrsm('L', 'U', 'N', 'N', n, nrhs, ONE, a, lda, b, ldb);
    } else {
        orc_rtrsm('L', 'U', 'T', 'N', n, nrhs, ONE, a, lda, b, ldb);
        orc_rtrsm('L', 'L', 'T', 'U', n, nrhs, ONE, a, lda, b, ldb);
        orc_rlaswp(nrhs, b, ldb, 1, n, ipiv, -1);
    }
    return 0;
}

/* ------------------------------------------------------------------------- */
/* Rpotrf2.cpp:32-151                                                        */
/* ------------------------------------------------------------------------- */
int orc_rpotrf2(char uplo, int n, uint32_t *a, int lda, int *info)
{
    *info = 0;
    int upper = same(uplo, 'U');
    if (!upper && !same(uplo, 'L'))
        *info = -1;
    else if (n < 0)
        *info = -2;
    else if (lda < imax(1, n))
        *info = -4;
    if (*info)
        return *info;
    if (n == 0)
        return 0;
    if (n == 1) {
        /* `a <= zero || Risnan(a)`; Risnan is always false (Risnan.cpp:32-37) */
        if (p_le(a[0], ZERO)) {
            *info = 1;
            return 0;
        }
        a[0] = orc_p32_sqrt(a[0]);
        return 0;
    }
    int n1 = n / 2, n2 = n - n1, iinfo = 0;
    orc_rpotrf2(uplo, n1, a, lda, &iinfo);
    if (iinfo) {
        *info = iinfo;
        return 0;
    }
    uint32_t *a22 = a + n1 + (size_t)n1 * lda;
    if (upper) {
        orc_rtrsm('L', 'U', 'T', 'N', n1, n2, ONE, a, lda, a + (size_t)n1 * lda, lda);
        orc_rsyrk(uplo, 'T', n2, n1, MONE, a + (size_t)n1 * lda, lda, ONE, a22, lda);
    } else {
        orc_rtrsm('R', 'L', 'T', 'N', n2, n1, ONE, a, lda, a + n1, lda);
        orc_rsyrk(uplo, 'N', n2, n1, MONE, a + n1, lda, ONE, a22, lda);
    }
    orc_rpotrf2(uplo, n2, a22, lda, &iinfo);
    if (iinfo)
        *info = iinfo + n1;
    return 0;
}

/* Rpotrf_opencl.cpp:32-166 / Rpotrf.cpp: blocked left-looking Cholesky, nb = 64 */
int orc_rpotrf(char uplo, int n, uint32_t *a, int lda, int *info, int nb, int nthreads)
{
    *info = 0;
    int upper = same(uplo, 'U');
    if (!upper && !same(uplo, 'L'))
        *info = -1;
    else if (n < 0)
        *info = -2;
    else if (lda < imax(1, n))
        *info = -4;
    if (*info)
        return *info;
    if (n == 0)
        return 0;
    if (nb <= 0)
        nb = 64; /* iMlaenv.cpp:223-229 */
    if (nb <= 1 || nb >= n)
        return orc_rpotrf2(uplo, n, a, lda, info);
#define A_(i, j) (a + (i) + (size_t)(j) * lda)
    for (int j = 1; j <= n; j += nb) {
        int jb = imin(nb, n - j + 1);
        if (upper) {
            orc_rsyrk('U', 'T', jb, j - 1, MONE, A_(0, j - 1), lda, ONE, A_(j - 1, j - 1), lda);
            orc_rpotrf2('U', jb, A_(j - 1, j - 1), lda, info);
            if (*info) {
                *info += j - 1;
                return 0;
            }
            if (j + jb <= n) {
                orc_rgemm('T', 'N', jb, n - j - jb + 1, j - 1, MONE, A_(0, j - 1), lda,
                          A_(0, j + jb - 1), lda, ONE, A_(j - 1, j + jb - 1), lda, nthreads);
                orc_rtrsm('L', 'U', 'T', 'N', jb, n - j - jb + 1, ONE, A_(j - 1, j - 1), lda,
                          A_(j - 1, j + jb - 1), lda);
            }
        } else {
            orc_rsyrk('L', 'N', jb, j - 1, MONE, A_(j - 1, 0), lda, ONE, A_(j - 1, j - 1), lda);
            orc_rpotrf2('L', jb, A_(j - 1, j - 1), lda, info);
            if (*info) {
                *info += j - 1;
                return 0;
            }
            if (j + jb <= n) {
                orc_rgemm('N', 'T', n - j - jb + 1, jb, j - 1, MONE, A_(j + jb - 1, 0), lda,
                          A_(j - 1, 0), lda, ONE, A_(j + jb - 1, j - 1), lda, nthreads);
                orc_rtrsm('R', 'L', 'T', 'N', n - j - jb + 1, jb, ONE, A_(j - 1, j - 1), lda,
                          A_(j + jb - 1, j - 1), lda);
            }
        }
    }
#undef A_
    return 0;
}

/* Rpotrs.cpp:32-110 */
int orc_rpotrs(char uplo, int n, int nrhs, const uint32_t *a, int lda, uint32_t *b, int ldb,
               int *info)
{
    *info = 0;
    int upper = same(uplo, 'U');
    if (!upper && !same(uplo, 'L'))
        *info = -1;
    else if (n < 0)
        *info = -2;
    else if (nrhs < 0)
        *info = -3;
    else if (lda < imax(1, n))
        *info = -5;
    else if (ldb < imax(1, n))
        *info = -7;
    if (*info)
        return *info;
    if (n == 0 || nrhs == 0)
        return 0;
    if (upper) {
        orc_rtrsm('L', 'U', 'T', 'N', n, nrhs, ONE, a, lda, b, ldb);
        orc_rtrsm('L', 'U', 'N', 'N', n, nrhs, ONE, a, lda, b, ldb);
    } else {
        orc_rtrsm('L', 'L', 'N', 'N', n, nrhs, ONE, a, lda, b, ldb);
        orc_rtrsm('L', 'L', 'T', 'N', n, nrhs, ONE, a, lda, b, ldb);
    }
    return 0;
}
