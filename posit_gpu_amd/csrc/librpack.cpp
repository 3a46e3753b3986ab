// librpack.cpp -- the libRpack C++ API (include/mpblas_posit.h, include/mplapack_posit.h)
// implemented on the C ABI.  Host arrays in and out; all matrix work runs on the GPU.
#include <cctype>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "../../include/mplapack_posit.h"
#include "../../include/posit_gpu_amd.h"

static_assert(sizeof(posit32) == 4, "posit32 must be one 32-bit pattern");

static uint32_t* P(posit32* p) { return reinterpret_cast<uint32_t*>(p); }
static const uint32_t* CP(const posit32* p) { return reinterpret_cast<const uint32_t*>(p); }

// Mlsame.cpp:109-113
bool Mlsame(const char* a, const char* b) { return toupper((unsigned char)*a) == toupper((unsigned char)*b); }

// Mxerbla.cpp:70-73: report the illegal parameter and exit with its number
void Mxerbla(const char* srname, int info) {
  fprintf(stderr, " ** On entry to %s parameter number %2d had an illegal value\n", srname, info);
  exit(info);
}

// Risnan.cpp:32-37 always answers false
bool Risnan(REAL const) { return false; }

// util.cpp:12-17
posit32 abs(posit32 x) { return ((int32_t)x.value < 0) ? posit32::from_bits(0u - x.value) : x; }

// iMlaenv.cpp: block size query (ispec 1) for the routines of this library; LU and
// Cholesky get 64 (iMlaenv.cpp:145-151, 223-229).
INTEGER iMlaenv(INTEGER const ispec, const char* name, const char*, INTEGER const, INTEGER const,
                INTEGER const, INTEGER const) {
  if (ispec != 1 || name == nullptr || strlen(name) < 6) return 1;
  char c2[3] = {(char)toupper((unsigned char)name[1]), (char)toupper((unsigned char)name[2]), 0};
  char c3[4] = {(char)toupper((unsigned char)name[3]), (char)toupper((unsigned char)name[4]),
                (char)toupper((unsigned char)name[5]), 0};
  if ((!strcmp(c2, "GE") || !strcmp(c2, "PO")) && !strcmp(c3, "TRF")) return 64;
  return 32;
}

static void fail(const char* name, pga_status st) {
  if (st == PGA_SUCCESS) return;
  if (st == PGA_ERROR_ILLEGAL_ARGUMENT) Mxerbla(name, pga_illegal_argument());
  fprintf(stderr, " ** %s: %s\n", name, pga_status_string(st));
  exit(1);
}

// initOpenCL (Rgemm_opencl.cpp:59-70): one-time device set-up.  OPENCL_DEVICE selects the
// device when PGA_DEVICE is not set, as the reference's environment did.
void initOpenCL(void) {
  int dev = -1;
  if (!getenv("PGA_DEVICE") && getenv("OPENCL_DEVICE")) dev = atoi(getenv("OPENCL_DEVICE"));
  fail("initOpenCL", pga_init(dev));
}

// ---------------------------------------------------------------------------------
// BLAS
// ---------------------------------------------------------------------------------
void Rgemm(const char* transa, const char* transb, mplapackint const m, mplapackint const n, mplapackint const k,
           posit32 const alpha, posit32* a, mplapackint const lda, posit32* b, mplapackint const ldb,
           posit32 const beta, posit32* c, mplapackint const ldc) {
  fail("Rgemm ", pga_rgemm(transa[0], transb[0], m, n, k, alpha.value, CP(a), lda, CP(b), ldb, beta.value, P(c),
                           ldc));
}

// RgemmG (Rgemm_opencl.cpp:199-250): below 32 rows or columns the reference calls the
// CPU Rgemm (which checks its arguments and reads 'C' as transposed); otherwise only
// 'T'/'t' means transposed.  Both run here on the GPU with Rgemm's bits.
void RgemmG(const char* transa, const char* transb, mplapackint const m, mplapackint const n, mplapackint const k,
            posit32 const alpha, posit32* A, mplapackint const lda, posit32* B, mplapackint const ldb,
            posit32 const beta, posit32* C, mplapackint const ldc) {
  if (m == 0 || n == 0 || ((alpha.value == 0 || k == 0) && beta.value == 0x40000000u)) return;
  if (m < 32 || n < 32) {
    Rgemm(transa, transb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
    return;
  }
  char ta = (transa[0] == 'T' || transa[0] == 't') ? 'T' : 'N';
  char tb = (transb[0] == 'T' || transb[0] == 't') ? 'T' : 'N';
  fail("RgemmG", pga_rgemm(ta, tb, m, n, k, alpha.value, CP(A), lda, CP(B), ldb, beta.value, P(C), ldc));
}

void Rtrsm(const char* side, const char* uplo, const char* transa, const char* diag, mplapackint const m,
           mplapackint const n, posit32 const alpha, posit32* a, mplapackint const lda, posit32* b,
           mplapackint const ldb) {
  fail("Rtrsm ", pga_rtrsm(side[0], uplo[0], transa[0], diag[0], m, n, alpha.value, CP(a), lda, P(b), ldb));
}

void Rsyrk(const char* uplo, const char* trans, mplapackint const n, mplapackint const k, posit32 const alpha,
           posit32* a, mplapackint const lda, posit32 const beta, posit32* c, mplapackint const ldc) {
  fail("Rsyrk ", pga_rsyrk(uplo[0], trans[0], n, k, alpha.value, CP(a), lda, beta.value, P(c), ldc));
}

void Rgemv(const char* trans, mplapackint const m, mplapackint const n, posit32 const alpha, posit32* a,
           mplapackint const lda, posit32* x, mplapackint const incx, posit32 const beta, posit32* y,
           mplapackint const incy) {
  fail("Rgemv ", pga_rgemv(trans[0], m, n, alpha.value, CP(a), lda, CP(x), incx, beta.value, P(y), incy));
}

void Rscal(mplapackint const n, posit32 const da, posit32* dx, mplapackint const incx) {
  fail("Rscal", pga_rscal(n, da.value, P(dx), incx));
}

mplapackint iRamax(mplapackint const n, posit32* dx, mplapackint const incx) {
  int r = 0;
  fail("iRamax", pga_iramax(n, CP(dx), incx, &r));
  return r;
}

void Rlaswp(INTEGER const n, REAL* a, INTEGER const lda, INTEGER const k1, INTEGER const k2, INTEGER* ipiv,
            INTEGER const incx) {
  fail("Rlaswp", pga_rlaswp(n, P(a), lda, k1, k2, ipiv, incx));
}

// ---------------------------------------------------------------------------------
// LAPACK
// ---------------------------------------------------------------------------------
void Rgetrf_opencl(INTEGER const m, INTEGER const n, REAL* a, INTEGER const lda, INTEGER* ipiv, INTEGER& info,
                   INTEGER nbb) {
  int nb = iMlaenv(1, "Rgetrf", " ", m, n, -1, -1);
  if (nbb > 0) nb = nbb;  // Rgetrf_opencl.cpp:89
  fail("Rgetrf", pga_rgetrf(m, n, P(a), lda, ipiv, &info, nb));
}

void Rgetrf(INTEGER const m, INTEGER const n, REAL* a, INTEGER const lda, INTEGER* ipiv, INTEGER& info) {
  Rgetrf_opencl(m, n, a, lda, ipiv, info, 0);
}

void Rgetrf2(INTEGER const m, INTEGER const n, REAL* a, INTEGER const lda, INTEGER* ipiv, INTEGER& info) {
  fail("Rgetrf2", pga_rgetrf2(m, n, P(a), lda, ipiv, &info));
}

void Rgetrf2_opencl(INTEGER const m, INTEGER const n, REAL* a, INTEGER const lda, INTEGER* ipiv, INTEGER& info) {
  Rgetrf2(m, n, a, lda, ipiv, info);
}

void Rgetrs(const char* trans, INTEGER const n, INTEGER const nrhs, REAL* a, INTEGER const lda, INTEGER* ipiv,
            REAL* b, INTEGER const ldb, INTEGER& info) {
  fail("Rgetrs", pga_rgetrs(trans[0], n, nrhs, CP(a), lda, ipiv, P(b), ldb, &info));
}

void Rpotrf_opencl(const char* uplo, INTEGER const n, REAL* a, INTEGER const lda, INTEGER& info) {
  int nb = iMlaenv(1, "Rpotrf", uplo, n, -1, -1, -1);  // Rpotrf_opencl.cpp:86
  fail("Rpotrf", pga_rpotrf(uplo[0], n, P(a), lda, &info, nb));
}

void Rpotrf(const char* uplo, INTEGER const n, REAL* a, INTEGER const lda, INTEGER& info) {
  Rpotrf_opencl(uplo, n, a, lda, info);
}

void Rpotrf2(const char* uplo, INTEGER const n, REAL* a, INTEGER const lda, INTEGER& info) {
  fail("Rpotrf2", pga_rpotrf2(uplo[0], n, P(a), lda, &info));
}

void Rpotrf2_opencl(const char* uplo, INTEGER const n, REAL* a, INTEGER const lda, INTEGER& info) {
  Rpotrf2(uplo, n, a, lda, info);
}

void Rpotrs(const char* uplo, INTEGER const n, INTEGER const nrhs, REAL* a, INTEGER const lda, REAL* b,
            INTEGER const ldb, INTEGER& info) {
  fail("Rpotrs", pga_rpotrs(uplo[0], n, nrhs, CP(a), lda, P(b), ldb, &info));
}
