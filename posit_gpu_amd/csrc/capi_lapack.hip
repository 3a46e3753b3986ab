// capi_lapack.hip -- C ABI of the posit32 BLAS-2/3 and LAPACK routines.
//
// *_device: device pointers, asynchronous on the given stream (NULL = library stream).
// plain names: host pointers, synchronous drop-ins for the reference routines; they copy
// the logical sub-matrices to the device, run there, and copy the results back.
// Argument checks follow the reference routine's order and Mxerbla parameter numbers.
#include <hip/hip_runtime.h>

#include <mutex>
#include <vector>

#include "pga_internal.hpp"
#include "posit32_core.hpp"

namespace pga {

hipError_t la_trsm(hipStream_t s, bool left, bool upper, bool trans, bool nounit, int m, int n, uint32_t alpha,
                   const uint32_t* A, int lda, uint32_t* B, int ldb);
hipError_t la_syrk(hipStream_t s, bool upper, bool trans, int n, int k, uint32_t alpha, const uint32_t* a, int lda,
                   uint32_t beta, uint32_t* c, int ldc);
hipError_t la_gemv(hipStream_t s, bool trans, int m, int n, uint32_t alpha, const uint32_t* A, int lda,
                   const uint32_t* x, int incx, uint32_t beta, uint32_t* y, int incy);
hipError_t la_scal(hipStream_t s, int n, uint32_t da, uint32_t* x, int incx);
hipError_t la_iamax(hipStream_t s, int n, const uint32_t* x, int incx, int* result);
hipError_t la_laswp(hipStream_t s, int n, uint32_t* a, int lda, int k1, int k2, const int* ipiv, int incx);
hipError_t la_getrf(hipStream_t s, int m, int n, uint32_t* a, int lda, int* ipiv, int* info, int nb);
hipError_t la_getrf2(hipStream_t s, int m, int n, uint32_t* a, int lda, int* ipiv, int* info);
hipError_t la_getrs(hipStream_t s, bool trans, int n, int nrhs, const uint32_t* a, int lda, const int* ipiv,
                    uint32_t* b, int ldb);
hipError_t la_potrf(hipStream_t s, bool upper, int n, uint32_t* a, int lda, int* info, int nb);
hipError_t la_potrf2(hipStream_t s, bool upper, int n, uint32_t* a, int lda, int* info);
hipError_t la_potrs(hipStream_t s, bool upper, int n, int nrhs, const uint32_t* a, int lda, uint32_t* b,
                    int ldb);

namespace {

__global__ void k_set_int(int* p, int v) {
  PGA_KERNEL_VGPR_FLOOR();
  if (threadIdx.x == 0) *p = v;
}

bool lsame(char a, char b) {
  auto up = [](char c) { return (c >= 'a' && c <= 'z') ? (char)(c - 'a' + 'A') : c; };
  return up(a) == up(b);
}
int imax(int a, int b) { return a > b ? a : b; }
hipStream_t as_stream(void* s) { return s ? (hipStream_t)s : runtime().stream; }

pga_status illegal(int i) {
  set_illegal_argument(i);
  return PGA_ERROR_ILLEGAL_ARGUMENT;
}

// Host<->device copies of a rows x cols column-major block with leading dimension ld
// into a packed device buffer (leading dimension rows); synchronous.
hipError_t put(uint32_t* d, const uint32_t* h, int rows, int cols, int ld, hipStream_t s) {
  return copy_h2d_block(d, h, rows, cols, ld, s);
}
hipError_t get(uint32_t* h, const uint32_t* d, int rows, int cols, int ld, hipStream_t s) {
  return copy_d2h_block(h, d, rows, cols, ld, s);
}

// Rtrsm's checks (Rtrsm.cpp:57-88)
int check_trsm(char side, char uplo, char transa, char diag, int m, int n, int lda, int ldb) {
  bool lside = lsame(side, 'L');
  int nrowa = lside ? m : n;
  bool upper = lsame(uplo, 'U');
  if (!lside && !lsame(side, 'R')) return 1;
  if (!upper && !lsame(uplo, 'L')) return 2;
  if (!lsame(transa, 'N') && !lsame(transa, 'T') && !lsame(transa, 'C')) return 3;
  if (!lsame(diag, 'U') && !lsame(diag, 'N')) return 4;
  if (m < 0) return 5;
  if (n < 0) return 6;
  if (lda < imax(1, nrowa)) return 9;
  if (ldb < imax(1, m)) return 11;
  return 0;
}

// Rsyrk's checks (Rsyrk.cpp:57-82)
int check_syrk(char uplo, char trans, int n, int k, int lda, int ldc) {
  int nrowa = lsame(trans, 'N') ? n : k;
  bool upper = lsame(uplo, 'U');
  if (!upper && !lsame(uplo, 'L')) return 1;
  if (!lsame(trans, 'N') && !lsame(trans, 'T') && !lsame(trans, 'C')) return 2;
  if (n < 0) return 3;
  if (k < 0) return 4;
  if (lda < imax(1, nrowa)) return 7;
  if (ldc < imax(1, n)) return 10;
  return 0;
}

// Rgemv's checks (Rgemv.cpp:32-52)
int check_gemv(char trans, int m, int n, int lda, int incx, int incy) {
  if (!lsame(trans, 'N') && !lsame(trans, 'T') && !lsame(trans, 'C')) return 1;
  if (m < 0) return 2;
  if (n < 0) return 3;
  if (lda < imax(1, m)) return 6;
  if (incx == 0) return 8;
  if (incy == 0) return 11;
  return 0;
}

}  // namespace
}  // namespace pga

using namespace pga;

extern "C" {

// ---------------------------------------------------------------------------------
// TRSM
// ---------------------------------------------------------------------------------
pga_status pga_rtrsm_device(char side, char uplo, char transa, char diag, int m, int n, uint32_t alpha,
                            const uint32_t* a, int lda, uint32_t* b, int ldb, void* stream) {
  PGA_CHECK_INIT();
  if (int bad = check_trsm(side, uplo, transa, diag, m, n, lda, ldb)) return illegal(bad);
  PGA_HIP(la_trsm(as_stream(stream), lsame(side, 'L'), lsame(uplo, 'U'), !lsame(transa, 'N'),
                  lsame(diag, 'N'), m, n, alpha, a, lda, b, ldb));
  return PGA_SUCCESS;
}

pga_status pga_rtrsm(char side, char uplo, char transa, char diag, int m, int n, uint32_t alpha,
                     const uint32_t* a, int lda, uint32_t* b, int ldb) {
  PGA_CHECK_INIT();
  if (int bad = check_trsm(side, uplo, transa, diag, m, n, lda, ldb)) return illegal(bad);
  if (m == 0 || n == 0) return PGA_SUCCESS;
  std::lock_guard<std::recursive_mutex> lk(host_call_mutex());
  hipStream_t s = runtime().stream;
  int na = lsame(side, 'L') ? m : n;
  pga_status st;
  uint32_t* dA = (uint32_t*)scratch(0, (size_t)na * na * 4, &st);
  if (st) return st;
  uint32_t* dB = (uint32_t*)scratch(1, (size_t)m * n * 4, &st);
  if (st) return st;
  PGA_HIP(put(dA, a, na, na, lda, s));
  PGA_HIP(put(dB, b, m, n, ldb, s));
  PGA_HIP(la_trsm(s, lsame(side, 'L'), lsame(uplo, 'U'), !lsame(transa, 'N'), lsame(diag, 'N'), m, n, alpha,
                  dA, na, dB, m));
  PGA_HIP(get(b, dB, m, n, ldb, s));
  PGA_HIP(hipStreamSynchronize(s));
  return PGA_SUCCESS;
}

// ---------------------------------------------------------------------------------
// SYRK
// ---------------------------------------------------------------------------------
pga_status pga_rsyrk_device(char uplo, char trans, int n, int k, uint32_t alpha, const uint32_t* a, int lda,
                            uint32_t beta, uint32_t* c, int ldc, void* stream) {
  PGA_CHECK_INIT();
  if (int bad = check_syrk(uplo, trans, n, k, lda, ldc)) return illegal(bad);
  PGA_HIP(la_syrk(as_stream(stream), lsame(uplo, 'U'), !lsame(trans, 'N'), n, k, alpha, a, lda, beta, c, ldc));
  return PGA_SUCCESS;
}

pga_status pga_rsyrk(char uplo, char trans, int n, int k, uint32_t alpha, const uint32_t* a, int lda,
                     uint32_t beta, uint32_t* c, int ldc) {
  PGA_CHECK_INIT();
  if (int bad = check_syrk(uplo, trans, n, k, lda, ldc)) return illegal(bad);
  if (n == 0) return PGA_SUCCESS;
  std::lock_guard<std::recursive_mutex> lk(host_call_mutex());
  hipStream_t s = runtime().stream;
  bool tr = !lsame(trans, 'N');
  int ar = tr ? k : n, ac = tr ? n : k;
  pga_status st;
  uint32_t* dA = (uint32_t*)scratch(0, (size_t)imax(ar, 1) * imax(ac, 1) * 4, &st);
  if (st) return st;
  uint32_t* dC = (uint32_t*)scratch(1, (size_t)n * n * 4, &st);
  if (st) return st;
  PGA_HIP(put(dA, a, ar, ac, lda, s));
  PGA_HIP(put(dC, c, n, n, ldc, s));
  PGA_HIP(la_syrk(s, lsame(uplo, 'U'), tr, n, k, alpha, dA, imax(ar, 1), beta, dC, n));
  PGA_HIP(get(c, dC, n, n, ldc, s));
  PGA_HIP(hipStreamSynchronize(s));
  return PGA_SUCCESS;
}

// ---------------------------------------------------------------------------------
// GEMV, SCAL, IAMAX, LASWP
// ---------------------------------------------------------------------------------
pga_status pga_rgemv_device(char trans, int m, int n, uint32_t alpha, const uint32_t* a, int lda,
                            const uint32_t* x, int incx, uint32_t beta, uint32_t* y, int incy, void* stream) {
  PGA_CHECK_INIT();
  if (int bad = check_gemv(trans, m, n, lda, incx, incy)) return illegal(bad);
  PGA_HIP(la_gemv(as_stream(stream), !lsame(trans, 'N'), m, n, alpha, a, lda, x, incx, beta, y, incy));
  return PGA_SUCCESS;
}

pga_status pga_rgemv(char trans, int m, int n, uint32_t alpha, const uint32_t* a, int lda, const uint32_t* x,
                     int incx, uint32_t beta, uint32_t* y, int incy) {
  PGA_CHECK_INIT();
  if (int bad = check_gemv(trans, m, n, lda, incx, incy)) return illegal(bad);
  if (m == 0 || n == 0) return PGA_SUCCESS;
  std::lock_guard<std::recursive_mutex> lk(host_call_mutex());
  hipStream_t s = runtime().stream;
  bool tr = !lsame(trans, 'N');
  int lenx = tr ? m : n, leny = tr ? n : m;
  size_t nx = (size_t)(lenx - 1) * (incx > 0 ? incx : -incx) + 1;
  size_t ny = (size_t)(leny - 1) * (incy > 0 ? incy : -incy) + 1;
  pga_status st;
  uint32_t* dA = (uint32_t*)scratch(0, (size_t)m * n * 4, &st);
  if (st) return st;
  uint32_t* dx = (uint32_t*)scratch(1, nx * 4, &st);
  if (st) return st;
  uint32_t* dy = (uint32_t*)scratch(2, ny * 4, &st);
  if (st) return st;
  PGA_HIP(put(dA, a, m, n, lda, s));
  PGA_HIP(hipMemcpyAsync(dx, x, nx * 4, hipMemcpyHostToDevice, s));
  PGA_HIP(hipMemcpyAsync(dy, y, ny * 4, hipMemcpyHostToDevice, s));
  PGA_HIP(la_gemv(s, tr, m, n, alpha, dA, m, dx, incx, beta, dy, incy));
  PGA_HIP(hipMemcpyAsync(y, dy, ny * 4, hipMemcpyDeviceToHost, s));
  PGA_HIP(hipStreamSynchronize(s));
  return PGA_SUCCESS;
}

pga_status pga_rscal_device(int n, uint32_t da, uint32_t* x, int incx, void* stream) {
  PGA_CHECK_INIT();
  PGA_HIP(la_scal(as_stream(stream), n, da, x, incx));
  return PGA_SUCCESS;
}

pga_status pga_rscal(int n, uint32_t da, uint32_t* x, int incx) {
  PGA_CHECK_INIT();
  if (n <= 0 || incx <= 0) return PGA_SUCCESS;
  std::lock_guard<std::recursive_mutex> lk(host_call_mutex());
  hipStream_t s = runtime().stream;
  size_t nx = (size_t)(n - 1) * incx + 1;
  pga_status st;
  uint32_t* dx = (uint32_t*)scratch(1, nx * 4, &st);
  if (st) return st;
  PGA_HIP(hipMemcpyAsync(dx, x, nx * 4, hipMemcpyHostToDevice, s));
  PGA_HIP(la_scal(s, n, da, dx, incx));
  PGA_HIP(hipMemcpyAsync(x, dx, nx * 4, hipMemcpyDeviceToHost, s));
  PGA_HIP(hipStreamSynchronize(s));
  return PGA_SUCCESS;
}

pga_status pga_iramax_device(int n, const uint32_t* x, int incx, int* result, void* stream) {
  PGA_CHECK_INIT();
  PGA_HIP(la_iamax(as_stream(stream), n, x, incx, result));
  return PGA_SUCCESS;
}

pga_status pga_iramax(int n, const uint32_t* x, int incx, int* result) {
  PGA_CHECK_INIT();
  *result = 0;
  if (n < 1 || incx <= 0) return PGA_SUCCESS;
  std::lock_guard<std::recursive_mutex> lk(host_call_mutex());
  hipStream_t s = runtime().stream;
  size_t nx = (size_t)(n - 1) * incx + 1;
  pga_status st;
  uint32_t* dx = (uint32_t*)scratch(1, nx * 4 + 16, &st);
  if (st) return st;
  int* dr = (int*)(dx + nx + 1);
  PGA_HIP(hipMemcpyAsync(dx, x, nx * 4, hipMemcpyHostToDevice, s));
  PGA_HIP(la_iamax(s, n, dx, incx, dr));
  PGA_HIP(hipMemcpyAsync(result, dr, sizeof(int), hipMemcpyDeviceToHost, s));
  PGA_HIP(hipStreamSynchronize(s));
  return PGA_SUCCESS;
}

pga_status pga_rlaswp_device(int n, uint32_t* a, int lda, int k1, int k2, const int* ipiv, int incx,
                             void* stream) {
  PGA_CHECK_INIT();
  PGA_HIP(la_laswp(as_stream(stream), n, a, lda, k1, k2, ipiv, incx));
  return PGA_SUCCESS;
}

pga_status pga_rlaswp(int n, uint32_t* a, int lda, int k1, int k2, const int* ipiv, int incx) {
  PGA_CHECK_INIT();
  if (n <= 0 || incx == 0 || k2 < k1) return PGA_SUCCESS;
  std::lock_guard<std::recursive_mutex> lk(host_call_mutex());
  hipStream_t s = runtime().stream;
  // rows touched: up to max(k2, max ipiv) in every column
  int rows = k2;
  int ix = incx > 0 ? k1 : k1 + (k1 - k2) * incx;
  for (int c = 0; c < k2 - k1 + 1; ++c, ix += incx) rows = imax(rows, ipiv[ix - 1]);
  int npiv = k1 + (k2 - k1) * (incx > 0 ? incx : -incx);
  pga_status st;
  uint32_t* dA = (uint32_t*)scratch(0, (size_t)rows * n * 4, &st);
  if (st) return st;
  int* dp = (int*)scratch(2, (size_t)npiv * 4, &st);
  if (st) return st;
  PGA_HIP(put(dA, a, rows, n, lda, s));
  PGA_HIP(hipMemcpyAsync(dp, ipiv, (size_t)npiv * 4, hipMemcpyHostToDevice, s));
  PGA_HIP(la_laswp(s, n, dA, rows, k1, k2, dp, incx));
  PGA_HIP(get(a, dA, rows, n, lda, s));
  PGA_HIP(hipStreamSynchronize(s));
  return PGA_SUCCESS;
}

// ---------------------------------------------------------------------------------
// LU
// ---------------------------------------------------------------------------------
static int check_getrf(int m, int n, int lda) {
  if (m < 0) return 1;
  if (n < 0) return 2;
  if (lda < imax(1, m)) return 4;
  return 0;
}

pga_status pga_rgetrf_device(int m, int n, uint32_t* a, int lda, int* ipiv, int* info, int nb, void* stream) {
  PGA_CHECK_INIT();
  if (int bad = check_getrf(m, n, lda)) return illegal(bad);
  hipStream_t s = as_stream(stream);
  hipLaunchKernelGGL(k_set_int, dim3(1), dim3(64), 0, s, info, 0);
  if (m == 0 || n == 0) return PGA_SUCCESS;
  if (nb <= 0) nb = 64;  // iMlaenv.cpp:145-151
  PGA_HIP(la_getrf(s, m, n, a, lda, ipiv, info, nb));
  return PGA_SUCCESS;
}

pga_status pga_rgetrf2_device(int m, int n, uint32_t* a, int lda, int* ipiv, int* info, void* stream) {
  PGA_CHECK_INIT();
  if (int bad = check_getrf(m, n, lda)) return illegal(bad);
  hipStream_t s = as_stream(stream);
  hipLaunchKernelGGL(k_set_int, dim3(1), dim3(64), 0, s, info, 0);
  if (m == 0 || n == 0) return PGA_SUCCESS;
  PGA_HIP(la_getrf2(s, m, n, a, lda, ipiv, info));
  return PGA_SUCCESS;
}

static pga_status host_getrf(bool recursive, int m, int n, uint32_t* a, int lda, int* ipiv, int* info, int nb) {
  PGA_CHECK_INIT();
  if (int bad = check_getrf(m, n, lda)) {
    *info = -bad;
    return illegal(bad);
  }
  *info = 0;
  if (m == 0 || n == 0) return PGA_SUCCESS;
  std::lock_guard<std::recursive_mutex> lk(host_call_mutex());
  hipStream_t s = runtime().stream;
  int mn = m < n ? m : n;
  pga_status st;
  uint32_t* dA = (uint32_t*)scratch(0, (size_t)m * n * 4, &st);
  if (st) return st;
  int* dp = (int*)scratch(1, (size_t)(mn + 1) * 4, &st);
  if (st) return st;
  int* dinfo = dp + mn;
  PGA_HIP(put(dA, a, m, n, lda, s));
  hipLaunchKernelGGL(k_set_int, dim3(1), dim3(64), 0, s, dinfo, 0);
  if (recursive)
    PGA_HIP(la_getrf2(s, m, n, dA, m, dp, dinfo));
  else
    PGA_HIP(la_getrf(s, m, n, dA, m, dp, dinfo, nb <= 0 ? 64 : nb));
  PGA_HIP(get(a, dA, m, n, lda, s));
  PGA_HIP(hipMemcpyAsync(ipiv, dp, (size_t)mn * 4, hipMemcpyDeviceToHost, s));
  PGA_HIP(hipMemcpyAsync(info, dinfo, 4, hipMemcpyDeviceToHost, s));
  PGA_HIP(hipStreamSynchronize(s));
  return PGA_SUCCESS;
}

pga_status pga_rgetrf(int m, int n, uint32_t* a, int lda, int* ipiv, int* info, int nb) {
  return host_getrf(false, m, n, a, lda, ipiv, info, nb);
}

pga_status pga_rgetrf2(int m, int n, uint32_t* a, int lda, int* ipiv, int* info) {
  return host_getrf(true, m, n, a, lda, ipiv, info, 0);
}

static int check_getrs(char trans, int n, int nrhs, int lda, int ldb) {
  if (!lsame(trans, 'N') && !lsame(trans, 'T') && !lsame(trans, 'C')) return 1;
  if (n < 0) return 2;
  if (nrhs < 0) return 3;
  if (lda < imax(1, n)) return 5;
  if (ldb < imax(1, n)) return 8;
  return 0;
}

pga_status pga_rgetrs_device(char trans, int n, int nrhs, const uint32_t* a, int lda, const int* ipiv,
                             uint32_t* b, int ldb, void* stream) {
  PGA_CHECK_INIT();
  if (int bad = check_getrs(trans, n, nrhs, lda, ldb)) return illegal(bad);
  PGA_HIP(la_getrs(as_stream(stream), !lsame(trans, 'N'), n, nrhs, a, lda, ipiv, b, ldb));
  return PGA_SUCCESS;
}

pga_status pga_rgetrs(char trans, int n, int nrhs, const uint32_t* a, int lda, const int* ipiv, uint32_t* b,
                      int ldb, int* info) {
  PGA_CHECK_INIT();
  if (int bad = check_getrs(trans, n, nrhs, lda, ldb)) {
    *info = -bad;
    return illegal(bad);
  }
  *info = 0;
  if (n == 0 || nrhs == 0) return PGA_SUCCESS;
  std::lock_guard<std::recursive_mutex> lk(host_call_mutex());
  hipStream_t s = runtime().stream;
  pga_status st;
  uint32_t* dA = (uint32_t*)scratch(0, (size_t)n * n * 4, &st);
  if (st) return st;
  uint32_t* dB = (uint32_t*)scratch(1, (size_t)n * nrhs * 4, &st);
  if (st) return st;
  int* dp = (int*)scratch(2, (size_t)n * 4, &st);
  if (st) return st;
  PGA_HIP(put(dA, a, n, n, lda, s));
  PGA_HIP(put(dB, b, n, nrhs, ldb, s));
  PGA_HIP(hipMemcpyAsync(dp, ipiv, (size_t)n * 4, hipMemcpyHostToDevice, s));
  PGA_HIP(la_getrs(s, !lsame(trans, 'N'), n, nrhs, dA, n, dp, dB, n));
  PGA_HIP(get(b, dB, n, nrhs, ldb, s));
  PGA_HIP(hipStreamSynchronize(s));
  return PGA_SUCCESS;
}

// ---------------------------------------------------------------------------------
// Cholesky
// ---------------------------------------------------------------------------------
static int check_potrf(char uplo, int n, int lda) {
  if (!lsame(uplo, 'U') && !lsame(uplo, 'L')) return 1;
  if (n < 0) return 2;
  if (lda < imax(1, n)) return 4;
  return 0;
}

pga_status pga_rpotrf_device(char uplo, int n, uint32_t* a, int lda, int* info, int nb, void* stream) {
  PGA_CHECK_INIT();
  if (int bad = check_potrf(uplo, n, lda)) return illegal(bad);
  hipStream_t s = as_stream(stream);
  hipLaunchKernelGGL(k_set_int, dim3(1), dim3(64), 0, s, info, 0);
  if (n == 0) return PGA_SUCCESS;
  if (nb <= 0) nb = 64;  // iMlaenv.cpp:223-229
  PGA_HIP(la_potrf(s, lsame(uplo, 'U'), n, a, lda, info, nb));
  return PGA_SUCCESS;
}

pga_status pga_rpotrf2_device(char uplo, int n, uint32_t* a, int lda, int* info, void* stream) {
  PGA_CHECK_INIT();
  if (int bad = check_potrf(uplo, n, lda)) return illegal(bad);
  hipStream_t s = as_stream(stream);
  hipLaunchKernelGGL(k_set_int, dim3(1), dim3(64), 0, s, info, 0);
  if (n == 0) return PGA_SUCCESS;
  PGA_HIP(la_potrf2(s, lsame(uplo, 'U'), n, a, lda, info));
  return PGA_SUCCESS;
}

static pga_status host_potrf(bool recursive, char uplo, int n, uint32_t* a, int lda, int* info, int nb) {
  PGA_CHECK_INIT();
  if (int bad = check_potrf(uplo, n, lda)) {
    *info = -bad;
    return illegal(bad);
  }
  *info = 0;
  if (n == 0) return PGA_SUCCESS;
  std::lock_guard<std::recursive_mutex> lk(host_call_mutex());
  hipStream_t s = runtime().stream;
  pga_status st;
  uint32_t* dA = (uint32_t*)scratch(0, (size_t)n * n * 4, &st);
  if (st) return st;
  int* dinfo = (int*)scratch(1, 64, &st);
  if (st) return st;
  PGA_HIP(put(dA, a, n, n, lda, s));
  hipLaunchKernelGGL(k_set_int, dim3(1), dim3(64), 0, s, dinfo, 0);
  if (recursive)
    PGA_HIP(la_potrf2(s, lsame(uplo, 'U'), n, dA, n, dinfo));
  else
    PGA_HIP(la_potrf(s, lsame(uplo, 'U'), n, dA, n, dinfo, nb <= 0 ? 64 : nb));
  PGA_HIP(get(a, dA, n, n, lda, s));
  PGA_HIP(hipMemcpyAsync(info, dinfo, 4, hipMemcpyDeviceToHost, s));
  PGA_HIP(hipStreamSynchronize(s));
  return PGA_SUCCESS;
}

pga_status pga_rpotrf(char uplo, int n, uint32_t* a, int lda, int* info, int nb) {
  return host_potrf(false, uplo, n, a, lda, info, nb);
}

pga_status pga_rpotrf2(char uplo, int n, uint32_t* a, int lda, int* info) {
  return host_potrf(true, uplo, n, a, lda, info, 0);
}

static int check_potrs(char uplo, int n, int nrhs, int lda, int ldb) {
  if (!lsame(uplo, 'U') && !lsame(uplo, 'L')) return 1;
  if (n < 0) return 2;
  if (nrhs < 0) return 3;
  if (lda < imax(1, n)) return 5;
  if (ldb < imax(1, n)) return 7;
  return 0;
}

pga_status pga_rpotrs_device(char uplo, int n, int nrhs, const uint32_t* a, int lda, uint32_t* b, int ldb,
                             void* stream) {
  PGA_CHECK_INIT();
  if (int bad = check_potrs(uplo, n, nrhs, lda, ldb)) return illegal(bad);
  PGA_HIP(la_potrs(as_stream(stream), lsame(uplo, 'U'), n, nrhs, a, lda, b, ldb));
  return PGA_SUCCESS;
}

pga_status pga_rpotrs(char uplo, int n, int nrhs, const uint32_t* a, int lda, uint32_t* b, int ldb, int* info) {
  PGA_CHECK_INIT();
  if (int bad = check_potrs(uplo, n, nrhs, lda, ldb)) {
    *info = -bad;
    return illegal(bad);
  }
  *info = 0;
  if (n == 0 || nrhs == 0) return PGA_SUCCESS;
  std::lock_guard<std::recursive_mutex> lk(host_call_mutex());
  hipStream_t s = runtime().stream;
  pga_status st;
  uint32_t* dA = (uint32_t*)scratch(0, (size_t)n * n * 4, &st);
  if (st) return st;
  uint32_t* dB = (uint32_t*)scratch(1, (size_t)n * nrhs * 4, &st);
  if (st) return st;
  PGA_HIP(put(dA, a, n, n, lda, s));
  PGA_HIP(put(dB, b, n, nrhs, ldb, s));
  PGA_HIP(la_potrs(s, lsame(uplo, 'U'), n, nrhs, dA, n, dB, n));
  PGA_HIP(get(b, dB, n, nrhs, ldb, s));
  PGA_HIP(hipStreamSynchronize(s));
  return PGA_SUCCESS;
}

}  // extern "C"
