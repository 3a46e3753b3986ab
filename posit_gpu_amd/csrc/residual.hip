// residual.hip -- FP64 backward error of a posit32 solve, computed on the GPU.
//
// berr_j = ||A x_j - b_j||_inf / (||A||_inf * ||x_j||_inf), with A, x, b decoded exactly
// from posit32 and the products summed in FP64.  This is the accuracy figure the
// reference's check programs print (main_lu_check.cpp:234-241 forms the FP64 residual
// of the posit solution); it is a reported metric, not part of any posit result.
#include <hip/hip_runtime.h>

#include "pga_internal.hpp"
#include "posit32_core.hpp"

namespace pga {
namespace {

constexpr int WG = 256;

__device__ double atomic_max_nonneg(double* addr, double v) {
  // non-negative doubles order like their bit patterns
  unsigned long long* a = (unsigned long long*)addr;
  unsigned long long nv = __double_as_longlong(v);
  unsigned long long old = *a;
  while (nv > old) {
    unsigned long long prev = atomicCAS(a, old, nv);
    if (prev == old) break;
    old = prev;
  }
  return __longlong_as_double(old);
}

// one lane per row: |r_i| and the row sum of |A|, reduced to maxima
__global__ __launch_bounds__(WG) void k_residual(int n, const uint32_t* __restrict__ A, int lda,
                                                 const uint32_t* __restrict__ x, const uint32_t* __restrict__ b,
                                                 double* out /* [0]=max|r|, [1]=||A||, [2]=||x|| */) {
  PGA_KERNEL_VGPR_FLOOR();
  int i = blockIdx.x * WG + threadIdx.x;
  if (i >= n) return;
  double r = -to_double(b[i]), arow = 0.0;
  for (int j = 0; j < n; ++j) {
    double a = to_double(A[i + (size_t)j * lda]);
    r = fma(a, to_double(x[j]), r);
    arow += fabs(a);
  }
  atomic_max_nonneg(out + 0, fabs(r));
  atomic_max_nonneg(out + 1, arow);
  atomic_max_nonneg(out + 2, fabs(to_double(x[i])));
}

__global__ void k_zero3(double* p) {
  PGA_KERNEL_VGPR_FLOOR();
  if (threadIdx.x < 3) p[threadIdx.x] = 0.0;
}

}  // namespace
}  // namespace pga

using namespace pga;

extern "C" pga_status pga_backward_error_device(int n, const uint32_t* a, int lda, const uint32_t* x,
                                                const uint32_t* b, double* result, void* stream) {
  PGA_CHECK_INIT();
  if (n < 0 || lda < (n > 1 ? n : 1)) {
    set_illegal_argument(n < 0 ? 1 : 3);
    return PGA_ERROR_ILLEGAL_ARGUMENT;
  }
  hipStream_t s = stream ? (hipStream_t)stream : runtime().stream;
  // the accumulator lives in a shared scratch slot: calls from several threads take turns
  // (the call is synchronous, so the lock is held until the result has been read back)
  std::lock_guard<std::recursive_mutex> lk(host_call_mutex());
  pga_status st;
  double* tmp = (double*)scratch(6, 3 * sizeof(double), &st);
  if (st) return st;
  hipLaunchKernelGGL(k_zero3, dim3(1), dim3(64), 0, s, tmp);
  if (n > 0)
    hipLaunchKernelGGL(k_residual, dim3((n + WG - 1) / WG), dim3(WG), 0, s, n, a, lda, x, b, tmp);
  double h[3];
  PGA_HIP(hipMemcpyAsync(h, tmp, sizeof h, hipMemcpyDeviceToHost, s));
  PGA_HIP(hipStreamSynchronize(s));
  *result = (h[1] > 0 && h[2] > 0) ? h[0] / (h[1] * h[2]) : h[0];
  return PGA_SUCCESS;
}
