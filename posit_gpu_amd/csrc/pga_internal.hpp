// pga_internal.hpp -- shared declarations of the HIP runtime layer (not part of the ABI).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <mutex>

#include "../../include/posit_gpu_amd.h"

namespace pga {

enum class GemmImpl { Tiled = 0, Exact = 1 };

// Options of the device GEMM used by the factorizations.
struct GemmOpts {
  bool skip_zero_b = false;  // skip l when the B-side multiplier is zero (Rsyrk 'N')
  int tri = 0;               // 0 full C, 1 store i >= j only, 2 store i <= j only
  const int* stop = nullptr; // device flag: when nonzero the launch does nothing
  int tile = 0;              // 0 automatic, else 64 / 32 / 16
  bool urgent = false;       // a step on a factorization's critical path (PGA_URGENT)
  int cap_per_cu = 0;        // at most this many workgroups per CU (0: no cap), through unused LDS
};

// Dynamic LDS that keeps a kernel at `cap` workgroups per CU or fewer (0: none): each
// workgroup then holds more than 1/(cap + 1) of the CU's LDS.  A factorization caps its
// trailing update so that the waves of its concurrent panel kernels find room on a CU.
unsigned lds_for_cap(const void* kernel, int cap);

// The tile edge pga_rgemm_autotune chose for this exact shape (0: none), and its record
int tuned_tile(bool ta, bool tb, int m, int n, int k);
void set_tuned_tile(bool ta, bool tb, int m, int n, int k, int tile);

// Device-pointer GEMM launch (rgemm.hip); arguments already validated.
hipError_t rgemm_launch(bool ta, bool tb, int m, int n, int k, uint32_t alpha, const uint32_t* A,
                        int lda, const uint32_t* B, int ldb, uint32_t beta, uint32_t* C, int ldc,
                        hipStream_t s, GemmImpl impl, const GemmOpts& o = GemmOpts());

// One problem of a batched GEMM (device-resident descriptor array).
struct BatchProblem {
  int m, n, k, lda, ldb, ldc;
  uint32_t alpha, beta;
  const uint32_t* a;
  const uint32_t* b;
  uint32_t* c;
};

// One task of a batched factorization plan (device-resident descriptor array).
struct FactorTask {
  int kind, uplo, n, nrhs, nb, lda, ldb;
  uint32_t* a;
  int* ipiv;
  uint32_t* b;
  int* info;
};

// Looks up this device's address of the no-allocation batched GEMM's fixed device
// array (batched.hip); the device must be current.
hipError_t batched_noalloc_setup(int device);

// Persistent task kernel of batched_factor.hip
int factor_tasks_grid(int count);
hipError_t factor_tasks_launch(const FactorTask* d_tasks, int count, int* d_counter, int grid, hipStream_t s);

// Blocked lockstep batches of tasks larger than one workgroup handles (blocked_batch.hip).
// build: host descriptors (device pointers); upload: the plan image into a device buffer
// of blocked_plan_bytes(); launch: the whole sequence on one stream (graph-capturable).
struct BlockedPlan;
BlockedPlan* blocked_plan_build(const FactorTask* tasks, int count);
size_t blocked_plan_bytes(const BlockedPlan* p);
hipError_t blocked_plan_upload(BlockedPlan* p, void* dev, hipStream_t s);
hipError_t blocked_plan_launch(const BlockedPlan* p, hipStream_t s);
void blocked_plan_destroy(BlockedPlan* p);

// Per-call device drivers of lapack.hip (stream-ordered; arguments already validated)
hipError_t la_getrf(hipStream_t s, int m, int n, uint32_t* a, int lda, int* ipiv, int* info, int nb);
hipError_t la_getrs(hipStream_t s, bool trans, int n, int nrhs, const uint32_t* a, int lda, const int* ipiv,
                    uint32_t* b, int ldb);
hipError_t la_potrf(hipStream_t s, bool upper, int n, uint32_t* a, int lda, int* info, int nb);
hipError_t la_potrs(hipStream_t s, bool upper, int n, int nrhs, const uint32_t* a, int lda, uint32_t* b,
                    int ldb);

// Runtime state (runtime.hip)
struct Runtime {
  bool initialized = false;
  int device = -1;
  hipStream_t stream = nullptr;
  GemmImpl gemm_impl = GemmImpl::Tiled;
  int cus = 256;
  int lds_per_cu = 160 * 1024;  // LDS bytes of one CU (the device's maxSharedMemoryPerMultiProcessor)
  int gemm_tile = 0;  // PGA_GEMM_TILE: 0 per shape, else the tile edge of the plain GEMM
  bool gemm_panels = true;         // PGA_GEMM_PANELS=0: large GEMMs decode in every tile
  long long panel_min_work = 1LL << 30;  // m*n*k from which the 64 x 32 tile uses panels
  hipMemPool_t pool = nullptr;     // stream-ordered workspace of the GEMM panels
};
Runtime& runtime();

void set_illegal_argument(int index);
pga_status from_hip(hipError_t e);

// Growable device scratch buffers used by the host-pointer entry points and the backward
// error.  Every caller holds host_call_mutex() from before its first scratch() call until
// its last use of the slots has completed on the device, so no two calls share a slot and
// no slot is freed while another call is filling or reading it.
void* scratch(int slot, size_t bytes, pga_status* st);
std::recursive_mutex& host_call_mutex();
// Synchronous host <-> device copies of a column-major block through pinned staging
// (host leading dimension ld, device leading dimension rows).
hipError_t copy_h2d_block(uint32_t* d, const uint32_t* h, int rows, int cols, int ld, hipStream_t s);
hipError_t copy_d2h_block(uint32_t* h, const uint32_t* d, int rows, int cols, int ld, hipStream_t s);

}  // namespace pga

#define PGA_CHECK_INIT()                                         \
  do {                                                           \
    if (!::pga::runtime().initialized) {                         \
      pga_status _s = pga_init(-1);                              \
      if (_s != PGA_SUCCESS) return _s;                          \
    }                                                            \
  } while (0)

#define PGA_HIP(call)                                            \
  do {                                                           \
    hipError_t _e = (call);                                      \
    if (_e != hipSuccess) return ::pga::from_hip(_e);            \
  } while (0)

// First statement of every kernel: the empty asm marks v31 as written, so each wave
// allocates at least 32 VGPRs.  On MI355X (gfx950), waves of kernels with smaller
// allocations (16 or 24 VGPRs) intermittently computed wrong values while long-running
// compute kernels of another stream or process shared their SIMDs; the same code with
// 32 or more VGPRs allocated computed correct values in every run.  Occupancy is not
// affected (32 VGPRs still allow the maximum number of waves per SIMD).
#define PGA_KERNEL_VGPR_FLOOR() asm volatile("" ::: "v31")

// Raises the issue priority of the calling wave (s_setprio 3): latency-bound steps on a
// factorization's critical path (panel leaves, diagonal blocks, their small updates)
// then win the SIMD's instruction arbitration against the large trailing-update GEMMs
// that a second stream runs beside them, which fill the cycles these waves leave idle.
#define PGA_URGENT() __builtin_amdgcn_s_setprio(3)
