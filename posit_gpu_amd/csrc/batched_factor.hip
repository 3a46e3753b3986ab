// batched_factor.hip -- many independent small posit32 LU / Cholesky factorizations and
// solves in one persistent launch.
//
// Each workgroup takes whole problems from a global counter and runs the reference
// routine on its problem with the workgroup-wide routines of wg_factor.hpp, in place of
// the kernel launches of the per-call drivers (lapack.hip); every result equals the
// per-call result bit for bit.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <vector>

#include "pga_internal.hpp"
#include "wg_factor.hpp"

namespace pga {
namespace wg {

// ---------------------------------------------------------------------------------
// The persistent kernel
// ---------------------------------------------------------------------------------
__device__ void run_task(const FactorTask& t, Shared& sh) {
  int info = 0;
  const int nb = t.nb > 0 ? t.nb : 64;  // iMlaenv for xGETRF / xPOTRF
  const bool upper = (t.uplo == 'U');
  const bool factor = (t.kind == PGA_TASK_GETRF || t.kind == PGA_TASK_POTRF || t.kind == PGA_TASK_GESV ||
                       t.kind == PGA_TASK_POSV);
  const bool lu = (t.kind == PGA_TASK_GETRF || t.kind == PGA_TASK_GESV || t.kind == PGA_TASK_GETRS);
  if (factor && t.n > 0) {
    if (lu) getrf(t.n, t.a, t.lda, t.ipiv, nb, info, sh);
    else info = upper ? potrf<true>(t.n, t.a, t.lda, nb, sh) : potrf<false>(t.n, t.a, t.lda, nb, sh);
  }
  // factor-and-solve tasks solve only after a successful factorization
  const bool solve = (t.kind == PGA_TASK_GETRS || t.kind == PGA_TASK_POTRS ||
                      ((t.kind == PGA_TASK_GESV || t.kind == PGA_TASK_POSV) && info == 0));
  if (solve && t.n > 0 && t.nrhs > 0) {
    if (lu) getrs(t.n, t.nrhs, t.a, t.lda, t.ipiv, t.b, t.ldb, sh);
    else potrs(upper, t.n, t.nrhs, t.a, t.lda, t.b, t.ldb);
  }
  if (threadIdx.x == 0 && t.info) *t.info = info;
  __syncthreads();
}

// two workgroups per CU: the tile code then keeps its registers (248) without spills in
// its loops, and one workgroup's barrier-bound steps overlap the other's tile work
__global__ __launch_bounds__(NT, 3) void k_factor_tasks(const FactorTask* __restrict__ tasks, int ntasks,
                                                     int* __restrict__ counter) {
  PGA_KERNEL_VGPR_FLOOR();
  __shared__ Shared sh;
  for (;;) {
    if (threadIdx.x == 0) sh.bcast = atomicAdd(counter, 1);
    __syncthreads();
    const int t = sh.bcast;
    __syncthreads();
    if (t >= ntasks) return;  // every workgroup ends here once the list is exhausted
    run_task(tasks[t], sh);
  }
}

__global__ void k_reset_counter(int* counter) {
  PGA_KERNEL_VGPR_FLOOR();
  if (threadIdx.x == 0) *counter = 0;
}

// The autotuner's restore: task t's operands (a, b, and ipiv of a re-solve) from src[t]
// into dst[t], one workgroup per task.
__global__ __launch_bounds__(NT) void k_copy_task_operands(const FactorTask* __restrict__ src,
                                                           const FactorTask* __restrict__ dst, int count) {
  PGA_KERNEL_VGPR_FLOOR();
  const int t = blockIdx.x;
  if (t >= count) return;
  const FactorTask s = src[t], d = dst[t];
  for (long e = threadIdx.x; e < (long)s.lda * s.n; e += NT) d.a[e] = s.a[e];
  if (s.b)
    for (long e = threadIdx.x; e < (long)s.ldb * s.nrhs; e += NT) d.b[e] = s.b[e];
  if (s.ipiv && s.kind == PGA_TASK_GETRS)
    for (int e = threadIdx.x; e < s.n; e += NT) d.ipiv[e] = s.ipiv[e];
}

}  // namespace wg
}  // namespace pga

// ===================================================================================
// Plans (C ABI)
// ===================================================================================
struct pga_factor_plan {
  int count = 0;                       // tasks run one workgroup each
  int grid = 0;
  int sched = 0;                       // task order: 0 longest first, 1 the caller's, 2 long and short alternating
  std::vector<pga::FactorTask> host;  // scheduling order
  std::vector<int> order;             // scheduling order -> the caller's index
  pga::FactorTask* d_tasks = nullptr;
  int* d_counter = nullptr;
  pga::BlockedPlan* blocked = nullptr;  // tasks larger than one workgroup handles
  int blocked_count = 0;
  void* d_blocked = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

namespace pga {

// Workgroups for a persistent launch of `count` tasks: as many as fit on the device.
int factor_tasks_grid(int count) {
  static const int per_cu = [] {
    int v = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&v, wg::k_factor_tasks, wg::NT, 0) != hipSuccess || v < 1) v = 1;
    return v;
  }();
  return std::max(1, std::min(count, per_cu * runtime().cus));
}

// One persistent launch over a device task list (the counter is cleared by a kernel, so
// the sequence can be captured in a graph).
hipError_t factor_tasks_launch(const FactorTask* d_tasks, int count, int* d_counter, int grid, hipStream_t s) {
  hipLaunchKernelGGL(wg::k_reset_counter, dim3(1), dim3(64), 0, s, d_counter);
  hipLaunchKernelGGL(wg::k_factor_tasks, dim3(grid), dim3(wg::NT), 0, s, d_tasks, count, d_counter);
  return hipGetLastError();
}

namespace {

hipError_t run(pga_factor_plan& p, hipStream_t s) {
  if (p.count) {
    hipError_t e = factor_tasks_launch(p.d_tasks, p.count, p.d_counter, p.grid, s);
    if (e != hipSuccess) return e;
  }
  return p.blocked ? blocked_plan_launch(p.blocked, s) : hipSuccess;
}

// relative cost of a task, for scheduling the longest first
double task_cost(const FactorTask& t) {
  const double n = t.n, r = t.nrhs;
  const bool factor = (t.kind == PGA_TASK_GETRF || t.kind == PGA_TASK_POTRF || t.kind == PGA_TASK_GESV ||
                       t.kind == PGA_TASK_POSV);
  const bool chol = (t.kind == PGA_TASK_POTRF || t.kind == PGA_TASK_POSV || t.kind == PGA_TASK_POTRS);
  double c = factor ? (chol ? n * n * n / 3 : 2 * n * n * n / 3) : 0.0;
  const bool solve = (t.kind != PGA_TASK_GETRF && t.kind != PGA_TASK_POTRF);
  // the lower Cholesky back substitution is one serial chain per column
  if (solve) c += (chol && t.uplo == 'L') ? 64.0 * n * n * r : 2 * n * n * r;
  return c;
}

}  // namespace
}  // namespace pga

using namespace pga;

extern "C" {

pga_status pga_factor_plan_create(pga_factor_plan** out, int count, const pga_factor_task* tasks, int grid) {
  PGA_CHECK_INIT();
  if (!out) {
    set_illegal_argument(1);
    return PGA_ERROR_ILLEGAL_ARGUMENT;
  }
  if (count < 0 || (count > 0 && !tasks)) {
    set_illegal_argument(2);
    return PGA_ERROR_ILLEGAL_ARGUMENT;
  }
  auto* p = new pga_factor_plan();
  std::vector<FactorTask> v, big;
  std::vector<int> v_idx;
  for (int i = 0; i < count; ++i) {
    const pga_factor_task& q = tasks[i];
    FactorTask t{};
    t.kind = q.kind;
    t.uplo = (q.uplo == 'u' || q.uplo == 'U') ? 'U' : (q.uplo == 'l' || q.uplo == 'L') ? 'L' : 0;
    t.n = q.n;
    t.nrhs = q.nrhs;
    t.nb = q.nb;
    t.lda = q.lda;
    t.ldb = q.ldb;
    t.a = q.a;
    t.ipiv = q.ipiv;
    t.b = q.b;
    t.info = q.info;
    const bool lu = (t.kind == PGA_TASK_GETRF || t.kind == PGA_TASK_GESV || t.kind == PGA_TASK_GETRS);
    const bool solve = (t.kind == PGA_TASK_GETRS || t.kind == PGA_TASK_POTRS || t.kind == PGA_TASK_GESV ||
                        t.kind == PGA_TASK_POSV);
    bool ok = t.kind >= PGA_TASK_GETRF && t.kind <= PGA_TASK_POSV;
    ok = ok && t.n >= 0 && t.n <= PGA_BLOCKED_TASK_MAX_N && t.lda >= std::max(1, t.n) && t.a;
    ok = ok && (lu ? t.ipiv != nullptr : t.uplo != 0);
    if (solve) ok = ok && t.nrhs >= 0 && (t.nrhs == 0 || (t.b && t.ldb >= std::max(1, t.n)));
    if (!ok) {
      delete p;
      set_illegal_argument(3);
      return PGA_ERROR_ILLEGAL_ARGUMENT;
    }
    if (!solve) t.nrhs = 0;
    if (t.n <= PGA_FACTOR_TASK_MAX_N) {
      v.push_back(t);
      v_idx.push_back(i);
    } else {
      big.push_back(t);
    }
  }
  std::vector<int> idx(v.size());
  for (size_t i = 0; i < idx.size(); ++i) idx[i] = (int)i;
  std::stable_sort(idx.begin(), idx.end(), [&](int x, int y) { return task_cost(v[x]) > task_cost(v[y]); });
  for (int i : idx) {
    p->host.push_back(v[i]);
    p->order.push_back(v_idx[i]);
  }
  p->count = (int)p->host.size();
  p->grid = grid > 0 ? std::max(1, std::min(p->count, grid)) : factor_tasks_grid(p->count);
  hipError_t e = hipMalloc(&p->d_counter, sizeof(int));
  if (e == hipSuccess && p->count) e = hipMalloc(&p->d_tasks, sizeof(FactorTask) * p->count);
  if (e == hipSuccess && p->count)
    e = hipMemcpy(p->d_tasks, p->host.data(), sizeof(FactorTask) * p->count, hipMemcpyHostToDevice);
  if (e == hipSuccess && !big.empty()) {
    p->blocked = blocked_plan_build(big.data(), (int)big.size());
    p->blocked_count = (int)big.size();
    e = hipMalloc(&p->d_blocked, blocked_plan_bytes(p->blocked));
    if (e == hipSuccess) e = blocked_plan_upload(p->blocked, p->d_blocked, runtime().stream);
    if (e == hipSuccess) e = hipStreamSynchronize(runtime().stream);
  }
  if (e != hipSuccess) {
    if (p->d_counter) (void)hipFree(p->d_counter);
    if (p->d_tasks) (void)hipFree(p->d_tasks);
    if (p->d_blocked) (void)hipFree(p->d_blocked);
    if (p->blocked) blocked_plan_destroy(p->blocked);
    delete p;
    return from_hip(e);
  }
  *out = p;
  return PGA_SUCCESS;
}

pga_status pga_factor_plan_execute(pga_factor_plan* p, void* stream) {
  PGA_CHECK_INIT();
  if (!p) return PGA_ERROR_ILLEGAL_ARGUMENT;
  if (p->count == 0 && !p->blocked) return PGA_SUCCESS;
  PGA_HIP(run(*p, stream ? (hipStream_t)stream : runtime().stream));
  return PGA_SUCCESS;
}

pga_status pga_factor_plan_capture_graph(pga_factor_plan* p, void* stream) {
  PGA_CHECK_INIT();
  if (!p) return PGA_ERROR_ILLEGAL_ARGUMENT;
  hipStream_t s = stream ? (hipStream_t)stream : runtime().stream;
  if (p->exec) {
    (void)hipGraphExecDestroy(p->exec);
    (void)hipGraphDestroy(p->graph);
    p->exec = nullptr;
    p->graph = nullptr;
  }
  PGA_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  hipError_t e = (p->count || p->blocked) ? run(*p, s) : hipSuccess;
  hipError_t e2 = hipStreamEndCapture(s, &p->graph);
  if (e != hipSuccess) return from_hip(e);
  PGA_HIP(e2);
  PGA_HIP(hipGraphInstantiate(&p->exec, p->graph, nullptr, nullptr, 0));
  return PGA_SUCCESS;
}

pga_status pga_factor_plan_launch_graph(pga_factor_plan* p, void* stream) {
  PGA_CHECK_INIT();
  if (!p || !p->exec) return PGA_ERROR_ILLEGAL_ARGUMENT;
  PGA_HIP(hipGraphLaunch(p->exec, stream ? (hipStream_t)stream : runtime().stream));
  return PGA_SUCCESS;
}

// Tunes the persistent launch of a plan: times it on scratch copies of every task's
// operands for each workgroup count per CU that fits (1 .. the occupancy limit) and each
// task order (longest first, the caller's, long and short alternating), and keeps the
// fastest.  Every configuration gives the same bits (each task runs the same code).
pga_status pga_factor_plan_autotune(pga_factor_plan* p, int* grid, int* sched, double* tune_ms) {
  PGA_CHECK_INIT();
  if (!p) return PGA_ERROR_ILLEGAL_ARGUMENT;
  const auto t_start = std::chrono::steady_clock::now();
  if (p->count > 0) {
    hipStream_t s = runtime().stream;
    // back to the longest-first list (ties in the caller's order), which every order below
    // is taken from, so that tuning a plan again starts from the same list
    {
      std::vector<int> k(p->count);
      for (int i = 0; i < p->count; ++i) k[i] = i;
      std::sort(k.begin(), k.end(), [&](int x, int y) {
        const double cx = task_cost(p->host[x]), cy = task_cost(p->host[y]);
        return cx != cy ? cx > cy : p->order[x] < p->order[y];
      });
      std::vector<FactorTask> h(p->count);
      std::vector<int> o(p->count);
      for (int i = 0; i < p->count; ++i) {
        h[i] = p->host[k[i]];
        o[i] = p->order[k[i]];
      }
      p->host = h;
      p->order = o;
    }
    // scratch operands, one block per task
    std::vector<FactorTask> scratch = p->host;
    size_t words = 0;
    for (const FactorTask& t : p->host) {
      words += (size_t)t.lda * t.n + (t.b ? (size_t)t.ldb * t.nrhs : 0) + (t.ipiv ? (size_t)t.n : 0) + 1;
      words = (words + 63) & ~(size_t)63;
    }
    uint32_t* buf = nullptr;
    FactorTask *d_src = nullptr, *d_dst = nullptr, *d_try = nullptr;
    hipError_t e = hipMalloc((void**)&buf, words * 4);
    if (e == hipSuccess) e = hipMalloc((void**)&d_src, sizeof(FactorTask) * p->count);
    if (e == hipSuccess) e = hipMalloc((void**)&d_dst, sizeof(FactorTask) * p->count);
    if (e == hipSuccess) e = hipMalloc((void**)&d_try, sizeof(FactorTask) * p->count);
    if (e == hipSuccess) {
      size_t off = 0;
      for (FactorTask& t : scratch) {
        t.a = buf + off;
        off += (size_t)t.lda * t.n;
        if (t.b) {
          t.b = buf + off;
          off += (size_t)t.ldb * t.nrhs;
        }
        if (t.ipiv) {
          t.ipiv = reinterpret_cast<int*>(buf + off);
          off += t.n;
        }
        t.info = reinterpret_cast<int*>(buf + off);
        off = (off + 1 + 63) & ~(size_t)63;
      }
      e = hipMemcpyAsync(d_src, p->host.data(), sizeof(FactorTask) * p->count, hipMemcpyHostToDevice, s);
      if (e == hipSuccess)
        e = hipMemcpyAsync(d_dst, scratch.data(), sizeof(FactorTask) * p->count, hipMemcpyHostToDevice, s);
    }
    int per_cu = 1;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, wg::k_factor_tasks, wg::NT, 0) != hipSuccess ||
        per_cu < 1)
      per_cu = 1;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    float best = 1e30f;
    int best_grid = p->grid, best_sched = 0;
    std::vector<FactorTask> ordered(p->count);
    for (int sc = 0; sc < 3 && e == hipSuccess; ++sc) {
      // positions in the longest-first list for this order
      std::vector<int> pos(p->count);
      for (int i = 0; i < p->count; ++i) pos[i] = i;
      if (sc == 1) std::sort(pos.begin(), pos.end(), [&](int x, int y) { return p->order[x] < p->order[y]; });
      if (sc == 2)
        for (int i = 0, lo = 0, hi = p->count - 1; i < p->count; ++i) pos[i] = (i % 2 == 0) ? lo++ : hi--;
      for (int i = 0; i < p->count; ++i) ordered[i] = scratch[pos[i]];
      if ((e = hipMemcpyAsync(d_try, ordered.data(), sizeof(FactorTask) * p->count, hipMemcpyHostToDevice, s)) !=
          hipSuccess)
        break;
      for (int g = 1; g <= per_cu && e == hipSuccess; ++g) {
        const int gr = std::max(1, std::min(p->count, g * runtime().cus));
        float total = 0.f;
        for (int r = 0; r < 2 && e == hipSuccess; ++r) {
          hipLaunchKernelGGL(wg::k_copy_task_operands, dim3((unsigned)p->count), dim3(wg::NT), 0, s, d_src, d_dst,
                             p->count);
          (void)hipEventRecord(e0, s);
          e = factor_tasks_launch(d_try, p->count, p->d_counter, gr, s);
          (void)hipEventRecord(e1, s);
          if (e == hipSuccess) e = hipEventSynchronize(e1);
          float ms = 0.f;
          (void)hipEventElapsedTime(&ms, e0, e1);
          if (r > 0) total += ms;  // the first run warms up
        }
        if (e == hipSuccess && total < best) {
          best = total;
          best_grid = gr;
          best_sched = sc;
        }
      }
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipStreamSynchronize(s);
    if (buf) (void)hipFree(buf);
    if (d_src) (void)hipFree(d_src);
    if (d_dst) (void)hipFree(d_dst);
    if (d_try) (void)hipFree(d_try);
    if (e != hipSuccess) {  // the plan keeps its grid, in longest-first order
      p->sched = 0;
      (void)hipMemcpy(p->d_tasks, p->host.data(), sizeof(FactorTask) * p->count, hipMemcpyHostToDevice);
      return from_hip(e);
    }
    // the chosen order on the plan's own tasks
    std::vector<int> pos(p->count);
    for (int i = 0; i < p->count; ++i) pos[i] = i;
    if (best_sched == 1) std::sort(pos.begin(), pos.end(), [&](int x, int y) { return p->order[x] < p->order[y]; });
    if (best_sched == 2)
      for (int i = 0, lo = 0, hi = p->count - 1; i < p->count; ++i) pos[i] = (i % 2 == 0) ? lo++ : hi--;
    std::vector<FactorTask> host(p->count);
    std::vector<int> order(p->count);
    for (int i = 0; i < p->count; ++i) {
      host[i] = p->host[pos[i]];
      order[i] = p->order[pos[i]];
    }
    p->host = host;
    p->order = order;
    p->grid = best_grid;
    p->sched = best_sched;
    PGA_HIP(hipMemcpy(p->d_tasks, p->host.data(), sizeof(FactorTask) * p->count, hipMemcpyHostToDevice));
    if (p->exec) {  // a captured graph holds the old launch: capture again to use the new one
      (void)hipGraphExecDestroy(p->exec);
      (void)hipGraphDestroy(p->graph);
      p->exec = nullptr;
      p->graph = nullptr;
    }
  }
  if (grid) *grid = p->grid;
  if (sched) *sched = p->sched;
  if (tune_ms) *tune_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start).count();
  return PGA_SUCCESS;
}

pga_status pga_factor_plan_info(const pga_factor_plan* p, int* count, int* grid) {
  if (!p) return PGA_ERROR_ILLEGAL_ARGUMENT;
  if (count) *count = p->count + p->blocked_count;
  if (grid) *grid = p->grid;
  return PGA_SUCCESS;
}

pga_status pga_factor_plan_destroy(pga_factor_plan* p) {
  if (!p) return PGA_SUCCESS;
  (void)hipStreamSynchronize(runtime().stream);
  if (p->exec) (void)hipGraphExecDestroy(p->exec);
  if (p->graph) (void)hipGraphDestroy(p->graph);
  if (p->d_tasks) (void)hipFree(p->d_tasks);
  if (p->d_counter) (void)hipFree(p->d_counter);
  if (p->d_blocked) (void)hipFree(p->d_blocked);
  if (p->blocked) blocked_plan_destroy(p->blocked);
  delete p;
  return PGA_SUCCESS;
}

}  // extern "C"
