// batched.hip -- batched posit32 GEMM: one persistent launch over many problems,
// tile shape and grid size chosen by an autotuner, launch sequence captured in a hipGraph.
//
// A plan holds the problem descriptors (device), the per-problem tile offsets, a work
// counter and the chosen configuration.  Executing the plan launches a one-lane kernel
// that clears the counter and the persistent kernel (rgemm.hip:
// rgemm_batched_persistent); both can be captured in a graph and replayed.  Every element
// is computed by the same tile code as the plain GEMM, so results equal per-call results
// bit for bit.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <numeric>
#include <string>
#include <vector>

#include "pga_internal.hpp"

namespace pga {

hipError_t batched_launch(bool ta, bool tb, int tile, int grid, const BatchProblem* probs,
                          const int* tile_start, int nprob, int total_tiles, int* counter,
                          hipStream_t s);
int batched_tiles(int m, int n, int tile);

}  // namespace pga

// ---------------------------------------------------------------------------------
// Device memory of pga_batch_gemm_noalloc: the descriptors, tile offsets and work
// counter of one batch, held by every device the library runs on until the process
// ends.  Host code never names this array as a device address: kernels and copies get
// the address hipGetSymbolAddress returned for the current device (g_noalloc[dev].ptr).
// ---------------------------------------------------------------------------------
constexpr size_t NOALLOC_PROBS_BYTES = sizeof(pga::BatchProblem) * PGA_NOALLOC_MAX_PROBLEMS;
constexpr size_t NOALLOC_START_OFF = (NOALLOC_PROBS_BYTES + 255) & ~(size_t)255;
constexpr size_t NOALLOC_COUNTER_OFF = (NOALLOC_START_OFF + sizeof(int) * (PGA_NOALLOC_MAX_PROBLEMS + 1) + 255) &
                                       ~(size_t)255;
constexpr size_t NOALLOC_BYTES = NOALLOC_COUNTER_OFF + 256;
static __device__ __align__(256) unsigned char scratch[NOALLOC_BYTES];

struct pga_batch_plan {
  bool ta = false, tb = false;
  int count = 0;
  int tile = 64;
  int grid = 0;
  int sched = 0;  // problem order (pga::schedule_positions): 0 largest first, 1 the caller's, 2 large and small alternating
  int total_tiles = 0;
  double tune_ms = 0.0;
  std::vector<pga::BatchProblem> host;  // in scheduling order
  std::vector<int> order;               // scheduling order -> caller's index
  pga::BatchProblem* d_probs = nullptr;
  int* d_tile_start = nullptr;
  int* d_counter = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

namespace pga {
namespace {

std::mutex g_tune_mutex;
struct TuneEntry {
  int tile, grid, sched;
};
std::map<std::string, TuneEntry> g_tune_cache;  // signature -> tuned configuration
bool g_tune_loaded = false;

const char* tune_file() { return getenv("PGA_TUNE_CACHE"); }

void load_tune_cache() {
  if (g_tune_loaded) return;
  g_tune_loaded = true;
  const char* f = tune_file();
  if (!f) return;
  FILE* fp = fopen(f, "r");
  if (!fp) return;
  char line[640], sig[512];
  int tile, grid, sched;
  // one entry per line: <signature> <tile> <grid> [<schedule>]
  while (fgets(line, sizeof line, fp)) {
    sched = 0;
    if (sscanf(line, "%511s %d %d %d", sig, &tile, &grid, &sched) >= 3) g_tune_cache[sig] = {tile, grid, sched};
  }
  fclose(fp);
}

void save_tune_cache() {
  const char* f = tune_file();
  if (!f) return;
  FILE* fp = fopen(f, "w");
  if (!fp) return;
  for (auto& kv : g_tune_cache)
    fprintf(fp, "%s %d %d %d\n", kv.first.c_str(), kv.second.tile, kv.second.grid, kv.second.sched);
  fclose(fp);
}

// Signature of a batch for the tuning cache: transposes, problem count and a coarse
// histogram of log2 sizes of (m, n, k).
std::string signature(const pga_batch_plan& p) {
  int hist[3][12] = {};
  for (auto& q : p.host) {
    int v[3] = {q.m, q.n, q.k};
    for (int d = 0; d < 3; ++d) {
      int b = 0;
      while ((1 << (b + 1)) <= v[d] && b < 11) ++b;
      hist[d][b]++;
    }
  }
  std::string s = std::string(p.ta ? "T" : "N") + (p.tb ? "T" : "N") + ":" + std::to_string(p.count);
  for (int d = 0; d < 3; ++d) {
    s += ":";
    for (int b = 0; b < 12; ++b) s += std::to_string((hist[d][b] * 16 + p.count / 2) / std::max(1, p.count)) + ",";
  }
  return s;
}

hipError_t build_device_tables(pga_batch_plan& p) {
  std::vector<int> start(p.count + 1, 0);
  for (int i = 0; i < p.count; ++i) start[i + 1] = start[i] + batched_tiles(p.host[i].m, p.host[i].n, p.tile);
  p.total_tiles = start[p.count];
  hipStream_t s = runtime().stream;
  hipError_t e;
  if (!p.d_probs) {
    if ((e = hipMalloc(&p.d_probs, sizeof(BatchProblem) * std::max(1, p.count))) != hipSuccess) return e;
    if ((e = hipMalloc(&p.d_tile_start, sizeof(int) * (p.count + 1))) != hipSuccess) return e;
    if ((e = hipMalloc(&p.d_counter, sizeof(int))) != hipSuccess) return e;
  }
  if ((e = hipMemcpyAsync(p.d_probs, p.host.data(), sizeof(BatchProblem) * p.count, hipMemcpyHostToDevice, s)))
    return e;
  if ((e = hipMemcpyAsync(p.d_tile_start, start.data(), sizeof(int) * (p.count + 1), hipMemcpyHostToDevice, s)))
    return e;
  return hipStreamSynchronize(s);
}

hipError_t run(pga_batch_plan& p, hipStream_t s) {
  return batched_launch(p.ta, p.tb, p.tile, p.grid, p.d_probs, p.d_tile_start, p.count, p.total_tiles,
                        p.d_counter, s);
}

// Per device: the address of `scratch` and an event marking the end of the last call's
// descriptor copies (the host staging below is rewritten only after it).
constexpr int NOALLOC_MAX_DEVICES = 64;
struct NoallocDevice {
  unsigned char* ptr = nullptr;
  hipEvent_t copied = nullptr;
  bool pending = false;  // `copied` has been recorded and may not have completed yet
};
NoallocDevice g_noalloc[NOALLOC_MAX_DEVICES];
// Host staging of the descriptors; one call at a time fills it (g_noalloc_mu).
std::mutex g_noalloc_mu;
BatchProblem g_noalloc_probs[PGA_NOALLOC_MAX_PROBLEMS];
int g_noalloc_start[PGA_NOALLOC_MAX_PROBLEMS + 1];
int g_noalloc_order[PGA_NOALLOC_MAX_PROBLEMS];

}  // namespace

// Set-up of the no-allocation entry point on the current device `device` (called by
// pga_init once the device is current; later calls for the same device do nothing).
hipError_t batched_noalloc_setup(int device) {
  if (device < 0 || device >= NOALLOC_MAX_DEVICES) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lk(g_noalloc_mu);
  NoallocDevice& d = g_noalloc[device];
  if (d.ptr) return hipSuccess;
  void* p = nullptr;
  hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(::scratch));
  if (e != hipSuccess) return e;
  if ((e = hipEventCreateWithFlags(&d.copied, hipEventDisableTiming)) != hipSuccess) return e;
  d.ptr = static_cast<unsigned char*>(p);
  return hipSuccess;
}

namespace {

// Positions in the largest-first list for schedule `sched`: 0 that list, 1 the caller's
// order (order[] holds each position's caller index), 2 largest and smallest alternating
// (long tiles and short ones mixed through the persistent launch).
std::vector<int> schedule_positions(const std::vector<int>& order, int sched) {
  const int n = (int)order.size();
  std::vector<int> pos(n);
  for (int i = 0; i < n; ++i) pos[i] = i;
  if (sched == 1) {
    std::sort(pos.begin(), pos.end(), [&](int x, int y) { return order[x] < order[y]; });
  } else if (sched == 2) {
    for (int i = 0, lo = 0, hi = n - 1; i < n; ++i) pos[i] = (i % 2 == 0) ? lo++ : hi--;
  }
  return pos;
}

// Puts the plan's problems (kept in largest-first order in lpt / lpt_order) into the order
// of schedule `sched`.
void apply_schedule(pga_batch_plan& p, const std::vector<BatchProblem>& lpt, const std::vector<int>& lpt_order,
                    int sched) {
  const std::vector<int> pos = schedule_positions(lpt_order, sched);
  for (size_t i = 0; i < pos.size(); ++i) {
    p.host[i] = lpt[pos[i]];
    p.order[i] = lpt_order[pos[i]];
  }
  p.sched = sched;
}

// Time every (tile, grid) candidate on scratch copies of the output blocks, then the
// other problem orders at the best of them; keep the best.
hipError_t autotune(pga_batch_plan& p) {
  const int cus = runtime().cus;
  const int tiles[3] = {64, 32, 16};
  // largest grid first: workgroups beyond the queue's tiles end at once, so a larger grid
  // is rarely much slower, while a too-small one leaves CUs idle
  const int grids_per_cu[4] = {8, 4, 2, 1};
  // scratch outputs so that tuning does not touch the caller's C
  size_t elems = 0;
  for (auto& q : p.host) elems += (size_t)q.m * q.n;
  uint32_t* scratch_c = nullptr;
  hipError_t e = hipMalloc(&scratch_c, std::max<size_t>(4, elems * 4));
  if (e != hipSuccess) return e;
  hipStream_t s = runtime().stream;
  // defined contents, so the timed runs take the same paths on every configuration
  if ((e = hipMemsetAsync(scratch_c, 0, std::max<size_t>(4, elems * 4), s)) != hipSuccess) {
    (void)hipFree(scratch_c);
    return e;
  }
  std::vector<BatchProblem> saved = p.host;
  size_t off = 0;
  for (auto& q : p.host) {
    q.c = scratch_c + off;
    q.ldc = std::max(1, q.m);
    off += (size_t)q.m * q.n;
  }
  hipEvent_t e0, e1;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  // ms per run of `reps` runs back to back
  auto timed = [&](int reps, float& ms) -> hipError_t {
    (void)hipEventRecord(e0, s);
    for (int r = 0; r < reps; ++r)
      if (hipError_t er = run(p, s)) return er;
    (void)hipEventRecord(e1, s);
    if (hipError_t er = hipEventSynchronize(e1)) return er;
    (void)hipEventElapsedTime(&ms, e0, e1);
    ms /= reps;
    return hipSuccess;
  };
  // a candidate whose first (warm-up) run is already well behind the best so far is
  // dropped without the timed runs, and so are the tile's smaller grids after it (a
  // tile's first run may include loading its kernel: a wider margin there)
  constexpr float PRUNE = 1.15f, PRUNE_FIRST = 1.5f;
  float best = 1e30f;
  int best_tile = 64, best_grid = cus;
  auto t0 = std::chrono::steady_clock::now();
  for (int tile : tiles) {
    p.tile = tile;
    if ((e = build_device_tables(p)) != hipSuccess) break;
    for (int g : grids_per_cu) {
      p.grid = g * cus;
      float warm = 0.f, ms = 0.f;
      if ((e = timed(1, warm)) != hipSuccess) break;
      if (warm > (g == grids_per_cu[0] ? PRUNE_FIRST : PRUNE) * best) break;
      if ((e = timed(2, ms)) != hipSuccess) break;
      if (ms < best) {
        best = ms;
        best_tile = tile;
        best_grid = p.grid;
      }
    }
    if (e != hipSuccess) break;
  }
  // the problem order, at the best tile and grid
  const std::vector<BatchProblem> lpt_scratch = p.host;
  const std::vector<int> lpt_order = p.order;
  int best_sched = 0;
  p.tile = best_tile;
  p.grid = best_grid;
  for (int sched = 1; sched <= 2 && e == hipSuccess; ++sched) {
    apply_schedule(p, lpt_scratch, lpt_order, sched);
    if ((e = build_device_tables(p)) != hipSuccess) break;
    float warm = 0.f, ms = 0.f;
    if ((e = timed(1, warm)) != hipSuccess) break;
    if (warm > PRUNE * best) continue;
    if ((e = timed(2, ms)) != hipSuccess) break;
    if (ms < best) {
      best = ms;
      best_sched = sched;
    }
  }
  p.host = lpt_scratch;
  p.order = lpt_order;
  p.tune_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipStreamSynchronize(s);
  (void)hipFree(scratch_c);
  p.host = saved;
  p.tile = best_tile;
  p.grid = best_grid;
  if (e != hipSuccess) return e;
  apply_schedule(p, saved, lpt_order, best_sched);
  return build_device_tables(p);
}

}  // namespace
}  // namespace pga

using namespace pga;

extern "C" {

pga_status pga_batch_plan_create(pga_batch_plan** out, char transa, char transb, int count, const int* m,
                                 const int* n, const int* k, const uint32_t* alpha, const uint32_t* const* a,
                                 const int* lda, const uint32_t* const* b, const int* ldb, const uint32_t* beta,
                                 uint32_t* const* c, const int* ldc, int tile, int grid) {
  PGA_CHECK_INIT();
  auto up = [](char ch) { return (ch >= 'a' && ch <= 'z') ? (char)(ch - 32) : ch; };
  char A = up(transa), B = up(transb);
  if (A != 'N' && A != 'T' && A != 'C') { set_illegal_argument(2); return PGA_ERROR_ILLEGAL_ARGUMENT; }
  if (B != 'N' && B != 'T' && B != 'C') { set_illegal_argument(3); return PGA_ERROR_ILLEGAL_ARGUMENT; }
  if (count < 0 || !out) { set_illegal_argument(4); return PGA_ERROR_ILLEGAL_ARGUMENT; }
  if (tile != 0 && tile != 16 && tile != 32 && tile != 64) { set_illegal_argument(16); return PGA_ERROR_ILLEGAL_ARGUMENT; }
  auto* p = new pga_batch_plan();
  p->ta = (A != 'N');
  p->tb = (B != 'N');
  // keep problems with work; schedule the most expensive first
  std::vector<int> idx;
  for (int i = 0; i < count; ++i) {
    int nra = p->ta ? k[i] : m[i], nrb = p->tb ? n[i] : k[i];
    if (m[i] < 0 || n[i] < 0 || k[i] < 0 || lda[i] < std::max(1, nra) || ldb[i] < std::max(1, nrb) ||
        ldc[i] < std::max(1, m[i])) {
      delete p;
      set_illegal_argument(5);
      return PGA_ERROR_ILLEGAL_ARGUMENT;
    }
    // the batched kernel covers alpha != 0, k > 0 (other problems are quick returns or
    // pure scalings; they are rejected so that a batch never silently differs from Rgemm)
    if (m[i] == 0 || n[i] == 0) continue;
    if (alpha[i] == 0 || k[i] == 0) {
      if (beta[i] == 0x40000000u) continue;
      delete p;
      set_illegal_argument(8);
      return PGA_ERROR_UNSUPPORTED;
    }
    idx.push_back(i);
  }
  std::stable_sort(idx.begin(), idx.end(), [&](int x, int y) {
    return (double)m[x] * n[x] * k[x] > (double)m[y] * n[y] * k[y];
  });
  for (int i : idx) {
    BatchProblem q;
    q.m = m[i];
    q.n = n[i];
    q.k = k[i];
    q.lda = lda[i];
    q.ldb = ldb[i];
    q.ldc = ldc[i];
    q.alpha = alpha[i];
    q.beta = beta[i];
    q.a = a[i];
    q.b = b[i];
    q.c = c[i];
    p->host.push_back(q);
    p->order.push_back(i);
  }
  p->count = (int)p->host.size();
  hipError_t e;
  if (tile == 0) {
    std::lock_guard<std::mutex> lk(g_tune_mutex);
    load_tune_cache();
    std::string sig = signature(*p);
    auto it = g_tune_cache.find(sig);
    if (it != g_tune_cache.end()) {
      p->tile = it->second.tile;
      p->grid = it->second.grid;
      const std::vector<BatchProblem> lpt = p->host;
      const std::vector<int> lpt_order = p->order;
      apply_schedule(*p, lpt, lpt_order, it->second.sched);
      e = build_device_tables(*p);
    } else {
      e = p->count ? autotune(*p) : hipSuccess;
      if (e == hipSuccess && p->count) {
        g_tune_cache[sig] = {p->tile, p->grid, p->sched};
        save_tune_cache();
      }
    }
  } else {
    p->tile = tile;
    p->grid = grid > 0 ? grid : 2 * runtime().cus;
    e = build_device_tables(*p);
  }
  if (e != hipSuccess) {
    delete p;
    return from_hip(e);
  }
  *out = p;
  return PGA_SUCCESS;
}

pga_status pga_batch_plan_execute(pga_batch_plan* p, void* stream) {
  PGA_CHECK_INIT();
  if (!p) return PGA_ERROR_ILLEGAL_ARGUMENT;
  if (p->count == 0) return PGA_SUCCESS;
  PGA_HIP(run(*p, stream ? (hipStream_t)stream : runtime().stream));
  return PGA_SUCCESS;
}

pga_status pga_batch_plan_capture_graph(pga_batch_plan* p, void* stream) {
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
  hipError_t e = p->count ? run(*p, s) : hipSuccess;
  hipError_t e2 = hipStreamEndCapture(s, &p->graph);
  if (e != hipSuccess) return from_hip(e);
  PGA_HIP(e2);
  PGA_HIP(hipGraphInstantiate(&p->exec, p->graph, nullptr, nullptr, 0));
  return PGA_SUCCESS;
}

pga_status pga_batch_plan_launch_graph(pga_batch_plan* p, void* stream) {
  PGA_CHECK_INIT();
  if (!p || !p->exec) return PGA_ERROR_ILLEGAL_ARGUMENT;
  PGA_HIP(hipGraphLaunch(p->exec, stream ? (hipStream_t)stream : runtime().stream));
  return PGA_SUCCESS;
}

pga_status pga_batch_plan_schedule(const pga_batch_plan* p, int* sched) {
  if (!p) return PGA_ERROR_ILLEGAL_ARGUMENT;
  if (sched) *sched = p->sched;
  return PGA_SUCCESS;
}

pga_status pga_batch_plan_info(const pga_batch_plan* p, int* tile, int* grid, int* total_tiles, double* tune_ms) {
  if (!p) return PGA_ERROR_ILLEGAL_ARGUMENT;
  if (tile) *tile = p->tile;
  if (grid) *grid = p->grid;
  if (total_tiles) *total_tiles = p->total_tiles;
  if (tune_ms) *tune_ms = p->tune_ms;
  return PGA_SUCCESS;
}

pga_status pga_batch_plan_destroy(pga_batch_plan* p) {
  if (!p) return PGA_SUCCESS;
  (void)hipStreamSynchronize(runtime().stream);
  if (p->exec) (void)hipGraphExecDestroy(p->exec);
  if (p->graph) (void)hipGraphDestroy(p->graph);
  if (p->d_probs) (void)hipFree(p->d_probs);
  if (p->d_tile_start) (void)hipFree(p->d_tile_start);
  if (p->d_counter) (void)hipFree(p->d_counter);
  delete p;
  return PGA_SUCCESS;
}

// No-allocation form of pga_batch_plan_create + pga_batch_plan_execute: the same
// problems, the same persistent kernel, the same bits, with the descriptors in the
// library's fixed device array instead of memory allocated per plan.
pga_status pga_batch_gemm_noalloc(char transa, char transb, int count, const int* m, const int* n, const int* k,
                                  const uint32_t* alpha, const uint32_t* const* a, const int* lda,
                                  const uint32_t* const* b, const int* ldb, const uint32_t* beta,
                                  uint32_t* const* c, const int* ldc, int tile, int grid, void* stream) {
  PGA_CHECK_INIT();
  auto up = [](char ch) { return (ch >= 'a' && ch <= 'z') ? (char)(ch - 32) : ch; };
  const char A = up(transa), B = up(transb);
  if (A != 'N' && A != 'T' && A != 'C') { set_illegal_argument(1); return PGA_ERROR_ILLEGAL_ARGUMENT; }
  if (B != 'N' && B != 'T' && B != 'C') { set_illegal_argument(2); return PGA_ERROR_ILLEGAL_ARGUMENT; }
  if (count < 0) { set_illegal_argument(3); return PGA_ERROR_ILLEGAL_ARGUMENT; }
  if (count > PGA_NOALLOC_MAX_PROBLEMS) return PGA_ERROR_EXCEEDS_SCRATCH;
  if (tile != 0 && tile != 16 && tile != 32 && tile != 64) { set_illegal_argument(15); return PGA_ERROR_ILLEGAL_ARGUMENT; }
  const bool ta = (A != 'N'), tb = (B != 'N');
  const int dev = runtime().device;
  if (dev < 0 || dev >= NOALLOC_MAX_DEVICES || !g_noalloc[dev].ptr) return PGA_ERROR_NOT_INITIALIZED;
  NoallocDevice& d = g_noalloc[dev];
  std::lock_guard<std::mutex> lk(g_noalloc_mu);
  // the checks and quick returns of pga_batch_plan_create, problem by problem
  int live = 0;
  for (int i = 0; i < count; ++i) {
    const int nra = ta ? k[i] : m[i], nrb = tb ? n[i] : k[i];
    if (m[i] < 0 || n[i] < 0 || k[i] < 0 || lda[i] < std::max(1, nra) || ldb[i] < std::max(1, nrb) ||
        ldc[i] < std::max(1, m[i])) {
      set_illegal_argument(4);
      return PGA_ERROR_ILLEGAL_ARGUMENT;
    }
    if (m[i] == 0 || n[i] == 0) continue;
    if (alpha[i] == 0 || k[i] == 0) {
      if (beta[i] == 0x40000000u) continue;  // posit32 1.0: C unchanged
      set_illegal_argument(7);
      return PGA_ERROR_UNSUPPORTED;
    }
    g_noalloc_order[live++] = i;
  }
  if (live == 0) return PGA_SUCCESS;
  // most expensive first; ties in the caller's order (std::sort allocates nothing)
  std::sort(g_noalloc_order, g_noalloc_order + live, [&](int x, int y) {
    const double wx = (double)m[x] * n[x] * k[x], wy = (double)m[y] * n[y] * k[y];
    return wx != wy ? wx > wy : x < y;
  });
  if (tile == 0) {  // the largest tile that still gives every CU four tiles
    tile = 16;
    for (int t : {64, 32}) {
      long long tiles = 0;
      for (int j = 0; j < live; ++j) tiles += batched_tiles(m[g_noalloc_order[j]], n[g_noalloc_order[j]], t);
      if (tiles >= 4LL * runtime().cus) {
        tile = t;
        break;
      }
    }
  }
  if (grid <= 0) grid = 2 * runtime().cus;
  hipStream_t s = stream ? (hipStream_t)stream : runtime().stream;
  // the previous call's copies read the staging: wait for them before rewriting it
  if (d.pending) PGA_HIP(hipEventSynchronize(d.copied));
  d.pending = false;
  g_noalloc_start[0] = 0;
  for (int j = 0; j < live; ++j) {
    const int i = g_noalloc_order[j];
    BatchProblem& q = g_noalloc_probs[j];
    q.m = m[i];
    q.n = n[i];
    q.k = k[i];
    q.lda = lda[i];
    q.ldb = ldb[i];
    q.ldc = ldc[i];
    q.alpha = alpha[i];
    q.beta = beta[i];
    q.a = a[i];
    q.b = b[i];
    q.c = c[i];
    g_noalloc_start[j + 1] = g_noalloc_start[j] + batched_tiles(q.m, q.n, tile);
  }
  auto* probs = reinterpret_cast<BatchProblem*>(d.ptr);
  auto* start = reinterpret_cast<int*>(d.ptr + NOALLOC_START_OFF);
  auto* counter = reinterpret_cast<int*>(d.ptr + NOALLOC_COUNTER_OFF);
  PGA_HIP(hipMemcpyAsync(probs, g_noalloc_probs, sizeof(BatchProblem) * live, hipMemcpyHostToDevice, s));
  PGA_HIP(hipMemcpyAsync(start, g_noalloc_start, sizeof(int) * (live + 1), hipMemcpyHostToDevice, s));
  PGA_HIP(hipEventRecord(d.copied, s));
  d.pending = true;
  // the counter is cleared by a kernel of the launch sequence, not by a memset
  PGA_HIP(batched_launch(ta, tb, tile, grid, probs, start, live, g_noalloc_start[live], counter, s));
  return PGA_SUCCESS;
}

}  // extern "C"
