// serve.cpp -- a request server for posit32 solves with continuous batching.
//
// Requests (factor-and-solve of a general or SPD matrix, or a re-solve against a factor
// kept on the device) are admitted in arrival order while their device memory fits an
// HBM budget; resident factors not in use are evicted least-recently-used when memory is
// short.  Admitted requests up to `batch_max_n` are batched: whenever one of the batch
// slots is free, every admitted small request is launched together as one persistent
// task launch (batched_factor.hip) on that slot's stream, so requests that arrive while
// others run join the next launch and the launches overlap on the device.  Larger
// requests are taken by worker threads (a stream each): a worker runs every large
// request ready at that moment as one blocked batch (blocked_batch.hip: all of them
// advance one column block per step, a few launches per step for the whole batch), and
// requests that arrive meanwhile go to the next free worker.  All paths execute the
// reference routines' operations in the reference order, so every result equals the
// per-request plain result bit for bit.  In plain mode one worker runs every request
// alone through the per-call drivers, one after the other: the baseline the batched
// form is measured against.
//
// All device memory of requests and factors comes from one arena allocated at creation
// (its size is the budget), so admission is exact.  Only the workers' small plan
// descriptor buffers are allocated later (grown when a larger batch needs more).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <list>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "pga_internal.hpp"

namespace pga {
namespace {

constexpr size_t ALIGN = 256;
constexpr size_t NPOS = ~(size_t)0;
size_t round_up(size_t x) { return (x + ALIGN - 1) / ALIGN * ALIGN; }

// First-fit allocator over [0, cap) with coalescing free ranges.
class Arena {
 public:
  void init(size_t cap) {
    cap_ = cap;
    free_.clear();
    free_[0] = cap;
    used_ = peak_ = 0;
  }
  size_t alloc(size_t bytes) {
    bytes = round_up(std::max<size_t>(bytes, 1));
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second >= bytes) {
        size_t off = it->first, rest = it->second - bytes;
        free_.erase(it);
        if (rest) free_[off + bytes] = rest;
        used_ += bytes;
        peak_ = std::max(peak_, used_);
        return off;
      }
    }
    return NPOS;
  }
  void release(size_t off, size_t bytes) {
    bytes = round_up(std::max<size_t>(bytes, 1));
    used_ -= bytes;
    auto it = free_.emplace(off, bytes).first;
    auto nx = std::next(it);
    if (nx != free_.end() && it->first + it->second == nx->first) {
      it->second += nx->second;
      free_.erase(nx);
    }
    if (it != free_.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first) {
        pv->second += it->second;
        free_.erase(it);
      }
    }
  }
  size_t used() const { return used_; }
  size_t peak() const { return peak_; }
  size_t cap() const { return cap_; }

 private:
  std::map<size_t, size_t> free_;
  size_t cap_ = 0, used_ = 0, peak_ = 0;
};

double now_s(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

// A factor on the device: A (n x n, factored in place), ipiv (LU), info.
struct Factor {
  int64_t id = 0;
  bool lu = true;
  char uplo = 'U';
  int n = 0;
  size_t off = NPOS, bytes = 0;
  bool ready = false;   // factorization finished
  bool keep = false;    // stays resident after its request
  int info = 0;
  int refs = 0;         // requests using it (in flight or waiting)
  uint64_t last_use = 0;
};

enum State { PENDING, ADMITTED, RUNNING, DONE };

struct Request {
  int64_t id = 0;
  pga_request r{};
  State state = PENDING;
  std::shared_ptr<Factor> f;  // new or resident factor
  size_t sol_off = NPOS, sol_bytes = 0;
  size_t stage_off = NPOS;  // pinned staging of the solution (batched requests)
  bool large = false;
  pga_request_result res{};
};

}  // namespace
}  // namespace pga

using namespace pga;

struct pga_server {
  pga_server_config cfg{};
  std::chrono::steady_clock::time_point t0;
  char* base = nullptr;  // arena memory
  Arena arena;
  char* hbase = nullptr;  // pinned host staging for batched solutions
  Arena harena;
  int* h_info = nullptr;  // pinned info mailboxes, one per request slot
  int info_slots = 0;

  std::mutex mu;
  std::condition_variable cv_sched, cv_work, cv_done;
  bool stop = false;
  int64_t next_id = 1;
  uint64_t tick = 0;
  std::unordered_map<int64_t, std::unique_ptr<Request>> reqs;
  std::unordered_map<int64_t, std::shared_ptr<Factor>> factors;  // resident or in production
  std::deque<Request*> pending;   // arrival order
  std::deque<Request*> small_ready, large_ready;
  double t_last_submit = -1.0;    // seconds from t0
  int inflight = 0;
  pga_server_stats stats{};

  struct Slot {
    hipStream_t s = nullptr;
    hipEvent_t ev = nullptr;
    FactorTask* h_tasks = nullptr;  // pinned
    FactorTask* d_tasks = nullptr;
    int* d_counter = nullptr;
    std::vector<Request*> batch;
    bool busy = false;
    hipError_t err = hipSuccess;  // the launch of `batch` failed to enqueue
  };
  std::vector<Slot> slots;
  std::vector<hipStream_t> worker_streams;  // one per worker thread
  // PGA_SERVE_FAIL_EVERY=k (tests of the failure paths): every k-th launch of the batch
  // slots and of the workers reports an error before it enqueues anything
  int fail_every = 0;
  std::atomic<int64_t> launches{0};
  bool inject_failure() { return fail_every > 0 && ++launches % fail_every == 0; }
  std::vector<std::thread> workers;
  std::thread sched;

  uint32_t* dev(size_t off) { return (uint32_t*)(base + off); }
};

namespace pga {
namespace {

size_t factor_bytes(bool lu, int n) {
  return round_up((size_t)n * n * 4) + (lu ? round_up((size_t)n * 4) : 0) + ALIGN;
}
uint32_t* factor_a(pga_server* S, const Factor& f) { return S->dev(f.off); }
int* factor_ipiv(pga_server* S, const Factor& f) { return (int*)S->dev(f.off + round_up((size_t)f.n * f.n * 4)); }
int* factor_info(pga_server* S, const Factor& f) {
  return (int*)S->dev(f.off + f.bytes - ALIGN);
}

bool is_new(const pga_request& r) { return r.kind == PGA_REQ_LU_SOLVE || r.kind == PGA_REQ_CHOL_SOLVE; }

// Caller holds S->mu.  Frees the request's solve block and, unless kept, its factor.
void finish(pga_server* S, Request* q, int status, int info) {
  const double t = now_s(S->t0);
  q->res.status = status;
  q->res.info = info;
  q->res.t_done = t;
  q->state = DONE;
  if (q->sol_off != NPOS) {
    S->arena.release(q->sol_off, q->sol_bytes);
    q->sol_off = NPOS;
  }
  if (q->stage_off != NPOS) {
    // x is written only by a solve that ran (a factorization with info != 0 skips it)
    if (status == 0 && info == 0 && q->r.nrhs > 0)
      memcpy(q->r.x, S->hbase + q->stage_off, (size_t)q->r.n * q->r.nrhs * 4);
    S->harena.release(q->stage_off, q->sol_bytes);
    q->stage_off = NPOS;
  }
  if (q->f) {
    Factor& f = *q->f;
    if (is_new(q->r)) {
      f.ready = true;
      f.info = info;
      if (status != PGA_REQ_OK) f.keep = false;  // a failed factorization is never re-used
    }
    f.refs--;
    f.last_use = ++S->tick;
    if (!f.keep && f.refs == 0 && f.off != NPOS) {
      S->arena.release(f.off, f.bytes);
      f.off = NPOS;
      S->factors.erase(f.id);
    }
  }
  S->stats.completed++;
  if (status == 0) S->stats.solved++;
  S->inflight--;
  S->cv_done.notify_all();
  S->cv_sched.notify_all();
}

// Evict resident factors nobody uses, least recently used first, until `bytes` fit.
size_t alloc_evicting(pga_server* S, size_t bytes) {
  size_t off = S->arena.alloc(bytes);
  while (off == NPOS) {
    std::shared_ptr<Factor> victim;
    for (auto& kv : S->factors) {
      Factor& f = *kv.second;
      if (f.keep && f.ready && f.refs == 0 && f.off != NPOS && (!victim || f.last_use < victim->last_use))
        victim = kv.second;
    }
    if (!victim) return NPOS;
    S->arena.release(victim->off, victim->bytes);
    victim->off = NPOS;
    S->factors.erase(victim->id);
    S->stats.evictions++;
    off = S->arena.alloc(bytes);
  }
  return off;
}

// Caller holds S->mu.  Admit pending requests in arrival order while memory allows.
void admit(pga_server* S) {
  for (auto it = S->pending.begin(); it != S->pending.end();) {
    Request* q = *it;
    const pga_request& r = q->r;
    if (!is_new(r)) {
      auto fit = S->factors.find(r.factor);
      if (fit == S->factors.end() || fit->second->off == NPOS) {  // unknown or evicted
        it = S->pending.erase(it);
        finish(S, q, PGA_REQ_NOT_RESIDENT, 0);
        continue;
      }
      if (!fit->second->ready) {  // its factorization is still running: later requests may pass
        ++it;
        continue;
      }
      if (fit->second->info != 0) {  // nothing to solve with
        it = S->pending.erase(it);
        finish(S, q, 0, fit->second->info);
        continue;
      }
    }
    const bool lu = (r.kind == PGA_REQ_LU_SOLVE) || (!is_new(r) && S->factors[r.factor]->lu);
    const size_t fb = is_new(r) ? factor_bytes(lu, r.n) : 0;
    const size_t sb = round_up((size_t)r.n * r.nrhs * 4);
    if (fb + sb > S->arena.cap() || sb > S->harena.cap()) {  // can never fit
      it = S->pending.erase(it);
      finish(S, q, PGA_REQ_TOO_LARGE, 0);
      continue;
    }
    const bool large = S->cfg.mode == PGA_SERVER_PLAIN || r.n > S->cfg.batch_max_n;
    size_t foff = NPOS, soff = NPOS, hoff = NPOS;
    if (!large) hoff = S->harena.alloc(sb);
    if (large || hoff != NPOS) {
      if (fb) foff = alloc_evicting(S, fb);
      if (!fb || foff != NPOS) soff = alloc_evicting(S, sb);
    }
    if ((!large && hoff == NPOS) || (fb && foff == NPOS) || soff == NPOS) {
      if (hoff != NPOS) S->harena.release(hoff, sb);
      if (foff != NPOS) S->arena.release(foff, fb);
      break;  // wait for memory; keep arrival order for the rest
    }
    if (is_new(r)) {
      auto f = std::make_shared<Factor>();
      f->id = q->id;
      f->lu = lu;
      f->uplo = S->cfg.chol_uplo;
      f->n = r.n;
      f->off = foff;
      f->bytes = fb;
      f->keep = r.keep_factor != 0;
      f->refs = 1;
      q->f = f;
      q->res.factor = f->id;
      S->factors[f->id] = f;
    } else {
      q->f = S->factors[r.factor];
      q->f->refs++;
      q->res.factor = r.factor;
    }
    q->sol_off = soff;
    q->sol_bytes = sb;
    q->stage_off = hoff;
    q->state = ADMITTED;
    q->res.t_admit = now_s(S->t0);
    q->large = large;
    (q->large ? S->large_ready : S->small_ready).push_back(q);
    it = S->pending.erase(it);
    if (q->large) S->cv_work.notify_one();
  }
  S->stats.bytes_in_use = S->arena.used();
  S->stats.bytes_peak = S->arena.peak();
}

// Enqueue one request's copies and per-call drivers on stream s; returns info.
pga_status run_alone(pga_server* S, Request* q, hipStream_t s, int* info_out) {
  const pga_request& r = q->r;
  Factor& f = *q->f;
  uint32_t* a = factor_a(S, f);
  uint32_t* b = S->dev(q->sol_off);
  int* dinfo = factor_info(S, f);
  int* hinfo = S->h_info + (q->id % S->info_slots);
  const size_t nn = (size_t)r.n * r.n * 4, nb = (size_t)r.n * r.nrhs * 4;
  int info = 0;
  if (is_new(r)) {
    PGA_HIP(hipMemcpyAsync(a, r.a, nn, hipMemcpyHostToDevice, s));
    pga_status st = f.lu ? pga_rgetrf_device(r.n, r.n, a, r.n, factor_ipiv(S, f), dinfo, 64, s)
                         : pga_rpotrf_device(f.uplo, r.n, a, r.n, dinfo, 64, s);
    if (st != PGA_SUCCESS) return st;
    PGA_HIP(hipMemcpyAsync(hinfo, dinfo, sizeof(int), hipMemcpyDeviceToHost, s));
    PGA_HIP(hipStreamSynchronize(s));
    info = *hinfo;
  }
  if (info == 0 && r.nrhs > 0) {
    PGA_HIP(hipMemcpyAsync(b, r.b, nb, hipMemcpyHostToDevice, s));
    pga_status st = f.lu ? pga_rgetrs_device('N', r.n, r.nrhs, a, r.n, factor_ipiv(S, f), b, r.n, s)
                         : pga_rpotrs_device(f.uplo, r.n, r.nrhs, a, r.n, b, r.n, s);
    if (st != PGA_SUCCESS) return st;
    PGA_HIP(hipMemcpyAsync(r.x, b, nb, hipMemcpyDeviceToHost, s));
    PGA_HIP(hipStreamSynchronize(s));
  }
  *info_out = info;
  return PGA_SUCCESS;
}

// Device buffer of a worker's blocked-batch plans (grown, never shrunk while running).
struct PlanBuf {
  void* dev = nullptr;
  size_t bytes = 0;
};

// Run several large requests together as one blocked batch (blocked_batch.hip) on
// stream s: inputs in, all factorizations advancing one column block at a time, the
// solves, solutions and infos out.  infos[i] receives request i's info.
pga_status run_blocked(pga_server* S, const std::vector<Request*>& batch, hipStream_t s, PlanBuf& buf,
                       std::vector<int>& infos) {
  const int cnt = (int)batch.size();
  std::vector<FactorTask> tasks(cnt);
  for (int i = 0; i < cnt; ++i) {
    Request* q = batch[i];
    const pga_request& r = q->r;
    Factor& f = *q->f;
    const bool fresh = is_new(r);
    if (fresh) PGA_HIP(hipMemcpyAsync(factor_a(S, f), r.a, (size_t)r.n * r.n * 4, hipMemcpyHostToDevice, s));
    if (r.nrhs > 0)
      PGA_HIP(hipMemcpyAsync(S->dev(q->sol_off), r.b, (size_t)r.n * r.nrhs * 4, hipMemcpyHostToDevice, s));
    FactorTask& t = tasks[i];
    t.kind = f.lu ? (fresh ? PGA_TASK_GESV : PGA_TASK_GETRS) : (fresh ? PGA_TASK_POSV : PGA_TASK_POTRS);
    t.uplo = f.uplo;
    t.n = r.n;
    t.nrhs = r.nrhs;
    t.nb = 64;
    t.lda = t.ldb = std::max(1, r.n);
    t.a = factor_a(S, f);
    t.ipiv = f.lu ? factor_ipiv(S, f) : nullptr;
    t.b = S->dev(q->sol_off);
    t.info = fresh ? factor_info(S, f) : nullptr;  // a re-solve must not touch its factor's info
  }
  BlockedPlan* p = blocked_plan_build(tasks.data(), cnt);
  const size_t need = blocked_plan_bytes(p);
  hipError_t e = hipSuccess;
  if (need > buf.bytes) {
    if (buf.dev) {
      (void)hipStreamSynchronize(s);
      (void)hipFree(buf.dev);
      buf.dev = nullptr;
      buf.bytes = 0;
    }
    const size_t want = std::max(2 * need, (size_t)1 << 20);
    e = hipMalloc(&buf.dev, want);
    if (e == hipSuccess) buf.bytes = want;
  }
  if (e == hipSuccess) e = blocked_plan_upload(p, buf.dev, s);
  if (e == hipSuccess) e = blocked_plan_launch(p, s);
  for (int i = 0; i < cnt && e == hipSuccess; ++i) {
    Request* q = batch[i];
    if (is_new(q->r))
      e = hipMemcpyAsync(S->h_info + (q->id % S->info_slots), factor_info(S, *q->f), sizeof(int),
                         hipMemcpyDeviceToHost, s);
  }
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  // solutions of the requests whose solve ran (factor-and-solve requests solve only
  // when their factorization succeeded)
  infos.assign(cnt, 0);
  for (int i = 0; i < cnt && e == hipSuccess; ++i) {
    Request* q = batch[i];
    const pga_request& r = q->r;
    infos[i] = is_new(r) ? S->h_info[q->id % S->info_slots] : 0;
    if (infos[i] == 0 && r.nrhs > 0)
      e = hipMemcpy(r.x, S->dev(q->sol_off), (size_t)r.n * r.nrhs * 4, hipMemcpyDeviceToHost);
  }
  blocked_plan_destroy(p);
  return e == hipSuccess ? PGA_SUCCESS : from_hip(e);
}

// Larger requests.  Plain mode: one request at a time through the per-call drivers.
// Batched mode: every ready large request (up to max_batch) runs as one blocked batch;
// requests beyond the blocked path's largest order, and a request that is ready alone,
// run through the per-call drivers.
void worker_main(pga_server* S, hipStream_t s) {
  (void)hipSetDevice(runtime().device);
  PlanBuf buf;
  std::vector<int> infos;
  std::unique_lock<std::mutex> lk(S->mu);
  for (;;) {
    S->cv_work.wait(lk, [&] { return S->stop || !S->large_ready.empty(); });
    if (S->stop) break;
    if (S->cfg.mode != PGA_SERVER_PLAIN) {
      // a burst of submissions still arriving (the last one under 1 ms ago): let it gather
      // for at most 4 ms, so that it forms a few large blocked plans instead of many small
      // ones; requests that arrive alone are taken after at most 1 ms
      const double t_wait0 = now_s(S->t0);
      while (!S->stop && now_s(S->t0) - S->t_last_submit < 1e-3 && now_s(S->t0) - t_wait0 < 4e-3)
        S->cv_work.wait_for(lk, std::chrono::microseconds(500));
      if (S->stop) break;
      if (S->large_ready.empty()) continue;  // another worker took them meanwhile
    }
    Request* first = S->large_ready.front();
    if (S->cfg.mode == PGA_SERVER_PLAIN || first->r.n > PGA_BLOCKED_TASK_MAX_N) {
      S->large_ready.pop_front();
      first->state = RUNNING;
      lk.unlock();
      int info = 0;
      pga_status st = S->inject_failure() ? PGA_ERROR_HIP : run_alone(S, first, s, &info);
      // a failed request's memory returns to the arena: let whatever was queued finish first
      if (st != PGA_SUCCESS) (void)hipStreamSynchronize(s);
      lk.lock();
      finish(S, first, st == PGA_SUCCESS ? 0 : PGA_REQ_FAILED, info);
      continue;
    }
    std::vector<Request*> batch;
    for (auto it = S->large_ready.begin(); it != S->large_ready.end() && (int)batch.size() < S->cfg.max_batch;) {
      if ((*it)->r.n <= PGA_BLOCKED_TASK_MAX_N) {
        (*it)->state = RUNNING;
        batch.push_back(*it);
        it = S->large_ready.erase(it);
      } else {
        ++it;
      }
    }
    if (batch.size() == 1) {  // alone: the per-call drivers (same bits, no plan to build)
      Request* q = batch[0];
      lk.unlock();
      int info = 0;
      pga_status st = S->inject_failure() ? PGA_ERROR_HIP : run_alone(S, q, s, &info);
      if (st != PGA_SUCCESS) (void)hipStreamSynchronize(s);
      lk.lock();
      finish(S, q, st == PGA_SUCCESS ? 0 : PGA_REQ_FAILED, info);
      continue;
    }
    S->stats.batches++;
    S->stats.batched_requests += (int64_t)batch.size();
    lk.unlock();
    pga_status st = S->inject_failure() ? PGA_ERROR_HIP : run_blocked(S, batch, s, buf, infos);
    if (st != PGA_SUCCESS) (void)hipStreamSynchronize(s);
    lk.lock();
    for (size_t i = 0; i < batch.size(); ++i)
      finish(S, batch[i], st == PGA_SUCCESS ? 0 : PGA_REQ_FAILED, st == PGA_SUCCESS ? infos[i] : 0);
  }
  lk.unlock();
  (void)hipStreamSynchronize(s);
  if (buf.dev) (void)hipFree(buf.dev);
}

// Launch every admitted small request as one persistent launch on a free slot.  Returns
// the first HIP error; after an error nothing more is enqueued and the work already
// queued has finished, so the requests can be failed and their memory released.
hipError_t launch_batch(pga_server* S, pga_server::Slot& sl) {
  hipStream_t s = sl.s;
  const int cnt = (int)sl.batch.size();
  if (S->inject_failure()) return hipErrorLaunchFailure;
  hipError_t e = hipSuccess;
  for (int i = 0; i < cnt && e == hipSuccess; ++i) {
    Request* q = sl.batch[i];
    const pga_request& r = q->r;
    Factor& f = *q->f;
    const bool fresh = is_new(r);
    if (fresh) e = hipMemcpyAsync(factor_a(S, f), r.a, (size_t)r.n * r.n * 4, hipMemcpyHostToDevice, s);
    if (e == hipSuccess && r.nrhs > 0)
      e = hipMemcpyAsync(S->dev(q->sol_off), r.b, (size_t)r.n * r.nrhs * 4, hipMemcpyHostToDevice, s);
    FactorTask& t = sl.h_tasks[i];
    t = FactorTask{};
    t.kind = f.lu ? (fresh ? PGA_TASK_GESV : PGA_TASK_GETRS) : (fresh ? PGA_TASK_POSV : PGA_TASK_POTRS);
    t.uplo = f.uplo;
    t.n = r.n;
    t.nrhs = r.nrhs;
    t.nb = 64;
    t.lda = std::max(1, r.n);
    t.ldb = std::max(1, r.n);
    t.a = factor_a(S, f);
    t.ipiv = f.lu ? factor_ipiv(S, f) : nullptr;
    t.b = S->dev(q->sol_off);
    t.info = factor_info(S, f);
    if (!fresh) t.info = nullptr;
  }
  if (e == hipSuccess)
    e = hipMemcpyAsync(sl.d_tasks, sl.h_tasks, sizeof(FactorTask) * cnt, hipMemcpyHostToDevice, s);
  if (e == hipSuccess) e = factor_tasks_launch(sl.d_tasks, cnt, sl.d_counter, factor_tasks_grid(cnt), s);
  for (int i = 0; i < cnt && e == hipSuccess; ++i) {
    Request* q = sl.batch[i];
    const pga_request& r = q->r;
    if (r.nrhs > 0)
      e = hipMemcpyAsync(S->hbase + q->stage_off, S->dev(q->sol_off), (size_t)r.n * r.nrhs * 4,
                         hipMemcpyDeviceToHost, s);
    if (e == hipSuccess && is_new(r))
      e = hipMemcpyAsync(S->h_info + (q->id % S->info_slots), factor_info(S, *q->f), sizeof(int),
                         hipMemcpyDeviceToHost, s);
  }
  if (e == hipSuccess) e = hipEventRecord(sl.ev, s);
  if (e != hipSuccess) (void)hipStreamSynchronize(s);
  return e;
}

void sched_main(pga_server* S) {
  (void)hipSetDevice(runtime().device);
  std::unique_lock<std::mutex> lk(S->mu);
  while (!S->stop) {
    // completions; a launch that failed to enqueue, or whose event reports an error
    // other than not-ready, finishes its requests as failed
    for (auto& sl : S->slots) {
      if (!sl.busy) continue;
      hipError_t done = sl.err;
      if (done == hipSuccess) {
        done = hipEventQuery(sl.ev);
        if (done == hipErrorNotReady) continue;
      }
      for (Request* q : sl.batch) {
        if (done != hipSuccess) {
          finish(S, q, PGA_REQ_FAILED, 0);
          continue;
        }
        // a factor-and-solve writes the info of its new factor; a re-solve reports
        // the resident factor's info (zero, or it would not have been launched)
        const int info = is_new(q->r) ? S->h_info[q->id % S->info_slots] : q->f->info;
        finish(S, q, PGA_REQ_OK, info);
      }
      sl.batch.clear();
      sl.busy = false;
      sl.err = hipSuccess;
    }
    admit(S);
    // launches
    for (auto& sl : S->slots) {
      if (sl.busy || S->small_ready.empty()) continue;
      const int take = std::min<int>((int)S->small_ready.size(), S->cfg.max_batch);
      sl.batch.assign(S->small_ready.begin(), S->small_ready.begin() + take);
      S->small_ready.erase(S->small_ready.begin(), S->small_ready.begin() + take);
      for (Request* q : sl.batch) q->state = RUNNING;
      sl.busy = true;
      S->stats.batches++;
      S->stats.batched_requests += take;
      lk.unlock();
      const hipError_t e = launch_batch(S, sl);
      lk.lock();
      sl.err = e;
    }
    // Every wait releases the mutex: workers need it to finish their requests, which is
    // what blocked requests (memory, a factor still in production) are waiting for.
    // Batch completions are polled; everything else notifies cv_sched.
    bool busy = false;
    for (auto& sl : S->slots) busy = busy || sl.busy;
    if (busy) S->cv_sched.wait_for(lk, std::chrono::microseconds(50));
    else if (!S->pending.empty()) S->cv_sched.wait_for(lk, std::chrono::milliseconds(1));
    else S->cv_sched.wait(lk, [&] { return S->stop || !S->pending.empty(); });
  }
}

}  // namespace
}  // namespace pga

extern "C" {

pga_status pga_server_create(pga_server** out, const pga_server_config* cfg) {
  PGA_CHECK_INIT();
  if (!out || !cfg) return PGA_ERROR_ILLEGAL_ARGUMENT;
  pga_server_config c = *cfg;
  if (c.hbm_budget_bytes < (1u << 20)) return PGA_ERROR_ILLEGAL_ARGUMENT;
  if (c.mode != PGA_SERVER_BATCHED && c.mode != PGA_SERVER_PLAIN) return PGA_ERROR_ILLEGAL_ARGUMENT;
  if (c.batch_max_n < 0 || c.batch_max_n > PGA_FACTOR_TASK_MAX_N) return PGA_ERROR_ILLEGAL_ARGUMENT;
  if (c.chol_uplo == 'u') c.chol_uplo = 'U';
  if (c.chol_uplo == 'l') c.chol_uplo = 'L';
  if (c.chol_uplo != 'U' && c.chol_uplo != 'L') return PGA_ERROR_ILLEGAL_ARGUMENT;
  if (c.max_batch <= 0) c.max_batch = 1024;
  if (c.batch_slots <= 0) c.batch_slots = 2;
  if (c.workers <= 0) c.workers = 4;
  // concurrency this server has been run with on the device
  if (c.batch_slots > 2 || c.workers > 8) return PGA_ERROR_ILLEGAL_ARGUMENT;
  if (c.mode == PGA_SERVER_PLAIN) {
    c.batch_slots = 0;
    c.workers = 1;
  }
  auto* S = new pga_server();
  S->cfg = c;
  if (const char* fe = getenv("PGA_SERVE_FAIL_EVERY")) S->fail_every = atoi(fe);
  S->t0 = std::chrono::steady_clock::now();
  hipError_t e = hipMalloc((void**)&S->base, c.hbm_budget_bytes);
  if (e != hipSuccess) {
    delete S;
    return from_hip(e);
  }
  S->arena.init(c.hbm_budget_bytes);
  S->info_slots = 1 << 16;
  const size_t staging = c.staging_bytes ? c.staging_bytes : ((size_t)256 << 20);
  if ((e = hipHostMalloc((void**)&S->h_info, sizeof(int) * S->info_slots, 0)) != hipSuccess ||
      (e = hipHostMalloc((void**)&S->hbase, staging, 0)) != hipSuccess) {
    if (S->h_info) (void)hipHostFree(S->h_info);
    (void)hipFree(S->base);
    delete S;
    return from_hip(e);
  }
  S->harena.init(staging);
  S->slots.resize(c.batch_slots);
  for (auto& sl : S->slots) {
    if ((e = hipStreamCreateWithFlags(&sl.s, hipStreamNonBlocking)) != hipSuccess ||
        (e = hipEventCreateWithFlags(&sl.ev, hipEventDisableTiming)) != hipSuccess ||
        (e = hipHostMalloc((void**)&sl.h_tasks, sizeof(FactorTask) * c.max_batch, 0)) != hipSuccess ||
        (e = hipMalloc((void**)&sl.d_tasks, sizeof(FactorTask) * c.max_batch)) != hipSuccess ||
        (e = hipMalloc((void**)&sl.d_counter, sizeof(int))) != hipSuccess) {
      (void)hipGetLastError();
      pga_server_destroy(S);
      return from_hip(e);
    }
  }
  S->worker_streams.assign(c.workers, nullptr);
  for (auto& ws : S->worker_streams) {
    if ((e = hipStreamCreateWithFlags(&ws, hipStreamNonBlocking)) != hipSuccess) {
      (void)hipGetLastError();
      pga_server_destroy(S);
      return from_hip(e);
    }
  }
  S->sched = std::thread(sched_main, S);
  for (int i = 0; i < c.workers; ++i) S->workers.emplace_back(worker_main, S, S->worker_streams[i]);
  *out = S;
  return PGA_SUCCESS;
}

pga_status pga_server_submit(pga_server* S, const pga_request* r, int64_t* id) {
  if (!S || !r || !id) return PGA_ERROR_ILLEGAL_ARGUMENT;
  const bool fresh = r->kind == PGA_REQ_LU_SOLVE || r->kind == PGA_REQ_CHOL_SOLVE;
  if (r->kind != PGA_REQ_LU_SOLVE && r->kind != PGA_REQ_CHOL_SOLVE && r->kind != PGA_REQ_RESOLVE)
    return PGA_ERROR_ILLEGAL_ARGUMENT;
  if (r->n < 1 || r->nrhs < 0 || (fresh && !r->a) || (r->nrhs > 0 && (!r->b || !r->x)))
    return PGA_ERROR_ILLEGAL_ARGUMENT;
  std::lock_guard<std::mutex> lk(S->mu);
  if (!fresh) {
    auto it = S->factors.find(r->factor);
    if (it != S->factors.end() && it->second->n != r->n) return PGA_ERROR_ILLEGAL_ARGUMENT;
  }
  if (S->inflight + 1 >= S->info_slots) return PGA_ERROR_UNSUPPORTED;
  auto q = std::make_unique<Request>();
  q->id = S->next_id++;
  q->r = *r;
  q->res.id = q->id;
  q->res.factor = fresh ? 0 : r->factor;
  q->res.t_submit = now_s(S->t0);
  *id = q->id;
  S->pending.push_back(q.get());
  S->t_last_submit = q->res.t_submit;
  S->reqs[q->id] = std::move(q);
  S->inflight++;
  S->stats.submitted++;
  S->cv_sched.notify_all();
  return PGA_SUCCESS;
}

pga_status pga_server_wait(pga_server* S, int64_t id, pga_request_result* out) {
  if (!S) return PGA_ERROR_ILLEGAL_ARGUMENT;
  std::unique_lock<std::mutex> lk(S->mu);
  auto it = S->reqs.find(id);
  if (it == S->reqs.end()) return PGA_ERROR_ILLEGAL_ARGUMENT;
  Request* q = it->second.get();
  S->cv_done.wait(lk, [&] { return q->state == DONE; });
  if (out) *out = q->res;
  return PGA_SUCCESS;
}

pga_status pga_server_drain(pga_server* S) {
  if (!S) return PGA_ERROR_ILLEGAL_ARGUMENT;
  std::unique_lock<std::mutex> lk(S->mu);
  S->cv_done.wait(lk, [&] { return S->inflight == 0; });
  return PGA_SUCCESS;
}

pga_status pga_server_release_factor(pga_server* S, int64_t factor) {
  if (!S) return PGA_ERROR_ILLEGAL_ARGUMENT;
  std::lock_guard<std::mutex> lk(S->mu);
  auto it = S->factors.find(factor);
  if (it == S->factors.end()) return PGA_SUCCESS;
  Factor& f = *it->second;
  f.keep = false;
  if (f.refs == 0 && f.ready && f.off != NPOS) {
    S->arena.release(f.off, f.bytes);
    f.off = NPOS;
    S->factors.erase(it);
  }
  S->cv_sched.notify_all();
  return PGA_SUCCESS;
}

pga_status pga_server_get_stats(pga_server* S, pga_server_stats* out) {
  if (!S || !out) return PGA_ERROR_ILLEGAL_ARGUMENT;
  std::lock_guard<std::mutex> lk(S->mu);
  *out = S->stats;
  out->bytes_in_use = S->arena.used();
  out->bytes_peak = S->arena.peak();
  int resident = 0;
  for (auto& kv : S->factors) resident += (kv.second->ready && kv.second->off != NPOS) ? 1 : 0;
  out->resident_factors = resident;
  return PGA_SUCCESS;
}

pga_status pga_server_destroy(pga_server* S) {
  if (!S) return PGA_SUCCESS;
  {
    std::unique_lock<std::mutex> lk(S->mu);
    S->cv_done.wait(lk, [&] { return S->inflight == 0; });
    S->stop = true;
  }
  S->cv_sched.notify_all();
  S->cv_work.notify_all();
  if (S->sched.joinable()) S->sched.join();
  for (auto& t : S->workers)
    if (t.joinable()) t.join();
  for (hipStream_t ws : S->worker_streams)
    if (ws) (void)hipStreamDestroy(ws);
  for (auto& sl : S->slots) {
    if (sl.s) (void)hipStreamSynchronize(sl.s);
    if (sl.s) (void)hipStreamDestroy(sl.s);
    if (sl.ev) (void)hipEventDestroy(sl.ev);
    if (sl.h_tasks) (void)hipHostFree(sl.h_tasks);
    if (sl.d_tasks) (void)hipFree(sl.d_tasks);
    if (sl.d_counter) (void)hipFree(sl.d_counter);
  }
  if (S->h_info) (void)hipHostFree(S->h_info);
  if (S->hbase) (void)hipHostFree(S->hbase);
  if (S->base) (void)hipFree(S->base);
  delete S;
  return PGA_SUCCESS;
}

}  // extern "C"
