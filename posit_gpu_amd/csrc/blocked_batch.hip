// blocked_batch.hip -- batched LU / Cholesky factorizations and solves of problems too
// large for one workgroup.
//
// Every problem follows the blocking of the per-call driver it mirrors
// (Rgetrf_opencl.cpp:93-146, Rpotrf_opencl.cpp:86-160) and all problems advance together,
// one column block per step.  A step is at most five launches for the whole batch:
//   1. the Cholesky diagonal-block SYRKs (one GEMM launch over all their tiles),
//   2. the panels: Rpotrf2 on each Cholesky diagonal block by one workgroup (wg_factor.hpp,
//      the code of the one-workgroup task kernel); the LU column blocks walked through the
//      Rgetrf2 recursion node by node for all problems at once (two-column nodes in one
//      launch, lu_leaf.hpp; larger nodes' GEMMs as one launch over every problem's tiles),
//   3. the LU row-block triangular solves (U12 := L11^-1 A12), a few columns per
//      workgroup, each column first taking the block's row interchanges (Rlaswp),
//   4. the trailing GEMMs of every problem (LU A22 -= A21 U12; Cholesky panel updates),
//   5. the Cholesky panel triangular solves.
// The columns left of each LU block take that block's interchanges at the end, in one
// pass (nothing reads them in between, as in the reference).  The solves of all problems
// follow, one workgroup per right-hand side.  Each element
// receives the reference routines' operations in their order (the GEMM tiles are the
// plain kernel's), so every factor, pivot, info code and solution equals the per-call
// result bit for bit.  A failing Cholesky block sets info, which stops every later step
// of that problem exactly where the reference stops.  The plan deals its tasks into
// groups that run the sequence side by side on their own streams (blocked_plan_launch).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <mutex>
#include <vector>

#include "lu_leaf.hpp"
#include "pga_internal.hpp"
#include "wg_factor.hpp"

namespace pga {
namespace bb {

constexpr int NT = wg::NT;
constexpr int RECUR_W = 64;    // LU panels of this width are walked leaf by leaf for all problems
constexpr int TRSM_COLS = 4;   // columns per workgroup of the row-block solves (63 x 4 pairs)
constexpr int TRSM_ROWS = 64;  // rows per workgroup of the lower-Cholesky panel solves

// A GEMM of one problem at one step: C = alpha op(A) op(B) + beta C in Rgemm's order.
struct Gemm {
  int m, n, k, lda, ldb, ldc;
  int ta, tb, skip0, tri;
  uint32_t alpha, beta;
  const uint32_t* a;
  const uint32_t* b;
  uint32_t* c;
  const int* stop;  // nonzero: the problem stopped (a Cholesky block failed)
  int task;         // owning task (the plan points stop at its info when uploaded)
};

// A unit of panel / solve work: task, column block (j, jb), and a column (or row) range.
struct Item {
  int task, j, jb, off, cnt;
};

__device__ __forceinline__ bool is_lu(int kind) {
  return kind == PGA_TASK_GETRF || kind == PGA_TASK_GESV || kind == PGA_TASK_GETRS;
}

// info := 0 for every task (the per-call drivers start that way; a solve reports 0), and
// the tile counters of the group's launches := 0
__global__ void k_bb_init(const FactorTask* __restrict__ tasks, int ntasks, int* __restrict__ counters,
                          int ncounters) {
  PGA_KERNEL_VGPR_FLOOR();
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < ntasks) *tasks[i].info = 0;
  if (i < ncounters) counters[i] = 0;
}

// Panels: Rgetrf2 on columns j..j+jb-1 (rows j..n-1), or Rpotrf2 on the diagonal block.
__global__ __launch_bounds__(NT, 2) void k_bb_panel(const FactorTask* __restrict__ tasks,
                                                    const Item* __restrict__ items) {
  PGA_KERNEL_VGPR_FLOOR();
  PGA_URGENT();
  __shared__ wg::Shared sh;
  const Item it = items[blockIdx.x];
  const FactorTask& t = tasks[it.task];
  if (threadIdx.x == 0) sh.bcast = *t.info;
  __syncthreads();
  int info = sh.bcast;
  __syncthreads();
  if (is_lu(t.kind)) {
    wg::lu_panel(t.n, t.a, t.lda, it.j, it.jb, t.ipiv, info, sh, true);
    if (threadIdx.x == 0) *t.info = info;
    return;
  }
  if (info) return;  // an earlier block failed: the reference has stopped
  uint32_t* d = t.a + it.j + (size_t)it.j * t.lda;
  const int f = t.uplo == 'U' ? wg::potrf2<true>(it.jb, d, t.lda, sh) : wg::potrf2<false>(it.jb, d, t.lda, sh);
  if (f && threadIdx.x == 0) *t.info = it.j + f;
}

// Triangular solves on a block: LU U12 := L11^-1 U12 (Rtrsm LLNU, columns), upper
// Cholesky U12 := U11^-T U12 (LUTN, columns), lower Cholesky L21 := L21 L11^-T (RLTN, rows).
__global__ __launch_bounds__(NT) void k_bb_trsm(const FactorTask* __restrict__ tasks,
                                                const Item* __restrict__ items) {
  PGA_KERNEL_VGPR_FLOOR();
  PGA_URGENT();
  __shared__ uint32_t sA[64 * 64];  // the diagonal block for the one-wave-per-column solves
  const Item it = items[blockIdx.x];
  const FactorTask& t = tasks[it.task];
  const int lda = t.lda, j = it.j, jb = it.jb;
  uint32_t* d = t.a + j + (size_t)j * lda;
  if (is_lu(t.kind)) {
    // Rlaswp(k1 = j+1, k2 = j+jb) on these columns (Rgetrf_opencl.cpp:125-127)
    if (threadIdx.x < it.cnt) {
      uint32_t* c = t.a + (size_t)(j + jb + it.off + threadIdx.x) * lda;
      for (int k = j; k < j + jb; ++k) {
        const int p = t.ipiv[k] - 1;
        if (p != k) {
          const uint32_t x = c[k];
          c[k] = c[p];
          c[p] = x;
        }
      }
    }
    __syncthreads();
    if (jb <= 64)
      wg::trsm_waves<wg::WS_LLNU>(jb, it.cnt, d, lda, d + (size_t)(jb + it.off) * lda, lda, (wg::lds_u32*)sA);
    else wg::trsm_llnu(jb, it.cnt, d, lda, d + (size_t)(jb + it.off) * lda, lda);
    return;
  }
  if (*t.info) return;
  if (t.uplo == 'U') {
    if (jb <= 64)
      wg::trsm_waves<wg::WS_LUTN>(jb, it.cnt, d, lda, d + (size_t)(jb + it.off) * lda, lda, (wg::lds_u32*)sA);
    else wg::trsm_lutn(jb, it.cnt, d, lda, d + (size_t)(jb + it.off) * lda, lda);
  } else {
    wg::trsm_rltn(it.cnt, jb, d, lda, d + jb + it.off, lda);
  }
}

// LU panels of RECUR_W columns, all problems at once, walked node by node in the
// reference's depth-first order (the per-call driver runs the same steps one problem at a
// time).
// Two-column nodes [jr, jr+2) of every problem's panel: leaf jr, the node's 1 x 1 Rgemm,
// leaf jr+1, then the Rtrsm of the node whose left half ends at jr+1 (lu_leaf.hpp).
constexpr int PAIR_NT = 512;

__global__ __launch_bounds__(PAIR_NT) void k_bb_pair(const FactorTask* __restrict__ tasks,
                                                     const Item* __restrict__ items, int jr, int ts, int tm, int te) {
  PGA_KERNEL_VGPR_FLOOR();
  PGA_URGENT();
  const Item it = items[blockIdx.x];
  const FactorTask& t = tasks[it.task];
  uint32_t* root = t.a + it.j + (size_t)it.j * t.lda;
  leaf::getrf2_pair<RECUR_W / 2, RECUR_W / 2, PAIR_NT>(t.n - it.j, it.jb, root, t.lda, jr, t.ipiv + it.j, t.info,
                                                       it.j, it.j, ts, tm, te);
}

// The node's Rgemm (A22 -= A21 A12, Rgetrf2.cpp:155-159) of every problem's panel in one
// launch of T x T tiles (workgroups take tiles from a counter).
template <int T>
__global__ __launch_bounds__(NT, (gemm::square_tile_wgs_per_cu<T>())) void k_bb_node_gemm(const FactorTask* __restrict__ tasks,
                                                     const Item* __restrict__ items,
                                                     const int* __restrict__ tile_start, int nitems, int total,
                                                     int* __restrict__ counter, int s, int mid, int e) {
  PGA_KERNEL_VGPR_FLOOR();
  __shared__ __align__(16) uint32_t lds[gemm::square_tile_lds_words<T>()];
  __shared__ int s_tile;
  for (;;) {
    if (threadIdx.x == 0) s_tile = atomicAdd(counter, 1);
    __syncthreads();
    const int t = s_tile;
    __syncthreads();
    if (t >= total) return;
    int lo = 0, hi = nitems - 1;
    while (lo < hi) {
      const int md = (lo + hi + 1) >> 1;
      if (tile_start[md] <= t) lo = md;
      else hi = md - 1;
    }
    const Item it = items[lo];
    const FactorTask& tk = tasks[it.task];
    const int lda = tk.lda, j = it.j;
    const int m = tk.n - j - mid, n2 = e - mid, k = mid - s;
    const uint32_t* A = tk.a + (j + mid) + (size_t)(j + s) * lda;
    const uint32_t* B = tk.a + (j + s) + (size_t)(j + mid) * lda;
    uint32_t* C = tk.a + (j + mid) + (size_t)(j + mid) * lda;
    const int tmr = (m + T - 1) / T, tl = t - tile_start[lo];
    const int i0 = (tl % tmr) * T, j0 = (tl / tmr) * T;
    if constexpr (T == 64) {
      for (int jh = j0; jh < min(j0 + 64, n2); jh += 32)  // two 64 x 32 halves
        gemm::gemm_tile_in<64, 32, 16, 4, 2, false, false, false>(m, n2, k, 0xC0000000u, A, lda, B, lda, 0x40000000u,
                                                                  C, lda, i0, jh, 0, lds);
    } else {
      gemm::gemm_tile_in<T, T, 16, T / 16, T / 16, false, false, false>(m, n2, k, 0xC0000000u, A, lda, B, lda,
                                                                      0x40000000u, C, lda, i0, j0, 0, lds);
    }
  }
}

// The interchanges of every later block on the LU columns of one block, [off, off + cnt)
// (Rgetrf_opencl.cpp:123 applies them block by block to the columns on the left; item.j
// is the task's block size).  The columns of a block all take the same interchanges,
// rows k0 = (off / j + 1) * j .. n - 1 in order: lane 0 composes them once into a
// permutation of those rows (LDS), and every column is then permuted by a gather.
constexpr int LASWP_LEFT_REGS = PGA_BLOCKED_TASK_MAX_N / NT;  // gathered rows per lane
static_assert(PGA_BLOCKED_TASK_MAX_N <= 65536 && PGA_BLOCKED_TASK_MAX_N % NT == 0, "16-bit row indices");

__global__ __launch_bounds__(NT) void k_bb_laswp_left(const FactorTask* __restrict__ tasks,
                                                      const Item* __restrict__ items) {
  PGA_KERNEL_VGPR_FLOOR();
  PGA_URGENT();
  extern __shared__ uint16_t perm[];  // row k0 + r of the result is row k0 + perm[r] before
  const Item it = items[blockIdx.x];
  const FactorTask& t = tasks[it.task];
  const int tid = threadIdx.x, nb = it.j, k0 = (it.off / nb + 1) * nb, R = t.n - k0;
  if (R <= 0) return;
  for (int r = tid; r < R; r += NT) perm[r] = (uint16_t)r;
  __syncthreads();
  if (tid < 64) {  // wave 0 reads the pivots 64 at a time, lane 0 applies them in order
    for (int i0 = k0; i0 < t.n; i0 += 64) {
      const int p = i0 + tid < t.n ? t.ipiv[i0 + tid] - 1 : 0;
#pragma unroll
      for (int j = 0; j < 64; ++j) {
        const int k = i0 + j, pk = __shfl(p, j);
        if (tid == 0 && k < t.n && pk != k && pk >= k0 && pk < t.n) {
          const uint16_t x = perm[k - k0];
          perm[k - k0] = perm[pk - k0];
          perm[pk - k0] = x;
        }
      }
    }
  }
  __syncthreads();
  for (int c = it.off; c < it.off + it.cnt; ++c) {
    uint32_t* col = t.a + (size_t)c * t.lda + k0;
    uint32_t v[LASWP_LEFT_REGS];
#pragma unroll
    for (int q = 0; q < LASWP_LEFT_REGS; ++q) {
      const int r = q * NT + tid;
      v[q] = r < R ? col[perm[r]] : 0u;
    }
    __syncthreads();  // the column's reads are done before any lane writes it
#pragma unroll
    for (int q = 0; q < LASWP_LEFT_REGS; ++q) {
      const int r = q * NT + tid;
      if (r < R) col[r] = v[q];
    }
  }
}

// GEMM over the tiles of many problems, workgroups taking tiles from a counter: T x T tiles (64: 4 x 4 elements per
// lane; 32: 2 x 2; 16: 1 x 1), the size chosen per launch so that the tiles fill the
// device (a step of a few problems has few large tiles; K is never split).
template <int T, bool TA, bool TB, bool SKIP0>
__device__ __forceinline__ void bb_tile(const Gemm& p, int i0, int j0, uint32_t* lds) {
  if constexpr (T == 64) {
    // two 64 x 32 halves (4 x 2 per lane: three waves per SIMD)
    for (int jh = j0; jh < min(j0 + 64, p.n); jh += 32) {
      if (p.tri == 1 && i0 + 63 < jh) continue;
      gemm::gemm_tile_in<64, 32, 16, 4, 2, TA, TB, SKIP0>(p.m, p.n, p.k, p.alpha, p.a, p.lda, p.b, p.ldb, p.beta, p.c,
                                                          p.ldc, i0, jh, p.tri, lds);
    }
  } else {
    gemm::gemm_tile_in<T, T, 16, T / 16, T / 16, TA, TB, SKIP0>(p.m, p.n, p.k, p.alpha, p.a, p.lda, p.b, p.ldb,
                                                                p.beta, p.c, p.ldc, i0, j0, p.tri, lds);
  }
}

template <int T>
__global__ __launch_bounds__(NT, (gemm::square_tile_wgs_per_cu<T>())) void k_bb_gemm(const Gemm* __restrict__ probs, const int* __restrict__ tile_start,
                                                int nprob, int total, int* __restrict__ counter) {
  PGA_KERNEL_VGPR_FLOOR();
  __shared__ __align__(16) uint32_t lds[gemm::square_tile_lds_words<T>()];
  __shared__ int s_tile;
  for (;;) {
    if (threadIdx.x == 0) s_tile = atomicAdd(counter, 1);
    __syncthreads();
    const int t = s_tile;
    __syncthreads();
    if (t >= total) return;  // every workgroup ends here once the tiles are taken
    int lo = 0, hi = nprob - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (tile_start[mid] <= t) lo = mid;
      else hi = mid - 1;
    }
    const Gemm& p = probs[lo];
    if (p.stop && *p.stop) continue;  // the whole workgroup reads the same flag
    const int tm = (p.m + T - 1) / T;
    const int i0 = ((t - tile_start[lo]) % tm) * T, j0 = ((t - tile_start[lo]) / tm) * T;
    if (p.tri == 1 && i0 + T - 1 < j0) continue;
    if (p.tri == 2 && j0 + T - 1 < i0) continue;
    if (!p.ta && !p.tb) bb_tile<T, false, false, false>(p, i0, j0, lds);
    else if (p.ta && !p.tb) bb_tile<T, true, false, false>(p, i0, j0, lds);
    else if (p.skip0) bb_tile<T, false, true, true>(p, i0, j0, lds);
    else bb_tile<T, false, true, false>(p, i0, j0, lds);
  }
}

// Solves, one workgroup per right-hand side or pair of them, the columns staged in LDS:
// Rgetrs 'N' (Rlaswp, L, U), Rpotrs upper (U^T, U) or lower (L, L^T).  The L, U and U^T
// sweeps run in FP64 with the column in registers (wg::col_solve_fp64, 64 rows of the
// sweep per barrier), falling back to the pattern-level sweeps (1024 lanes: each step
// updates up to n rows of the column in LDS at once) when a value leaves the FP64 range.
constexpr int SOLVE_NT = 1024;
static_assert(PGA_BLOCKED_TASK_MAX_N <= wg::COL_SOLVE_SLOTS * SOLVE_NT, "a column fits the solve's registers");

template <int MODE, int C>
__device__ __forceinline__ void bb_sweep(int n, const uint32_t* a, int lda, uint32_t* col, wg::lds_f64* sb) {
  if (wg::col_solve_fp64<MODE, SOLVE_NT, C>(n, a, lda, (wg::lds_u32*)col, n, sb)) return;
  if (MODE == wg::CS_LLNU) wg::trsm_llnu<SOLVE_NT>(n, C, a, lda, col, n);
  else if (MODE == wg::CS_LLNN) wg::trsm_llnn<SOLVE_NT>(n, C, a, lda, col, n);
  else if (MODE == wg::CS_LUNN) wg::trsm_lunn<SOLVE_NT>(n, C, a, lda, col, n);
  else wg::trsm_lutn<SOLVE_NT>(n, C, a, lda, col, n);
}

// the solves of C columns held in col (LDS, n apart)
template <int C>
__device__ __forceinline__ void bb_solve_cols(const FactorTask& t, int n, uint32_t* col, wg::lds_f64* sb) {
  if (is_lu(t.kind)) {
    // Rlaswp(1, n, ipiv, 1): the interchanges in order by lane 0, wave 0 reading the
    // pivots 64 at a time
    if (threadIdx.x < 64) {
      for (int i0 = 0; i0 < n; i0 += 64) {
        const int p = i0 + (int)threadIdx.x < n ? t.ipiv[i0 + threadIdx.x] - 1 : 0;
#pragma unroll
        for (int j = 0; j < 64; ++j) {
          const int i = i0 + j, pj = __shfl(p, j);
          if (threadIdx.x == 0 && i < n && pj != i) {
#pragma unroll
            for (int c = 0; c < C; ++c) {
              const uint32_t x = col[c * n + i];
              col[c * n + i] = col[c * n + pj];
              col[c * n + pj] = x;
            }
          }
        }
      }
    }
    __syncthreads();
    bb_sweep<wg::CS_LLNU, C>(n, t.a, t.lda, col, sb);
    bb_sweep<wg::CS_LUNN, C>(n, t.a, t.lda, col, sb);
  } else if (t.uplo == 'U') {
    bb_sweep<wg::CS_LUTN, C>(n, t.a, t.lda, col, sb);
    bb_sweep<wg::CS_LUNN, C>(n, t.a, t.lda, col, sb);
  } else {
    bb_sweep<wg::CS_LLNN, C>(n, t.a, t.lda, col, sb);
    wg::trsm_lltn<SOLVE_NT>(n, C, t.a, t.lda, col, n);
  }
}

// Items: (task, first column off, cnt = 1 or SOLVE_PAIR columns); a pair's columns share
// each factor entry's decoding (orders up to SOLVE_PAIR_MAX_N)
constexpr int SOLVE_PAIR = 2, SOLVE_PAIR_MAX_N = wg::COL_SOLVE_SLOTS / SOLVE_PAIR * SOLVE_NT;

__global__ __launch_bounds__(SOLVE_NT) void k_bb_solve(const FactorTask* __restrict__ tasks,
                                                       const Item* __restrict__ items) {
  PGA_KERNEL_VGPR_FLOOR();
  PGA_URGENT();
  extern __shared__ uint32_t col[];  // cnt columns, n apart
  __shared__ double sb[SOLVE_PAIR * 128];
  const Item it = items[blockIdx.x];
  const FactorTask& t = tasks[it.task];
  const bool factor = (t.kind == PGA_TASK_GESV || t.kind == PGA_TASK_POSV);
  if (factor && *t.info) return;  // factor-and-solve tasks solve only after success
  const int n = t.n, cnt = it.cnt;
  uint32_t* b = t.b + (size_t)it.off * t.ldb;
  for (int e = threadIdx.x; e < n * cnt; e += SOLVE_NT) col[e] = b[e % n + (size_t)(e / n) * t.ldb];
  gemm::round_table_fill(threadIdx.x, SOLVE_NT);
  __syncthreads();
  wg::lds_f64* sbl = (wg::lds_f64*)sb;
  if (cnt == SOLVE_PAIR) bb_solve_cols<SOLVE_PAIR>(t, n, col, sbl);
  else bb_solve_cols<1>(t, n, col, sbl);
  for (int e = threadIdx.x; e < n * cnt; e += SOLVE_NT) b[e % n + (size_t)(e / n) * t.ldb] = col[e];
}

}  // namespace bb

// ===================================================================================
// Host side: the plan image (descriptors of every step) and the launch sequence
// ===================================================================================
// One group of tasks: its own lockstep sequence of steps on one stream.
struct BlockedGroup {
  struct Node {  // a recursion node of the RECUR_W-wide LU panels: its Rgemm launch
    int jr, s, mid, e, tile, total, counter;
    size_t tiles;
  };
  struct Step {
    size_t syrk, syrk_tiles, panel, trsm_lu, gemm, gemm_tiles, trsm_ch, recur;  // byte offsets in the image
    int n_syrk = 0, syrk_total = 0, n_panel = 0, n_trsm_lu = 0, n_gemm = 0, gemm_total = 0, n_trsm_ch = 0;
    int syrk_tile = 64, gemm_tile = 64, n_recur = 0;
    std::vector<Node> nodes;
  };
  std::vector<unsigned char> image;
  std::vector<Step> steps;
  size_t tasks = 0, solve = 0, counters = 0, spare_info = 0, swaps = 0;
  int ntasks = 0, n_solve = 0, ncounters = 0, n_swaps = 0, swap_rows = 0, node_counter_base = 0;
  size_t max_solve_words = 0;  // LDS words of the largest solve item
  unsigned char* dev = nullptr;  // device copy of the image
};

namespace {

size_t align16(size_t x) { return (x + 15) & ~(size_t)15; }

template <class T>
size_t put(std::vector<unsigned char>& img, const std::vector<T>& v) {
  const size_t off = align16(img.size());
  img.resize(off + v.size() * sizeof(T));
  if (!v.empty()) std::memcpy(img.data() + off, v.data(), v.size() * sizeof(T));
  return off;
}

int gemm_tiles(int m, int n, int t) { return ((m + t - 1) / t) * ((n + t - 1) / t); }

// largest-first order of a step's GEMMs, the tile size (the largest that still gives
// every CU a tile: the other group's launches fill the device beside a group's GEMMs, so
// larger tiles, with fewer instructions per multiply-add, beat more of them; 0.232 s ->
// 0.223 s on the 100-task mix against two tiles per CU) and the tile prefix sums
size_t put_gemms(std::vector<unsigned char>& img, std::vector<bb::Gemm>& g, size_t& tiles_off, int& total, int& tile) {
  std::stable_sort(g.begin(), g.end(), [](const bb::Gemm& x, const bb::Gemm& y) {
    return (double)x.m * x.n * x.k > (double)y.m * y.n * y.k;
  });
  long t64 = 0, t32 = 0;
  for (const bb::Gemm& x : g) {
    t64 += gemm_tiles(x.m, x.n, 64);
    t32 += gemm_tiles(x.m, x.n, 32);
  }
  const long want = runtime().cus;
  tile = t64 >= want ? 64 : (t32 >= want ? 32 : 16);
  std::vector<int> start(g.size());
  total = 0;
  for (size_t i = 0; i < g.size(); ++i) {
    start[i] = total;
    total += gemm_tiles(g[i].m, g[i].n, tile);
  }
  const size_t off = put(img, g);
  tiles_off = put(img, start);
  return off;
}

// one GEMM launch over the step's tiles, at the step's tile size
hipError_t launch_gemm(const BlockedGroup* p, int tile, int total, int nprob, const unsigned char* probs,
                       const unsigned char* starts, int* counter, hipStream_t s) {
  // one workgroup per tile: they retire as tiles finish, so another group's launches on
  // another stream get workgroup slots meanwhile
  const unsigned grid = (unsigned)total;
  (void)p;
  const bb::Gemm* g = reinterpret_cast<const bb::Gemm*>(probs);
  const int* st = reinterpret_cast<const int*>(starts);
  if (tile == 64) hipLaunchKernelGGL(bb::k_bb_gemm<64>, dim3(grid), dim3(bb::NT), 0, s, g, st, nprob, total, counter);
  else if (tile == 32)
    hipLaunchKernelGGL(bb::k_bb_gemm<32>, dim3(grid), dim3(bb::NT), 0, s, g, st, nprob, total, counter);
  else hipLaunchKernelGGL(bb::k_bb_gemm<16>, dim3(grid), dim3(bb::NT), 0, s, g, st, nprob, total, counter);
  return hipGetLastError();
}

// the node of a recursion on w columns whose left half ends at column jr (the split of
// Rgetrf2.cpp:137 with rows >= columns: n1 = columns / 2)
bool node_after_host(int jr, int w, int& s, int& mid, int& e) {
  s = 0;
  e = w;
  while (e - s > 1) {
    mid = s + (e - s) / 2;
    if (jr + 1 == mid) return true;
    if (jr < mid) e = mid;
    else s = mid;
  }
  return false;
}

hipError_t launch_node_gemm(const BlockedGroup::Node& nd, const FactorTask* tasks, const bb::Item* items, int nitems,
                            const int* starts, int* counter, hipStream_t s) {
  const unsigned g = (unsigned)nd.total;  // one workgroup per tile
#define PGA_NODE(T)                                                                                          \
  hipLaunchKernelGGL(bb::k_bb_node_gemm<T>, dim3(g), dim3(bb::NT), 0, s, tasks, items, starts, nitems, nd.total, \
                     counter, nd.s, nd.mid, nd.e)
  if (nd.tile == 64) PGA_NODE(64);
  else if (nd.tile == 32) PGA_NODE(32);
  else PGA_NODE(16);
#undef PGA_NODE
  return hipGetLastError();
}

bool factor_kind(int k) {
  return k == PGA_TASK_GETRF || k == PGA_TASK_POTRF || k == PGA_TASK_GESV || k == PGA_TASK_POSV;
}
bool lu_kind(int k) { return k == PGA_TASK_GETRF || k == PGA_TASK_GESV || k == PGA_TASK_GETRS; }

}  // namespace

BlockedGroup* group_build(const FactorTask* tasks, int count) {
  auto* p = new BlockedGroup();
  p->ntasks = count;
  std::vector<FactorTask> tv(tasks, tasks + count);
  p->tasks = put(p->image, tv);
  // column blocks of every task: unblocked (one block of n columns) as Rgetrf_opencl /
  // Rpotrf_opencl do when nb <= 1 or nb >= n
  int nsteps = 0, node_counters = 0;
  std::vector<int> nbt(count), steps_of(count);
  for (int i = 0; i < count; ++i) {
    const FactorTask& t = tasks[i];
    const int nb = t.nb > 0 ? t.nb : 64;
    nbt[i] = (nb <= 1 || nb >= t.n) ? t.n : nb;
    steps_of[i] = factor_kind(t.kind) && t.n > 0 ? (t.n + nbt[i] - 1) / nbt[i] : 0;
    nsteps = std::max(nsteps, steps_of[i]);
  }
  const uint32_t ONE = 0x40000000u, MONE = 0xC0000000u;
  for (int s = 0; s < nsteps; ++s) {
    std::vector<bb::Gemm> syrk, gemm;
    std::vector<bb::Item> panel, trsm_lu, trsm_ch, recur;
    for (int i = 0; i < count; ++i) {
      if (s >= steps_of[i]) continue;
      const FactorTask& t = tasks[i];
      const int n = t.n, lda = t.lda, j = s * nbt[i], jb = std::min(nbt[i], n - j);
      auto A = [&](int r, int c) { return t.a + r + (size_t)c * lda; };
      const int rest = n - j - jb;
      if (lu_kind(t.kind)) {
        // full-width panels of a blocked factorization go leaf by leaf with all others
        if (jb == bb::RECUR_W && nbt[i] < n) recur.push_back({i, j, jb, 0, 0});
        else panel.push_back({i, j, jb, 0, 0});
        if (rest > 0) {
          for (int c = 0; c < rest; c += bb::TRSM_COLS) trsm_lu.push_back({i, j, jb, c, std::min(bb::TRSM_COLS, rest - c)});
          gemm.push_back({rest, rest, jb, lda, lda, lda, 0, 0, 0, 0, MONE, ONE, A(j + jb, j), A(j, j + jb),
                          A(j + jb, j + jb), nullptr, -1});
        }
        continue;
      }
      const bool upper = (t.uplo == 'U');
      if (j > 0) {  // Rpotrf_opencl.cpp:103-106 / 133-136 (Rsyrk, K = j)
        if (upper)
          syrk.push_back({jb, jb, j, lda, lda, lda, 1, 0, 0, 2, MONE, ONE, A(0, j), A(0, j), A(j, j), nullptr, i});
        else
          syrk.push_back({jb, jb, j, lda, lda, lda, 0, 1, 1, 1, MONE, ONE, A(j, 0), A(j, 0), A(j, j), nullptr, i});
      }
      panel.push_back({i, j, jb, 0, 0});
      if (rest > 0) {
        if (j > 0) {  // Rgemm with K = j (Rgemm quick-returns on K = 0)
          if (upper)
            gemm.push_back({jb, rest, j, lda, lda, lda, 1, 0, 0, 0, MONE, ONE, A(0, j), A(0, j + jb), A(j, j + jb),
                            nullptr, i});
          else
            gemm.push_back({rest, jb, j, lda, lda, lda, 0, 1, 0, 0, MONE, ONE, A(j + jb, 0), A(j, 0), A(j + jb, j),
                            nullptr, i});
        }
        if (upper) {
          for (int c = 0; c < rest; c += bb::TRSM_COLS) trsm_ch.push_back({i, j, jb, c, std::min(bb::TRSM_COLS, rest - c)});
        } else {
          for (int r = 0; r < rest; r += bb::TRSM_ROWS) trsm_ch.push_back({i, j, jb, r, std::min(bb::TRSM_ROWS, rest - r)});
        }
      }
    }
    BlockedGroup::Step st;
    st.syrk = put_gemms(p->image, syrk, st.syrk_tiles, st.syrk_total, st.syrk_tile);
    st.n_syrk = (int)syrk.size();
    st.panel = put(p->image, panel);
    st.n_panel = (int)panel.size();
    st.trsm_lu = put(p->image, trsm_lu);
    st.n_trsm_lu = (int)trsm_lu.size();
    st.gemm = put_gemms(p->image, gemm, st.gemm_tiles, st.gemm_total, st.gemm_tile);
    st.n_gemm = (int)gemm.size();
    st.trsm_ch = put(p->image, trsm_ch);
    st.n_trsm_ch = (int)trsm_ch.size();
    st.recur = put(p->image, recur);
    st.n_recur = (int)recur.size();
    if (!recur.empty()) {
      for (int jr = 0; jr < bb::RECUR_W; ++jr) {
        int ns, nmid, ne;
        if (!node_after_host(jr, bb::RECUR_W, ns, nmid, ne)) continue;
        if (ne - ns == 2) continue;  // two-column nodes run inside k_bb_pair
        BlockedGroup::Node nd{jr, ns, nmid, ne, 16, 0, 0, 0};
        long t64 = 0, t32 = 0;
        for (const bb::Item& it : recur) {
          const int m = tasks[it.task].n - it.j - nmid;
          t64 += gemm_tiles(m, ne - nmid, 64);
          t32 += gemm_tiles(m, ne - nmid, 32);
        }
        const long want = runtime().cus;  // as put_gemms
        nd.tile = t64 >= want ? 64 : (t32 >= want ? 32 : 16);
        std::vector<int> start(recur.size());
        for (size_t q = 0; q < recur.size(); ++q) {
          start[q] = nd.total;
          nd.total += gemm_tiles(tasks[recur[q].task].n - recur[q].j - nmid, ne - nmid, nd.tile);
        }
        nd.tiles = put(p->image, start);
        nd.counter = node_counters++;
        st.nodes.push_back(nd);
      }
    }
    p->steps.push_back(st);
  }
  // the left columns' interchanges of the blocked LU tasks
  std::vector<bb::Item> swaps;
  for (int i = 0; i < count; ++i) {
    const FactorTask& t = tasks[i];
    if (!lu_kind(t.kind) || steps_of[i] < 2) continue;
    const int last = (steps_of[i] - 1) * nbt[i];  // columns of the last block take nothing
    for (int c = 0; c < last; c += nbt[i]) swaps.push_back({i, nbt[i], 0, c, nbt[i]});
    p->swap_rows = std::max(p->swap_rows, t.n - nbt[i]);
  }
  p->swaps = put(p->image, swaps);
  p->n_swaps = (int)swaps.size();
  std::vector<bb::Item> solve;
  auto solves = [&](const FactorTask& t) { return !(t.kind == PGA_TASK_GETRF || t.kind == PGA_TASK_POTRF) && t.n > 0; };
  int solve_n_max = 0;
  for (int i = 0; i < count; ++i)
    if (solves(tasks[i])) solve_n_max = std::max(solve_n_max, tasks[i].n);
  for (int i = 0; i < count; ++i) {
    const FactorTask& t = tasks[i];
    if (!solves(t)) continue;
    // pairs for orders up to half the largest: a pair's workgroup then stays shorter than
    // the largest order's single columns, which end the launch
    const int per = (t.n <= bb::SOLVE_PAIR_MAX_N && 2 * t.n <= solve_n_max) ? bb::SOLVE_PAIR : 1;
    for (int c = 0; c < t.nrhs; c += per) solve.push_back({i, 0, 0, c, std::min(per, t.nrhs - c)});
    p->max_solve_words = std::max(p->max_solve_words, (size_t)t.n * std::min(per, t.nrhs));
  }
  // the largest orders first: the launch ends with its longest workgroups otherwise
  std::stable_sort(solve.begin(), solve.end(),
                   [&](const bb::Item& x, const bb::Item& y) { return tasks[x.task].n > tasks[y.task].n; });
  p->solve = put(p->image, solve);
  p->n_solve = (int)solve.size();
  p->node_counter_base = 2 * (int)p->steps.size();
  p->ncounters = p->node_counter_base + node_counters;
  p->counters = put(p->image, std::vector<int>(std::max(1, p->ncounters), 0));
  // info of tasks given none: the steps read it as their stop flag
  p->spare_info = put(p->image, std::vector<int>(std::max(1, count), 0));
  return p;
}

size_t group_bytes(const BlockedGroup* p) { return align16(p->image.size()); }

hipError_t group_upload(BlockedGroup* p, void* dev, hipStream_t s) {
  p->dev = static_cast<unsigned char*>(dev);
  FactorTask* t = reinterpret_cast<FactorTask*>(p->image.data() + p->tasks);
  for (int i = 0; i < p->ntasks; ++i)
    if (!t[i].info) t[i].info = reinterpret_cast<int*>(p->dev + p->spare_info) + i;
  // the Cholesky GEMMs stop with their task
  for (const BlockedGroup::Step& st : p->steps) {
    for (auto [off, cnt] : {std::pair<size_t, int>(st.syrk, st.n_syrk), std::pair<size_t, int>(st.gemm, st.n_gemm)}) {
      bb::Gemm* g = reinterpret_cast<bb::Gemm*>(p->image.data() + off);
      for (int k = 0; k < cnt; ++k) g[k].stop = g[k].task >= 0 ? t[g[k].task].info : nullptr;
    }
  }
  return hipMemcpyAsync(dev, p->image.data(), p->image.size(), hipMemcpyHostToDevice, s);
}

hipError_t group_launch(const BlockedGroup* p, hipStream_t s) {
  using namespace bb;
  unsigned char* d = p->dev;
  const FactorTask* tasks = reinterpret_cast<const FactorTask*>(d + p->tasks);
  int* counters = reinterpret_cast<int*>(d + p->counters);
  const int init_n = std::max(p->ntasks, p->ncounters);
  if (init_n > 0)
    hipLaunchKernelGGL(k_bb_init, dim3((init_n + 255) / 256), dim3(256), 0, s, tasks, p->ntasks, counters,
                       p->ncounters);
  for (size_t i = 0; i < p->steps.size(); ++i) {
    const BlockedGroup::Step& st = p->steps[i];
    if (st.syrk_total > 0)
      (void)launch_gemm(p, st.syrk_tile, st.syrk_total, st.n_syrk, d + st.syrk, d + st.syrk_tiles, counters + 2 * i, s);
    if (st.n_panel > 0)
      hipLaunchKernelGGL(k_bb_panel, dim3(st.n_panel), dim3(NT), 0, s, tasks,
                         reinterpret_cast<const Item*>(d + st.panel));
    if (st.n_recur > 0) {
      const Item* ri = reinterpret_cast<const Item*>(d + st.recur);
      size_t nn = 0;
      for (int jr = 0; jr < RECUR_W; jr += 2) {
        const BlockedGroup::Node* nd =
            (nn < st.nodes.size() && st.nodes[nn].jr == jr + 1) ? &st.nodes[nn++] : nullptr;
        // the pair's second leaf ends the left half of node nd: its Rtrsm rides along
        hipLaunchKernelGGL(k_bb_pair, dim3(st.n_recur), dim3(PAIR_NT), 0, s, tasks, ri, jr, nd ? nd->s : 0,
                           nd ? nd->mid : 0, nd ? nd->e : 0);
        if (nd && nd->total > 0)
          (void)launch_node_gemm(*nd, tasks, ri, st.n_recur, reinterpret_cast<const int*>(d + nd->tiles),
                                 counters + p->node_counter_base + nd->counter, s);
      }
    }
    if (st.n_trsm_lu > 0)
      hipLaunchKernelGGL(k_bb_trsm, dim3(st.n_trsm_lu), dim3(NT), 0, s, tasks,
                         reinterpret_cast<const Item*>(d + st.trsm_lu));
    if (st.gemm_total > 0)
      (void)launch_gemm(p, st.gemm_tile, st.gemm_total, st.n_gemm, d + st.gemm, d + st.gemm_tiles, counters + 2 * i + 1,
                        s);
    if (st.n_trsm_ch > 0)
      hipLaunchKernelGGL(k_bb_trsm, dim3(st.n_trsm_ch), dim3(NT), 0, s, tasks,
                         reinterpret_cast<const Item*>(d + st.trsm_ch));
  }
  if (p->n_swaps > 0)
    hipLaunchKernelGGL(k_bb_laswp_left, dim3(p->n_swaps), dim3(NT), (size_t)p->swap_rows * sizeof(uint16_t), s, tasks,
                       reinterpret_cast<const Item*>(d + p->swaps));
  if (p->n_solve > 0)
    hipLaunchKernelGGL(k_bb_solve, dim3(p->n_solve), dim3(SOLVE_NT), p->max_solve_words * sizeof(uint32_t), s, tasks,
                       reinterpret_cast<const Item*>(d + p->solve));
  return hipGetLastError();
}

// ---------------------------------------------------------------------------------
// The plan: the tasks dealt into groups that run side by side on their own streams.
// A group's panel steps are latency bound and leave most of the device idle; another
// group's GEMMs fill it meanwhile.  Tasks are independent, so the bits do not change.
// ---------------------------------------------------------------------------------
struct BlockedPlan {
  std::vector<BlockedGroup*> groups;
  std::vector<size_t> offsets;  // of each group's image in the device buffer
  size_t bytes = 0;
};

#define BB_TRY(x)                    \
  do {                               \
    hipError_t _e = (x);             \
    if (_e != hipSuccess) return _e; \
  } while (0)

namespace {

// a stream and an event for each extra group of a launch (kept for reuse)
struct Lane {
  hipStream_t s = nullptr;
  hipEvent_t done = nullptr;
};
std::mutex g_lane_mu;
std::vector<Lane> g_lanes;
hipEvent_t g_fork = nullptr;


}  // namespace

BlockedPlan* blocked_plan_build(const FactorTask* tasks, int count) {
  auto* p = new BlockedPlan();
  // LU tasks in one group, Cholesky tasks in the other: their per-step chains (the
  // Rgetrf2 panel walk; the Rpotrf2 blocks) then run side by side instead of one after the
  // other on one stream.  (Measured on the 100-task mix: 0.59 s against 0.75 s for two
  // groups balanced by work regardless of kind; a third group for LU gained nothing.)
  std::vector<FactorTask> part[2];
  for (int i = 0; i < count; ++i) part[lu_kind(tasks[i].kind) ? 0 : 1].push_back(tasks[i]);
  for (const std::vector<FactorTask>& v : part) {
    if (v.empty()) continue;
    p->groups.push_back(group_build(v.data(), (int)v.size()));
    p->offsets.push_back(p->bytes);
    p->bytes += (group_bytes(p->groups.back()) + 255) & ~(size_t)255;
  }
  if (p->groups.empty()) {  // no tasks: one empty group keeps launch and upload uniform
    p->groups.push_back(group_build(nullptr, 0));
    p->offsets.push_back(0);
    p->bytes = (group_bytes(p->groups.back()) + 255) & ~(size_t)255;
  }
  return p;
}

size_t blocked_plan_bytes(const BlockedPlan* p) { return p->bytes; }

hipError_t blocked_plan_upload(BlockedPlan* p, void* dev, hipStream_t s) {
  for (size_t g = 0; g < p->groups.size(); ++g) {
    hipError_t e = group_upload(p->groups[g], static_cast<unsigned char*>(dev) + p->offsets[g], s);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

hipError_t blocked_plan_launch(const BlockedPlan* p, hipStream_t s) {
  const int G = (int)p->groups.size();
  if (G == 1) return group_launch(p->groups[0], s);
  std::lock_guard<std::mutex> lk(g_lane_mu);  // the lanes and the fork event serve one launch at a time
  while ((int)g_lanes.size() < G - 1) {
    Lane l;
    BB_TRY(hipStreamCreateWithFlags(&l.s, hipStreamNonBlocking));
    BB_TRY(hipEventCreateWithFlags(&l.done, hipEventDisableTiming));
    g_lanes.push_back(l);
  }
  if (!g_fork) BB_TRY(hipEventCreateWithFlags(&g_fork, hipEventDisableTiming));
  BB_TRY(hipEventRecord(g_fork, s));
  for (int g = 1; g < G; ++g) {
    const Lane& l = g_lanes[g - 1];
    BB_TRY(hipStreamWaitEvent(l.s, g_fork, 0));
    BB_TRY(group_launch(p->groups[g], l.s));
    BB_TRY(hipEventRecord(l.done, l.s));
  }
  BB_TRY(group_launch(p->groups[0], s));
  for (int g = 1; g < G; ++g) BB_TRY(hipStreamWaitEvent(s, g_lanes[g - 1].done, 0));
  return hipSuccess;
}

void blocked_plan_destroy(BlockedPlan* p) {
  if (!p) return;
  for (BlockedGroup* g : p->groups) delete g;
  delete p;
}

}  // namespace pga
