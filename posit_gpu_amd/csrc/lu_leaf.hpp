// lu_leaf.hpp -- the Rgetrf2 leaf of the per-call and the blocked-batch drivers, run by
// one workgroup (256 lanes unless a kernel asks for more).
#pragma once

#include <hip/hip_runtime.h>

#include "pga_internal.hpp"
#include "posit32_core.hpp"

namespace pga {
namespace leaf {

constexpr int NT = 256;

// IAMAX (iRamax.cpp:31-88): first index of the largest |x| in signed pattern order.
// The largest value wins, ties go to the smaller index (a total order, so the reduction
// order does not matter): the sequential first maximum.
// With out_val, the element at the index (its pattern) is broadcast too, read before the
// closing barrier, so callers may overwrite it once the call returns.
template <int T = NT>
static __device__ void block_iamax(int n, const uint32_t* x, long incx, int* out_idx, uint32_t* out_val = nullptr) {
  __shared__ int sv[T / 64], si[T / 64];
  const int tid = threadIdx.x;
  int bv = INT32_MIN, bi = INT32_MAX;  // bi == INT32_MAX: nothing seen
  for (int i = tid; i < n; i += T) {
    const int v = (int)abs_p(x[(long)i * incx]);
    if (bi == INT32_MAX || v > bv) {
      bv = v;
      bi = i;
    }
  }
  auto take = [&](int ov, int oi) {
    if (oi != INT32_MAX && (bi == INT32_MAX || ov > bv || (ov == bv && oi < bi))) {
      bv = ov;
      bi = oi;
    }
  };
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) take(__shfl_xor(bv, off), __shfl_xor(bi, off));
  if ((tid & 63) == 0) {
    sv[tid >> 6] = bv;
    si[tid >> 6] = bi;
  }
  __syncthreads();
  bv = sv[0];
  bi = si[0];
#pragma unroll
  for (int w = 1; w < T / 64; ++w) take(sv[w], si[w]);
  if (tid == 0) {
    *out_idx = (bi == INT32_MAX) ? -1 : bi;  // 0-based, -1 when n == 0
    if (out_val) *out_val = (bi == INT32_MAX) ? 0u : x[(long)bi * incx];
  }
  __syncthreads();
}

// Leaf for column j of a Rgetrf2 call on the m_root x w_root block at `root` (rows
// j..m_root-1 of column j): the pivot's row interchange is applied at once to every
// column of the block (the reference applies it to the other columns later, with the
// Rlaswp calls of Rgetrf2.cpp:143,169; nothing reads those rows in between); the pivot
// column itself is swapped only for a nonzero pivot (Rgetrf2.cpp:109-117).  Column j is
// final when the call returns; the caller places a barrier before anything reads it.
template <int T>
static __device__ void leaf_step(int m_root, int w_root, uint32_t* root, int lda, int j, int* ipiv, int* info,
                                 int row_off, int col_off) {
  __shared__ int piv;
  __shared__ uint32_t pval;
  uint32_t* col = root + (size_t)j * lda;
  // the pivot's value comes from the search, not from col[p]: the lane that swaps the
  // pivot column may overwrite col[p] before a later wave would have read it
  block_iamax<T>(m_root - j, col + j, 1, &piv, &pval);
  const int p = j + piv;
  const uint32_t ap = pval;
  if (threadIdx.x == 0) ipiv[j] = row_off + p + 1;
  if (p != j) {
    for (int c = threadIdx.x; c < w_root; c += T) {
      if (c == j && ap == 0) continue;
      uint32_t* x = root + (size_t)c * lda;
      const uint32_t t = x[j];
      x[j] = x[p];
      x[p] = t;
    }
  }
  __syncthreads();
  if (ap == 0) {
    if (threadIdx.x == 0 && *info == 0) *info = col_off + j + 1;
  } else if (!lt(abs_p(ap), P_MINPOS)) {
    // sfmin = FLT_MIN as posit32 = minpos: every pivot but NaR takes the Rscal branch
    const uint32_t r = div(P_ONE, ap);
    for (int i = j + 1 + threadIdx.x; i < m_root; i += T) col[i] = mul(r, col[i]);
  } else {
    for (int i = j + 1 + threadIdx.x; i < m_root; i += T) col[i] = div(col[i], ap);
  }
}

// The Rtrsm of a small recursion node [ts, te) split at tm (A12 := L11^-1 A12,
// Rgetrf2.cpp:151-154; tm - ts <= N1, te - tm <= N2), staged in LDS.  It is run by the
// workgroup whose last leaf ends the node's left half: the next step of the reference's
// order either way, and it saves a launch.
template <int N1, int N2, int T>
static __device__ void node_trsm(uint32_t* root, int lda, int ts, int tm, int te) {
  __shared__ uint32_t sL[N1][N1], sB[N1][N2];
  const int n1 = tm - ts, n2 = te - tm;
  __syncthreads();
  for (int e = threadIdx.x; e < n1 * n1; e += T) sL[e % n1][e / n1] = root[ts + e % n1 + (size_t)(ts + e / n1) * lda];
  for (int e = threadIdx.x; e < n1 * n2; e += T) sB[e % n1][e / n1] = root[ts + e % n1 + (size_t)(tm + e / n1) * lda];
  __syncthreads();
  for (int k = 0; k + 1 < n1; ++k) {  // Rtrsm.cpp:139-155: a zero b_k skips step k
    const int rows = n1 - 1 - k;
    for (int e = threadIdx.x; e < rows * n2; e += T) {
      const int i = k + 1 + e % rows, c = e / rows;
      const uint32_t bk = sB[k][c];
      if (bk != 0) sB[i][c] = sub(sB[i][c], mul(bk, sL[i][k]));
    }
    __syncthreads();
  }
  for (int e = threadIdx.x; e < n1 * n2; e += T) root[ts + e % n1 + (size_t)(tm + e / n1) * lda] = sB[e % n1][e / n1];
}

// One leaf, then (tm > ts) the pending node's Rtrsm.
template <int N1, int N2, int T = NT>
static __device__ void getrf2_leaf(int m_root, int w_root, uint32_t* root, int lda, int j, int* ipiv, int* info,
                                   int row_off, int col_off, int ts, int tm, int te) {
  leaf_step<T>(m_root, w_root, root, lda, j, ipiv, info, row_off, col_off);
  if (tm > ts) node_trsm<N1, N2, T>(root, lda, ts, tm, te);
}

// A whole two-column node [j, j+2) of the recursion (rows >= 2, so it splits 1 + 1):
// leaf j, the node's Rgemm (column j+1 below row j -= column j * a(j, j+1), the NN form
// of Rgemm.cpp:163-185 with alpha = -1, beta = 1: c + (-u) * l, no skip), leaf j+1,
// then the pending node's Rtrsm.  The node's own Rtrsm is 1 x 1 unit lower: nothing.
template <int N1, int N2, int T>
static __device__ void getrf2_pair(int m_root, int w_root, uint32_t* root, int lda, int j, int* ipiv, int* info,
                                   int row_off, int col_off, int ts, int tm, int te) {
  leaf_step<T>(m_root, w_root, root, lda, j, ipiv, info, row_off, col_off);
  __syncthreads();
  const uint32_t* l = root + (size_t)j * lda;
  uint32_t* c = root + (size_t)(j + 1) * lda;
  const uint32_t t = neg(c[j]);
  for (int i = j + 1 + threadIdx.x; i < m_root; i += T) c[i] = add(c[i], mul(t, l[i]));
  __syncthreads();
  leaf_step<T>(m_root, w_root, root, lda, j + 1, ipiv, info, row_off, col_off);
  if (tm > ts) node_trsm<N1, N2, T>(root, lda, ts, tm, te);
}

// ---------------------------------------------------------------------------------
// A two-column node on several workgroups at once
// ---------------------------------------------------------------------------------
// The state the workgroups of one launch exchange through device memory: a ticket
// counter for the grid barrier, a failure word, and per pivot search one record per
// workgroup (its first maximum and that row's two values) plus the two values of the
// pivot search's own row.
struct PairRec {
  int key, idx;      // |x| as a signed pattern (NaR lowest), global row (INT32_MAX: none)
  uint32_t v0, v1;   // the row's values in the node's two columns
};
constexpr int PAIR_MC_MAX_G = 16;
// info after a launch whose grid barrier gave up waiting (its results are not valid)
constexpr int PAIR_SYNC_FAILED_INFO = -1000;
// One workgroup's candidate for a pivot search of the W-column node kernel (lu_panel.hpp):
// the magnitude of its first maximum (-1 for NaR), that row (INT32_MAX: none), and the
// row's values in the node's W columns, as FP64 numbers.
constexpr int PANEL_MC_MAX_G = 32;
constexpr int PANEL_MC_MAX_W = 64;
constexpr int PANEL_MC_MAX_LAUNCHES = 1024;  // W-column node launches per factorization
struct PanelRec {
  double key;
  int idx;
  int nan;  // an FP64 result of this workgroup's previous step was NaN
  double v[PANEL_MC_MAX_W];
};
struct PairSync {
  unsigned ctr;
  unsigned fail;
  unsigned pad[14];
  PairRec cand[2][PAIR_MC_MAX_G];
  PairRec own[2];  // row j before the first search, row j+1 before the second
  // the W-column node kernel: per step parity, the workgroups' candidates and row j
  PanelRec pcand[2][PANEL_MC_MAX_G];
  double pown[2][PANEL_MC_MAX_W];
  unsigned pctr[PANEL_MC_MAX_LAUNCHES];  // one ticket counter per W-column node launch
  // the fast form's exchange: 64-bit words that each carry the tag of the step they
  // belong to, so that a reader needs no barrier: per step parity, each workgroup's
  // header (key, row, NaN report) and candidate row, and row j + 1
  unsigned long long thead[2][PANEL_MC_MAX_G];
  unsigned long long tval[2][PANEL_MC_MAX_G][PANEL_MC_MAX_W];
  unsigned long long town[2][PANEL_MC_MAX_W];
};

__device__ __forceinline__ void rec_store(PairRec* r, const PairRec& v) {
  __hip_atomic_store(&r->key, v.key, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __hip_atomic_store(&r->idx, v.idx, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __hip_atomic_store(&r->v0, v.v0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __hip_atomic_store(&r->v1, v.v1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ PairRec rec_load(PairRec* r) {
  PairRec v;
  v.key = __hip_atomic_load(&r->key, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  v.idx = __hip_atomic_load(&r->idx, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  v.v0 = __hip_atomic_load(&r->v0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  v.v1 = __hip_atomic_load(&r->v1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  return v;
}

// Barrier across the launch's workgroups: every workgroup adds one to the ticket counter
// and waits until it reaches `target`.  The wait is bounded (about 0.1 s of sleeps): past
// it, or once another workgroup has given up, the failure word is set and the wait
// ends, so every wave exits; the caller reports the failure.
__device__ __forceinline__ void pair_grid_barrier(PairSync* sy, unsigned target) {
  __syncthreads();
  if (threadIdx.x == 0) {
    __threadfence();
    __hip_atomic_fetch_add(&sy->ctr, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    unsigned spins = 0;
    while (__hip_atomic_load(&sy->ctr, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) < target) {
      __builtin_amdgcn_s_sleep(1);
      if (++spins > (1u << 22) || __hip_atomic_load(&sy->fail, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
        __hip_atomic_fetch_or(&sy->fail, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        break;
      }
    }
    __threadfence();
  }
  __syncthreads();
}

// first maximum of a (key, idx) pair: larger key, then smaller index
__device__ __forceinline__ void rec_take(PairRec& b, const PairRec& o) {
  if (o.idx != INT32_MAX && (b.idx == INT32_MAX || o.key > b.key || (o.key == b.key && o.idx < b.idx))) b = o;
}

// The workgroup's first maximum over its rows (keys and indices per lane; values carried).
template <int RPL>
__device__ __forceinline__ PairRec wg_first_max(const int* key, const int* idx, const uint32_t* v0,
                                                const uint32_t* v1, PairRec* s_part) {
  PairRec b{INT32_MIN, INT32_MAX, 0u, 0u};
#pragma unroll
  for (int q = 0; q < RPL; ++q) rec_take(b, PairRec{key[q], idx[q], v0[q], v1[q]});
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    PairRec o;
    o.key = __shfl_xor(b.key, off);
    o.idx = __shfl_xor(b.idx, off);
    o.v0 = (uint32_t)__shfl_xor((int)b.v0, off);
    o.v1 = (uint32_t)__shfl_xor((int)b.v1, off);
    rec_take(b, o);
  }
  if ((threadIdx.x & 63) == 0) s_part[threadIdx.x >> 6] = b;
  __syncthreads();
  b = s_part[0];
#pragma unroll
  for (int w = 1; w < 4; ++w) rec_take(b, s_part[w]);
  __syncthreads();  // s_part is free again
  return b;
}

// The launch's first maximum from the G records (every workgroup computes the same).
__device__ __forceinline__ PairRec grid_first_max(PairRec* cand, int G, PairRec* s_part) {
  if (threadIdx.x < 64) {
    PairRec b{INT32_MIN, INT32_MAX, 0u, 0u};
    if ((int)threadIdx.x < G) b = rec_load(&cand[threadIdx.x]);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      PairRec o;
      o.key = __shfl_xor(b.key, off);
      o.idx = __shfl_xor(b.idx, off);
      o.v0 = (uint32_t)__shfl_xor((int)b.v0, off);
      o.v1 = (uint32_t)__shfl_xor((int)b.v1, off);
      rec_take(b, o);
    }
    if (threadIdx.x == 0) s_part[0] = b;
  }
  __syncthreads();
  const PairRec b = s_part[0];
  __syncthreads();
  return b;
}

// A whole two-column node [j, j+2) of the recursion, as getrf2_pair, with the rows
// [j, m_root) dealt in G contiguous blocks to the launch's G workgroups of 256 lanes
// (RPL rows per lane; block 0 holds rows j and j+1).  Each workgroup keeps its rows of
// the two columns in registers; the two pivot searches meet at a grid barrier each
// (records in `sy`); the pivot row's values travel in the records, so every workgroup
// makes the interchange in its own rows.  The same operations as getrf2_pair reach every
// element in the same order: the interchange of rows j and p in both columns (in column
// j only for a nonzero pivot), the scaling of column j below row j, the 1 x 1 update of
// column j+1 (c + (-u) * l), then the same for row j+1.  The interchanges of the block's
// other columns follow, both in order, spread over the workgroups by column; a pending
// node's Rtrsm runs on workgroup 0 after a third barrier.  `base` is the ticket count
// before this launch; the launch adds G per barrier.
template <int RPL, int N1, int N2>
__device__ void getrf2_pair_mc(int m_root, int w_root, uint32_t* root, int lda, int j, int* ipiv, int* info,
                               int row_off, int col_off, int ts, int tm, int te, PairSync* sy, unsigned base) {
  __shared__ PairRec s_part[4];
  const int G = gridDim.x, w = blockIdx.x, tid = threadIdx.x;
  const int rows = m_root - j, R = (rows + G - 1) / G;
  const int r0 = j + w * R, r1 = min(m_root, r0 + R);
  uint32_t* c0p = root + (size_t)j * lda;
  uint32_t* c1p = root + (size_t)(j + 1) * lda;
  uint32_t u0[RPL], u1[RPL];
  int key[RPL], idx[RPL];
#pragma unroll
  for (int q = 0; q < RPL; ++q) {
    const int i = r0 + tid + q * 256;
    const bool in = i < r1;
    u0[q] = in ? c0p[i] : 0u;
    u1[q] = in ? c1p[i] : 0u;
  }
  // ---- leaf j: pivot search over rows j .. m_root-1 of column j
#pragma unroll
  for (int q = 0; q < RPL; ++q) {
    const int i = r0 + tid + q * 256;
    key[q] = (int)abs_p(u0[q]);
    idx[q] = i < r1 ? i : INT32_MAX;
  }
  PairRec mine = wg_first_max<RPL>(key, idx, u0, u1, s_part);
  if (tid == 0) {
    rec_store(&sy->cand[0][w], mine);
    if (w == 0) rec_store(&sy->own[0], PairRec{0, j, u0[0], u1[0]});  // row j: lane 0, slot 0
  }
  pair_grid_barrier(sy, base + (unsigned)G);
  const PairRec pv = grid_first_max(sy->cand[0], G, s_part);
  const PairRec rj = rec_load(&sy->own[0]);
  const int p = pv.idx == INT32_MAX ? j : pv.idx;
  const uint32_t ap = pv.idx == INT32_MAX ? rj.v0 : pv.v0;
  if (w == 0 && tid == 0) {
    ipiv[j] = row_off + p + 1;
    if (ap == 0 && *info == 0) *info = col_off + j + 1;
  }
  // interchange rows j and p (column j only for a nonzero pivot), then scale and update
  const uint32_t r = (ap != 0 && !lt(abs_p(ap), P_MINPOS)) ? div(P_ONE, ap) : 0u;
  const uint32_t cj = p != j ? pv.v1 : rj.v1;  // a(j, j+1) after the interchange
  const uint32_t t = neg(cj);
#pragma unroll
  for (int q = 0; q < RPL; ++q) {
    const int i = r0 + tid + q * 256;
    if (i >= r1) continue;
    if (p != j) {
      if (i == j) {
        if (ap != 0) u0[q] = pv.v0;
        u1[q] = pv.v1;
      } else if (i == p) {
        if (ap != 0) u0[q] = rj.v0;
        u1[q] = rj.v1;
      }
    }
    if (i > j) {
      if (ap != 0) u0[q] = lt(abs_p(ap), P_MINPOS) ? div(u0[q], ap) : mul(r, u0[q]);
      u1[q] = add(u1[q], mul(t, u0[q]));
    }
  }
  // ---- leaf j+1: pivot search over rows j+1 .. m_root-1 of column j+1
#pragma unroll
  for (int q = 0; q < RPL; ++q) {
    const int i = r0 + tid + q * 256;
    key[q] = (int)abs_p(u1[q]);
    idx[q] = (i < r1 && i > j) ? i : INT32_MAX;
  }
  mine = wg_first_max<RPL>(key, idx, u0, u1, s_part);
  if (tid == 1 && w == 0) rec_store(&sy->own[1], PairRec{0, j + 1, u0[0], u1[0]});  // row j+1: lane 1, slot 0
  if (tid == 0) rec_store(&sy->cand[1][w], mine);
  pair_grid_barrier(sy, base + 2u * (unsigned)G);
  const PairRec pv2 = grid_first_max(sy->cand[1], G, s_part);
  const PairRec rj1 = rec_load(&sy->own[1]);
  const int j1 = j + 1;
  const int p2 = pv2.idx == INT32_MAX ? j1 : pv2.idx;
  const uint32_t ap2 = pv2.idx == INT32_MAX ? rj1.v1 : pv2.v1;
  if (w == 0 && tid == 0) {
    ipiv[j1] = row_off + p2 + 1;
    if (ap2 == 0 && *info == 0) *info = col_off + j1 + 1;
  }
  const uint32_t r2 = (ap2 != 0 && !lt(abs_p(ap2), P_MINPOS)) ? div(P_ONE, ap2) : 0u;
#pragma unroll
  for (int q = 0; q < RPL; ++q) {
    const int i = r0 + tid + q * 256;
    if (i >= r1) continue;
    if (p2 != j1) {
      if (i == j1) {
        u0[q] = pv2.v0;
        if (ap2 != 0) u1[q] = pv2.v1;
      } else if (i == p2) {
        u0[q] = rj1.v0;
        if (ap2 != 0) u1[q] = rj1.v1;
      }
    }
    if (i > j1 && ap2 != 0) u1[q] = lt(abs_p(ap2), P_MINPOS) ? div(u1[q], ap2) : mul(r2, u1[q]);
    c0p[i] = u0[q];
    c1p[i] = u1[q];
  }
  // the two interchanges in the block's other columns, in order
  for (int c = w * 256 + tid; c < w_root; c += G * 256) {
    if (c == j || c == j1) continue;
    uint32_t* x = root + (size_t)c * lda;
    if (p != j) {
      const uint32_t a = x[j];
      x[j] = x[p];
      x[p] = a;
    }
    if (p2 != j1) {
      const uint32_t a = x[j1];
      x[j1] = x[p2];
      x[p2] = a;
    }
  }
  if (tm > ts) {
    pair_grid_barrier(sy, base + 3u * (unsigned)G);
    if (w == 0) node_trsm<N1, N2, 256>(root, lda, ts, tm, te);
  }
  if (w == 0 && tid == 0 && __hip_atomic_load(&sy->fail, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
    *info = PAIR_SYNC_FAILED_INFO;
}

}  // namespace leaf
}  // namespace pga
