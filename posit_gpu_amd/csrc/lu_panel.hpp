// lu_panel.hpp -- a W-column node of the Rgetrf2 recursion (Rgetrf2.cpp:32-185; W = 32 or
// 64 columns, rows [s, m)) in one launch of G workgroups, one column at a time.
//
// The recursion gives element (i, c) of the node, for every column j < c, the term
// a(i, c) + (-u(j, c)) * l(i, j) in ascending j, then (for i > c) the scaling of its own
// column by the reciprocal of the pivot; a column's pivot search runs once all its terms
// are in.  Which routine applies a term decides one detail: the node whose halves split
// j from c (its split point `mid`) applies it with Rtrsm (Rtrsm.cpp:139-155: skipped when
// u(j, c) is zero) to its rows i < mid, and with Rgemm (Rgemm.cpp:163-185, alpha = -1,
// beta = 1: never skipped, so a zero multiplier still spreads a NaR) to the rows below.
// Row interchanges only move values: each is applied to every column of the node as
// soon as its pivot is known (the recursion applies it to the other half later, with
// Rlaswp; nothing reads those rows in between).  Right-looking order with that skip rule
// therefore gives every element the same operations in the same order: the same bits.
//
// For a node of W columns whose rows are at least W (every node of a tall panel), the
// splits are the halves n1 = cols / 2 (Rgetrf2.cpp:137), so for a power-of-two W the node
// that splits relative columns j < c splits at c with its bits below the highest bit of
// j ^ c cleared.
//
// Workgroup w owns rows [s + w R, s + (w + 1) R) of the node's W columns, RPL rows per lane.
// Per column j:
//   1. each workgroup finds its first maximum of |a(i, j)|, i >= j, and publishes it with
//      that row's W values (and workgroup 0 publishes row j) through agent-scope atomic
//      stores, then all meet at a grid barrier;
//   2. wave 0 of every workgroup reads all candidates and takes the first maximum, the
//      pivot row p (ties: the smaller row, as iRamax.cpp:67 keeps the first);
//   3. each workgroup interchanges rows j and p where it owns them (column j itself only
//      for a nonzero pivot, Rgetrf2.cpp:109-117), scales its rows of column j below j
//      (by round(1/pivot), Rgetrf2.cpp:121-122; by division for a NaR pivot) and applies
//      column j to its rows of the later columns.
// The fast form holds the rows in registers as FP64 values and rounds products and sums
// to the posit grid with the GEMM's constants (gemm::fp64_mac).  A scale outside the
// table or a NaR gives NaN; every workgroup reports whether it has seen one with its next
// candidate (and once more after the last column), and if any has, all of them start the
// node again from its inputs in memory, which the fast form does not touch before its
// end, with the pattern-level operations on memory (`panel_exact`).  The registers are
// kept rotated so that their indices stay static: before chunk jb (U columns a chunk),
// register column c holds node column (jb + c) mod W.
// At the end every workgroup writes its rows back and applies the node's W interchanges,
// in order, to a share of the block's other columns.
#pragma once

#include <hip/hip_runtime.h>

#include "lu_leaf.hpp"
#include "rgemm_tile.hpp"

namespace pga {
namespace leaf {

PGA_D void st_rec_f64(double* p, double v) {
  __hip_atomic_store(reinterpret_cast<unsigned long long*>(p), (unsigned long long)__double_as_longlong(v),
                     __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
PGA_D double ld_rec_f64(double* p) {
  return __longlong_as_double(
      (long long)__hip_atomic_load(reinterpret_cast<unsigned long long*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
}
PGA_D void st_rec_i32(int* p, int v) { __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
PGA_D int ld_rec_i32(int* p) { return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }

// Barrier across the launch's workgroups for records that are written and read only with
// agent-scope atomic stores and loads (coherent across the XCDs without cache flushes):
// after wave 0 has stored its records, its lane 0 waits for the stores to complete,
// adds one to the launch's ticket counter and waits until the counter reaches `target`.
// The wait is bounded (about 2^21 polls); past it, or once another workgroup has given
// up, the failure word is set and the wait ends, so every wave exits.  With one
// workgroup only the completion of wave 0's stores is needed.
PGA_D void rec_grid_barrier(PairSync* sy, unsigned* ctr, unsigned target) {
  __syncthreads();
  if (threadIdx.x < 64) {
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __builtin_amdgcn_s_waitcnt(0);  // wave 0's record stores have completed
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
  }
  if (gridDim.x > 1 && threadIdx.x == 0) {
    __hip_atomic_fetch_add(ctr, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    unsigned spins = 0;
    while (__hip_atomic_load(ctr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
      __builtin_amdgcn_s_sleep(1);
      if (++spins > (1u << 21) || __hip_atomic_load(&sy->fail, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
        __hip_atomic_fetch_or(&sy->fail, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        break;
      }
    }
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
  }
  __syncthreads();
}

// first maximum of (key, row): larger key, then smaller row; INT32_MAX marks no row
PGA_D void panel_take(double& bk, int& bi, double ok, int oi) {
  if (oi != INT32_MAX && (bi == INT32_MAX || ok > bk || (ok == bk && oi < bi))) {
    bk = ok;
    bi = oi;
  }
}

// the pivot search key of a value: |x|, NaR (NaN) below every posit
PGA_D double panel_key(double x) { return (x != x) ? -1.0 : __builtin_fabs(x); }

// The workgroup's first maximum from each lane's (bk, bi): every lane gets it.
PGA_D void panel_wg_max(double& bk, int& bi, double* s_wkey, int* s_widx) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) panel_take(bk, bi, __shfl_xor(bk, off), __shfl_xor(bi, off));
  if (lane == 0) {
    s_wkey[wave] = bk;
    s_widx[wave] = bi;
  }
  __syncthreads();
  bk = s_wkey[0];
  bi = s_widx[0];
#pragma unroll
  for (int v = 1; v < 4; ++v) panel_take(bk, bi, s_wkey[v], s_widx[v]);
}

// Wave 0: the first maximum over the G candidates of one step (all lanes get it), the
// winner's W values into s_row and row j's into s_rowj; lane 0 stores the pivot row in
// *s_p and whether any workgroup reported a NaN in *s_nan.  Element e = g * W + c of the
// candidates' values is loaded by lane e % 64 into slot e / 64 (the same slot for all W
// values of one candidate), so every load is in flight at once.
template <int W>
PGA_D void panel_pick(PairSync* sy, int par, int G, double* s_row, double* s_rowj, int* s_p, int* s_nan) {
  constexpr int K = PANEL_MC_MAX_G * W / 64;
  const int lane = threadIdx.x & 63;
  double ck = -2.0;
  int ci = INT32_MAX, cg = lane, cn = 0;
  if (lane < G) {
    ck = ld_rec_f64(&sy->pcand[par][lane].key);
    ci = ld_rec_i32(&sy->pcand[par][lane].idx);
    cn = ld_rec_i32(&sy->pcand[par][lane].nan);
  }
  double cv[K];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int e = lane + 64 * k, g = e / W, c = e % W;
    cv[k] = g < G ? ld_rec_f64(&sy->pcand[par][g].v[c]) : 0.0;
  }
  const double rj = lane < W ? ld_rec_f64(&sy->pown[par][lane]) : 0.0;
  const bool any_nan = __any(cn != 0);
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const double ok = __shfl_xor(ck, off);
    const int oi = __shfl_xor(ci, off), og = __shfl_xor(cg, off);
    if (oi != INT32_MAX && (ci == INT32_MAX || ok > ck || (ok == ck && oi < ci))) {
      ck = ok;
      ci = oi;
      cg = og;
    }
  }
  const int slot = (cg * W) >> 6;
  double pv = cv[0];
#pragma unroll
  for (int k = 1; k < K; ++k)
    if (k == slot) pv = cv[k];
  pv = __shfl(pv, (cg * W + lane) & 63);
  if (lane < W) {
    s_row[lane] = pv;
    s_rowj[lane] = rj;
  }
  if (lane == 0) {
    *s_p = ci;
    *s_nan = any_nan ? 1 : 0;
  }
}

// The pivot row of step j from *s_p: a row outside [j, m) can only come from a failed
// barrier (then the failure word is set and the result is not used); p = j keeps every
// access in bounds.
PGA_D int panel_pivot(PairSync* sy, int sp, int j, int m_root) {
  if (sp < j || sp >= m_root) {
    if (threadIdx.x == 0) __hip_atomic_fetch_or(&sy->fail, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return j;
  }
  return sp;
}

// The node with the pattern-level operations, on the matrix in memory: the same steps as
// the fast form, each workgroup on its own rows (one lane per row).
template <int W>
__device__ __forceinline__ void panel_exact(int m_root, uint32_t* root, int lda, int s, int* ipiv, int* info, int row_off,
                            int col_off, PairSync* sy, unsigned* ctr, unsigned& nbar, int r0, int r1,
                            double* s_row, double* s_rowj, double* s_wkey, int* s_widx, int* s_piv, int* s_p,
                            int* s_nan) {
  const int G = gridDim.x, w = blockIdx.x, tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  for (int jj = 0; jj < W; ++jj) {
    const int j = s + jj, par = (int)(nbar & 1u);
    uint32_t* colj = root + (size_t)j * lda;
    double bk = -2.0;
    int bi = INT32_MAX;
    for (int i = max(r0, j) + tid; i < r1; i += 256) panel_take(bk, bi, panel_key(to_double(colj[i])), i);
    panel_wg_max(bk, bi, s_wkey, s_widx);
    if (wave == 0) {
      PanelRec* rc = &sy->pcand[par][w];
      if (lane < W) {
        st_rec_f64(&rc->v[lane], bi != INT32_MAX ? to_double(root[bi + (size_t)(s + lane) * lda]) : 0.0);
        if (w == 0) st_rec_f64(&sy->pown[par][lane], to_double(root[j + (size_t)(s + lane) * lda]));
      }
      if (lane == 0) {
        st_rec_f64(&rc->key, bk);
        st_rec_i32(&rc->idx, bi);
        st_rec_i32(&rc->nan, 0);
      }
    }
    rec_grid_barrier(sy, ctr, ++nbar * (unsigned)G);
    if (wave == 0) panel_pick<W>(sy, par, G, s_row, s_rowj, s_p, s_nan);
    __syncthreads();
    const int p = panel_pivot(sy, *s_p, j, m_root);
    const double ap = s_row[jj];
    if (tid == 0) {
      s_piv[jj] = p;
      if (w == 0) {
        ipiv[j] = row_off + p + 1;
        if (ap == 0.0 && *info == 0) *info = col_off + j + 1;
      }
    }
    // interchange (where this workgroup owns the rows), one lane per column
    if (p != j && tid < W && (tid != jj || ap != 0.0)) {
      uint32_t* x = root + (size_t)(s + tid) * lda;
      if (j >= r0 && j < r1) x[j] = from_double(s_row[tid]);
      if (p >= r0 && p < r1) x[p] = from_double(s_rowj[tid]);
    }
    __syncthreads();
    const uint32_t apat = from_double(ap);
    const double* urow = p != j ? s_row : s_rowj;
    for (int i = max(r0, j + 1) + tid; i < r1; i += 256) {
      uint32_t l = colj[i];
      if (apat != 0) {
        l = lt(abs_p(apat), P_MINPOS) ? div(l, apat) : mul(div(P_ONE, apat), l);
        colj[i] = l;
      }
      for (int lc = jj + 1; lc < W; ++lc) {
        const uint32_t uc = from_double(urow[lc]);
        const int mid = lc & ~((1 << (31 - __builtin_clz(jj ^ lc))) - 1);
        if (uc != 0 || i >= s + mid) {
          uint32_t* x = root + i + (size_t)(s + lc) * lda;
          *x = add(*x, mul(neg(uc), l));
        }
      }
    }
    __syncthreads();
  }
}

template <int W, int RPL, int U>
__device__ __forceinline__ void panel_mc(int m_root, int w_root, uint32_t* root, int lda, int s, int* ipiv, int* info, int row_off,
                         int col_off, PairSync* sy, int launch) {
  static_assert(W <= PANEL_MC_MAX_W && W <= 64 && W % U == 0 && (W & (W - 1)) == 0, "node width");
  __shared__ double s_row[W], s_rowj[W];
  __shared__ double s_wkey[4];
  __shared__ int s_widx[4];
  __shared__ int s_piv[W];
  __shared__ int s_p, s_nan;
  const int G = gridDim.x, w = blockIdx.x, tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  unsigned* ctr = &sy->pctr[launch];
  unsigned nbar = 0;  // barriers passed in this launch
  const int info0 = (w == 0 && tid == 0) ? *info : 0;
  gemm::round_table_fill(tid, 256);
  const int rows = m_root - s, R = (rows + G - 1) / G;
  const int r0 = s + w * R, r1 = min(m_root, r0 + R);
  double a[RPL][W];
#pragma unroll
  for (int q = 0; q < RPL; ++q) {
    const int i = r0 + q * 256 + tid;
#pragma unroll
    for (int c = 0; c < W; ++c) a[q][c] = (i < r1) ? to_double(root[i + (size_t)(s + c) * lda]) : 0.0;
  }
  __syncthreads();  // the rounding table
  double chk = 0.0;  // sum of every FP64 result since the last report: NaN if any was
  bool exact = false;
  for (int jb = 0; jb < W && !exact; jb += U) {
#pragma unroll
    for (int u = 0; u < U; ++u) {
      if (exact) continue;  // (no break: the loop must unroll, so that register indices stay static)
      const int jj = jb + u, j = s + jj, par = (int)(nbar & 1u);
      // 1. this workgroup's first maximum of column j over its rows i >= j
      double bk = -2.0;
      int bi = INT32_MAX;
#pragma unroll
      for (int q = 0; q < RPL; ++q) {
        const int i = r0 + q * 256 + tid;
        if (i < r1 && i >= j) panel_take(bk, bi, panel_key(a[q][u]), i);
      }
      panel_wg_max(bk, bi, s_wkey, s_widx);
      const bool nan_here = __syncthreads_or(chk != chk);
      chk = 0.0;
      // that row's values, and row j's, to LDS; wave 0 publishes them
#pragma unroll
      for (int q = 0; q < RPL; ++q) {
        const int i = r0 + q * 256 + tid;
        if (i == bi) {
#pragma unroll
          for (int c = 0; c < W; ++c) s_row[c] = a[q][c];
        }
      }
      if (w == 0 && tid == jj) {  // row j = s + jj: workgroup 0, slot 0, lane jj
#pragma unroll
        for (int c = 0; c < W; ++c) s_rowj[c] = a[0][c];
      }
      __syncthreads();
      if (wave == 0) {
        PanelRec* rc = &sy->pcand[par][w];
        if (lane < W) st_rec_f64(&rc->v[lane], s_row[lane]);
        if (w == 0 && lane < W) st_rec_f64(&sy->pown[par][lane], s_rowj[lane]);
        if (lane == 0) {
          st_rec_f64(&rc->key, bk);
          st_rec_i32(&rc->idx, bi);
          st_rec_i32(&rc->nan, nan_here ? 1 : 0);
        }
      }
      rec_grid_barrier(sy, ctr, ++nbar * (unsigned)G);
      // 2. the first maximum over every workgroup's candidate
      if (wave == 0) panel_pick<W>(sy, par, G, s_row, s_rowj, &s_p, &s_nan);
      __syncthreads();
      if (s_nan) {  // a NaN somewhere: the whole node again, exactly
        exact = true;
        continue;
      }
      const int p = panel_pivot(sy, s_p, j, m_root);
      const double ap = s_row[u];
      const bool swapped = p != j;
      if (tid == 0) {
        s_piv[jj] = p;
        if (w == 0) {
          ipiv[j] = row_off + p + 1;
          if (ap == 0.0 && *info == 0) *info = col_off + j + 1;
        }
      }
      // 3. interchange rows j and p, scale column j, apply it to the later columns
      if (swapped) {
#pragma unroll
        for (int q = 0; q < RPL; ++q) {
          const int i = r0 + q * 256 + tid;
          if (i == j || i == p) {
            const double* src = (i == j) ? s_row : s_rowj;
#pragma unroll
            for (int c = 0; c < W; ++c)
              if (c != u || ap != 0.0) a[q][c] = src[c];
          }
        }
      }
      const double r = (ap != 0.0 && ap == ap) ? to_double(div(P_ONE, from_double(ap))) : 0.0;
      const double* urow = swapped ? s_row : s_rowj;  // row j after the interchange
#pragma unroll
      for (int q = 0; q < RPL; ++q) {
        const int i = r0 + q * 256 + tid;
        if (i > j && i < r1) {
          double l = a[q][u];
          if (ap != 0.0) {  // a NaR pivot makes the column NaR (division by NaR)
            l = (ap == ap) ? gemm::fp64_posit_mul(r, l) : ap;
            a[q][u] = l;
            chk += l;
          }
#pragma unroll
          for (int c = u + 1; c < W; ++c) {
            if (c < W - jb) {  // node column jb + c, not yet final
              const double uc = urow[c];
              const int lc = jb + c;
              const int mid = lc & ~((1 << (31 - __builtin_clz(jj ^ lc))) - 1);
              if (uc != 0.0 || i >= s + mid) {
                const double y = gemm::fp64_mac(a[q][c], -uc, l);
                a[q][c] = y;
                chk += y;
              }
            }
          }
        }
      }
    }
    if (exact) break;
    // rotate by U: register column c takes node column (jb + U + c) mod W
#pragma unroll
    for (int q = 0; q < RPL; ++q) {
      double t[U];
#pragma unroll
      for (int k = 0; k < U; ++k) t[k] = a[q][k];
#pragma unroll
      for (int c = 0; c + U < W; ++c) a[q][c] = a[q][c + U];
#pragma unroll
      for (int k = 0; k < U; ++k) a[q][W - U + k] = t[k];
    }
  }
  if (!exact) {
    // the last column's scaling: report once more
    const bool nan_here = __syncthreads_or(chk != chk);
    if (wave == 0 && lane == 0) st_rec_i32(&sy->pcand[nbar & 1u][w].nan, nan_here ? 1 : 0);
    rec_grid_barrier(sy, ctr, ++nbar * (unsigned)G);
    if (wave == 0) {
      const int cn = lane < G ? ld_rec_i32(&sy->pcand[(nbar - 1) & 1u][lane].nan) : 0;
      const bool any_nan = __any(cn != 0);
      if (lane == 0) s_nan = any_nan ? 1 : 0;
    }
    __syncthreads();
    exact = s_nan != 0;
  }
  if (exact) {
    if (w == 0 && tid == 0) *info = info0;  // the fast form's info is discarded
    panel_exact<W>(m_root, root, lda, s, ipiv, info, row_off, col_off, sy, ctr, nbar, r0, r1, s_row, s_rowj, s_wkey,
                   s_widx, s_piv, &s_p, &s_nan);
  } else {
#pragma unroll
    for (int q = 0; q < RPL; ++q) {
      const int i = r0 + q * 256 + tid;
      if (i < r1) {
#pragma unroll
        for (int c = 0; c < W; ++c) root[i + (size_t)(s + c) * lda] = from_double(a[q][c]);
      }
    }
  }
  __syncthreads();  // s_piv
  // the node's interchanges, in order, in the block's other columns
  for (int col = w * 256 + tid; col < w_root; col += G * 256) {
    if (col >= s && col < s + W) continue;
    uint32_t* x = root + (size_t)col * lda;
    for (int jj = 0; jj < W; ++jj) {
      const int j = s + jj, p = s_piv[jj];
      if (p != j) {
        const uint32_t t = x[j];
        x[j] = x[p];
        x[p] = t;
      }
    }
  }
  if (w == 0 && tid == 0 && __hip_atomic_load(&sy->fail, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
    *info = PAIR_SYNC_FAILED_INFO;
}

}  // namespace leaf
}  // namespace pga
