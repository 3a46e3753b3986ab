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
// to the posit grid as the GEMM does (rgemm_tile.hpp: fp64_mac), with the constants
// computed from the exponent (fp64_grid_c_arith).  A scale outside -108 .. 107 or a NaR
// gives NaN; every workgroup reports whether it has seen one with its next
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
PGA_D void rec_grid_barrier(PairSync* sy, unsigned* ctr, unsigned target, int G) {
  __syncthreads();
  if (threadIdx.x < 64) {
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __builtin_amdgcn_s_waitcnt(0);  // wave 0's record stores have completed
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
  }
  if (G > 1 && threadIdx.x == 0) {
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

// round(t * a) on the posit grid, rounded once: fma(t, a, C) - C with C from the binade of
// t * a rounded to FP64 (the GEMM's product rounding, rgemm_tile.hpp, with the constant
// computed instead of read from a table)
PGA_D double panel_mul(double t, double a) {
  const double c = fp64_grid_c_arith(t * a);
  return __builtin_fma(t, a, c) - c;
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

// Wave 0: the first maximum over the G candidates of one step (every lane gets it); lane
// c < W receives the winner's value c in pv and row j's in rj.  Element e = g * W + c of
// the candidates' values is loaded by lane e % 64 into slot e / 64 (the same slot for all
// W values of one candidate), so every load is in flight at once.
template <int W>
PGA_D void panel_gather(PairSync* sy, int par, int G, double& pv, double& rj, int& pidx, bool& any_nan) {
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
  rj = lane < W ? ld_rec_f64(&sy->pown[par][lane]) : 0.0;
  any_nan = __any(cn != 0);
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
  pv = cv[0];
#pragma unroll
  for (int k = 1; k < K; ++k)
    if (k == slot) pv = cv[k];
  pv = __shfl(pv, (cg * W + lane) & 63);
  pidx = ci;
}

// Wave 0: panel_gather, with the winner's row and row j into LDS and the pivot row and
// the NaN report into *s_p and *s_nan.
template <int W>
PGA_D void panel_pick(PairSync* sy, int par, int G, double* s_row, double* s_rowj, int* s_p, int* s_nan) {
  const int lane = threadIdx.x & 63;
  double pv, rj;
  int pidx;
  bool any_nan;
  panel_gather<W>(sy, par, G, pv, rj, pidx, any_nan);
  if (lane < W) {
    s_row[lane] = pv;
    s_rowj[lane] = rj;
  }
  if (lane == 0) {
    *s_p = pidx;
    *s_nan = any_nan ? 1 : 0;
  }
}

// The split point of the recursion node that separates node columns jj < lc: lc with its
// bits below the highest bit of jj ^ lc cleared (W a power of two, every node split in
// halves).  Rows above it receive column jj's term through Rtrsm, which skips a zero
// multiplier.
PGA_HD int panel_mid(int jj, int lc) { return lc & ~((1 << (31 - __builtin_clz(jj ^ lc))) - 1); }

// One term a + (-u) * l rounded as the GEMM rounds it (product, then sum); NaN where a
// scale falls outside -108 .. 107 or an operand is NaR.
PGA_D double panel_mac(double a, double u, double l) {
  const double t = -u;
  const double cp = fp64_grid_c_arith(t * l);
  const double x = a + (__builtin_fma(t, l, cp) - cp);
  const double cs = fp64_grid_c_arith(x);
  return (x + cs) - cs;
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
                            int col_off, PairSync* sy, unsigned* ctr, unsigned& nbar, int G, int w, int r0, int r1,
                            double* s_row, double* s_rowj, double* s_wkey, int* s_widx, int* s_piv, int* s_p,
                            int* s_nan) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
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
    rec_grid_barrier(sy, ctr, ++nbar * (unsigned)G, G);
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
        const int mid = panel_mid(jj, lc);
        if (uc != 0 || i >= s + mid) {
          uint32_t* x = root + i + (size_t)(s + lc) * lda;
          *x = add(*x, mul(neg(uc), l));
        }
      }
    }
    __syncthreads();
  }
}

// ---- the fast form's exchange: tagged words, no barrier ----
// A step's words carry a tag (1 .. 2047) made of the launch and the step's exchange
// number; a reader polls the words it needs until each carries the tag it expects, so it
// never waits on a counter, and a writer needs no ordering between its stores.
constexpr unsigned long long PANEL_ROW_NONE = 0xFFFFFull;  // row field of "no candidate"
PGA_D unsigned panel_tag(int launch, unsigned n) { return (unsigned)(((unsigned)launch * 128u + n) % 2047u) + 1u; }
PGA_D void st_word(unsigned long long* p, unsigned long long v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
PGA_D unsigned long long ld_word(unsigned long long* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Sort key of a value in row i for the pivot search: larger |x| first, then the smaller
// row (the first maximum); NaR below zero.  |x|'s FP64 bits keep the order of posit
// magnitudes, and a posit's 27 fraction bits all lie above bit 25.
PGA_D unsigned long long panel_sortkey(double x, int i) {
  const unsigned long long mag =
      (x != x) ? 0ull : (((unsigned long long)__double_as_longlong(__builtin_fabs(x)) >> 25) + 1ull);
  return (mag << 20) | (PANEL_ROW_NONE - (unsigned long long)i);
}
PGA_D unsigned long long shfl_xor_u64(unsigned long long v, int off) {
  const unsigned lo = (unsigned)__shfl_xor((int)(unsigned)v, off), hi = (unsigned)__shfl_xor((int)(unsigned)(v >> 32), off);
  return ((unsigned long long)hi << 32) | lo;
}
// the workgroup's largest key (every lane gets it)
PGA_D unsigned long long panel_wg_best(unsigned long long v, unsigned long long* s_wv) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const unsigned long long o = shfl_xor_u64(v, off);
    v = o > v ? o : v;
  }
  if (lane == 0) s_wv[wave] = v;
  __syncthreads();
  v = s_wv[0];
#pragma unroll
  for (int k = 1; k < 4; ++k) v = s_wv[k] > v ? s_wv[k] : v;
  return v;
}

// Wave 0 publishes this workgroup's candidate (row `bi`, its values in s_cand; key from
// column kc) and, for workgroup 0, row j + 1 (s_row1), with tag t.
template <int W>
PGA_D void panel_publish(PairSync* sy, int par, int w, unsigned t, const double* s_cand, const double* s_row1, int bi,
                         int kc, bool nan_here) {
  const int lane = threadIdx.x & 63;
  const unsigned long long th = (unsigned long long)t << 32;
  uint32_t pc = 0;
  if (lane < W) {
    pc = from_double(s_cand[lane]);
    st_word(&sy->tval[par][w][lane], th | pc);
    if (w == 0) st_word(&sy->town[par][lane], th | from_double(s_row1[lane]));
  }
  const uint32_t pk = (uint32_t)__shfl((int)pc, kc);
  if (lane == 0) {
    const bool none = bi == (int)PANEL_ROW_NONE;
    const unsigned long long key = (none || pk == P_NAR) ? 0ull : (unsigned long long)abs_p(pk) + 1ull;
    st_word(&sy->thead[par][w], key | ((unsigned long long)(unsigned)bi << 32) | ((unsigned long long)nan_here << 52) |
                                    ((unsigned long long)t << 53));
  }
}

// Waves 0 and 1 poll a step's words (tag t) until all have arrived, then take the first
// maximum over the G candidates (both from the headers).  Element e = g * W + c of the
// candidates' values is held by lane e % 64 in slot e / 64 (the same slot for all W values
// of one candidate); wave `half` polls half of the slots, so that neither holds all of
// them.  The wave whose half holds the winner gets its value c in lane c's pv (have_pv);
// wave 0 also gets row j + 1's value c in rj.  pidx is the winner's row, any_nan whether a
// workgroup reported a NaN.  Returns false if the wait gave up (bounded, as
// rec_grid_barrier; the failure word is set).
template <int W>
PGA_D bool panel_poll(PairSync* sy, int par, int G, unsigned t, int half, double& pv, bool& have_pv, double& rj,
                      int& pidx, bool& any_nan) {
  constexpr int K = PANEL_MC_MAX_G * W / 64, KH = K / 2;
  static_assert(K % 2 == 0, "two halves");
  const int lane = threadIdx.x & 63;
  unsigned long long h = 0, o = 0;
  uint32_t v[KH] = {};  // the value patterns (their tags checked as they arrive)
  unsigned spins = 0;
  for (;;) {
    bool ok = true;
    if (lane < G) {
      h = ld_word(&sy->thead[par][lane]);
      ok = (unsigned)(h >> 53) == t;
    }
#pragma unroll
    for (int k = 0; k < KH; ++k) {
      const int e = lane + 64 * (half * KH + k), g = e / W, c = e % W;
      if (g < G) {
        const unsigned long long x = ld_word(&sy->tval[par][g][c]);
        v[k] = (uint32_t)x;
        ok = ok && (unsigned)(x >> 32) == t;
      }
    }
    if (half == 0 && lane < W) {
      o = ld_word(&sy->town[par][lane]);
      ok = ok && (unsigned)(o >> 32) == t;
    }
    if (__all(ok)) break;
    __builtin_amdgcn_s_sleep(1);
    if (++spins > (1u << 21) || __hip_atomic_load(&sy->fail, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
      if (lane == 0) __hip_atomic_fetch_or(&sy->fail, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      pv = rj = 0.0;
      have_pv = half == 0;
      pidx = -1;
      any_nan = true;
      return false;
    }
  }
  any_nan = __any(lane < G && ((h >> 52) & 1ull));
  // first maximum: key, then the smaller row
  const unsigned row = (unsigned)(h >> 32) & 0xFFFFFu;
  unsigned long long best = lane < G ? (((h & 0xFFFFFFFFull) << 20) | (PANEL_ROW_NONE - row)) : 0ull;
  int bg = lane;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const unsigned long long ob = shfl_xor_u64(best, off);
    const int og = __shfl_xor(bg, off);
    if (ob > best) {
      best = ob;
      bg = og;
    }
  }
  pidx = (int)(PANEL_ROW_NONE - (best & PANEL_ROW_NONE));
  const int slot = ((bg * W) >> 6) - half * KH;
  have_pv = slot >= 0 && slot < KH;
  uint32_t pw = v[0];
#pragma unroll
  for (int k = 1; k < KH; ++k)
    if (k == slot) pw = v[k];
  const uint32_t pp = (uint32_t)__shfl((int)pw, (bg * W + lane) & 63);
  pv = to_double(pp);
  rj = to_double((uint32_t)o);
  return true;
}

template <int W, int RPL, int U, int GS>
__device__ __forceinline__ void panel_mc(int m_root, int w_root, uint32_t* root, int lda, int s, int* ipiv, int* info,
                                         int row_off, int col_off, PairSync* sy, int launch) {
  static_assert(W <= PANEL_MC_MAX_W && W <= 64 && W % U == 0 && U % GS == 0 && (W & (W - 1)) == 0, "node width");
  // per step parity: the pivot row and row j of the step (register column order)
  __shared__ double s_prow[2][W], s_jrow[2][W];
  __shared__ double s_cand[W], s_row1[W];  // this workgroup's candidate and row j + 1, to publish
  __shared__ double s_wkey[4];
  __shared__ int s_widx[4];
  __shared__ unsigned long long s_wv[4];
  __shared__ int s_piv[W];
  __shared__ int s_p[2], s_nan;
  const int G = gridDim.x, w = blockIdx.x, tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  unsigned* ctr = &sy->pctr[launch];
  // workgroup 0, lane 0: the factorization's info at the start and as this node sets it
  const int info_start = (w == 0 && tid == 0) ? *info : 0;
  int info_now = info_start;
  unsigned nbar = 0;  // exchanges passed in this launch
  const int rows = m_root - s, R = (rows + G - 1) / G;
  const int r0 = s + w * R, r1 = min(m_root, r0 + R);
  double a[RPL][W];
#pragma unroll
  for (int q = 0; q < RPL; ++q) {
    const int i = r0 + q * 256 + tid;
#pragma unroll
    for (int c = 0; c < W; ++c) a[q][c] = (i < r1) ? to_double(root[i + (size_t)(s + c) * lda]) : 0.0;
  }
  double chk = 0.0;  // sum of every FP64 result since the last report: NaN if any was
  bool exact = false;
  // the pivot search of column 0
  {
    unsigned long long best = 0;
#pragma unroll
    for (int q = 0; q < RPL; ++q) {
      const int i = r0 + q * 256 + tid;
      if (i < r1) {
        const unsigned long long k = panel_sortkey(a[q][0], i);
        best = k > best ? k : best;
      }
    }
    best = panel_wg_best(best, s_wv);
    const int bi = best ? (int)(PANEL_ROW_NONE - (best & PANEL_ROW_NONE)) : (int)PANEL_ROW_NONE;
#pragma unroll
    for (int q = 0; q < RPL; ++q)
      if (r0 + q * 256 + tid == bi) {
#pragma unroll
        for (int c = 0; c < W; ++c) s_cand[c] = a[q][c];
      }
    if (w == 0 && tid == 0) {
#pragma unroll
      for (int c = 0; c < W; ++c) s_row1[c] = a[0][c];
    }
    __syncthreads();
    if (wave == 0) panel_publish<W>(sy, 0, w, panel_tag(launch, 0), s_cand, s_row1, bi, 0, false);
    if (wave < 2) {
      double pv, rj;
      int pidx;
      bool have_pv, any_nan;
      panel_poll<W>(sy, 0, G, panel_tag(launch, 0), wave, pv, have_pv, rj, pidx, any_nan);
      if (have_pv && lane < W) s_prow[0][lane] = pv;
      if (wave == 0 && lane < W) s_jrow[0][lane] = rj;
      if (wave == 0 && lane == 0) s_p[0] = pidx;
    }
    ++nbar;
    __syncthreads();
  }
  // Step jj: the interchange, scaling and update of column jj; on the way, column jj + 1
  // is updated first and its candidates published, so that the exchange runs while the
  // workgroups update their remaining columns.  A candidate is published as its row stands
  // before its remaining columns' terms of this step; every workgroup applies those terms
  // to the winner (and to row jj + 1) itself, with the same operations as the owner.
  for (int jb = 0; jb < W && !exact; jb += U) {
#pragma unroll
    for (int u = 0; u < U; ++u) {
      if (exact) continue;  // (no break: the loop must unroll, so that register indices stay static)
      const int jj = jb + u, j = s + jj, cur = jj & 1, nxt = cur ^ 1;
      const bool more = jj + 1 < W;
      const int p = panel_pivot(sy, s_p[cur], j, m_root);
      const double ap = s_prow[cur][u];
      const bool swapped = p != j;
      if (tid == 0) {
        s_piv[jj] = p;
        if (w == 0) {
          ipiv[j] = row_off + p + 1;
          if (ap == 0.0 && info_now == 0) {
            info_now = col_off + j + 1;
            *info = info_now;
          }
        }
      }
      // 1. interchange rows j and p (column j only for a nonzero pivot)
      if (swapped) {
#pragma unroll
        for (int q = 0; q < RPL; ++q) {
          const int i = r0 + q * 256 + tid;
          if (i == j || i == p) {
            const double* src = (i == j) ? s_prow[cur] : s_jrow[cur];
#pragma unroll
            for (int c = 0; c < W; ++c)
              if (c != u || ap != 0.0) a[q][c] = src[c];
          }
        }
      }
      const double r = (ap != 0.0 && ap == ap) ? fp64_posit_recip(ap) : 0.0;
      const double* urow = swapped ? s_prow[cur] : s_jrow[cur];  // row j after the interchange
      // 2. scale column j below row j (a NaR pivot makes it NaR: division by NaR), and
      //    apply it to column j + 1
#pragma unroll
      for (int q = 0; q < RPL; ++q) {
        const int i = r0 + q * 256 + tid;
        if (i > j && i < r1) {
          if (ap != 0.0) {
            a[q][u] = (ap == ap) ? panel_mul(r, a[q][u]) : ap;
            chk += a[q][u];
          }
          if (more) {
            const double uc = urow[u + 1];
            const double y = panel_mac(a[q][u + 1], uc, a[q][u]);
            chk += y;
            if (uc != 0.0 || i >= s + panel_mid(jj, jj + 1)) a[q][u + 1] = y;
          }
        }
      }
      // 3. the candidates of column j + 1, published
      if (more) {
        unsigned long long best = 0;
#pragma unroll
        for (int q = 0; q < RPL; ++q) {
          const int i = r0 + q * 256 + tid;
          if (i < r1 && i > j) {
            const unsigned long long k = panel_sortkey(a[q][u + 1], i);
            best = k > best ? k : best;
          }
        }
        best = panel_wg_best(best, s_wv);
        const int bi = best ? (int)(PANEL_ROW_NONE - (best & PANEL_ROW_NONE)) : (int)PANEL_ROW_NONE;
        const bool nan_here = __syncthreads_or(chk != chk);
        chk = 0.0;
#pragma unroll
        for (int q = 0; q < RPL; ++q)
          if (r0 + q * 256 + tid == bi) {
#pragma unroll
            for (int c = 0; c < W; ++c) s_cand[c] = a[q][c];
          }
        if (w == 0 && tid == jj + 1) {  // row j + 1 = s + jj + 1: workgroup 0, slot 0
#pragma unroll
          for (int c = 0; c < W; ++c) s_row1[c] = a[0][c];
        }
        __syncthreads();
        if (wave == 0) panel_publish<W>(sy, (int)(nbar & 1u), w, panel_tag(launch, nbar), s_cand, s_row1, bi, u + 1,
                                        nan_here);
      }
      // 4. the remaining columns, in groups of GS: a group is wholly live or wholly final
      //    (W - jb is a multiple of U, and U of GS), so one uniform branch a group; within
      //    it the multiply-adds run side by side
#pragma unroll
      for (int q = 0; q < RPL; ++q) {
        const int i = r0 + q * 256 + tid;
        if (i > j && i < r1) {
          const double l = a[q][u];
#pragma unroll
          for (int g0 = 0; g0 < W; g0 += GS) {
            if (g0 + GS - 1 <= u + 1) continue;  // columns up to j + 1: done
            if (g0 >= W - jb) continue;          // node columns before jb: final
#pragma unroll
            for (int k = 0; k < GS; ++k) {
              const int c = g0 + k;
              if (c <= u + 1) continue;
              const double uc = urow[c];
              const double y = panel_mac(a[q][c], uc, l);
              chk += y;
              if (uc != 0.0 || i >= s + panel_mid(jj, jb + c)) a[q][c] = y;
            }
          }
        }
      }
      // 5. the next step's pivot: wave 0 takes the first maximum and applies this step's
      //    remaining terms to the winner and to row j + 1
      if (more) {
        if (wave < 2) {
          double pv, rj;
          int pidx;
          bool have_pv, any_nan;
          panel_poll<W>(sy, (int)(nbar & 1u), G, panel_tag(launch, nbar), wave, pv, have_pv, rj, pidx, any_nan);
          const double lp = __shfl(pv, u), lj = __shfl(rj, u);
          const bool live = lane > u + 1 && lane < W - jb;
          const double uc = live ? urow[lane] : 1.0;
          const int mid = live ? s + panel_mid(jj, jb + lane) : 0;
          if (have_pv) {
            if (live && (uc != 0.0 || pidx >= mid)) pv = panel_mac(pv, uc, lp);
            if (lane < W) s_prow[nxt][lane] = pv;
          }
          if (wave == 0) {
            if (live && (uc != 0.0 || j + 1 >= mid)) rj = panel_mac(rj, uc, lj);
            if (lane < W) s_jrow[nxt][lane] = rj;
            if (lane == 0) {
              s_p[nxt] = pidx;
              s_nan = any_nan ? 1 : 0;
            }
          }
        }
        ++nbar;
        __syncthreads();
        if (s_nan) exact = true;  // a NaN somewhere: the whole node again, exactly
      }
    }
    if (exact) break;
    // rotate by U: register column c takes node column (jb + U + c) mod W; so do the next
    // step's rows in LDS
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
    if (jb + U < W) {
      const int nx = (jb + U) & 1;
      double pv = 0.0, rj = 0.0;
      if (tid < W) {
        pv = s_prow[nx][(tid + U) % W];
        rj = s_jrow[nx][(tid + U) % W];
      }
      __syncthreads();
      if (tid < W) {
        s_prow[nx][tid] = pv;
        s_jrow[nx][tid] = rj;
      }
      __syncthreads();
    }
  }
  if (!exact) {
    // the last column's scaling: report once more (headers only)
    const bool nan_here = __syncthreads_or(chk != chk);
    const unsigned t = panel_tag(launch, nbar);
    const int par = (int)(nbar & 1u);
    if (tid == 0)
      st_word(&sy->thead[par][w], ((unsigned long long)nan_here << 52) | ((unsigned long long)t << 53));
    if (wave == 0) {
      unsigned long long h = 0;
      unsigned spins = 0;
      for (;;) {
        if (lane < G) h = ld_word(&sy->thead[par][lane]);
        if (__all(lane >= G || (unsigned)(h >> 53) == t)) break;
        __builtin_amdgcn_s_sleep(1);
        if (++spins > (1u << 21) || __hip_atomic_load(&sy->fail, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
          if (lane == 0) __hip_atomic_fetch_or(&sy->fail, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          break;
        }
      }
      const bool any_nan = __any(lane < G && ((h >> 52) & 1ull));
      if (lane == 0) s_nan = any_nan ? 1 : 0;
    }
    ++nbar;
    __syncthreads();
    exact = s_nan != 0;
  }
  if (exact) {
    if (w == 0 && tid == 0) *info = info_start;  // the fast form's info is discarded
    unsigned nb_exact = 0;  // the launch's ticket counter serves the exact form alone
    panel_exact<W>(m_root, root, lda, s, ipiv, info, row_off, col_off, sy, ctr, nb_exact, G, w, r0, r1, s_prow[0], s_jrow[0],
                   s_wkey, s_widx, s_piv, &s_p[0], &s_nan);
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
