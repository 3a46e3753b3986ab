// lapack.hip -- posit32 BLAS-2/3 kernels and LU / Cholesky drivers on the GPU.
//
// Every routine keeps the per-element operation order of the reference routine cited
// beside it, so results are bit-identical; all arithmetic runs on the device.  The
// factorization drivers follow the reference's blocking and recursion exactly
// (Rgetrf_opencl.cpp, Rgetrf2.cpp, Rpotrf_opencl.cpp, Rpotrf2.cpp), which also keeps
// the reference's zero-multiplier skips of Rtrsm / Rsyrk in the same places; the
// recursion is driven from the host, the work is launched on one stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <mutex>
#include <vector>

#include "pga_internal.hpp"
#include "posit32_core.hpp"
#include "rgemm_kernels.hpp"
#include "rgemm_tile.hpp"
#include "lu_leaf.hpp"
#include "lu_panel.hpp"
#include "wg_factor.hpp"

namespace pga {
namespace la {

constexpr int WG = 256;

__device__ __forceinline__ bool stopped(const int* stop) { return stop && *stop; }

// ---------------------------------------------------------------------------------
// TRSM, left side: one workgroup per column of B (Rtrsm.cpp:115-188)
// ---------------------------------------------------------------------------------
// (the not-transposed forms and the upper transposed one; k_trsm_lltn is the fourth)
template <bool UPPER, bool TRANS>
__global__ __launch_bounds__(WG) void k_trsm_left(int m, int n, uint32_t alpha, int nounit,
                                                  const uint32_t* __restrict__ A, int lda,
                                                  uint32_t* __restrict__ B, int ldb, const int* stop) {
  static_assert(UPPER || !TRANS, "the lower transposed form is k_trsm_lltn");
  PGA_KERNEL_VGPR_FLOOR();
  PGA_URGENT();
  const int j = blockIdx.x;
  if (j >= n || stopped(stop)) return;
  uint32_t* b = B + (size_t)j * ldb;
  const int tid = threadIdx.x;
  auto a = [&](int r, int c) { return A[r + (size_t)c * lda]; };

  if (!TRANS) {
    // B := alpha*B, then column sweep; a zero b_k skips step k (Rtrsm.cpp:128,146)
    if (alpha != P_ONE) {
      for (int i = tid; i < m; i += WG) b[i] = mul(alpha, b[i]);
      __syncthreads();
    }
    for (int s = 0; s < m; ++s) {
      const int k = UPPER ? m - 1 - s : s;
      uint32_t bk = b[k];
      bool active = (bk != 0);
      if (active) {
        if (nounit) bk = div(bk, a(k, k));
        if (UPPER) {
          for (int i = tid; i < k; i += WG) b[i] = sub(b[i], mul(bk, a(i, k)));
        } else {
          for (int i = k + 1 + tid; i < m; i += WG) b[i] = sub(b[i], mul(bk, a(i, k)));
        }
      }
      __syncthreads();
      if (active && nounit && tid == 0) b[k] = bk;
    }
    __syncthreads();
  } else {
    // temp_i = alpha*b_i - sum_{k<i} a_ki*b_k in ascending k (Rtrsm.cpp:161-173); b_k is
    // final before any term that needs it, so the sums advance as a sweep over k
    for (int i = tid; i < m; i += WG) b[i] = mul(alpha, b[i]);
    __syncthreads();
    for (int k = 0; k < m; ++k) {
      uint32_t bk = b[k];
      if (nounit) bk = div(bk, a(k, k));
      for (int i = k + 1 + tid; i < m; i += WG) b[i] = sub(b[i], mul(a(k, i), bk));
      __syncthreads();
      if (nounit && tid == 0) b[k] = bk;
    }
    __syncthreads();
  }
}

// Left, lower, transposed (Rtrsm.cpp:174-186), one right-hand side per workgroup:
// temp_i = alpha*b_i - sum_{k>i} a_ki*b_k in ascending k, i descending.  Row i's first
// term needs b_{i+1}, so the rows form one chain of m^2/2 rounded additions.  Per row
// the lanes form the negated products (posits, exact in FP64) in LDS, CHAIN_CHUNK terms
// at a time; wave 0 then runs the additions 64 terms a block, one term per lane.
//
// While the running sum stays in one binade the additions are integer steps on that
// binade's posit grid (chain_grid_*, posit32_core.hpp): each lane turns its term into a
// step d (and a tie flag), one wave-wide prefix sum gives every lane the grid value
// after its term, ties are settled in order (each rounds to the even value), and every
// lane checks that its value lies strictly inside the binade.  The values before the
// first lane that fails are the posit sums; that lane's sum is formed on its own with
// fp64_posit_add (the exact decoded addition where that does not round it), the grid
// is taken anew from the result, and the next block starts with the following term.
// A NaR term or start value makes the row NaR, as NaR absorbs every addition.
constexpr int CHAIN_CHUNK = 4096;

// t <- t + p exactly for posit values t and p (finite), through the decoded form
__device__ __forceinline__ void chain_add_exact(double& t, double p) {
  dec::Val v = dec::decode(from_double(t));
  gemm::add_checked(v, dec::decode(from_double(p)));
  t = to_double(dec::encode(v));
}

// one posit addition of the chain, every lane alike
__device__ __forceinline__ double chain_add_one(double t, double p) {
  uint32_t reach = 0;
  const double r = fp64_posit_add(t, p, reach);
  if (__builtin_amdgcn_readfirstlane(reach) <= FP64_ADD_REACH_MAX) return r;
  chain_add_exact(t, p);
  return t;
}

// Inclusive prefix sum over the 64 lanes of a wave: shifts by 1, 2, 4 and 8 lanes within
// each row of 16, then the last lane of row 0 added to row 1 and of row 2 to row 3, then
// the last lane of row 1 to rows 2 and 3 (DPP row_shr / row_bcast; lanes without a source
// add the 0 passed as the old value).
__device__ __forceinline__ int wave_prefix_sum(int v) {
  v += __builtin_amdgcn_update_dpp(0, v, 0x111, 0xf, 0xf, false);  // row_shr:1
  v += __builtin_amdgcn_update_dpp(0, v, 0x112, 0xf, 0xf, false);  // row_shr:2
  v += __builtin_amdgcn_update_dpp(0, v, 0x114, 0xf, 0xf, false);  // row_shr:4
  v += __builtin_amdgcn_update_dpp(0, v, 0x118, 0xf, 0xf, false);  // row_shr:8
  v += __builtin_amdgcn_update_dpp(0, v, 0x142, 0xa, 0xf, false);  // row_bcast:15
  v += __builtin_amdgcn_update_dpp(0, v, 0x143, 0xc, 0xf, false);  // row_bcast:31
  return v;
}

// The chain's additions over the kn terms in sp (LDS), by one full wave.
__device__ __forceinline__ void chain_terms(double& t, const double* __restrict__ sp, int kn) {
  const int lane = (int)__builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
  ChainGrid g;
  bool live = __builtin_amdgcn_readfirstlane((uint32_t)chain_grid_setup(t, g)) != 0;
  int k = 0;
  while (k < kn) {
    if (!live) {  // zero or a scale outside the grids' range: one term on its own
      t = chain_add_one(t, sp[k]);
      ++k;
      live = __builtin_amdgcn_readfirstlane((uint32_t)chain_grid_setup(t, g)) != 0;
      continue;
    }
    const int cnt = min(64, kn - k);
    const bool valid = lane < cnt;
    const double p = valid ? sp[k + lane] : 0.0;
    int d;
    bool tie;
    chain_grid_term(p, g.qinv, d, tie);
    const int D = wave_prefix_sum(valid ? d : 0);
    // ties in lane order: tie w rounds m0 + D_w + (the ties before it) up to even
    int c = 0, cs = 0;
    for (uint64_t T = __builtin_amdgcn_ballot_w64(tie && valid); T; T &= T - 1) {
      const int w = __builtin_ctzll(T);
      const int up = (g.m0 + __builtin_amdgcn_readlane(D, w) + cs) & 1;
      cs += up;
      c += (up && lane >= w) ? 1 : 0;
    }
    const int m = g.m0 + D + c;
    const uint64_t F = __builtin_amdgcn_ballot_w64(valid && !(m > g.lo && m < g.hi));
    if (F == 0) {
      g.m0 = __builtin_amdgcn_readlane(m, cnt - 1);
      k += cnt;
      continue;
    }
    const int us = __builtin_ctzll(F);  // first term whose sum left the binade
    if (us > 0) g.m0 = __builtin_amdgcn_readlane(m, us - 1);
    t = chain_add_one(chain_grid_value(g, g.m0), sp[k + us]);
    k += us + 1;
    live = __builtin_amdgcn_readfirstlane((uint32_t)chain_grid_setup(t, g)) != 0;
  }
  if (live) t = chain_grid_value(g, g.m0);
}

__global__ __launch_bounds__(WG) void k_trsm_lltn(int m, int n, uint32_t alpha, int nounit,
                                                  const uint32_t* __restrict__ A, int lda,
                                                  uint32_t* __restrict__ B, int ldb, const int* stop) {
  PGA_KERNEL_VGPR_FLOOR();
  __shared__ double sp[CHAIN_CHUNK];
  const int j = blockIdx.x;
  if (j >= n || stopped(stop)) return;
  uint32_t* b = B + (size_t)j * ldb;
  const int tid = threadIdx.x;
  for (int i = m - 1; i >= 0; --i) {
    const uint32_t t0 = mul(alpha, b[i]);  // every lane reads the same values
    bool nar = t0 == P_NAR;
    double t = nar ? 0.0 : to_double(t0);
    for (int k0 = i + 1; k0 < m; k0 += CHAIN_CHUNK) {
      const int kn = min(CHAIN_CHUNK, m - k0);
      __syncthreads();  // wave 0 has used the previous chunk
      bool term_nar = false;
      for (int u = tid; u < kn; u += WG) {
        const uint32_t v = neg(mul(A[k0 + u + (size_t)i * lda], b[k0 + u]));
        term_nar |= v == P_NAR;
        sp[u] = term_nar ? 0.0 : to_double(v);
      }
      nar |= __syncthreads_or(term_nar) != 0;
      if (tid < 64 && !nar) chain_terms(t, sp, kn);
    }
    if (tid == 0) {
      uint32_t tp = nar ? P_NAR : from_double(t);
      if (nounit) tp = div(tp, A[i + (size_t)i * lda]);
      b[i] = tp;
    }
    __syncthreads();  // b_i is final before row i-1 forms its products
  }
}

// The same solve with the products of the next row formed while wave 0 runs a row's
// chain, for rows of up to LLTN_PIPE_CAP terms (m <= LLTN_PIPE_CAP + 1).  In the step
// for row i, wave 0 adds row i's terms: the first (k = i+1, whose b_{i+1} the step
// before made final) formed by itself, the rest (k = i+2 .. m-1) from the LDS buffer
// the other waves filled in the step before; meanwhile the other three waves form row
// i-1's terms k = i+1 .. m-1 into the other buffer (b_k final for k > i).  One barrier
// a row hands the buffers over.
constexpr int LLTN_PIPE_CAP = 8192;
__global__ __launch_bounds__(WG) void k_trsm_lltn_pipe(int m, int n, uint32_t alpha, int nounit,
                                                       const uint32_t* __restrict__ A, int lda,
                                                       uint32_t* __restrict__ B, int ldb, const int* stop) {
  PGA_KERNEL_VGPR_FLOOR();
  __shared__ double sp[2][LLTN_PIPE_CAP];
  __shared__ int row_nar[2];  // a NaR among the buffer's terms
  const int j = blockIdx.x;
  if (j >= n || stopped(stop)) return;
  uint32_t* b = B + (size_t)j * ldb;
  const int tid = threadIdx.x;
  if (tid < 2) row_nar[tid] = 0;
  __syncthreads();
  for (int i = m - 1; i >= 0; --i) {
    if (tid < 64) {
      const uint32_t t0 = mul(alpha, b[i]);  // every lane of wave 0 reads the same values
      bool nar = t0 == P_NAR || row_nar[i & 1] != 0;
      double t = nar ? 0.0 : to_double(t0);
      if (i + 1 < m) {
        const uint32_t v = neg(mul(A[i + 1 + (size_t)i * lda], b[i + 1]));
        nar |= v == P_NAR;
        if (!nar) t = chain_add_one(t, to_double(v));
      }
      if (!nar && m - i - 2 > 0) chain_terms(t, sp[i & 1], m - i - 2);
      if (tid == 0) {
        uint32_t tp = nar ? P_NAR : from_double(t);
        if (nounit) tp = div(tp, A[i + (size_t)i * lda]);
        b[i] = tp;
        row_nar[i & 1] = 0;  // the buffer's next row (i-2) is formed after the barrier
      }
    } else if (i >= 1) {
      const int r = i - 1, cnt = m - r - 2;  // row r's terms k = r+2 .. m-1
      double* out = sp[r & 1];
      const uint32_t* ar = A + (size_t)r * lda + r + 2;
      const uint32_t* bk = b + r + 2;
      bool term_nar = false;
      for (int u = tid - 64; u < cnt; u += WG - 64) {
        const uint32_t v = neg(mul(ar[u], bk[u]));
        term_nar |= v == P_NAR;
        out[u] = v == P_NAR ? 0.0 : to_double(v);
      }
      if (term_nar) row_nar[r & 1] = 1;
    }
    __syncthreads();  // b_i and row i-1's terms are final
  }
}

// TRSM, left side, not transposed (Rtrsm.cpp:115-155), `cpw` columns of B per
// workgroup: the same column sweep with the (row, column) pairs of a step spread over
// the lanes, so wide right-hand sides take fewer workgroup-steps.
template <bool UPPER>
__global__ __launch_bounds__(WG) void k_trsm_left_cols(int m, int n, uint32_t alpha, int nounit,
                                                       const uint32_t* __restrict__ A, int lda,
                                                       uint32_t* __restrict__ B, int ldb, int cpw,
                                                       const int* stop) {
  PGA_KERNEL_VGPR_FLOOR();
  PGA_URGENT();
  const int c0 = blockIdx.x * cpw;
  if (c0 >= n || stopped(stop)) return;
  const int C = min(cpw, n - c0);
  uint32_t* b = B + (size_t)c0 * ldb;
  const int tid = threadIdx.x;
  if (alpha != P_ONE) {
    for (int e = tid; e < m * C; e += WG) b[e % m + (size_t)(e / m) * ldb] = mul(alpha, b[e % m + (size_t)(e / m) * ldb]);
    __syncthreads();
  }
  for (int s = 0; s < m; ++s) {
    const int k = UPPER ? m - 1 - s : s;
    const int lo = UPPER ? 0 : k + 1, cnt = UPPER ? k : m - 1 - k;
    const uint32_t akk = nounit ? A[k + (size_t)k * lda] : P_ONE;
    for (int e = tid; e < cnt * C; e += WG) {
      const int i = lo + e % cnt, c = e / cnt;
      uint32_t bk = b[k + (size_t)c * ldb];
      if (bk != 0) {  // a zero b_k skips step k (Rtrsm.cpp:128,146)
        if (nounit) bk = div(bk, akk);
        b[i + (size_t)c * ldb] = sub(b[i + (size_t)c * ldb], mul(bk, A[i + (size_t)k * lda]));
      }
    }
    __syncthreads();
    if (nounit)  // b_k = b_k / a_kk once every lane has used the old b_k
      for (int c = tid; c < C; c += WG) {
        const uint32_t bk = b[k + (size_t)c * ldb];
        if (bk != 0) b[k + (size_t)c * ldb] = div(bk, akk);
      }
  }
}

// ---------------------------------------------------------------------------------
// TRSM, right side: rows of B are independent; one lane per row (Rtrsm.cpp:189-286)
// ---------------------------------------------------------------------------------
template <bool UPPER, bool TRANS>
__global__ __launch_bounds__(WG) void k_trsm_right(int m, int n, uint32_t alpha, int nounit,
                                                   const uint32_t* __restrict__ A, int lda,
                                                   uint32_t* __restrict__ B, int ldb,
                                                   const int* stop) {
  PGA_KERNEL_VGPR_FLOOR();
  PGA_URGENT();
  const int i = blockIdx.x * WG + threadIdx.x;
  if (i >= m || stopped(stop)) return;
  auto a = [&](int r, int c) { return A[r + (size_t)c * lda]; };
  auto b = [&](int c) -> uint32_t& { return B[i + (size_t)c * ldb]; };
  if (!TRANS) {
    if (UPPER) {  // :194-214
      for (int j = 0; j < n; ++j) {
        if (alpha != P_ONE) b(j) = mul(alpha, b(j));
        for (int k = 0; k < j; ++k)
          if (a(k, j) != 0) b(j) = sub(b(j), mul(a(k, j), b(k)));
        if (nounit) {
          uint32_t t = div(P_ONE, a(j, j));
          b(j) = mul(t, b(j));
        }
      }
    } else {  // :215-236
      for (int j = n - 1; j >= 0; --j) {
        if (alpha != P_ONE) b(j) = mul(alpha, b(j));
        for (int k = j + 1; k < n; ++k)
          if (a(k, j) != 0) b(j) = sub(b(j), mul(a(k, j), b(k)));
        if (nounit) {
          uint32_t t = div(P_ONE, a(j, j));
          b(j) = mul(t, b(j));
        }
      }
    }
  } else {
    if (UPPER) {  // :241-262
      for (int k = n - 1; k >= 0; --k) {
        if (nounit) {
          uint32_t t = div(P_ONE, a(k, k));
          b(k) = mul(t, b(k));
        }
        for (int j = 0; j < k; ++j) {
          uint32_t t = a(j, k);
          if (t != 0) b(j) = sub(b(j), mul(t, b(k)));
        }
        if (alpha != P_ONE) b(k) = mul(alpha, b(k));
      }
    } else {  // :263-285
      for (int k = 0; k < n; ++k) {
        if (nounit) {
          uint32_t t = div(P_ONE, a(k, k));
          b(k) = mul(t, b(k));
        }
        for (int j = k + 1; j < n; ++j) {
          uint32_t t = a(j, k);
          if (t != 0) b(j) = sub(b(j), mul(t, b(k)));
        }
        if (alpha != P_ONE) b(k) = mul(alpha, b(k));
      }
    }
  }
}

// TRSM, right side, transposed (Rtrsm.cpp:241-285) as a sweep over k: step k scales
// column k, then subtracts it from the columns it feeds (j < k upper, j > k lower), then
// applies alpha to column k.  Every element sees its terms in the reference's order; the
// (row, column) pairs of a step are independent.  A workgroup owns `rows` rows.
template <bool UPPER>
__device__ void right_sweep_rows(int R, int n, uint32_t alpha, int nounit, const uint32_t* __restrict__ A, int lda,
                                 uint32_t* __restrict__ b, int ldb) {
  const int tid = threadIdx.x;
  for (int s = 0; s < n; ++s) {
    const int k = UPPER ? n - 1 - s : s;
    if (nounit) {
      const uint32_t t = div(P_ONE, A[k + (size_t)k * lda]);
      for (int i = tid; i < R; i += WG) b[i + (size_t)k * ldb] = mul(t, b[i + (size_t)k * ldb]);
      __syncthreads();
    }
    const int cnt = UPPER ? k : n - 1 - k;
    for (int e = tid; e < R * cnt; e += WG) {
      const int i = e % R, j = UPPER ? e / R : k + 1 + e / R;
      const uint32_t t = A[j + (size_t)k * lda];
      if (t != 0) b[i + (size_t)j * ldb] = sub(b[i + (size_t)j * ldb], mul(t, b[i + (size_t)k * ldb]));
    }
    if (alpha != P_ONE) {
      __syncthreads();
      for (int i = tid; i < R; i += WG) b[i + (size_t)k * ldb] = mul(alpha, b[i + (size_t)k * ldb]);
    }
    __syncthreads();
  }
}

template <bool UPPER>
__global__ __launch_bounds__(WG) void k_trsm_right_sweep(int m, int n, uint32_t alpha, int nounit,
                                                         const uint32_t* __restrict__ A, int lda,
                                                         uint32_t* __restrict__ B, int ldb, int rows,
                                                         const int* stop) {
  PGA_KERNEL_VGPR_FLOOR();
  PGA_URGENT();
  const int r0 = blockIdx.x * rows;
  if (r0 >= m || stopped(stop)) return;
  right_sweep_rows<UPPER>(min(rows, m - r0), n, alpha, nounit, A, lda, B + r0, ldb);
}

// The lower Cholesky panel solve, B := B L^-T with L non-unit lower n x n, alpha = 1
// (Rtrsm.cpp:263-285), on RSW_ROWS rows of B per workgroup with B's rows and L held in
// LDS as FP64 values: step k scales column k by 1/l_kk (the reciprocals formed once, in
// parallel), then subtracts l_jk * b_k from every later column j (skipped when l_jk is
// zero), each product and sum rounded to the posit grid as in the GEMM (fp64_mac_arith:
// the GEMM's roundings with computed constants); L's strictly lower triangle is packed, so
// that the kernel's 25 KB of LDS leave it room beside the concurrent trailing update.  A
// scale outside -108 .. 107 or a NaR leaves the block untouched in memory and the rows are
// solved again with the pattern-level sweep.
constexpr int RSW_ROWS = 16, RSW_NMAX = 64;

__global__ __launch_bounds__(WG) void k_trsm_rltn_fp64(int m, int n, const uint32_t* __restrict__ A, int lda,
                                                       uint32_t* __restrict__ B, int ldb, const int* stop) {
  PGA_KERNEL_VGPR_FLOOR();
  PGA_URGENT();
  // the strictly lower triangle of L, packed by columns: l_jk (j > k) at tri(k) + j - k - 1
  __shared__ double sL[RSW_NMAX * (RSW_NMAX - 1) / 2];
  __shared__ double sB[RSW_NMAX][RSW_ROWS];  // sB[column][row]
  __shared__ double sR[RSW_NMAX];            // 1 / l_kk
  const int r0 = blockIdx.x * RSW_ROWS;
  if (r0 >= m || stopped(stop)) return;
  const int R = min(RSW_ROWS, m - r0);
  const int tid = threadIdx.x;
  uint32_t* b = B + r0;
  auto tri = [n](int k) { return k * n - k * (k + 1) / 2; };
  for (int k = 0; k + 1 < n; ++k)
    for (int j = k + 1 + tid; j < n; j += WG) sL[tri(k) + j - k - 1] = to_double(A[j + (size_t)k * lda]);
  for (int e = tid; e < R * n; e += WG) sB[e / R][e % R] = to_double(b[e % R + (size_t)(e / R) * ldb]);
  if (tid < n) {
    const double d = to_double(A[tid + (size_t)tid * lda]);
    sR[tid] = (d != 0.0 && d == d) ? fp64_posit_recip(d) : __builtin_nan("");  // NaR or zero: redo
  }
  __syncthreads();
  for (int k = 0; k < n; ++k) {
    if (tid < R) sB[k][tid] = fp64_posit_mul_arith(sR[k], sB[k][tid]);
    __syncthreads();
    const int cnt = n - 1 - k;
    for (int e = tid; e < R * cnt; e += WG) {
      const int i = e % R, j = k + 1 + e / R;
      const double t = sL[tri(k) + j - k - 1];
      if (t != 0.0) sB[j][i] = fp64_mac_arith(sB[j][i], -t, sB[k][i]);
    }
    __syncthreads();
  }
  bool bad = false;
  for (int e = tid; e < R * n; e += WG) bad |= sB[e / R][e % R] != sB[e / R][e % R];  // NaN
  if (__syncthreads_or(bad)) {
    right_sweep_rows<false>(R, n, P_ONE, 1, A, lda, b, ldb);
    return;
  }
  for (int e = tid; e < R * n; e += WG) b[e % R + (size_t)(e / R) * ldb] = from_double(sB[e / R][e % R]);
}

// B := U^-T B with U non-unit upper m x m, alpha = 1, at the pattern level on C columns
// of B, one after another (Rtrsm.cpp:161-173; the sweep of k_trsm_left<true, true>).
static __device__ void lutn_cols_pattern(int m, int C, const uint32_t* __restrict__ A, int lda,
                                         uint32_t* __restrict__ b, int ldb) {
  const int tid = threadIdx.x;
  for (int c = 0; c < C; ++c) {
    uint32_t* bc = b + (size_t)c * ldb;
    for (int k = 0; k < m; ++k) {
      const uint32_t bk = div(bc[k], A[k + (size_t)k * lda]);
      for (int i = k + 1 + tid; i < m; i += WG) bc[i] = sub(bc[i], mul(A[k + (size_t)i * lda], bk));
      __syncthreads();
      if (tid == 0) bc[k] = bk;
      __syncthreads();
    }
  }
}

// The upper Cholesky's row-block solve, B := U^-T B with U non-unit upper m x m, alpha =
// 1 (Rtrsm.cpp:161-173: temp = b_i - sum over k < i of u_ki * b_k in ascending k, then
// temp / u_ii), on LUTN_COLS columns of B per workgroup with B's columns and U held in
// LDS as FP64 values: step k divides b_k by u_kk (fp64_posit_div), then subtracts
// u_ki * b_k from every later row i (fp64_mac_arith: the GEMM's roundings with computed
// constants; this form skips no zero term).  A scale outside -108 .. 107, a NaR or a zero
// pivot leaves the block untouched in memory and the columns are solved again with the
// pattern-level sweep.
constexpr int LUTN_COLS = 16, LUTN_MMAX = 64;

// With St, B first takes the upper Cholesky's right-looking sums, b := -t + b with the
// sum of element (i, c) at St[c + i * lds] (k_apply_sums on the block row), written back
// before the solve.
__global__ __launch_bounds__(WG) void k_trsm_lutn_fp64(int m, int n, const uint32_t* __restrict__ A, int lda,
                                                       uint32_t* __restrict__ B, int ldb, const int* stop,
                                                       const uint32_t* __restrict__ St = nullptr, int lds = 0) {
  PGA_KERNEL_VGPR_FLOOR();
  PGA_URGENT();
  // U's strictly upper triangle, packed by columns: u_ki (k < i) at i * (i - 1) / 2 + k
  __shared__ double sU[LUTN_MMAX * (LUTN_MMAX - 1) / 2];
  __shared__ double sB[LUTN_MMAX][LUTN_COLS];  // sB[row][column]
  __shared__ double sD[LUTN_MMAX];
  const int c0 = blockIdx.x * LUTN_COLS;
  if (c0 >= n || stopped(stop)) return;
  const int C = min(LUTN_COLS, n - c0);
  const int tid = threadIdx.x;
  uint32_t* b = B + (size_t)c0 * ldb;
  for (int e = tid; e < m * m; e += WG) {
    const int k = e % m, i = e / m;
    if (k < i) sU[i * (i - 1) / 2 + k] = to_double(A[k + (size_t)i * lda]);
  }
  for (int e = tid; e < m * C; e += WG) {
    const int r = e % m, cl = e / m;
    uint32_t v = b[r + (size_t)cl * ldb];
    if (St) {
      v = add(mul(P_MONE, St[c0 + cl + (size_t)r * lds]), mul(P_ONE, v));
      b[r + (size_t)cl * ldb] = v;  // the pattern-level sweep below starts from it too
    }
    sB[r][cl] = to_double(v);
  }
  if (tid < m) {
    const double d = to_double(A[tid + (size_t)tid * lda]);
    sD[tid] = (d != 0.0) ? d : __builtin_nan("");  // a zero pivot: redo (NaR is NaN already)
  }
  __syncthreads();
  for (int k = 0; k < m; ++k) {
    if (tid < C) sB[k][tid] = fp64_posit_div(sB[k][tid], sD[k]);
    __syncthreads();
    const int cnt = m - 1 - k;
    for (int e = tid; e < cnt * C; e += WG) {
      const int c = e % C, i = k + 1 + e / C;
      sB[i][c] = fp64_mac_arith(sB[i][c], -sU[i * (i - 1) / 2 + k], sB[k][c]);
    }
    __syncthreads();
  }
  bool bad = false;
  for (int e = tid; e < m * C; e += WG) bad |= sB[e % m][e / m] != sB[e % m][e / m];  // NaN
  if (__syncthreads_or(bad)) {
    lutn_cols_pattern(m, C, A, lda, b, ldb);
    return;
  }
  for (int e = tid; e < m * C; e += WG) b[e % m + (size_t)(e / m) * ldb] = from_double(sB[e % m][e / m]);
}

// Left solves with alpha = 1 of up to TRSM_COL_MAX rows, one 1024-lane workgroup per
// column of B: B := L^-1 B (unit or not), U^-1 B or U^-T B (non-unit), the column staged
// in LDS and solved in FP64 registers 64 rows of the sweep per barrier
// (wg::col_solve_fp64: the Rtrsm.cpp:120-173 loop orders and zero skips, term for term);
// a column that leaves the FP64 range is solved again at the pattern level.  One column's
// work stays on one CU, so the per-call solves use it for U^T only (whose pattern-level
// sweep, k_trsm_left<true, true>, has the same one-workgroup-per-column shape).
constexpr int TRSM_COL_NT = 1024, TRSM_COL_MAX = wg::COL_SOLVE_SLOTS * TRSM_COL_NT;

template <int MODE>
__global__ __launch_bounds__(TRSM_COL_NT) void k_trsm_col(int m, const uint32_t* __restrict__ A, int lda,
                                                          uint32_t* __restrict__ B, int ldb, const int* stop) {
  PGA_KERNEL_VGPR_FLOOR();
  PGA_URGENT();
  extern __shared__ uint32_t col[];
  __shared__ double sb[128];
  if (stopped(stop)) return;
  uint32_t* b = B + (size_t)blockIdx.x * ldb;
  for (int i = threadIdx.x; i < m; i += TRSM_COL_NT) col[i] = b[i];
  gemm::round_table_fill(threadIdx.x, TRSM_COL_NT);
  __syncthreads();
  if (!wg::col_solve_fp64<MODE, TRSM_COL_NT>(m, A, lda, (wg::lds_u32*)col, m, (wg::lds_f64*)sb)) {
    if (MODE == wg::CS_LLNU) wg::trsm_llnu<TRSM_COL_NT>(m, 1, A, lda, col, m);
    else if (MODE == wg::CS_LLNN) wg::trsm_llnn<TRSM_COL_NT>(m, 1, A, lda, col, m);
    else if (MODE == wg::CS_LUNN) wg::trsm_lunn<TRSM_COL_NT>(m, 1, A, lda, col, m);
    else wg::trsm_lutn<TRSM_COL_NT>(m, 1, A, lda, col, m);
  }
  for (int i = threadIdx.x; i < m; i += TRSM_COL_NT) b[i] = col[i];
}

// B = 0 (alpha == 0, Rtrsm.cpp:101-108)
__global__ __launch_bounds__(WG) void k_zero(int m, int n, uint32_t* B, int ldb, const int* stop) {
  PGA_KERNEL_VGPR_FLOOR();
  const int i = blockIdx.x * WG + threadIdx.x;
  if (i >= m || stopped(stop)) return;
  for (int j = 0; j < n; ++j) B[i + (size_t)j * ldb] = 0;
}

// ---------------------------------------------------------------------------------
// GEMV (Rgemv.cpp:31-204), one lane per element of y
// ---------------------------------------------------------------------------------
__global__ __launch_bounds__(WG) void k_gemv(int trans, int m, int n, uint32_t alpha,
                                             const uint32_t* __restrict__ A, int lda,
                                             const uint32_t* __restrict__ x, int incx, uint32_t beta,
                                             uint32_t* __restrict__ y, int incy) {
  PGA_KERNEL_VGPR_FLOOR();
  const int leny = trans ? n : m, lenx = trans ? m : n;
  const int r = blockIdx.x * WG + threadIdx.x;
  if (r >= leny) return;
  const long kx = incx > 0 ? 0 : -(long)(lenx - 1) * incx;
  const long ky = incy > 0 ? 0 : -(long)(leny - 1) * incy;
  uint32_t& yr = y[ky + (long)r * incy];
  uint32_t v = yr;
  if (beta != P_ONE) v = (beta == 0) ? 0u : mul(beta, v);
  if (alpha != 0) {
    if (!trans) {  // y_i += (alpha*x_j)*a_ij, ascending j (:150-173)
      for (int j = 0; j < n; ++j) v = add(v, mul(mul(alpha, x[kx + (long)j * incx]), A[r + (size_t)j * lda]));
    } else {  // y_j += alpha*(sum_i a_ij*x_i) (:174-201)
      uint32_t t = 0;
      for (int i = 0; i < m; ++i) t = add(t, mul(A[i + (size_t)r * lda], x[kx + (long)i * incx]));
      v = add(v, mul(alpha, t));
    }
  }
  yr = v;
}

// SCAL (Rscal.cpp:62-119)
__global__ __launch_bounds__(WG) void k_scal(int n, uint32_t da, uint32_t* x, int incx) {
  PGA_KERNEL_VGPR_FLOOR();
  for (long i = blockIdx.x * (long)WG + threadIdx.x; i < n; i += (long)gridDim.x * WG)
    x[i * incx] = mul(da, x[i * incx]);
}

__global__ __launch_bounds__(WG) void k_iamax(int n, const uint32_t* x, int incx, int* result) {
  PGA_KERNEL_VGPR_FLOOR();
  __shared__ int idx;
  if (n < 1 || incx <= 0) {
    if (threadIdx.x == 0) *result = 0;
    return;
  }
  leaf::block_iamax(n, x, incx, &idx);
  if (threadIdx.x == 0) *result = idx + 1;
}

// LASWP (Rlaswp.cpp:32-111): one lane per column; ipiv entries minus `ipoff` are rows
// (1-based) of the block starting at `a`.
__global__ __launch_bounds__(WG) void k_laswp(int n, uint32_t* a, int lda, int k1, int k2,
                                              const int* ipiv, int incx, int ipoff, const int* stop) {
  PGA_KERNEL_VGPR_FLOOR();
  PGA_URGENT();
  const int col = blockIdx.x * WG + threadIdx.x;
  if (col >= n || stopped(stop)) return;
  int ix0, i1, i2, inc;
  if (incx > 0) {
    ix0 = k1;
    i1 = k1;
    i2 = k2;
    inc = 1;
  } else {
    ix0 = k1 + (k1 - k2) * incx;
    i1 = k2;
    i2 = k1;
    inc = -1;
  }
  uint32_t* ac = a + (size_t)col * lda;
  int ix = ix0;
  for (int i = i1; inc > 0 ? i <= i2 : i >= i2; i += inc) {
    int ip = ipiv[ix - 1] - ipoff;
    if (ip != i) {
      uint32_t t = ac[i - 1];
      ac[i - 1] = ac[ip - 1];
      ac[ip - 1] = t;
    }
    ix += incx;
  }
}

// LASWP (Rlaswp.cpp:70-92, incx = 1) of one column per workgroup, for interchanges that
// stay within the `rows` rows from row k1 (1-based) down, as an LU's do: the lanes stage
// those rows in LDS, one lane applies the interchanges in order there, the lanes write
// the rows back.  The same moves in the same order as k_laswp, without a global-memory
// round trip per interchange on the one lane's path.
__global__ __launch_bounds__(WG) void k_laswp_lds(uint32_t* a, int lda, int k1, int k2, const int* ipiv, int rows,
                                                  const int* stop) {
  PGA_KERNEL_VGPR_FLOOR();
  PGA_URGENT();
  extern __shared__ uint32_t seg[];  // rows k1 .. k1 + rows - 1 of this column
  if (stopped(stop)) return;
  uint32_t* col = a + (size_t)blockIdx.x * lda + (k1 - 1);
  for (int r = threadIdx.x; r < rows; r += WG) seg[r] = col[r];
  __syncthreads();
  if (threadIdx.x == 0) {
    constexpr int U = 16;  // the pivots of U interchanges read ahead
    for (int i0 = k1; i0 <= k2; i0 += U) {
      int ipc[U];
#pragma unroll
      for (int u = 0; u < U; ++u) ipc[u] = i0 + u <= k2 ? ipiv[i0 + u - 1] : i0 + u;
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int i = i0 + u, ip = ipc[u];
        if (ip != i && ip >= k1 && ip - k1 < rows) {  // (an LU's pivots always are in range)
          const uint32_t t = seg[i - k1];
          seg[i - k1] = seg[ip - k1];
          seg[ip - k1] = t;
        }
      }
    }
  }
  __syncthreads();
  for (int r = threadIdx.x; r < rows; r += WG) col[r] = seg[r];
}

// ---------------------------------------------------------------------------------
// Rgetrf2 leaves (Rgetrf2.cpp:86-132); ipiv is written as a global 1-based row index
// (row_off + local), info as a global column index (col_off + 1), first one only.
// ---------------------------------------------------------------------------------
__global__ __launch_bounds__(WG) void k_getrf2_row(uint32_t* a, int* ipiv, int* info, int row_off,
                                                   int col_off) {
  PGA_KERNEL_VGPR_FLOOR();
  PGA_URGENT();
  if (threadIdx.x != 0) return;
  ipiv[0] = row_off + 1;
  if (a[0] == 0 && *info == 0) *info = col_off + 1;
}

constexpr int LEAF_TRSM_N1 = 16, LEAF_TRSM_N2 = 32;

// One Rgetrf2 leaf (and a small node's Rtrsm), lu_leaf.hpp.  The leaves are latency
// bound.  512 lanes (two waves per SIMD) measured fastest beside the concurrent trailing
// GEMM: each lane has half the column that 256 would give it, and the workgroup still fits
// in the registers that the GEMM's waves leave free on a CU (1024 lanes often do not, and
// then wait for a CU to drain).
constexpr int LEAF_NT = 512;

__global__ __launch_bounds__(LEAF_NT) void k_getrf2_leaf(int m_root, int w_root, uint32_t* root, int lda, int j,
                                                         int* ipiv, int* info, int row_off, int col_off, int ts,
                                                         int tm, int te) {
  PGA_KERNEL_VGPR_FLOOR();
  PGA_URGENT();
  leaf::getrf2_leaf<LEAF_TRSM_N1, LEAF_TRSM_N2, LEAF_NT>(m_root, w_root, root, lda, j, ipiv, info, row_off, col_off,
                                                         ts, tm, te);
}

// A whole two-column node of the recursion in one launch (leaf, 1 x 1 Rgemm, leaf, and a
// pending small node's Rtrsm), lu_leaf.hpp
__global__ __launch_bounds__(LEAF_NT) void k_getrf2_pair(int m_root, int w_root, uint32_t* root, int lda, int j,
                                                         int* ipiv, int* info, int row_off, int col_off, int ts,
                                                         int tm, int te) {
  PGA_KERNEL_VGPR_FLOOR();
  PGA_URGENT();
  leaf::getrf2_pair<LEAF_TRSM_N1, LEAF_TRSM_N2, LEAF_NT>(m_root, w_root, root, lda, j, ipiv, info, row_off, col_off,
                                                         ts, tm, te);
}

// The same node on G workgroups (lu_leaf.hpp: getrf2_pair_mc), RPL rows per lane.
template <int RPL>
__global__ __launch_bounds__(256) void k_getrf2_pair_mc(int m_root, int w_root, uint32_t* root, int lda, int j,
                                                        int* ipiv, int* info, int row_off, int col_off, int ts,
                                                        int tm, int te, leaf::PairSync* sy, unsigned base) {
  PGA_KERNEL_VGPR_FLOOR();
  PGA_URGENT();
  leaf::getrf2_pair_mc<RPL, LEAF_TRSM_N1, LEAF_TRSM_N2>(m_root, w_root, root, lda, j, ipiv, info, row_off, col_off,
                                                         ts, tm, te, sy, base);
}

// A W-column node of the recursion on G workgroups (lu_panel.hpp: panel_mc), RPL rows per
// lane.  At most 256 VGPRs, so that a workgroup fits beside the trailing update's waves.
template <int W, int RPL>
__global__ __launch_bounds__(256, 2) void k_getrf2_panel_mc(int m_root, int w_root, uint32_t* root, int lda, int s,
                                                            int* ipiv, int* info, int row_off, int col_off,
                                                            leaf::PairSync* sy, int launch) {
  PGA_KERNEL_VGPR_FLOOR();
  PGA_URGENT();
  leaf::panel_mc<W, RPL, 4, 4>(m_root, w_root, root, lda, s, ipiv, info, row_off, col_off, sy, launch);
}

__global__ void k_zero_words(unsigned* p, int n) {
  PGA_KERNEL_VGPR_FLOOR();
  for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0u;
}

// ---------------------------------------------------------------------------------
// Rpotrf2 leaf (Rpotrf2.cpp:86-100): not positive (or NaR) -> info, else sqrt
// ---------------------------------------------------------------------------------
__global__ void k_potrf2_leaf(uint32_t* a, int* info, int col_off) {
  PGA_KERNEL_VGPR_FLOOR();
  PGA_URGENT();
  if (threadIdx.x != 0 || *info) return;
  uint32_t v = a[0];
  if (le(v, 0)) {
    *info = col_off + 1;
    return;
  }
  a[0] = pga::sqrt(v);
}

// ---------------------------------------------------------------------------------
// Rpotrf2 on a whole diagonal block (n <= 64) by one workgroup, lower and upper: the
// recursion's operations in its order (as the launches of lah::potrf2 give them), in the
// FP64 form first and at the pattern level where that gives up.  info doubles as the stop
// flag, as for every step of lah::potrf.
// ---------------------------------------------------------------------------------
constexpr int POTRF2_ONE_WG_MAX = 64;

// The same recursion on the block in memory with the pattern-level operations, one
// element per lane: neither LDS nor many registers.  A failing step returns its 1-based
// column and leaves the block as the reference leaves it.
template <int T>
static __device__ int potrf2_upper_plain(int n, uint32_t* a, int lda) {
  const int tid = threadIdx.x;
  auto A = [&](int r, int c) -> uint32_t& { return a[r + (size_t)c * lda]; };
  for (int j = 0; j < n; ++j) {
    const uint32_t v = A(j, j);
    if (le(v, 0)) return j + 1;  // Rpotrf2.cpp:90 (NaR is below zero in pattern order)
    __syncthreads();
    if (tid == 0) A(j, j) = pga::sqrt(v);
    __syncthreads();
    int s, mid, e;
    if (wg::node_after(j, n, s, mid, e)) {
      const int n2 = e - mid;
      for (int k = s; k < mid; ++k) {
        for (int c = mid + tid; c < e; c += T) A(k, c) = div(A(k, c), A(k, k));
        __syncthreads();
        const int cnt = mid - 1 - k;
        for (int q = tid; q < cnt * n2; q += T) {
          const int i = k + 1 + q % cnt, c = mid + q / cnt;
          A(i, c) = sub(A(i, c), mul(A(k, i), A(k, c)));
        }
        __syncthreads();
      }
      for (int q = tid; q < n2 * n2; q += T) {
        const int i = mid + q % n2, jj = mid + q / n2;
        if (i <= jj) {
          uint32_t t = 0;
          for (int l = s; l < mid; ++l) t = add(t, mul(A(l, i), A(l, jj)));
          A(i, jj) = add(mul(P_MONE, t), mul(P_ONE, A(i, jj)));
        }
      }
      __syncthreads();
    }
  }
  return 0;
}

// The upper diagonal block on the upper Cholesky's critical path: the FP64 form
// (potrf2_upper_decoded, 17 KB of LDS), and where that gives up the recursion in memory
// (potrf2_upper_plain), so that the block finds room beside the concurrent trailing
// update.
// With St, the block first takes its right-looking sums, c := -t + c with the sum of
// element (i, c) at St[c + i * lds] (k_apply_sums on the diagonal block).
__global__ __launch_bounds__(WG, 2) void k_potrf2_upper_fast(int n, uint32_t* a, int lda, int* info, int col_off,
                                                             const uint32_t* __restrict__ St, int lds) {
  PGA_KERNEL_VGPR_FLOOR();
  PGA_URGENT();
  __shared__ double S[wg::UPPER_PACKED];
  if (*info) return;  // every lane reads the same flag
  if (St) {
    for (int e = threadIdx.x; e < n * n; e += WG) {
      const int i = e % n, c = e / n;
      if (i <= c) {
        uint32_t& v = a[i + (size_t)c * lda];
        v = add(mul(P_MONE, St[c + (size_t)i * lds]), mul(P_ONE, v));
      }
    }
    __syncthreads();
  }
  if (wg::potrf2_upper_decoded<WG>(n, a, lda, S) == 0) return;
  __syncthreads();
  const int f = potrf2_upper_plain<WG>(n, a, lda);
  if (f && threadIdx.x == 0) *info = col_off + f;
}


// The lower Rpotrf2 recursion (Rpotrf2.cpp:84-145) on a block in memory, as
// wg::potrf2<false> walks it, with its products one element per lane (wg::gemm_direct):
// neither LDS tiles nor many registers.  A failing step returns its 1-based column and
// leaves the block as the reference leaves it.
static __device__ int potrf2_lower_plain(int n, uint32_t* a, int lda) {
  for (int j = 0; j < n; ++j) {
    const uint32_t v = a[j + (size_t)j * lda];
    if (le(v, 0)) return j + 1;  // Rpotrf2.cpp:90 (NaR is below zero in pattern order)
    __syncthreads();
    if (threadIdx.x == 0) a[j + (size_t)j * lda] = pga::sqrt(v);
    __syncthreads();
    int s, mid, e;
    if (wg::node_after(j, n, s, mid, e)) {  // Rpotrf2.cpp:132-137
      const int n1 = mid - s, n2 = e - mid;
      uint32_t* a21 = wg::at(a, lda, mid, s);
      wg::trsm_rltn(n2, n1, wg::at(a, lda, s, s), lda, a21, lda);
      wg::gemm_direct<false, true, true>(n2, n2, n1, a21, lda, a21, lda, wg::at(a, lda, mid, mid), lda, 1);
    }
  }
  return 0;
}

// The lower block on the critical path of the lower Cholesky: the FP64 form in registers
// (potrf2_lower_decoded), and where that form gives up, the recursion itself on the block
// in memory (potrf2_lower_plain).  80 VGPRs and under 1 KB of LDS, so that it finds room
// on a CU beside the concurrent trailing update.
__global__ __launch_bounds__(WG, 2) void k_potrf2_lower_fast(int n, uint32_t* a, int lda, int* info, int col_off) {
  PGA_KERNEL_VGPR_FLOOR();
  PGA_URGENT();
  if (*info) return;  // every lane reads the same flag
  if (wg::potrf2_lower_decoded<WG>(n, a, lda) == 0) return;
  __syncthreads();
  const int f = potrf2_lower_plain(n, a, lda);
  if (f && threadIdx.x == 0) *info = col_off + f;
}

}  // namespace la

// ===================================================================================
// Host-side drivers (launch sequences).  All take device pointers and one stream.
// ===================================================================================
namespace lah {

using namespace la;

static unsigned blocks_for(long n) { return (unsigned)((n + WG - 1) / WG); }

#define PGA_TRY_E(x)                       \
  do {                                     \
    hipError_t _e = (x);                   \
    if (_e != hipSuccess) return _e;       \
  } while (0)

// The device state of the multi-workgroup panel pairs of one factorization, and the
// grid barrier's ticket count so far (the host adds each launch's arrivals).
struct PairState {
  leaf::PairSync* dev = nullptr;
  unsigned base = 0;
  int panel_launches = 0;  // W-column node launches so far (each has its own ticket counter)
};

struct Ctx {
  hipStream_t s;
  GemmImpl impl;
  PairState* pair = nullptr;  // set: two-column nodes of at least pair_mc_min() rows use G workgroups
};

static hipError_t trsm(const Ctx& c, bool left, bool upper, bool trans, bool nounit, int m, int n,
                       uint32_t alpha, const uint32_t* A, int lda, uint32_t* B, int ldb,
                       const int* stop);

// B := alpha * scale; elementwise (the first step of every Rtrsm branch with alpha != 1)
__global__ __launch_bounds__(WG) void k_scale_block(int m, int n, uint32_t alpha, uint32_t* B, int ldb,
                                                    const int* stop) {
  PGA_KERNEL_VGPR_FLOOR();
  if (stopped(stop)) return;
  for (long e = blockIdx.x * (long)WG + threadIdx.x; e < (long)m * n; e += (long)gridDim.x * WG)
    B[e % m + (size_t)(e / m) * ldb] = mul(alpha, B[e % m + (size_t)(e / m) * ldb]);
}

// The not-transposed left sweep (Rtrsm.cpp:120-155) on a diagonal block of ib <= 64 rows,
// one wave per right-hand side: lane r holds b_r, the block of A sits in LDS, and each
// step's b_k reaches every lane by a shuffle, so the steps need no barrier.  Lower: k
// ascending, rows below k; upper: k descending, rows above k.  A zero b_k skips its step;
// with nounit b_k is divided by a_kk before its terms.  alpha is applied by the caller.
constexpr int TRSM_DIAG_MAX = 64;

template <bool UPPER>
__global__ __launch_bounds__(64) void k_trsm_diag(int ib, int n, int nounit, const uint32_t* __restrict__ A,
                                                  int lda, uint32_t* __restrict__ B, int ldb, const int* stop) {
  PGA_KERNEL_VGPR_FLOOR();
  PGA_URGENT();
  __shared__ uint32_t sa[TRSM_DIAG_MAX][TRSM_DIAG_MAX + 1];  // sa[r][c] = a(r, c)
  const int j = blockIdx.x;
  if (j >= n || stopped(stop)) return;
  const int r = threadIdx.x;
  for (int c = 0; c < ib; ++c)
    if (r < ib) sa[r][c] = A[r + (size_t)c * lda];
  __syncthreads();
  uint32_t* bj = B + (size_t)j * ldb;
  const uint32_t b0 = r < ib ? bj[r] : 0;
  // FP64 form: each term rounded as the pattern-level operations round it (fp64_mac_arith,
  // fp64_posit_div) while every value stays in scales -108 .. 107; a NaN anywhere (a scale
  // outside them, a NaR) makes the wave redo its column at the pattern level
  {
    double b = to_double(b0);
    for (int s = 0; s < ib; ++s) {
      const int k = UPPER ? ib - 1 - s : s;
      double bk = __shfl(b, k);
      if (bk == 0.0) continue;  // the same value in every lane
      if (nounit) bk = fp64_posit_div(bk, to_double(sa[k][k]));
      if (UPPER ? r < k : (r > k && r < ib)) b = fp64_mac_arith(b, -bk, to_double(sa[r][k]));
      if (r == k) b = bk;
    }
    if (!__any(b != b)) {
      if (r < ib) bj[r] = from_double(b);
      return;
    }
  }
  uint32_t b = b0;
  for (int s = 0; s < ib; ++s) {
    const int k = UPPER ? ib - 1 - s : s;
    uint32_t bk = __shfl(b, k);
    if (bk == 0) continue;  // the same value in every lane
    if (nounit) bk = div(bk, sa[k][k]);
    if (UPPER ? r < k : (r > k && r < ib)) b = sub(b, mul(bk, sa[r][k]));
    if (r == k) b = bk;
  }
  if (r < ib) bj[r] = b;
}

// Rows [0, rows) of B after an upper diagonal block of kb rows is solved:
// b_i -= b_k * a_ik for the block's k in descending order, a zero b_k skipped
// (Rtrsm.cpp:127-133), one lane per (row, right-hand side).
__global__ __launch_bounds__(WG) void k_trsm_upper_update(int rows, int n, int kb, const uint32_t* __restrict__ A,
                                                          int lda, const uint32_t* __restrict__ Bk,
                                                          uint32_t* __restrict__ B, int ldb, const int* stop) {
  PGA_KERNEL_VGPR_FLOOR();
  PGA_URGENT();
  const long e = blockIdx.x * (long)WG + threadIdx.x;
  if (e >= (long)rows * n || stopped(stop)) return;
  const int i = (int)(e % rows), j = (int)(e / rows);
  uint32_t b = B[i + (size_t)j * ldb];
  for (int kk = kb - 1; kk >= 0; --kk) {
    const uint32_t bk = Bk[kk + (size_t)j * ldb];
    if (bk != 0) b = sub(b, mul(bk, A[i + (size_t)kk * lda]));
  }
  B[i + (size_t)j * ldb] = b;
}

// Left, not transposed (Rtrsm.cpp:120-155) in blocks of rows: each diagonal block is
// solved, then its rows update every row still to come -- below it through the GEMM with
// zero multipliers skipped (lower), above it in descending k (upper).  Each element
// still receives its terms in the reference's k order, and a zero b_k still skips its
// term, so the result is the same bits.  Narrow right-hand sides solve their diagonal
// blocks with k_trsm_diag (64 rows); wide ones with the column-sweep kernel (32 rows).
constexpr int TRSM_BLOCK = 32;

static bool trsm_narrow(int n) { return (long)n < 8L * runtime().cus; }

static hipError_t trsm_diag(const Ctx& c, bool upper, bool nounit, int ib, int n, const uint32_t* A, int lda,
                            uint32_t* B, int ldb, const int* stop) {
  if (!trsm_narrow(n)) return trsm(c, true, upper, false, nounit, ib, n, P_ONE, A, lda, B, ldb, stop);
  if (upper)
    hipLaunchKernelGGL(k_trsm_diag<true>, dim3(n), dim3(64), 0, c.s, ib, n, (int)nounit, A, lda, B, ldb, stop);
  else
    hipLaunchKernelGGL(k_trsm_diag<false>, dim3(n), dim3(64), 0, c.s, ib, n, (int)nounit, A, lda, B, ldb, stop);
  return hipGetLastError();
}

static hipError_t trsm_scale(const Ctx& c, int m, int n, uint32_t alpha, uint32_t* B, int ldb, const int* stop) {
  if (alpha == P_ONE) return hipSuccess;
  const long total = (long)m * n;
  hipLaunchKernelGGL(k_scale_block, dim3((unsigned)std::min<long>((total + WG - 1) / WG, 65536)), dim3(WG), 0, c.s,
                     m, n, alpha, B, ldb, stop);
  return hipGetLastError();
}

static hipError_t trsm_lln_blocked(const Ctx& c, bool nounit, int m, int n, uint32_t alpha, const uint32_t* A,
                                   int lda, uint32_t* B, int ldb, const int* stop) {
  PGA_TRY_E(trsm_scale(c, m, n, alpha, B, ldb, stop));
  GemmOpts o;
  o.skip_zero_b = true;
  o.stop = stop;
  const int tb = trsm_narrow(n) ? TRSM_DIAG_MAX : TRSM_BLOCK;
  for (int i0 = 0; i0 < m; i0 += tb) {
    const int ib = std::min(tb, m - i0);
    PGA_TRY_E(trsm_diag(c, false, nounit, ib, n, A + i0 + (size_t)i0 * lda, lda, B + i0, ldb, stop));
    if (i0 + ib < m)
      PGA_TRY_E(rgemm_launch(false, false, m - i0 - ib, n, ib, P_MONE, A + i0 + ib + (size_t)i0 * lda, lda, B + i0,
                             ldb, P_ONE, B + i0 + ib, ldb, c.s, c.impl, o));
  }
  return hipSuccess;
}

// Upper form, narrow right-hand sides only (the blocks go from the bottom up).
static hipError_t trsm_lun_blocked(const Ctx& c, bool nounit, int m, int n, uint32_t alpha, const uint32_t* A,
                                   int lda, uint32_t* B, int ldb, const int* stop) {
  PGA_TRY_E(trsm_scale(c, m, n, alpha, B, ldb, stop));
  for (int i1 = m; i1 > 0; i1 -= TRSM_DIAG_MAX) {
    const int i0 = std::max(0, i1 - TRSM_DIAG_MAX), ib = i1 - i0;
    PGA_TRY_E(trsm_diag(c, true, nounit, ib, n, A + i0 + (size_t)i0 * lda, lda, B + i0, ldb, stop));
    if (i0 > 0) {
      const long total = (long)i0 * n;
      hipLaunchKernelGGL(k_trsm_upper_update, dim3((unsigned)((total + WG - 1) / WG)), dim3(WG), 0, c.s, i0, n, ib,
                         A + (size_t)i0 * lda, lda, B + i0, B, ldb, stop);
      PGA_TRY_E(hipGetLastError());
    }
  }
  return hipSuccess;
}

static hipError_t trsm(const Ctx& c, bool left, bool upper, bool trans, bool nounit, int m, int n,
                       uint32_t alpha, const uint32_t* A, int lda, uint32_t* B, int ldb,
                       const int* stop) {
  if (m == 0 || n == 0) return hipSuccess;
  // U^T B with more rows than a diagonal block: one workgroup per right-hand side in
  // FP64 (k_trsm_col).  The other left forms keep their row-block paths, whose updates
  // spread one column's work over the whole device.
  if (left && upper && trans && nounit && alpha == P_ONE && m > TRSM_DIAG_MAX && m <= TRSM_COL_MAX) {
    hipLaunchKernelGGL(k_trsm_col<wg::CS_LUTN>, dim3(n), dim3(TRSM_COL_NT), (size_t)m * sizeof(uint32_t), c.s, m, A,
                       lda, B, ldb, stop);
    return hipGetLastError();
  }
  if (left && !upper && !trans && alpha != 0 && m > 2 * TRSM_BLOCK)
    return trsm_lln_blocked(c, nounit, m, n, alpha, A, lda, B, ldb, stop);
  if (left && upper && !trans && alpha != 0 && m > TRSM_DIAG_MAX && trsm_narrow(n))
    return trsm_lun_blocked(c, nounit, m, n, alpha, A, lda, B, ldb, stop);
  if (alpha == 0) {
    hipLaunchKernelGGL(k_zero, dim3(blocks_for(m)), dim3(WG), 0, c.s, m, n, B, ldb, stop);
    return hipGetLastError();
  }
  // a diagonal block's worth of rows: one wave per right-hand side (k_trsm_diag)
  if (left && !trans && alpha == P_ONE && m <= TRSM_DIAG_MAX && trsm_narrow(n))
    return trsm_diag(c, upper, nounit, m, n, A, lda, B, ldb, stop);
  if (left && !trans && (long)n >= 8L * runtime().cus) {
    const int cpw = (int)std::min<long>(8, (n + 4L * runtime().cus - 1) / (4L * runtime().cus));
    const unsigned grid = (unsigned)((n + cpw - 1) / cpw);
    if (upper)
      hipLaunchKernelGGL(k_trsm_left_cols<true>, dim3(grid), dim3(WG), 0, c.s, m, n, alpha, (int)nounit, A, lda,
                         B, ldb, cpw, stop);
    else
      hipLaunchKernelGGL(k_trsm_left_cols<false>, dim3(grid), dim3(WG), 0, c.s, m, n, alpha, (int)nounit, A, lda,
                         B, ldb, cpw, stop);
    return hipGetLastError();
  }
  if (left) {
#define PGA_TL(U, T)                                                                               \
  hipLaunchKernelGGL((k_trsm_left<U, T>), dim3(n), dim3(WG), 0, c.s, m, n, alpha, (int)nounit, A,  \
                     lda, B, ldb, stop)
    if (upper && !trans) PGA_TL(true, false);
    else if (!upper && !trans) PGA_TL(false, false);
    else if (upper && trans && nounit && alpha == P_ONE && m <= LUTN_MMAX)
      hipLaunchKernelGGL(k_trsm_lutn_fp64, dim3((unsigned)((n + LUTN_COLS - 1) / LUTN_COLS)), dim3(WG), 0, c.s, m, n,
                         A, lda, B, ldb, stop);
    else if (upper && trans) PGA_TL(true, true);
    else if (m <= LLTN_PIPE_CAP + 1)
      hipLaunchKernelGGL(k_trsm_lltn_pipe, dim3(n), dim3(WG), 0, c.s, m, n, alpha, (int)nounit, A, lda, B, ldb,
                         stop);
    else
      hipLaunchKernelGGL(k_trsm_lltn, dim3(n), dim3(WG), 0, c.s, m, n, alpha, (int)nounit, A, lda, B, ldb, stop);
#undef PGA_TL
    return hipGetLastError();
  } else if (trans) {
    // enough row blocks for every CU twice where m allows, at least 4 rows each
    const int rows = (int)std::min<long>(256, std::max<long>(4, (m + 2L * runtime().cus - 1) / (2L * runtime().cus)));
    const unsigned grid = (unsigned)((m + rows - 1) / rows);
    if (!upper && nounit && alpha == P_ONE && n <= RSW_NMAX) {
      hipLaunchKernelGGL(k_trsm_rltn_fp64, dim3((unsigned)((m + RSW_ROWS - 1) / RSW_ROWS)), dim3(WG), 0, c.s, m, n,
                         A, lda, B, ldb, stop);
      return hipGetLastError();
    }
    if (upper)
      hipLaunchKernelGGL((k_trsm_right_sweep<true>), dim3(grid), dim3(WG), 0, c.s, m, n, alpha, (int)nounit, A,
                         lda, B, ldb, rows, stop);
    else
      hipLaunchKernelGGL((k_trsm_right_sweep<false>), dim3(grid), dim3(WG), 0, c.s, m, n, alpha, (int)nounit, A,
                         lda, B, ldb, rows, stop);
  } else {
    // the non-transposed forms sum each element over earlier columns: one lane per row
#define PGA_TR(U)                                                                                 \
  hipLaunchKernelGGL((k_trsm_right<U, false>), dim3(blocks_for(m)), dim3(WG), 0, c.s, m, n, alpha, \
                     (int)nounit, A, lda, B, ldb, stop)
    if (upper) PGA_TR(true);
    else PGA_TR(false);
#undef PGA_TR
  }
  return hipGetLastError();
}

// Rows a column segment of k_laswp_lds may hold (64 KB of LDS)
constexpr int LASWP_LDS_ROWS = 16384;
// Columns up to which the interchanges of few columns go through LDS (beyond it the
// one-lane-per-column kernel has parallelism enough)
constexpr int LASWP_LDS_MAX_COLS = 1024;

// `rows` (optional): the rows from row k1 to the end of the matrix, which bound every
// interchange of an LU's ipiv (ip >= i).
static hipError_t laswp(const Ctx& c, int n, uint32_t* a, int lda, int k1, int k2, const int* ipiv,
                        int incx, int ipoff, const int* stop, int rows = 0) {
  if (n <= 0 || incx == 0) return hipSuccess;
  if (incx == 1 && ipoff == 0 && rows >= k2 - k1 + 1 && rows <= LASWP_LDS_ROWS && n <= LASWP_LDS_MAX_COLS) {
    hipLaunchKernelGGL(k_laswp_lds, dim3(n), dim3(WG), (unsigned)rows * 4u, c.s, a, lda, k1, k2, ipiv, rows, stop);
    return hipGetLastError();
  }
  hipLaunchKernelGGL(k_laswp, dim3(blocks_for(n)), dim3(WG), 0, c.s, n, a, lda, k1, k2, ipiv, incx, ipoff,
                     stop);
  return hipGetLastError();
}

#define PGA_TRY(x)                         \
  do {                                     \
    hipError_t _e = (x);                   \
    if (_e != hipSuccess) return _e;       \
  } while (0)

// Rows from which a two-column node runs on several workgroups (PGA_LU_PAIR_MC_MIN;
// 0 turns that off)
static int pair_mc_min() {
  static const int v = [] {
    const char* e = getenv("PGA_LU_PAIR_MC_MIN");
    const int x = e ? atoi(e) : 2048;
    return x > 0 ? std::max(x, 1024) : INT32_MAX;
  }();
  return v;
}

// Width of the recursion nodes that run as one multi-workgroup launch (PGA_LU_PANEL_W:
// 32 or 64; 0 turns that off), and the rows per workgroup it aims for (PGA_LU_PANEL_ROWS)
static int panel_w() {
  static const int v = [] {
    const char* e = getenv("PGA_LU_PANEL_W");
    const int x = e ? atoi(e) : 32;
    return (x == 32 || x == 64) ? x : 0;
  }();
  return v;
}
static int panel_rows() {
  static const int v = [] {
    const char* e = getenv("PGA_LU_PANEL_ROWS");
    const int x = e ? atoi(e) : 256;
    return std::max(256, std::min(x, 4096));
  }();
  return v;
}

// Workgroups per CU of the trailing update that runs beside the panel nodes
// (PGA_LU_GEMM_CAP; 0: no cap): three operand-panel GEMM workgroups leave a CU the
// registers and LDS of one panel-node workgroup.
static int panel_gemm_cap() {
  static const int v = [] {
    const char* e = getenv("PGA_LU_GEMM_CAP");
    return e ? std::max(0, atoi(e)) : 3;
  }();
  return v;
}

// Launches the W-column node [s, s + W) of the block on G workgroups; false when the node
// does not fit the kernel's shapes (the caller then recurses as usual).
static bool panel_mc_launch(const Ctx& c, int m, int n, uint32_t* a, int lda, int s, int* ipiv, int* info,
                            int row_off, int col_off, int W) {
  if (c.pair->panel_launches >= leaf::PANEL_MC_MAX_LAUNCHES) return false;
  const int rows = m - s;
  const int G = std::max(1, std::min(leaf::PANEL_MC_MAX_G, (rows + panel_rows() - 1) / panel_rows()));
  const int R = (rows + G - 1) / G, rpl = (R + 255) / 256;
  const int launch = c.pair->panel_launches;
  const dim3 g((unsigned)G), b(256);
  leaf::PairSync* sy = c.pair->dev;
  if (W == 32 && rpl <= 1)
    hipLaunchKernelGGL((k_getrf2_panel_mc<32, 1>), g, b, 0, c.s, m, n, a, lda, s, ipiv, info, row_off, col_off, sy,
                       launch);
  else if (W == 32 && rpl <= 2)
    hipLaunchKernelGGL((k_getrf2_panel_mc<32, 2>), g, b, 0, c.s, m, n, a, lda, s, ipiv, info, row_off, col_off, sy,
                       launch);
  else if (W == 64 && rpl <= 1)
    hipLaunchKernelGGL((k_getrf2_panel_mc<64, 1>), g, b, 0, c.s, m, n, a, lda, s, ipiv, info, row_off, col_off, sy,
                       launch);
  else
    return false;
  c.pair->panel_launches = launch + 1;
  return true;
}

// Rgetrf2 (Rgetrf2.cpp:32-185) on the m x n block at `a` (rows row_off.., columns
// col_off.. of the matrix that ipiv / info describe; ipiv points at the block's first
// entry).  The recursion's nodes cover columns [s, e) and rows [s, m) of the block; its
// row interchanges are made by the leaves across all n columns of the block, which
// replaces the Rlaswp calls between the recursive calls (same bits, fewer launches).
// pend: a small node [ps, pe) split at pm whose Rtrsm the last leaf of this subtree runs
// (pm <= ps: none).
static hipError_t getrf2_node(const Ctx& c, int m, int n, uint32_t* a, int lda, int* ipiv, int* info,
                              int row_off, int col_off, int s, int e, int ps, int pm, int pe) {
  const int rows = m - s, cols = e - s;
  if (rows <= 0 || cols <= 0) return hipSuccess;
  if (rows == 1) {  // Rgetrf2.cpp:86-95: nothing to interchange
    hipLaunchKernelGGL(k_getrf2_row, dim3(1), dim3(WG), 0, c.s, a + s + (size_t)s * lda, ipiv + s, info,
                       row_off + s, col_off + s);
    PGA_TRY(hipGetLastError());
    if (pm > ps)
      PGA_TRY(trsm(c, true, false, false, false, pm - ps, pe - pm, P_ONE, a + ps + (size_t)ps * lda, lda,
                   a + ps + (size_t)pm * lda, lda, nullptr));
    return hipSuccess;
  }
  if (cols == 1) {
    hipLaunchKernelGGL(k_getrf2_leaf, dim3(1), dim3(LEAF_NT), 0, c.s, m, n, a, lda, s, ipiv, info, row_off, col_off,
                       ps, pm, pe);
    return hipGetLastError();
  }
  if (c.pair && pm <= ps && cols == panel_w() && rows >= cols &&
      panel_mc_launch(c, m, n, a, lda, s, ipiv, info, row_off, col_off, cols))
    return hipGetLastError();
  if (cols == 2) {  // rows >= 2 here: the node splits 1 + 1
    if (c.pair && rows >= pair_mc_min()) {
      // G blocks of at least 512 rows, at most 8 rows per lane
      const int G = std::min(leaf::PAIR_MC_MAX_G, rows / 512), R = (rows + G - 1) / G, rpl = (R + 255) / 256;
      const unsigned base = c.pair->base;
      const dim3 g((unsigned)G), b(256);
      bool launched = true;
      if (rpl <= 2)
        hipLaunchKernelGGL(k_getrf2_pair_mc<2>, g, b, 0, c.s, m, n, a, lda, s, ipiv, info, row_off, col_off, ps, pm,
                           pe, c.pair->dev, base);
      else if (rpl <= 4)
        hipLaunchKernelGGL(k_getrf2_pair_mc<4>, g, b, 0, c.s, m, n, a, lda, s, ipiv, info, row_off, col_off, ps, pm,
                           pe, c.pair->dev, base);
      else if (rpl <= 8)
        hipLaunchKernelGGL(k_getrf2_pair_mc<8>, g, b, 0, c.s, m, n, a, lda, s, ipiv, info, row_off, col_off, ps, pm,
                           pe, c.pair->dev, base);
      else
        launched = false;
      if (launched) {
        c.pair->base = base + (unsigned)G * (pm > ps ? 3u : 2u);
        return hipGetLastError();
      }
    }
    hipLaunchKernelGGL(k_getrf2_pair, dim3(1), dim3(LEAF_NT), 0, c.s, m, n, a, lda, s, ipiv, info, row_off, col_off,
                       ps, pm, pe);
    return hipGetLastError();
  }
  const int n1 = (rows < cols ? rows : cols) / 2, mid = s + n1, n2 = cols - n1;
  uint32_t* a11 = a + s + (size_t)s * lda;
  uint32_t* a12 = a + s + (size_t)mid * lda;
  const bool fuse = (n1 <= LEAF_TRSM_N1 && n2 <= LEAF_TRSM_N2);
  if (fuse) {
    PGA_TRY(getrf2_node(c, m, n, a, lda, ipiv, info, row_off, col_off, s, mid, s, mid, e));
  } else {
    PGA_TRY(getrf2_node(c, m, n, a, lda, ipiv, info, row_off, col_off, s, mid, 0, 0, 0));
    PGA_TRY(trsm(c, true, false, false, false, n1, n2, P_ONE, a11, lda, a12, lda, nullptr));
  }
  GemmOpts o;
  o.urgent = true;
  PGA_TRY(rgemm_launch(false, false, rows - n1, n2, n1, P_MONE, a11 + n1, lda, a12, lda, P_ONE, a12 + n1, lda,
                       c.s, c.impl, o));
  return getrf2_node(c, m, n, a, lda, ipiv, info, row_off, col_off, mid, e, ps, pm, pe);
}

static hipError_t getrf2(const Ctx& c, int m, int n, uint32_t* a, int lda, int* ipiv, int* info,
                         int row_off, int col_off) {
  if (m == 0 || n == 0) return hipSuccess;
  return getrf2_node(c, m, n, a, lda, ipiv, info, row_off, col_off, 0, n, 0, 0, 0);
}

// A second stream with two events for one factorization at a time (kept for reuse; the
// pool grows to the number of factorizations running at once).
struct Aux {
  hipStream_t s = nullptr;
  hipEvent_t join = nullptr, done = nullptr;
};
static std::mutex g_aux_mu;
static std::vector<Aux> g_aux_free;

static hipError_t aux_take(Aux& x) {
  {
    std::lock_guard<std::mutex> lk(g_aux_mu);
    if (!g_aux_free.empty()) {
      x = g_aux_free.back();
      g_aux_free.pop_back();
      return hipSuccess;
    }
  }
  PGA_TRY_E(hipStreamCreateWithFlags(&x.s, hipStreamNonBlocking));
  PGA_TRY_E(hipEventCreateWithFlags(&x.join, hipEventDisableTiming));
  return hipEventCreateWithFlags(&x.done, hipEventDisableTiming);
}
static void aux_give(const Aux& x) {
  std::lock_guard<std::mutex> lk(g_aux_mu);
  g_aux_free.push_back(x);
}

// The update of columns [c0, c0 + w) by block j of an LU (Rgetrf_opencl.cpp:123-146 on
// those columns): the block's row interchanges, the Rtrsm of its jb rows, the Rgemm of the
// rows below.  Columns are independent here, so any split of the right-hand columns
// gives the same bits.
static hipError_t getrf_update(const Ctx& c, int m, uint32_t* a, int lda, const int* ipiv, int j, int jb, int c0,
                               int w, bool urgent, int cap = 0) {
  if (w <= 0) return hipSuccess;
  uint32_t* ajj = a + j + (size_t)j * lda;
  uint32_t* a12 = a + j + (size_t)c0 * lda;
  PGA_TRY(laswp(c, w, a + (size_t)c0 * lda, lda, j + 1, j + jb, ipiv, 1, 0, nullptr, m - j));
  PGA_TRY(trsm(c, true, false, false, false, jb, w, P_ONE, ajj, lda, a12, lda, nullptr));
  GemmOpts o;
  o.urgent = urgent;
  o.cap_per_cu = cap;
  if (j + jb < m)
    PGA_TRY(rgemm_launch(false, false, m - j - jb, w, jb, P_MONE, ajj + jb, lda, a12, lda, P_ONE, a12 + jb, lda,
                         c.s, c.impl, o));
  return hipSuccess;
}

// Orders from which the LU runs with one panel of lookahead
constexpr int GETRF_LOOKAHEAD_MIN = 1024;

// Rgetrf_opencl with one panel of lookahead.  After panel j, the chain stream updates
// only block column j+1 and factors it as panel j+1, while a second stream updates the
// columns from j+2 on (the bulk of the work, one large GEMM) and applies panel j's
// interchanges to the columns left of it.  Each element sees the same operations in the
// same order as in the sequential driver: block column j+1 waits for the second stream's
// update by block j-1, and the left interchanges of panel j run after that update has
// read panel j-1.  `join` and `done` are two events owned by this call.
static hipError_t getrf_lookahead_on(const Ctx& c, hipStream_t wide, hipEvent_t join, hipEvent_t done, int m, int n,
                                     uint32_t* a, int lda, int* ipiv, int* info, int nb) {
  const int mn = m < n ? m : n;
  const Ctx cx{wide, c.impl, nullptr};
  // with multi-workgroup panel nodes, the bulk update leaves room on every CU for them
  const int wide_cap = (c.pair && panel_w()) ? panel_gemm_cap() : 0;
  hipError_t e = getrf2(c, m, std::min(nb, mn), a, lda, ipiv, info, 0, 0);
  bool recorded = false;  // the wide stream has recorded `done` at least once
  for (int j = 0; j < mn && e == hipSuccess; j += nb) {
    const int jb = std::min(nb, mn - j), c0 = j + jb, w1 = std::min(nb, n - c0), c1 = c0 + std::max(w1, 0);
    if ((e = hipEventRecord(join, c.s)) != hipSuccess) break;
    if ((e = hipStreamWaitEvent(wide, join, 0)) != hipSuccess) break;
    if (j > 0 && (e = laswp(cx, j, a, lda, j + 1, j + jb, ipiv, 1, 0, nullptr)) != hipSuccess) break;
    if ((e = getrf_update(cx, m, a, lda, ipiv, j, jb, c1, n - c1, false, wide_cap)) != hipSuccess) break;
    if (recorded && (e = hipStreamWaitEvent(c.s, done, 0)) != hipSuccess) break;
    if ((e = hipEventRecord(done, wide)) != hipSuccess) break;
    recorded = true;
    if ((e = getrf_update(c, m, a, lda, ipiv, j, jb, c0, w1, true)) != hipSuccess) break;
    if (c0 < mn) e = getrf2(c, m - c0, std::min(nb, mn - c0), a + c0 + (size_t)c0 * lda, lda, ipiv + c0, info, c0, c0);
  }
  if (recorded) {
    hipError_t w = hipStreamWaitEvent(c.s, done, 0);
    if (e == hipSuccess) e = w;
  }
  return e;
}

// The lookahead driver on the caller's stream and an Aux stream.
static hipError_t getrf_lookahead(const Ctx& c, int m, int n, uint32_t* a, int lda, int* ipiv, int* info, int nb) {
  Aux x;
  PGA_TRY_E(aux_take(x));
  hipError_t e = getrf_lookahead_on(c, x.s, x.join, x.done, m, n, a, lda, ipiv, info, nb);
  aux_give(x);
  return e;
}

// Rgetrf_opencl (Rgetrf_opencl.cpp:32-154): blocked right-looking LU, nb columns per
// panel; ipiv holds global 1-based rows, info the first zero pivot column.
static hipError_t getrf_blocked(const Ctx& c, int m, int n, uint32_t* a, int lda, int* ipiv, int* info, int nb) {
  const int mn = m < n ? m : n;
  if (nb <= 1 || nb >= mn) return getrf2(c, m, n, a, lda, ipiv, info, 0, 0);
  if (mn >= GETRF_LOOKAHEAD_MIN) return getrf_lookahead(c, m, n, a, lda, ipiv, info, nb);
  for (int j = 0; j < mn; j += nb) {
    const int jb = (mn - j < nb) ? mn - j : nb;
    uint32_t* ajj = a + j + (size_t)j * lda;
    PGA_TRY(getrf2(c, m - j, jb, ajj, lda, ipiv + j, info, j, j));
    // rows j..j+jb-1 of the columns left of the panel (Rgetrf_opencl.cpp:123)
    PGA_TRY(laswp(c, j, a, lda, j + 1, j + jb, ipiv, 1, 0, nullptr));
    if (j + jb < n) {
      uint32_t* a12 = a + j + (size_t)(j + jb) * lda;
      PGA_TRY(laswp(c, n - j - jb, a + (size_t)(j + jb) * lda, lda, j + 1, j + jb, ipiv, 1, 0, nullptr));
      PGA_TRY(trsm(c, true, false, false, false, jb, n - j - jb, P_ONE, ajj, lda, a12, lda, nullptr));
      if (j + jb < m)
        PGA_TRY(rgemm_launch(false, false, m - j - jb, n - j - jb, jb, P_MONE, ajj + jb, lda, a12, lda,
                             P_ONE, a12 + jb, lda, c.s, c.impl));
    }
  }
  return hipSuccess;
}

// The blocked LU with the device state of its multi-workgroup panel pairs, taken from the
// library's memory pool on the caller's stream for the call (without a pool, or for
// short panels, every pair runs on one workgroup).
static hipError_t getrf(const Ctx& c, int m, int n, uint32_t* a, int lda, int* ipiv, int* info, int nb) {
  hipMemPool_t pool = runtime().pool;
  if (!pool || (m < pair_mc_min() && panel_w() == 0)) return getrf_blocked(c, m, n, a, lda, ipiv, info, nb);
  void* p = nullptr;
  if (hipMallocFromPoolAsync(&p, sizeof(leaf::PairSync), pool, c.s) != hipSuccess) {
    (void)hipGetLastError();
    return getrf_blocked(c, m, n, a, lda, ipiv, info, nb);
  }
  hipLaunchKernelGGL(k_zero_words, dim3(1), dim3(256), 0, c.s, (unsigned*)p,
                     (int)(sizeof(leaf::PairSync) / sizeof(unsigned)));
  PairState st;
  st.dev = (leaf::PairSync*)p;
  Ctx cp = c;
  cp.pair = &st;
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = getrf_blocked(cp, m, n, a, lda, ipiv, info, nb);
  const hipError_t f = hipFreeAsync(p, c.s);
  return e != hipSuccess ? e : f;
}

// Rgetrs (Rgetrs.cpp:32-118)
static hipError_t getrs(const Ctx& c, bool trans, int n, int nrhs, const uint32_t* a, int lda,
                        const int* ipiv, uint32_t* b, int ldb) {
  if (n == 0 || nrhs == 0) return hipSuccess;
  if (!trans) {
    PGA_TRY(laswp(c, nrhs, b, ldb, 1, n, ipiv, 1, 0, nullptr));
    PGA_TRY(trsm(c, true, false, false, false, n, nrhs, P_ONE, a, lda, b, ldb, nullptr));
    PGA_TRY(trsm(c, true, true, false, true, n, nrhs, P_ONE, a, lda, b, ldb, nullptr));
  } else {
    PGA_TRY(trsm(c, true, true, true, true, n, nrhs, P_ONE, a, lda, b, ldb, nullptr));
    PGA_TRY(trsm(c, true, false, true, false, n, nrhs, P_ONE, a, lda, b, ldb, nullptr));
    PGA_TRY(laswp(c, nrhs, b, ldb, 1, n, ipiv, -1, 0, nullptr));
  }
  return hipSuccess;
}

// Rsyrk (Rsyrk.cpp:31-215) as a triangular GEMM: 'N' is C op= alpha*A*A^T with zero
// multipliers skipped, 'T' is the sum-first form alpha*A^T*A + beta*C.
static hipError_t syrk(const Ctx& c, bool upper, bool trans, int n, int k, uint32_t alpha,
                       const uint32_t* a, int lda, uint32_t beta, uint32_t* cc, int ldc, const int* stop) {
  if (n == 0 || ((alpha == 0 || k == 0) && beta == P_ONE)) return hipSuccess;
  GemmOpts o;
  o.skip_zero_b = !trans;
  o.tri = upper ? 2 : 1;
  o.stop = stop;
  if (!trans) return rgemm_launch(false, true, n, n, k, alpha, a, lda, a, lda, beta, cc, ldc, c.s, c.impl, o);
  return rgemm_launch(true, false, n, n, k, alpha, a, lda, a, lda, beta, cc, ldc, c.s, c.impl, o);
}

// Rpotrf2 (Rpotrf2.cpp:32-151); info (device) doubles as the stop flag
// Orders up to POTRF2_ONE_WG_MAX run as one single-workgroup launch (the recursion's
// steps are then barriers instead of ~3n kernel launches); larger ones recurse with
// device-wide GEMMs down to that size.
static hipError_t potrf2(const Ctx& c, bool upper, int n, uint32_t* a, int lda, int* info, int col_off) {
  if (n == 0) return hipSuccess;
  if (n <= POTRF2_ONE_WG_MAX) {
    if (upper)
      hipLaunchKernelGGL(k_potrf2_upper_fast, dim3(1), dim3(WG), 0, c.s, n, a, lda, info, col_off, nullptr, 0);
    else hipLaunchKernelGGL(k_potrf2_lower_fast, dim3(1), dim3(WG), 0, c.s, n, a, lda, info, col_off);
    return hipGetLastError();
  }
  if (n == 1) {
    hipLaunchKernelGGL(k_potrf2_leaf, dim3(1), dim3(64), 0, c.s, a, info, col_off);
    return hipGetLastError();
  }
  const int n1 = n / 2, n2 = n - n1;
  uint32_t* a22 = a + n1 + (size_t)n1 * lda;
  PGA_TRY(potrf2(c, upper, n1, a, lda, info, col_off));
  if (upper) {
    PGA_TRY(trsm(c, true, true, true, true, n1, n2, P_ONE, a, lda, a + (size_t)n1 * lda, lda, info));
    PGA_TRY(syrk(c, true, true, n2, n1, P_MONE, a + (size_t)n1 * lda, lda, P_ONE, a22, lda, info));
  } else {
    PGA_TRY(trsm(c, false, false, true, true, n2, n1, P_ONE, a, lda, a + n1, lda, info));
    PGA_TRY(syrk(c, false, false, n2, n1, P_MONE, a + n1, lda, P_ONE, a22, lda, info));
  }
  return potrf2(c, upper, n2, a22, lda, info, col_off + n1);
}

// B := A on an m x n block; with a stop flag set, nothing (the panel update of a block
// whose diagonal factorization failed is not applied: the reference stops before it)
__global__ __launch_bounds__(WG) void k_copy_block(int m, int n, const uint32_t* A, int lda, uint32_t* B, int ldb,
                                                   const int* stop) {
  PGA_KERNEL_VGPR_FLOOR();
  if (stopped(stop)) return;
  for (long e = blockIdx.x * (long)WG + threadIdx.x; e < (long)m * n; e += (long)gridDim.x * WG)
    B[e % m + (size_t)(e / m) * ldb] = A[e % m + (size_t)(e / m) * lda];
}

static hipError_t copy_block(hipStream_t s, int m, int n, const uint32_t* A, int lda, uint32_t* B, int ldb,
                             const int* stop) {
  const long total = (long)m * n;
  if (total <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_copy_block, dim3((unsigned)std::min<long>((total + WG - 1) / WG, 4096)), dim3(WG), 0, s, m, n,
                     A, lda, B, ldb, stop);
  return hipGetLastError();
}

// Orders from which the panel GEMM of each block runs beside the block's diagonal steps
constexpr int POTRF_OVERLAP_MIN_N = 1024;

// Rpotrf_opencl with the panel GEMM of block j on a second stream, into a workspace W:
// it needs only the columns before j, so it runs while the same block's Rsyrk and
// Rpotrf2 run on the caller's stream.  W is copied over the panel once Rpotrf2 has
// succeeded (info doubles as the stop flag), then the panel's Rtrsm runs.  Every
// operation sees the same operands as in the sequential order, so the bits are the same.
static hipError_t potrf_overlap(const Ctx& c, bool upper, int n, uint32_t* a, int lda, int* info, int nb) {
  auto A = [&](int r, int col) { return a + r + (size_t)col * lda; };
  Aux x;
  PGA_TRY_E(aux_take(x));
  uint32_t* W = nullptr;
  hipError_t e = hipMallocAsync((void**)&W, (size_t)(n - nb) * nb * sizeof(uint32_t), c.s);
  bool pending = false;  // the aux stream has work that c.s has not joined
  for (int j = 0; j < n && e == hipSuccess; j += nb) {
    const int jb = (n - j < nb) ? n - j : nb, rest = n - j - jb;
    const int wm = upper ? jb : rest, wn = upper ? rest : jb;  // W's shape (the panel)
    uint32_t* P = upper ? A(j, j + jb) : A(j + jb, j);
    const bool gemm = rest > 0 && j > 0;  // Rgemm quick-returns on K = 0
    if (gemm) {
      if ((e = hipEventRecord(x.join, c.s)) != hipSuccess) break;
      if ((e = hipStreamWaitEvent(x.s, x.join, 0)) != hipSuccess) break;
      if ((e = copy_block(x.s, wm, wn, P, lda, W, wm, nullptr)) != hipSuccess) break;
      e = upper ? rgemm_launch(true, false, jb, rest, j, P_MONE, A(0, j), lda, A(0, j + jb), lda, P_ONE, W, wm, x.s,
                               c.impl)
                : rgemm_launch(false, true, rest, jb, j, P_MONE, A(j + jb, 0), lda, A(j, 0), lda, P_ONE, W, wm, x.s,
                               c.impl);
      if (e != hipSuccess || (e = hipEventRecord(x.done, x.s)) != hipSuccess) break;
      pending = true;
    }
    if (upper) e = syrk(c, true, true, jb, j, P_MONE, A(0, j), lda, P_ONE, A(j, j), lda, info);
    else e = syrk(c, false, false, jb, j, P_MONE, A(j, 0), lda, P_ONE, A(j, j), lda, info);
    if (e == hipSuccess) e = potrf2(c, upper, jb, A(j, j), lda, info, j);
    if (e != hipSuccess || rest == 0) break;
    if (gemm) {
      if ((e = hipStreamWaitEvent(c.s, x.done, 0)) != hipSuccess) break;
      pending = false;
      if ((e = copy_block(c.s, wm, wn, W, wm, P, lda, info)) != hipSuccess) break;
    }
    e = upper ? trsm(c, true, true, true, true, jb, rest, P_ONE, A(j, j), lda, P, lda, info)
              : trsm(c, false, false, true, true, rest, jb, P_ONE, A(j, j), lda, P, lda, info);
  }
  if (pending) (void)hipStreamWaitEvent(c.s, x.done, 0);
  if (W) {
    hipError_t f = hipFreeAsync(W, c.s);
    if (e == hipSuccess) e = f;
  }
  aux_give(x);
  return e;
}

// A Cholesky trailing update by one finished block of kb columns, L (rows from the
// region's origin down): C(i, c) = beta*C(i, c) + sum of (alpha*L(c, l)) * L(i, l) for
// l = 0 .. kb-1 in order, on the lower trapezoid i >= c of an m x ncols region whose
// origin lies on the nb-block grid.  The lower factor (alpha = -1, beta = 1, SKIP): an
// element whose row and column fall in one nb x nb diagonal block receives the Rsyrk
// 'L','N' terms (zero multipliers L(c, l) skipped, Rsyrk.cpp:149-169), any other element
// the RgemmG 'N','T' terms (no skip, Rpotrf_opencl.cpp:146) -- the same per-element steps
// as the left-looking order's Rsyrk on the diagonal block and RgemmG on the panel below
// it.  The upper factor's running sums (alpha = 1, no SKIP) take every term.
// T divides nb, so each T x T tile is wholly one or the other.  One workgroup per tile of
// the lower trapezoid, tiles numbered column by column.
template <int T, bool SKIP>
__global__ __launch_bounds__(WG, (gemm::square_tile_wgs_per_cu<T>())) void k_potrf_trail(
    int m, int ncols, int kb, const uint32_t* __restrict__ L, int lda, uint32_t* __restrict__ C, int ldc,
    uint32_t alpha, uint32_t beta, int nb, int urgent, const int* stop) {
  PGA_KERNEL_VGPR_FLOOR();
  if (urgent) PGA_URGENT();
  __shared__ __align__(16) uint32_t lds[gemm::square_tile_lds_words<T>()];
  if (stopped(stop)) return;
  const int tm = (m + T - 1) / T, tc = (ncols + T - 1) / T, t = blockIdx.x;
  // column q holds tiles q .. tm-1: it starts at q*tm - q*(q-1)/2
  int lo = 0, hi = tc - 1;
  while (lo < hi) {
    const int q = (lo + hi + 1) >> 1;
    if (q * tm - q * (q - 1) / 2 <= t) lo = q;
    else hi = q - 1;
  }
  const int tj = lo, ti = tj + (t - (tj * tm - tj * (tj - 1) / 2));
  const int i0 = ti * T, j0 = tj * T;
  if constexpr (T == 64) {
    // a 64 x 64 tile as two 64 x 32 halves (4 x 2 per lane)
    for (int jh = j0; jh < min(j0 + 64, ncols); jh += 32) {
      if (i0 / nb == jh / nb)
        gemm::gemm_tile_in<64, 32, 16, 4, 2, false, true, SKIP>(m, ncols, kb, alpha, L, lda, L, lda, beta, C, ldc,
                                                                i0, jh, 1, lds);
      else
        gemm::gemm_tile_in<64, 32, 16, 4, 2, false, true, false>(m, ncols, kb, alpha, L, lda, L, lda, beta, C, ldc,
                                                                 i0, jh, 0, lds);
    }
  } else {
    if (i0 / nb == j0 / nb)
      gemm::gemm_tile_in<T, T, 16, T / 16, T / 16, false, true, SKIP>(m, ncols, kb, alpha, L, lda, L, lda, beta, C,
                                                                     ldc, i0, j0, 1, lds);
    else
      gemm::gemm_tile_in<T, T, 16, T / 16, T / 16, false, true, false>(m, ncols, kb, alpha, L, lda, L, lda, beta,
                                                                      C, ldc, i0, j0, 0, lds);
  }
}

// k_potrf_trail<64> with its slices copied from FP64 operand panels of L (A side) and
// -L^T (B side) instead of decoded in every tile: 28 KB of LDS, so five workgroups share a
// CU as in the plain GEMM.  The same tiles, forms and order.
template <bool SKIP>
__global__ __launch_bounds__(WG, 3) void k_potrf_trail_panel(int m, int ncols, int kb, const uint32_t* __restrict__ L,
                                                          int lda, uint32_t* __restrict__ C, int ldc, uint32_t alpha,
                                                          uint32_t beta, int nb, int urgent, const int* stop,
                                                          gemm::Panels pn) {
  PGA_KERNEL_VGPR_FLOOR();
  if (urgent) PGA_URGENT();
  __shared__ __align__(16) double sA[16][64];
  __shared__ __align__(16) double sT[16][32];
  if (stopped(stop)) return;
  const int tm = (m + 63) / 64, tc = (ncols + 63) / 64, t = blockIdx.x;
  int lo = 0, hi = tc - 1;  // column q of the trapezoid holds tiles q .. tm-1
  while (lo < hi) {
    const int q = (lo + hi + 1) >> 1;
    if (q * tm - q * (q - 1) / 2 <= t) lo = q;
    else hi = q - 1;
  }
  const int tj = lo, ti = tj + (t - (tj * tm - tj * (tj - 1) / 2));
  const int i0 = ti * 64, j0 = tj * 64;
  for (int jh = j0; jh < min(j0 + 64, ncols); jh += 32) {
    if (i0 / nb == jh / nb)
      gemm::gemm_tile_impl_fp64<64, 32, 16, 4, 2, false, true, SKIP, true>(m, ncols, kb, alpha, L, lda, L, lda, beta,
                                                                          C, ldc, i0, jh, 1, sA, sT, pn);
    else
      gemm::gemm_tile_impl_fp64<64, 32, 16, 4, 2, false, true, false, true>(m, ncols, kb, alpha, L, lda, L, lda,
                                                                           beta, C, ldc, i0, jh, 0, sA, sT, pn);
  }
}

// Rpotrf_opencl returns at the first failing diagonal block (Rpotrf_opencl.cpp:137-139):
// the panel below that block and every later column keep their input values.  Once a
// right-looking factorization has failed, put the saved inputs back there: columns c from
// the failed block's start J, rows from max(c, J + nb).
__global__ __launch_bounds__(WG) void k_potrf_restore(int n, uint32_t* __restrict__ a, int lda,
                                                      const uint32_t* __restrict__ save, int nb,
                                                      const int* __restrict__ info) {
  PGA_KERNEL_VGPR_FLOOR();
  const int f = *info;
  if (f == 0) return;
  const int J = (f - 1) / nb * nb, r0 = min(n, J + nb);
  for (int c = J + blockIdx.y; c < n; c += gridDim.y)
    for (int i = max(c, r0) + blockIdx.x * WG + threadIdx.x; i < n; i += gridDim.x * WG)
      a[i + (size_t)c * lda] = save[i + (size_t)c * n];
}

// Tile edge for a trailing update of about m x ncols / 2 elements: the largest of 64, 32,
// 16 that divides nb and still gives every CU two tiles.
static int trail_tile(long m, long ncols, int nb) {
  const long want = 2L * runtime().cus;
  for (int T : {64, 32}) {
    if (nb % T) continue;
    const long tm = (m + T - 1) / T, tc = (ncols + T - 1) / T;
    if (tc * tm - tc * (tc - 1) / 2 >= want) return T;
  }
  return 16;
}

// Large lower-Cholesky trailing updates read FP64 operand panels (PGA_CHOL_TRAIL_PANELS=0
// turns that off)
static bool chol_trail_panels() {
  static const bool v = [] {
    const char* e = getenv("PGA_CHOL_TRAIL_PANELS");
    return !e || atoi(e) != 0;
  }();
  return v;
}

// Workgroups per CU of the lower Cholesky's bulk trailing update (PGA_CHOL_TRAIL_CAP; 0:
// no cap), so that the diagonal block and panel kernels of the next block find room
static int chol_trail_cap() {
  static const int v = [] {
    const char* e = getenv("PGA_CHOL_TRAIL_CAP");
    return e ? std::max(0, atoi(e)) : 0;
  }();
  return v;
}

template <bool SKIP>
static hipError_t potrf_trail_form(hipStream_t s, int m, int ncols, int kb, const uint32_t* L, int lda, uint32_t* C,
                                   int ldc, uint32_t alpha, uint32_t beta, int nb, bool urgent, const int* stop,
                                   int cap) {
  const int T = trail_tile(m, ncols, nb);
  const long tm = (m + T - 1) / T, tc = (ncols + T - 1) / T;
  const unsigned grid = (unsigned)(tc * tm - tc * (tc - 1) / 2);
  if (T == 64 && chol_trail_panels() && (long long)m * ncols * kb >= (1LL << 28)) {
    // the B-side panel is alpha*L^T; the Rsyrk tiles' zero skip is the tile's own (a zero stays zero)
    gemm::Panels pn;
    void* ws = nullptr;
    hipError_t e = gemm::panels_build(false, true, SKIP, m, ncols, kb, alpha, L, lda, L, lda, s, &pn, &ws);
    if (e != hipSuccess) return e;
    if (ws) {
      hipLaunchKernelGGL(k_potrf_trail_panel<SKIP>, dim3(grid), dim3(WG),
                         lds_for_cap((const void*)k_potrf_trail_panel<SKIP>, cap), s, m, ncols, kb, L, lda, C, ldc,
                         alpha, beta, nb, (int)urgent, stop, pn);
      e = hipGetLastError();
      const hipError_t f = hipFreeAsync(ws, s);
      return e != hipSuccess ? e : f;
    }
  }
#define PGA_TRAIL(TT)                                                                                               \
  hipLaunchKernelGGL((k_potrf_trail<TT, SKIP>), dim3(grid), dim3(WG),                                               \
                     lds_for_cap((const void*)k_potrf_trail<TT, SKIP>, cap), s, m, ncols, kb, L, lda, C, ldc, alpha, \
                     beta, nb, (int)urgent, stop)
  if (T == 64) PGA_TRAIL(64);
  else if (T == 32) PGA_TRAIL(32);
  else PGA_TRAIL(16);
#undef PGA_TRAIL
  return hipGetLastError();
}

// The Cholesky trailing updates on the lower trapezoid of an m x ncols region of C whose
// origin lies on the nb-block grid diagonal, L holding the region's rows:
// C(i, c) = beta*C(i, c) + sum over l = 0 .. kb-1, in ascending order, of (alpha*L(c, l)) *
// L(i, l), each step rounded (the GEMM's 'N','T' accumulation).  With skip (the lower
// factor: alpha = -1, beta = 1) an element of an nb x nb diagonal block takes the Rsyrk
// 'L','N' form instead, which skips zero multipliers L(c, l).
static hipError_t potrf_trail(hipStream_t s, int m, int ncols, int kb, const uint32_t* L, int lda, uint32_t* C,
                              int ldc, int nb, bool urgent, const int* stop, int cap = 0, uint32_t alpha = P_MONE,
                              uint32_t beta = P_ONE, bool skip = true) {
  if (m <= 0 || ncols <= 0 || kb <= 0) return hipSuccess;
  return skip ? potrf_trail_form<true>(s, m, ncols, kb, L, lda, C, ldc, alpha, beta, nb, urgent, stop, cap)
              : potrf_trail_form<false>(s, m, ncols, kb, L, lda, C, ldc, alpha, beta, nb, urgent, stop, cap);
}


// Lower Rpotrf_opencl in right-looking order with one block of lookahead.  Each element
// still receives its terms in ascending l with the Rsyrk or Rgemm form the left-looking
// order gives it (k_potrf_trail), then Rpotrf2 or the panel's Rtrsm, so the factor is
// the same bits.  Per block j, on the caller's stream: the update of block column j+1 by
// block j, then Rpotrf2 and the panel Rtrsm of block j+1; meanwhile a second stream
// applies block j to every later column -- the bulk of the work, as large GEMM launches.
// A failing block leaves info set (every later launch then does nothing) and the saved
// input is copied back over what Rpotrf_opencl never touches.  The (upper) 'T' forms sum
// each element's terms before adding them, so they cannot be split by block this way.
static hipError_t potrf_right_lower(const Ctx& c, int n, uint32_t* a, int lda, int* info, int nb) {
  auto A = [&](int r, int col) { return a + r + (size_t)col * lda; };
  uint32_t* save = nullptr;
  PGA_TRY_E(hipMallocAsync((void**)&save, (size_t)n * n * sizeof(uint32_t), c.s));
  Aux x;
  hipError_t e = aux_take(x);
  if (e != hipSuccess) {
    (void)hipFreeAsync(save, c.s);
    return e;
  }
  e = copy_block(c.s, n, n, a, lda, save, n, nullptr);
  const int jb0 = std::min(nb, n);
  if (e == hipSuccess) e = potrf2(c, false, jb0, a, lda, info, 0);
  if (e == hipSuccess && jb0 < n) e = trsm(c, false, false, true, true, n - jb0, jb0, P_ONE, a, lda, A(jb0, 0), lda, info);
  bool recorded = false;  // the aux stream has recorded x.done at least once
  for (int j = 0; j + nb < n && e == hipSuccess; j += nb) {
    const int c0 = j + nb, nb1 = std::min(nb, n - c0), c1 = c0 + nb1;
    // block j is final: the aux stream applies it to the columns from c1 on
    if (c1 < n) {
      if ((e = hipEventRecord(x.join, c.s)) != hipSuccess) break;
      if ((e = hipStreamWaitEvent(x.s, x.join, 0)) != hipSuccess) break;
      if ((e = potrf_trail(x.s, n - c1, n - c1, nb, A(c1, j), lda, A(c1, c1), lda, nb, false, info, chol_trail_cap())) !=
          hipSuccess)
        break;
    }
    // block column j+1 has had every earlier block's terms once the previous aux launch ends
    if (recorded && (e = hipStreamWaitEvent(c.s, x.done, 0)) != hipSuccess) break;
    if (c1 < n) {
      if ((e = hipEventRecord(x.done, x.s)) != hipSuccess) break;
      recorded = true;
    }
    if ((e = potrf_trail(c.s, n - c0, nb1, nb, A(c0, j), lda, A(c0, c0), lda, nb, true, info)) != hipSuccess) break;
    if ((e = potrf2(c, false, nb1, A(c0, c0), lda, info, c0)) != hipSuccess) break;
    if (c1 < n) e = trsm(c, false, false, true, true, n - c1, nb1, P_ONE, A(c0, c0), lda, A(c1, c0), lda, info);
  }
  if (recorded) {
    hipError_t w = hipStreamWaitEvent(c.s, x.done, 0);
    if (e == hipSuccess) e = w;
  }
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_potrf_restore, dim3((unsigned)std::min(16, (n + WG - 1) / WG), (unsigned)std::min(n, 256)),
                       dim3(WG), 0, c.s, n, a, lda, save, nb, info);
    e = hipGetLastError();
  }
  hipError_t f = hipFreeAsync(save, c.s);
  if (e == hipSuccess) e = f;
  aux_give(x);
  return e;
}

// dst (R x K) := src^T for src K x R (rows l, columns r), through 32 x 32 LDS tiles so
// that both sides stay coalesced; with a stop flag set, nothing.
__global__ __launch_bounds__(WG) void k_transpose(int K, int R, const uint32_t* __restrict__ src, int lds,
                                                  uint32_t* __restrict__ dst, int ldd, const int* stop) {
  PGA_KERNEL_VGPR_FLOOR();
  __shared__ uint32_t t[32][33];
  if (stopped(stop)) return;
  const int r0 = blockIdx.x * 32, l0 = blockIdx.y * 32, tx = threadIdx.x % 32, ty = threadIdx.x / 32;
  for (int q = ty; q < 32; q += WG / 32) {
    const int l = l0 + tx, r = r0 + q;
    if (l < K && r < R) t[q][tx] = src[l + (size_t)r * lds];
  }
  __syncthreads();
  for (int q = ty; q < 32; q += WG / 32) {
    const int r = r0 + tx, l = l0 + q;
    if (l < K && r < R) dst[r + (size_t)l * ldd] = t[tx][q];
  }
}

// The upper Cholesky's block row J takes its sums (Rsyrk 'U','T' on the diagonal block,
// RgemmG 'T','N' on the rest, Rpotrf_opencl.cpp:107,114-119; both end with
// c = alpha*t + beta*c, Rsyrk.cpp:188-192, Rgemm.cpp:197-201): A := -1*t + 1*A on an
// m x ncols region, its upper triangle only where `diag`, with the sum of element (i, c)
// at St[c + i * lds] (the sums are kept transposed); with a stop flag set, nothing.
__global__ __launch_bounds__(WG) void k_apply_sums(int m, int ncols, const uint32_t* __restrict__ St, int lds,
                                                   uint32_t* __restrict__ A, int lda, int diag, const int* stop) {
  PGA_KERNEL_VGPR_FLOOR();
  if (stopped(stop)) return;
  for (long e = blockIdx.x * (long)WG + threadIdx.x; e < (long)m * ncols; e += (long)gridDim.x * WG) {
    const int i = (int)(e % m), c = (int)(e / m);
    if (diag && i > c) continue;
    uint32_t& v = A[i + (size_t)c * lda];
    v = add(mul(P_MONE, St[c + (size_t)i * lds]), mul(P_ONE, v));
  }
}

static hipError_t apply_sums(hipStream_t s, int m, int ncols, const uint32_t* St, int lds, uint32_t* A, int lda,
                             bool diag, const int* stop) {
  const long total = (long)m * ncols;
  if (total <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_apply_sums, dim3((unsigned)std::min<long>((total + WG - 1) / WG, 4096)), dim3(WG), 0, s, m,
                     ncols, St, lds, A, lda, (int)diag, stop);
  return hipGetLastError();
}

// Upper Rpotrf_opencl (Rpotrf_opencl.cpp:100-124) in right-looking order.  The reference
// forms block row J from all earlier ones at once: Rsyrk 'U','T' on the diagonal block
// and RgemmG 'T','N' on the rest both compute t = sum over l < J of a_li * a_lc from zero
// in ascending l (each step t = t + a_li * a_lc rounded), then c = -t + c.  Those running
// sums live in a workspace, transposed (St(c, i) = t of element (i, c), a lower
// trapezoid): once block row j of U is final, its nb terms are added to every later
// element's sum in ascending l by the lower trailing update with alpha = 1 and no zero
// skip (potrf_trail, over P = the block row's transpose), beta = 0 for the first block so
// that every sum starts from zero; St keeps each sum as the posit the reference holds
// after that term.  Block row J then takes c := -t + c on its diagonal block, Rpotrf2,
// and, once that succeeded, c := -t + c and Rtrsm on the rest (the reference's RgemmG
// comes after its info check).  The caller's matrix is written only in the block row in
// hand, so a failing block leaves it as the reference does.  One block of lookahead: the
// caller's stream adds block j to block row j+1's sums and factors that row, while a
// second stream adds block j to every later row's sums.
static hipError_t potrf_right_upper(const Ctx& c, int n, uint32_t* a, int lda, int* info, int nb) {
  auto A = [&](int r, int col) { return a + r + (size_t)col * lda; };
  uint32_t *S = nullptr, *P = nullptr;
  PGA_TRY_E(hipMallocAsync((void**)&S, (size_t)n * n * sizeof(uint32_t), c.s));
  hipError_t e = hipMallocAsync((void**)&P, (size_t)2 * n * nb * sizeof(uint32_t), c.s);
  Aux x;
  if (e == hipSuccess) e = aux_take(x);
  if (e != hipSuccess) {
    (void)hipFreeAsync(S, c.s);
    if (P) (void)hipFreeAsync(P, c.s);
    return e;
  }
  auto St = [&](int r, int col) { return S + r + (size_t)col * n; };
  const int jb0 = std::min(nb, n);
  e = potrf2(c, true, jb0, a, lda, info, 0);
  if (e == hipSuccess && jb0 < n) e = trsm(c, true, true, true, true, jb0, n - jb0, P_ONE, a, lda, A(0, jb0), lda, info);
  bool recorded = false;  // the aux stream has recorded x.done at least once
  for (int j = 0; j + nb < n && e == hipSuccess; j += nb) {
    const int c0 = j + nb, nb1 = std::min(nb, n - c0), c1 = c0 + nb1, ldp = n - c0;
    const uint32_t beta = j == 0 ? 0u : P_ONE;  // block 0's terms start every sum
    // P (n - c0 x nb) = (rows j .. j+nb-1 of U, columns c0 ..)^T; the two buffers alternate,
    // and the aux launch that read this one two blocks ago has ended (x.done, waited below)
    uint32_t* Pj = P + (size_t)((j / nb) & 1) * n * nb;
    hipLaunchKernelGGL(k_transpose, dim3((unsigned)((n - c0 + 31) / 32), (unsigned)((nb + 31) / 32)), dim3(WG), 0,
                       c.s, nb, n - c0, A(j, c0), lda, Pj, ldp, info);
    if ((e = hipGetLastError()) != hipSuccess) break;
    if (c1 < n) {  // the aux stream adds block j to the sums of rows c1 ..
      if ((e = hipEventRecord(x.join, c.s)) != hipSuccess) break;
      if ((e = hipStreamWaitEvent(x.s, x.join, 0)) != hipSuccess) break;
      if ((e = potrf_trail(x.s, n - c1, n - c1, nb, Pj + (c1 - c0), ldp, St(c1, c1), n, nb, false, info,
                           chol_trail_cap(), P_ONE, beta, false)) != hipSuccess)
        break;
    }
    // block row c0's sums have every earlier block's terms once the previous aux launch ends
    if (recorded && (e = hipStreamWaitEvent(c.s, x.done, 0)) != hipSuccess) break;
    if (c1 < n) {
      if ((e = hipEventRecord(x.done, x.s)) != hipSuccess) break;
      recorded = true;
    }
    if ((e = potrf_trail(c.s, n - c0, nb1, nb, Pj, ldp, St(c0, c0), n, nb, true, info, 0, P_ONE, beta, false)) !=
        hipSuccess)
      break;
    if (nb1 <= POTRF2_ONE_WG_MAX && nb1 <= LUTN_MMAX) {
      // the sums applied inside the diagonal block's and the row solve's own launches
      hipLaunchKernelGGL(k_potrf2_upper_fast, dim3(1), dim3(WG), 0, c.s, nb1, A(c0, c0), lda, info, c0,
                         St(c0, c0), n);
      if ((e = hipGetLastError()) != hipSuccess) break;
      if (c1 < n) {
        hipLaunchKernelGGL(k_trsm_lutn_fp64, dim3((unsigned)((n - c1 + LUTN_COLS - 1) / LUTN_COLS)), dim3(WG), 0, c.s,
                           nb1, n - c1, A(c0, c0), lda, A(c0, c1), lda, info, St(c1, c0), n);
        e = hipGetLastError();
      }
      continue;
    }
    if ((e = apply_sums(c.s, nb1, nb1, St(c0, c0), n, A(c0, c0), lda, true, info)) != hipSuccess) break;
    if ((e = potrf2(c, true, nb1, A(c0, c0), lda, info, c0)) != hipSuccess) break;
    if (c1 < n) {
      if ((e = apply_sums(c.s, nb1, n - c1, St(c1, c0), n, A(c0, c1), lda, false, info)) != hipSuccess) break;
      e = trsm(c, true, true, true, true, nb1, n - c1, P_ONE, A(c0, c0), lda, A(c0, c1), lda, info);
    }
  }
  if (recorded) {
    hipError_t w = hipStreamWaitEvent(c.s, x.done, 0);
    if (e == hipSuccess) e = w;
  }
  hipError_t f = hipFreeAsync(S, c.s);
  if (e == hipSuccess) e = f;
  f = hipFreeAsync(P, c.s);
  if (e == hipSuccess) e = f;
  aux_give(x);
  return e;
}

// Rpotrf_opencl (Rpotrf_opencl.cpp:32-166): blocked left-looking Cholesky
static hipError_t potrf(const Ctx& c, bool upper, int n, uint32_t* a, int lda, int* info, int nb) {
  if (nb <= 1 || nb >= n) return potrf2(c, upper, n, a, lda, info, 0);
  if (!upper && n >= POTRF_OVERLAP_MIN_N && nb % 16 == 0) return potrf_right_lower(c, n, a, lda, info, nb);
  if (upper && n >= POTRF_OVERLAP_MIN_N && nb % 16 == 0) return potrf_right_upper(c, n, a, lda, info, nb);
  if (n >= POTRF_OVERLAP_MIN_N) return potrf_overlap(c, upper, n, a, lda, info, nb);
  auto A = [&](int r, int col) { return a + r + (size_t)col * lda; };
  for (int j = 0; j < n; j += nb) {
    const int jb = (n - j < nb) ? n - j : nb;
    if (upper) {
      PGA_TRY(syrk(c, true, true, jb, j, P_MONE, A(0, j), lda, P_ONE, A(j, j), lda, info));
      PGA_TRY(potrf2(c, true, jb, A(j, j), lda, info, j));
      if (j + jb < n) {
        GemmOpts o;
        o.stop = info;
        PGA_TRY(rgemm_launch(true, false, jb, n - j - jb, j, P_MONE, A(0, j), lda, A(0, j + jb), lda,
                             P_ONE, A(j, j + jb), lda, c.s, c.impl, o));
        PGA_TRY(trsm(c, true, true, true, true, jb, n - j - jb, P_ONE, A(j, j), lda, A(j, j + jb), lda,
                     info));
      }
    } else {
      PGA_TRY(syrk(c, false, false, jb, j, P_MONE, A(j, 0), lda, P_ONE, A(j, j), lda, info));
      PGA_TRY(potrf2(c, false, jb, A(j, j), lda, info, j));
      if (j + jb < n) {
        GemmOpts o;
        o.stop = info;
        PGA_TRY(rgemm_launch(false, true, n - j - jb, jb, j, P_MONE, A(j + jb, 0), lda, A(j, 0), lda,
                             P_ONE, A(j + jb, j), lda, c.s, c.impl, o));
        PGA_TRY(trsm(c, false, false, true, true, n - j - jb, jb, P_ONE, A(j, j), lda, A(j + jb, j), lda,
                     info));
      }
    }
  }
  return hipSuccess;
}

// Rpotrs (Rpotrs.cpp:32-110)
static hipError_t potrs(const Ctx& c, bool upper, int n, int nrhs, const uint32_t* a, int lda, uint32_t* b,
                        int ldb) {
  if (n == 0 || nrhs == 0) return hipSuccess;
  if (upper) {
    PGA_TRY(trsm(c, true, true, true, true, n, nrhs, P_ONE, a, lda, b, ldb, nullptr));
    PGA_TRY(trsm(c, true, true, false, true, n, nrhs, P_ONE, a, lda, b, ldb, nullptr));
  } else {
    PGA_TRY(trsm(c, true, false, false, true, n, nrhs, P_ONE, a, lda, b, ldb, nullptr));
    PGA_TRY(trsm(c, true, false, true, true, n, nrhs, P_ONE, a, lda, b, ldb, nullptr));
  }
  return hipSuccess;
}

#undef PGA_TRY

}  // namespace lah

// exported to the C ABI layer (capi_lapack.cpp)
hipError_t la_trsm(hipStream_t s, bool left, bool upper, bool trans, bool nounit, int m, int n, uint32_t alpha,
                   const uint32_t* A, int lda, uint32_t* B, int ldb) {
  return lah::trsm({s, runtime().gemm_impl}, left, upper, trans, nounit, m, n, alpha, A, lda, B, ldb, nullptr);
}
hipError_t la_syrk(hipStream_t s, bool upper, bool trans, int n, int k, uint32_t alpha, const uint32_t* a, int lda,
                   uint32_t beta, uint32_t* c, int ldc) {
  return lah::syrk({s, runtime().gemm_impl}, upper, trans, n, k, alpha, a, lda, beta, c, ldc, nullptr);
}
hipError_t la_gemv(hipStream_t s, bool trans, int m, int n, uint32_t alpha, const uint32_t* A, int lda,
                   const uint32_t* x, int incx, uint32_t beta, uint32_t* y, int incy) {
  if (m == 0 || n == 0 || (alpha == 0 && beta == P_ONE)) return hipSuccess;
  const int leny = trans ? n : m;
  hipLaunchKernelGGL(la::k_gemv, dim3(lah::blocks_for(leny)), dim3(la::WG), 0, s, (int)trans, m, n, alpha, A,
                     lda, x, incx, beta, y, incy);
  return hipGetLastError();
}
hipError_t la_scal(hipStream_t s, int n, uint32_t da, uint32_t* x, int incx) {
  if (n <= 0 || incx <= 0) return hipSuccess;
  unsigned g = lah::blocks_for(n);
  if (g > 65536) g = 65536;
  hipLaunchKernelGGL(la::k_scal, dim3(g), dim3(la::WG), 0, s, n, da, x, incx);
  return hipGetLastError();
}
hipError_t la_iamax(hipStream_t s, int n, const uint32_t* x, int incx, int* result) {
  hipLaunchKernelGGL(la::k_iamax, dim3(1), dim3(la::WG), 0, s, n, x, incx, result);
  return hipGetLastError();
}
hipError_t la_laswp(hipStream_t s, int n, uint32_t* a, int lda, int k1, int k2, const int* ipiv, int incx) {
  return lah::laswp({s, runtime().gemm_impl}, n, a, lda, k1, k2, ipiv, incx, 0, nullptr);
}
hipError_t la_getrf(hipStream_t s, int m, int n, uint32_t* a, int lda, int* ipiv, int* info, int nb) {
  return lah::getrf({s, runtime().gemm_impl}, m, n, a, lda, ipiv, info, nb);
}
hipError_t la_getrf2(hipStream_t s, int m, int n, uint32_t* a, int lda, int* ipiv, int* info) {
  return lah::getrf2({s, runtime().gemm_impl}, m, n, a, lda, ipiv, info, 0, 0);
}
hipError_t la_getrs(hipStream_t s, bool trans, int n, int nrhs, const uint32_t* a, int lda, const int* ipiv,
                    uint32_t* b, int ldb) {
  return lah::getrs({s, runtime().gemm_impl}, trans, n, nrhs, a, lda, ipiv, b, ldb);
}
hipError_t la_potrf(hipStream_t s, bool upper, int n, uint32_t* a, int lda, int* info, int nb) {
  return lah::potrf({s, runtime().gemm_impl}, upper, n, a, lda, info, nb);
}
hipError_t la_potrf2(hipStream_t s, bool upper, int n, uint32_t* a, int lda, int* info) {
  return lah::potrf2({s, runtime().gemm_impl}, upper, n, a, lda, info, 0);
}
hipError_t la_potrs(hipStream_t s, bool upper, int n, int nrhs, const uint32_t* a, int lda, uint32_t* b,
                    int ldb) {
  return lah::potrs({s, runtime().gemm_impl}, upper, n, nrhs, a, lda, b, ldb);
}

}  // namespace pga
