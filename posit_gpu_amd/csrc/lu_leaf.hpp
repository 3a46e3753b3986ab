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

}  // namespace leaf
}  // namespace pga
