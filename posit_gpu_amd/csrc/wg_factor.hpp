// wg_factor.hpp -- LU / Cholesky routines run by one workgroup of 256 lanes.
//
// Each routine performs the reference routine's operations in the reference's order
// with workgroup-wide steps separated by barriers: the blocking of Rgetrf_opencl
// (Rgetrf_opencl.cpp:93-146) and Rpotrf_opencl (Rpotrf_opencl.cpp:86-160), the recursion
// of Rgetrf2 (Rgetrf2.cpp:84-183) and Rpotrf2 (Rpotrf2.cpp:84-145), and the loop orders of
// Rtrsm, Rsyrk and Rgemm.  Factors, pivots, info codes and solutions therefore equal the
// per-call results bit for bit, also for matrices with zero pivots, NaR entries or a
// failing Cholesky step (where the zero-multiplier skips of Rtrsm / Rsyrk and the state
// left behind at the failure depend on that structure).
//
// Recursions are walked leaf by leaf: after leaf j of a Rgetrf2 / Rpotrf2 call on w
// columns, the one node whose left half ends at j runs its middle step (the Rtrsm and
// the Rgemm / Rsyrk between the two recursive calls).  That is the reference's
// depth-first order.  Row interchanges are applied to every other column as soon as a
// pivot is chosen; the reference applies them later with Rlaswp, and no step reads
// those rows in between.
//
// Kernels that call these routines must be declared __launch_bounds__(256, 2): the GEMM
// tile functions are not inlined and then keep their 248 registers without spills.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>

#include "pga_internal.hpp"
#include "rgemm_tile.hpp"

namespace pga {
namespace wg {

constexpr int NT = 256;            // lanes per workgroup
constexpr int MAX_N = PGA_FACTOR_TASK_MAX_N;

typedef __attribute__((address_space(3))) uint32_t lds_u32;

struct Shared {
  alignas(16) uint32_t tile[gemm::TileLds<64, 64, 16>::words];  // GEMM tiles of every shape
  int perm[MAX_N];                                 // row interchanges of right-hand sides
  int red_v[NT / 64];
  int red_i[NT / 64];
  int bcast;
  uint32_t pivot;  // the pattern at iamax's index
};

static __device__ __forceinline__ uint32_t* at(uint32_t* a, int lda, int i, int j) {
  return a + i + (size_t)j * lda;
}

// ---------------------------------------------------------------------------------
// C op= -op(A) op(B) with beta = 1 in the order of Rgemm (Rgemm.cpp:163-246) or, with
// SKIP0, of Rsyrk 'N' (Rsyrk.cpp:158-176: a zero B-side multiplier skips the step);
// tri 1 / 2 stores only the lower / upper triangle (Rsyrk).  The per-element order does
// not depend on the tile shape, so the shape is chosen for the least padding.
// ---------------------------------------------------------------------------------
template <int BM, int BN, bool TA, bool TB, bool SKIP0>
static __device__ __noinline__ void gemm_t(int m, int n, int k, const uint32_t* A, int lda, const uint32_t* B, int ldb,
                                    uint32_t* C, int ldc, int tri, lds_u32* lds) {
  const int tm = (m + BM - 1) / BM, tn = (n + BN - 1) / BN;
  for (int t = 0; t < tm * tn; ++t) {
    const int i0 = (t % tm) * BM, j0 = (t / tm) * BN;
    if (tri == 1 && i0 + BM - 1 < j0) continue;
    if (tri == 2 && j0 + BN - 1 < i0) continue;
    gemm::gemm_tile_in<BM, BN, 16, BM / 16, BN / 16, TA, TB, SKIP0>(m, n, k, P_MONE, A, lda, B, ldb, P_ONE, C, ldc,
                                                                    i0, j0, tri, (uint32_t*)lds);
  }
  __syncthreads();
}

// Small products one element per lane with the pattern-level operations, in the same
// order and with the same epilogue as gemm_tile (alpha = -1 is a negation, beta = 1).
template <bool TA, bool TB, bool SKIP0>
static __device__ __noinline__ void gemm_direct(int m, int n, int k, const uint32_t* A, int lda, const uint32_t* B,
                                                int ldb, uint32_t* C, int ldc, int tri) {
  for (int e = threadIdx.x; e < m * n; e += NT) {
    const int i = e % m, j = e / m;
    if ((tri == 1 && i < j) || (tri == 2 && i > j)) continue;
    uint32_t* cij = C + i + (size_t)j * ldc;
    if (!TA) {  // c = C; c = c + (alpha*b_lj)*a_il, l ascending (Rgemm.cpp:163-185)
      uint32_t c = *cij;
      for (int l = 0; l < k; ++l) {
        const uint32_t b = TB ? B[j + (size_t)l * ldb] : B[l + (size_t)j * ldb];
        if (SKIP0 && b == 0) continue;  // Rsyrk.cpp:169
        c = add(c, mul(neg(b), A[i + (size_t)l * lda]));
      }
      *cij = c;
    } else {  // t = sum a_li*b_lj from zero; C = alpha*t + beta*C (Rgemm.cpp:186-204)
      uint32_t t = 0;
      for (int l = 0; l < k; ++l) {
        const uint32_t b = TB ? B[j + (size_t)l * ldb] : B[l + (size_t)j * ldb];
        t = add(t, mul(A[l + (size_t)i * lda], b));
      }
      *cij = add(neg(t), *cij);
    }
  }
  __syncthreads();
}

template <bool TA, bool TB, bool SKIP0>
__device__ __forceinline__ void gemm(int m, int n, int k, const uint32_t* A, int lda, const uint32_t* B, int ldb,
                                     uint32_t* C, int ldc, int tri, Shared& sh) {
  if (m <= 0 || n <= 0 || k <= 0) return;  // Rgemm.cpp:134-136 (beta = 1)
  if ((long)m * n * k <= 4096) {  // a tile's staging would cost more than the work
    gemm_direct<TA, TB, SKIP0>(m, n, k, A, lda, B, ldb, C, ldc, tri);
    return;
  }
  auto cover = [&](int BM, int BN) { return (long)((m + BM - 1) / BM) * ((n + BN - 1) / BN) * BM * BN; };
  // relative cost per covered element: fewer elements per lane amortise the LDS
  // reads over fewer multiply-adds (64 x 32 rather than 64 x 64: 4 x 2 per lane keeps the
  // task kernels' registers at a level that lets more workgroups share a CU)
  const long c64 = cover(64, 32) * 8, c32 = cover(32, 32) * 9, c16 = cover(16, 16) * 11;
  lds_u32* lds = (lds_u32*)sh.tile;
  if (c64 <= c32 && c64 <= c16) gemm_t<64, 32, TA, TB, SKIP0>(m, n, k, A, lda, B, ldb, C, ldc, tri, lds);
  else if (c32 <= c16) gemm_t<32, 32, TA, TB, SKIP0>(m, n, k, A, lda, B, ldb, C, ldc, tri, lds);
  else gemm_t<16, 16, TA, TB, SKIP0>(m, n, k, A, lda, B, ldb, C, ldc, tri, lds);
}

// ---------------------------------------------------------------------------------
// Triangular solves (Rtrsm.cpp), alpha = 1.  Left no-transpose forms sweep over k with
// every (row, column) pair of a step in parallel; the transposed upper form (a row's
// sum runs over earlier rows in ascending order) is the same sweep; the right-side
// form sweeps over columns.
// ---------------------------------------------------------------------------------
// B := L^-1 B, L unit lower m x m (Rtrsm.cpp:139-155); b_k == 0 skips step k
template <int T = NT>
static __device__ __noinline__ void trsm_llnu(int m, int n, const uint32_t* L, int ldl, uint32_t* B, int ldb) {
  for (int k = 0; k + 1 < m; ++k) {
    const int rows = m - 1 - k;
    for (int e = threadIdx.x; e < rows * n; e += T) {
      const int i = k + 1 + e % rows, j = e / rows;
      const uint32_t bk = B[k + (size_t)j * ldb];
      if (bk != 0) B[i + (size_t)j * ldb] = sub(B[i + (size_t)j * ldb], mul(bk, L[i + (size_t)k * ldl]));
    }
    __syncthreads();
  }
}

// B := L^-1 B, L non-unit lower (Rtrsm.cpp:139-155 with the division of :147-149)
template <int T = NT>
static __device__ __noinline__ void trsm_llnn(int m, int n, const uint32_t* L, int ldl, uint32_t* B, int ldb) {
  for (int k = 0; k < m; ++k) {
    const uint32_t lkk = L[k + (size_t)k * ldl];
    for (int j = threadIdx.x; j < n; j += T) {
      const uint32_t bk = B[k + (size_t)j * ldb];
      if (bk != 0) B[k + (size_t)j * ldb] = div(bk, lkk);
    }
    __syncthreads();
    const int rows = m - 1 - k;
    for (int e = threadIdx.x; e < rows * n; e += T) {
      const int i = k + 1 + e % rows, j = e / rows;
      const uint32_t bk = B[k + (size_t)j * ldb];  // nonzero stays nonzero after the division
      if (bk != 0) B[i + (size_t)j * ldb] = sub(B[i + (size_t)j * ldb], mul(bk, L[i + (size_t)k * ldl]));
    }
    __syncthreads();
  }
}

// B := U^-1 B, U non-unit upper (Rtrsm.cpp:120-137)
template <int T = NT>
static __device__ __noinline__ void trsm_lunn(int m, int n, const uint32_t* U, int ldu, uint32_t* B, int ldb) {
  for (int k = m - 1; k >= 0; --k) {
    const uint32_t ukk = U[k + (size_t)k * ldu];
    for (int j = threadIdx.x; j < n; j += T) {
      const uint32_t bk = B[k + (size_t)j * ldb];
      if (bk != 0) B[k + (size_t)j * ldb] = div(bk, ukk);
    }
    __syncthreads();
    for (int e = threadIdx.x; e < k * n; e += T) {
      const int i = e % k, j = e / k;
      const uint32_t bk = B[k + (size_t)j * ldb];
      if (bk != 0) B[i + (size_t)j * ldb] = sub(B[i + (size_t)j * ldb], mul(bk, U[i + (size_t)k * ldu]));
    }
    __syncthreads();
  }
}

// B := U^-T B, U non-unit upper (Rtrsm.cpp:161-173): row i is b_i minus a_ki*b_k for
// k = 0..i-1 in ascending order, then divided by u_ii; no zero skips in this form
template <int T = NT>
static __device__ __noinline__ void trsm_lutn(int m, int n, const uint32_t* U, int ldu, uint32_t* B, int ldb) {
  for (int k = 0; k < m; ++k) {
    const uint32_t ukk = U[k + (size_t)k * ldu];
    for (int j = threadIdx.x; j < n; j += T) B[k + (size_t)j * ldb] = div(B[k + (size_t)j * ldb], ukk);
    __syncthreads();
    const int rows = m - 1 - k;
    for (int e = threadIdx.x; e < rows * n; e += T) {
      const int i = k + 1 + e % rows, j = e / rows;
      B[i + (size_t)j * ldb] = sub(B[i + (size_t)j * ldb], mul(U[k + (size_t)i * ldu], B[k + (size_t)j * ldb]));
    }
    __syncthreads();
  }
}

// B := L^-T B, L non-unit lower (Rtrsm.cpp:174-186): row i needs b_{i+1} for its first
// term, so each column is one chain; one lane per column
template <int T = NT>
static __device__ __noinline__ void trsm_lltn(int m, int n, const uint32_t* L, int ldl, uint32_t* B, int ldb) {
  for (int j = threadIdx.x; j < n; j += T) {
    uint32_t* b = B + (size_t)j * ldb;
    for (int i = m - 1; i >= 0; --i) {
      uint32_t t = b[i];
      for (int k = i + 1; k < m; ++k) t = sub(t, mul(L[k + (size_t)i * ldl], b[k]));
      b[i] = div(t, L[i + (size_t)i * ldl]);
    }
  }
  __syncthreads();
}

// B := B L^-T, B m x n, L non-unit lower n x n (Rtrsm.cpp:263-285)
template <int T = NT>
static __device__ __noinline__ void trsm_rltn(int m, int n, const uint32_t* L, int ldl, uint32_t* B, int ldb) {
  for (int k = 0; k < n; ++k) {
    const uint32_t r = div(P_ONE, L[k + (size_t)k * ldl]);
    for (int i = threadIdx.x; i < m; i += T) B[i + (size_t)k * ldb] = mul(r, B[i + (size_t)k * ldb]);
    __syncthreads();
    const int cols = n - 1 - k;
    for (int e = threadIdx.x; e < m * cols; e += T) {
      const int i = e % m, j = k + 1 + e / m;
      const uint32_t t = L[j + (size_t)k * ldl];
      if (t != 0) B[i + (size_t)j * ldb] = sub(B[i + (size_t)j * ldb], mul(t, B[i + (size_t)k * ldb]));
    }
    __syncthreads();
  }
}

// Left solves of up to 64 rows, one wave per right-hand side: lane r holds b_r as an
// FP64 value and each step's b_k reaches the other lanes by a shuffle, so the steps need
// no barrier.  The operations and order of trsm_llnu, trsm_llnn and trsm_lutn: the lower
// forms take k ascending, a zero b_k skipping its step (non-unit: b_k divided by l_kk
// before its terms); the transposed upper form takes k ascending with no skip, b_k
// divided by u_kk once its terms are in.  Products and sums are rounded as the
// pattern-level operations round them (fp64_mac_arith, fp64_posit_div) while every value
// stays in scales -108 .. 107; a NaN in any lane (out of range, NaR) makes the wave redo
// its column at the pattern level.  A (m x m) is staged in sA (LDS, m * m words); every
// lane of the block takes part.
enum WaveSolve { WS_LLNU, WS_LLNN, WS_LUTN };

template <int MODE, int T = NT>
static __device__ __noinline__ void trsm_waves(int m, int n, const uint32_t* A, int lda, uint32_t* B, int ldb,
                                               lds_u32* sA) {
  for (int e = threadIdx.x; e < m * m; e += T) sA[e] = A[e % m + (size_t)(e / m) * lda];
  __syncthreads();
  const int lane = threadIdx.x % 64, wave = threadIdx.x / 64;
  auto a = [&](int r, int c) -> uint32_t { return sA[r + c * m]; };
  for (int j = wave; j < n; j += T / 64) {
    uint32_t* bj = B + (size_t)j * ldb;
    const uint32_t b0 = lane < m ? bj[lane] : 0u;
    double b = to_double(b0);
    for (int k = 0; k < m; ++k) {
      if (MODE == WS_LUTN) {
        if (lane == k) b = fp64_posit_div(b, to_double(a(k, k)));
        const double bk = __shfl(b, k);
        if (lane > k && lane < m) b = fp64_mac_arith(b, -to_double(a(k, lane)), bk);
      } else {
        double bk = __shfl(b, k);
        if (bk == 0.0) continue;  // the same value in every lane
        if (MODE == WS_LLNN) bk = fp64_posit_div(bk, to_double(a(k, k)));
        if (lane > k && lane < m) b = fp64_mac_arith(b, -bk, to_double(a(lane, k)));
        if (lane == k) b = bk;
      }
    }
    if (!__any(b != b)) {
      if (lane < m) bj[lane] = from_double(b);
      continue;
    }
    uint32_t p = b0;  // the column again at the pattern level
    for (int k = 0; k < m; ++k) {
      if (MODE == WS_LUTN) {
        if (lane == k) p = div(p, a(k, k));
        const uint32_t pk = __shfl(p, k);
        if (lane > k && lane < m) p = sub(p, mul(a(k, lane), pk));
      } else {
        uint32_t pk = __shfl(p, k);
        if (pk == 0) continue;
        if (MODE == WS_LLNN) pk = div(pk, a(k, k));
        if (lane > k && lane < m) p = sub(p, mul(pk, a(lane, k)));
        if (lane == k) p = pk;
      }
    }
    if (lane < m) bj[lane] = p;
  }
  __syncthreads();
}

// C right-hand sides of n rows solved together by all T lanes of the block in FP64, the
// forms of trsm_llnu, trsm_llnn, trsm_lunn and trsm_lutn: row i of column c is held by
// lane i % T in register slot c * (COL_SOLVE_SLOTS / C) + i / T.  The sweep goes 64 rows
// at a time: the wave holding a block's rows runs the block's steps with shuffles (as
// trsm_waves does) and leaves the block's final values in LDS (sb: C x 128 values, two
// blocks alternating); after one barrier every lane applies the block's terms to its rows
// beyond the block.  The C columns share each factor entry's decoding, the block's wave
// and the barriers.  Every row so
// receives its terms in the pattern-level sweep's k order (ascending; descending for
// the upper form), zero values of b_k skipping their terms in the same places, each
// term rounded as the pattern-level operations round it (gemm::fp64_mac, its constants
// read from gemm::round_table, which the calling kernel fills first; fp64_posit_div).
// col holds the C columns as patterns (LDS, column c at col + c * ldc) and receives the
// solutions; when some value leaves scales -108 .. 107 or meets a NaR or a zero divisor
// (a NaN in some lane) the routine returns false with col untouched, and the caller runs
// the pattern-level sweep.  n <= COL_SOLVE_SLOTS / C * T.
enum ColSolve { CS_LLNU, CS_LLNN, CS_LUNN, CS_LUTN };
constexpr int COL_SOLVE_SLOTS = 16;
typedef __attribute__((address_space(3))) double lds_f64;
// the slots as one register vector: a slot chosen by a wave-uniform index is a register
// move, not a memory access
typedef double col_slots __attribute__((ext_vector_type(COL_SOLVE_SLOTS)));

// x of lane k, k the same in every lane
static __device__ __forceinline__ double lane_value(double x, int k) {
  const uint64_t u = __builtin_bit_cast(uint64_t, x);
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)u, k);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(u >> 32), k);
  return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}

// b / a at the pattern level, for quotients that land on a posit midpoint in FP64
static __device__ __noinline__ double col_div_exact(double b, double a) {
  return to_double(div(from_double(b), from_double(a)));
}

// fp64_posit_div with its rare exact case out of line
static __device__ __forceinline__ double col_div(double b, double a) {
  const double q = b / a;
  const double c = fp64_grid_c_arith(q);
  const double r = (q + c) - c;
  if (__builtin_fabs(q - r) == fp64_grid_half(c)) return col_div_exact(b, a);
  return r;
}

template <int MODE, int T, int C = 1>
static __device__ __noinline__ bool col_solve_fp64(int n, const uint32_t* __restrict__ Ag, int lda, lds_u32* col,
                                                   int ldc, lds_f64* sb) {
  constexpr int S = COL_SOLVE_SLOTS / C;  // slots of one column; column c holds b[c * S + s]
  constexpr bool BACK = MODE == CS_LUNN;
  static_assert(T % 64 == 0 && COL_SOLVE_SLOTS % C == 0, "whole waves, whole slots");
  typedef const __attribute__((address_space(1))) uint32_t gbl_u32;  // global loads, counted apart from LDS
  gbl_u32* A = (gbl_u32*)Ag;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  col_slots b;
#pragma unroll
  for (int c = 0; c < C; ++c)
#pragma unroll
    for (int s = 0; s < S; ++s) {
      const int i = s * T + tid;
      b[c * S + s] = i < n ? to_double(col[c * ldc + i]) : 0.0;
    }
  // a(r, c) of the sweep: LUTN reads u_kr (row k of U), the others l_rk / u_rk
  auto entry = [&](int r, int c) -> uint32_t {
    return MODE == CS_LUTN ? A[c + (size_t)r * lda] : A[r + (size_t)c * lda];
  };
  const int nblk = (n + 63) / 64;
  for (int q = 0; q < nblk; ++q) {
    const int kb = BACK ? nblk - 1 - q : q;
    const int k0 = kb * 64, m = min(64, n - k0);
    lds_f64* sk = sb + (q & 1) * 64;  // column c's values at sk[c * 128 + r]
    if (wave == (k0 % T) / 64) {  // the block's own steps
      const int hs = k0 / T;
      double x[C];
#pragma unroll
      for (int c = 0; c < C; ++c) x[c] = b[c * S + hs];
      const bool row_in = lane < m;
      for (int h = 0; h < 4; ++h) {
        const int kc = BACK ? 48 - 16 * h : 16 * h;  // 16 steps' entries at a time
        uint32_t av[16];
#pragma unroll
        for (int j = 0; j < 16; ++j) av[j] = (row_in && kc + j < m) ? entry(k0 + lane, k0 + kc + j) : 0u;
#pragma unroll
        for (int jj = 0; jj < 16; ++jj) {
          const int j = BACK ? 15 - jj : jj, k = kc + j;
          if (k >= m) continue;  // uniform
          const double a = to_double(av[j]);  // decoded once for the C columns
          const double akk = MODE == CS_LLNU ? 1.0 : lane_value(a, k);
#pragma unroll
          for (int c = 0; c < C; ++c) {
            if (MODE == CS_LUTN) {
              if (lane == k) x[c] = col_div(x[c], a);
              const double bk = lane_value(x[c], k);
              if (lane > k && row_in) x[c] = gemm::fp64_mac(x[c], -a, bk);
            } else {
              double bk = lane_value(x[c], k);
              if (bk != 0.0) {  // the same value in every lane; zero skips the step
                if (MODE != CS_LLNU) bk = col_div(bk, akk);
                if ((BACK ? lane < k : lane > k) && row_in) x[c] = gemm::fp64_mac(x[c], -bk, a);
                if (lane == k) x[c] = bk;
              }
            }
          }
        }
      }
#pragma unroll
      for (int c = 0; c < C; ++c) {
        b[c * S + hs] = x[c];
        if (row_in) sk[c * 128 + lane] = x[c];
      }
    }
    __syncthreads();
    // the block's terms on the rows beyond it: [k0 + m, n) forward, [0, k0) upward
    const int lo = BACK ? 0 : k0 + m, hi = BACK ? k0 : n;
    if (lo >= hi) continue;
    for (int s = lo / T; s <= (hi - 1) / T; ++s) {
      const int i = s * T + tid;
      if (i < lo || i >= hi) continue;
      double x[C];
#pragma unroll
      for (int c = 0; c < C; ++c) x[c] = b[c * S + s];
      for (int kc = 0; kc < m; kc += 16) {
        uint32_t av[16];
#pragma unroll
        for (int j = 0; j < 16; ++j) {
          const int t = kc + j, k = BACK ? m - 1 - t : t;
          av[j] = t < m ? entry(i, k0 + k) : 0u;
        }
#pragma unroll
        for (int j = 0; j < 16; ++j) {
          const int t = kc + j;
          if (t >= m) continue;  // uniform
          const double a = to_double(av[j]);
          const int r = BACK ? m - 1 - t : t;
#pragma unroll
          for (int c = 0; c < C; ++c) {
            const double bk = sk[c * 128 + r];
            if (MODE == CS_LUTN) x[c] = gemm::fp64_mac(x[c], -a, bk);
            else if (bk != 0.0) x[c] = gemm::fp64_mac(x[c], -bk, a);
          }
        }
      }
#pragma unroll
      for (int c = 0; c < C; ++c) b[c * S + s] = x[c];
    }
  }
  bool bad = false;
#pragma unroll
  for (int c = 0; c < C; ++c)
#pragma unroll
    for (int s = 0; s < S; ++s)
      if (s * T + tid < n) bad |= b[c * S + s] != b[c * S + s];
  if (__syncthreads_or(bad)) return false;
#pragma unroll
  for (int c = 0; c < C; ++c)
#pragma unroll
    for (int s = 0; s < S; ++s) {
      const int i = s * T + tid;
      if (i < n) col[c * ldc + i] = from_double(b[c * S + s]);
    }
  __syncthreads();
  return true;
}

// ---------------------------------------------------------------------------------
// LU (Rgetrf_opencl / Rgetrf2)
// ---------------------------------------------------------------------------------
// First index of the largest |x_i| in signed pattern order, i < n (iRamax.cpp:31-88);
// x at that index is left in sh.pivot, read before the closing barrier.
static __device__ int iamax(int n, const uint32_t* x, Shared& sh) {
  const int tid = threadIdx.x;
  int bv = INT32_MIN, bi = INT32_MAX;  // bi == INT32_MAX: no element yet
  for (int i = tid; i < n; i += NT) {
    const int v = (int)abs_p(x[i]);
    if (bi == INT32_MAX || v > bv) {
      bv = v;
      bi = i;
    }
  }
  // the largest value wins, ties go to the smaller index: the sequential first maximum
  auto take = [&](int ov, int oi) {
    if (oi != INT32_MAX && (bi == INT32_MAX || ov > bv || (ov == bv && oi < bi))) {
      bv = ov;
      bi = oi;
    }
  };
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const int ov = __shfl_xor(bv, off), oi = __shfl_xor(bi, off);
    take(ov, oi);
  }
  if ((tid & 63) == 0) {
    sh.red_v[tid >> 6] = bv;
    sh.red_i[tid >> 6] = bi;
  }
  __syncthreads();
  bv = sh.red_v[0];
  bi = sh.red_i[0];
#pragma unroll
  for (int w = 1; w < NT / 64; ++w) take(sh.red_v[w], sh.red_i[w]);
  if (threadIdx.x == 0) sh.pivot = (bi == INT32_MAX) ? 0u : x[bi];
  __syncthreads();
  return bi;
}

// Leaf of Rgetrf2 for column j, rows j..n-1 (Rgetrf2.cpp:98-132)
// Rows are interchanged in columns [c_lo, c_hi) (default: all n); the caller applies
// the interchanges to the other columns later, in the same order.
static __device__ void lu_leaf(int n, uint32_t* a, int lda, int j, int* ipiv, int& info, Shared& sh, int c_lo = 0,
                               int c_hi = -1) {
  if (c_hi < 0) c_hi = n;
  uint32_t* col = at(a, lda, 0, j);
  const int p = j + iamax(n - j, col + j, sh);
  // not col[p]: the lane that swaps the pivot column may overwrite it before a later
  // wave would read it
  const uint32_t ap = sh.pivot;
  if (threadIdx.x == 0) ipiv[j] = p + 1;
  if (p != j) {
    // the pivot column itself is swapped only for a nonzero pivot (Rgetrf2.cpp:109-117);
    // every other column follows ipiv (the Rlaswp calls of Rgetrf2 / Rgetrf_opencl)
    for (int c = c_lo + threadIdx.x; c < c_hi; c += NT) {
      if (c == j && ap == 0) continue;
      uint32_t* x = at(a, lda, 0, c);
      const uint32_t t = x[j];
      x[j] = x[p];
      x[p] = t;
    }
  }
  __syncthreads();
  if (ap != 0) {
    // sfmin = FLT_MIN is below minpos, so every pivot but NaR takes the Rscal branch
    if (!lt(abs_p(ap), P_MINPOS)) {
      const uint32_t r = div(P_ONE, ap);
      for (int i = j + 1 + threadIdx.x; i < n; i += NT) col[i] = mul(r, col[i]);
    } else {
      for (int i = j + 1 + threadIdx.x; i < n; i += NT) col[i] = div(col[i], ap);
    }
  } else if (info == 0) {
    info = j + 1;
  }
  __syncthreads();
}

// The node of a Rgetrf2 / Rpotrf2 recursion on w columns (split n1 = len / 2) whose left
// half ends at local column jr; false if jr ends no left half.
static __device__ __forceinline__ bool node_after(int jr, int w, int& s, int& mid, int& e) {
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

// Rgetrf2 on columns p0..p0+w-1, rows p0..n-1 of the n x n matrix (rows >= columns at
// every node, so n1 = columns / 2); with panel_swaps_only the row interchanges reach only
// the panel's columns
static __device__ void lu_panel(int n, uint32_t* a, int lda, int p0, int w, int* ipiv, int& info, Shared& sh,
                                bool panel_swaps_only = false) {
  for (int jr = 0; jr < w; ++jr) {
    if (panel_swaps_only) lu_leaf(n, a, lda, p0 + jr, ipiv, info, sh, p0, p0 + w);
    else lu_leaf(n, a, lda, p0 + jr, ipiv, info, sh);
    int s, mid, e;
    if (node_after(jr, w, s, mid, e)) {
      // Rgetrf2.cpp:151-159: A12 := L11^-1 A12; A22 -= A21 A12
      const int n1 = mid - s, n2 = e - mid;
      uint32_t* a12 = at(a, lda, p0 + s, p0 + mid);
      trsm_llnu(n1, n2, at(a, lda, p0 + s, p0 + s), lda, a12, lda);
      gemm<false, false, false>(n - p0 - mid, n2, n1, at(a, lda, p0 + mid, p0 + s), lda, a12, lda,
                                at(a, lda, p0 + mid, p0 + mid), lda, 0, sh);
    }
  }
}

// Rgetrf_opencl on an n x n matrix with block size nb (Rgetrf_opencl.cpp:93-146)
static __device__ void getrf(int n, uint32_t* a, int lda, int* ipiv, int nb, int& info, Shared& sh) {
  if (nb <= 1 || nb >= n) {
    lu_panel(n, a, lda, 0, n, ipiv, info, sh);
    return;
  }
  for (int j = 0; j < n; j += nb) {
    const int jb = min(nb, n - j);
    lu_panel(n, a, lda, j, jb, ipiv, info, sh);
    if (j + jb < n) {
      uint32_t* a12 = at(a, lda, j, j + jb);
      trsm_llnu(jb, n - j - jb, at(a, lda, j, j), lda, a12, lda);
      gemm<false, false, false>(n - j - jb, n - j - jb, jb, at(a, lda, j + jb, j), lda, a12, lda,
                                at(a, lda, j + jb, j + jb), lda, 0, sh);
    }
  }
}

// Rgetrs 'N' (Rgetrs.cpp:86-100): Rlaswp on B, then L and U solves
static __device__ void getrs(int n, int nrhs, uint32_t* a, int lda, const int* ipiv, uint32_t* b, int ldb, Shared& sh) {
  // the interchanges b_i <-> b_ipiv(i), i ascending, composed into one gather
  if (threadIdx.x == 0) {
    for (int r = 0; r < n; ++r) sh.perm[r] = r;
    for (int i = 0; i < n; ++i) {
      const int p = ipiv[i] - 1;
      if (p != i && p >= 0 && p < n) {
        const int t = sh.perm[i];
        sh.perm[i] = sh.perm[p];
        sh.perm[p] = t;
      }
    }
  }
  __syncthreads();
  for (int c = 0; c < nrhs; ++c) {
    uint32_t* bc = b + (size_t)c * ldb;
    uint32_t v[MAX_N / NT];
#pragma unroll
    for (int q = 0; q < MAX_N / NT; ++q) {
      const int r = threadIdx.x + q * NT;
      if (r < n) v[q] = bc[sh.perm[r]];
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < MAX_N / NT; ++q) {
      const int r = threadIdx.x + q * NT;
      if (r < n) bc[r] = v[q];
    }
    __syncthreads();
  }
  trsm_llnu(n, nrhs, a, lda, b, ldb);
  trsm_lunn(n, nrhs, a, lda, b, ldb);
}

// ---------------------------------------------------------------------------------
// Cholesky (Rpotrf_opencl / Rpotrf2); a failing step returns its 1-based column and
// leaves the matrix as the reference leaves it
// ---------------------------------------------------------------------------------
// Diagonal blocks of up to BLOCK_MAX columns in FP64 (the per-call, batched and blocked
// Cholesky factorizations all start a block with these and fall back to the
// pattern-level recursion where they give up)
constexpr int BLOCK_MAX = 64;

// Lower Rpotrf2 of an n x n block (n <= 64) with the lower triangle held in registers,
// T lanes, packed column-major elements dealt round-robin.  The recursion
// (Rpotrf2.cpp:102-145) gives element (i, j), i > j, the terms -L(j,l)*L(i,l) for
// l = 0 .. j-1 in ascending order -- from the Rsyrk 'L','N' of every node that has both
// columns in its right half (l over that node's left half), then from the Rtrsm
// right/lower/transposed of the node that splits them (l from the node's start to j) --
// each skipped when the multiplier L(j,l) is zero (Rsyrk.cpp:169, Rtrsm.cpp:272), and
// then the scale by 1/L(j,j) (Rtrsm.cpp:266).  The diagonal gets the same terms from the
// Rsyrk calls alone, then the test and the sqrt of the leaf (Rpotrf2.cpp:90-97).  An
// Rsyrk term c + (-m)*x and an Rtrsm term c - m*x round alike (posit rounding is
// symmetric in sign), so the right-looking order below -- column j final, then its terms
// applied to every later column -- computes the same bits.  The terms run in the GEMM's
// FP64 multiply-add form with computed constants (fp64_mac_arith), each rounded as the
// pattern-level operations round it as long as every product and sum stays in scales
// -108 .. 107; the root and the reciprocal come from FP64 rounded to the posit grid
// (fp64_posit_sqrt, fp64_posit_recip: the pattern-level operation decides where that could
// round wrongly).  A result outside that range (NaN), a NaR, or a failing diagonal (the
// state the recursion leaves after a failing leaf is not this one) makes the call return
// -1 with the block unchanged; the caller then walks the recursion itself.  Returns 0
// with the block written back otherwise.
template <int T>
static __device__ int potrf2_lower_decoded(int n, uint32_t* a, int lda) {
  constexpr int SLOTS = (BLOCK_MAX * (BLOCK_MAX + 1) / 2 + T - 1) / T;
  __shared__ double col[BLOCK_MAX];
  __shared__ double dg;
  double c[SLOTS];  // posit values, exact in FP64 (NaR and out-of-range results: NaN)
  int ri[SLOTS], ci[SLOTS];
  {
    int j = 0, off = 0;
#pragma unroll
    for (int q = 0; q < SLOTS; ++q) {
      const int e = threadIdx.x + q * T;
      while (j < n && off + (n - j) <= e) {
        off += n - j;
        ++j;
      }
      ri[q] = j + (e - off);
      ci[q] = j < n ? j : -1;
      c[q] = to_double(j < n ? a[ri[q] + (size_t)j * lda] : 0u);
    }
  }
  if (threadIdx.x == 0) dg = to_double(a[0]);
  __syncthreads();
  for (int j = 0; j < n; ++j) {
    const double v = dg;
    if (!(v > 0.0)) return -1;  // uniform (zero, negative or NaR); the pattern-level form decides
    const double d = fp64_posit_sqrt(v);
    const double r = j + 1 < n ? fp64_posit_recip(d) : 1.0;
#pragma unroll
    for (int q = 0; q < SLOTS; ++q) {
      if (ci[q] == j) {
        if (ri[q] == j) {
          c[q] = d;
        } else {
          c[q] = fp64_posit_mul_arith(r, c[q]);
          col[ri[q]] = c[q];
        }
      }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < SLOTS; ++q) {
      if (ci[q] > j) {
        const double m = col[ci[q]];
        // a zero multiplier skips the term (Rsyrk.cpp:169, Rtrsm.cpp:272); a zero factor
        // makes it zero, which leaves the sum unchanged
        if (m != 0.0) c[q] = fp64_mac_arith(c[q], -m, col[ri[q]]);
        if (ri[q] == j + 1 && ci[q] == j + 1) dg = c[q];  // NaN: redo
      }
    }
    __syncthreads();
  }
  bool bad = false;
#pragma unroll
  for (int q = 0; q < SLOTS; ++q) bad |= c[q] != c[q];  // NaN: a scale left the table, or NaR
  if (__syncthreads_or(bad)) return -1;
#pragma unroll
  for (int q = 0; q < SLOTS; ++q)
    if (ci[q] >= 0) a[ri[q] + (size_t)ci[q] * lda] = from_double(c[q]);
  return 0;
}

// The upper Rpotrf2 recursion (Rpotrf2.cpp:102-125) on an n x n block (n <= 64) held in
// LDS as FP64 values, its upper triangle packed by columns (element (r, c), r <= c, at
// c * (c + 1) / 2 + r): leaf j takes the root of its diagonal element (fp64_posit_sqrt);
// the one node whose left half ends at j then solves its A12 (Rtrsm 'L','U','T','N',
// Rtrsm.cpp:161-173: ascending k, then the quotient by u_ii, fp64_posit_div) and updates
// the upper triangle of its A22 (Rsyrk 'U','T', Rsyrk.cpp:182-195: t = sum over the left
// half's rows l of a_li * a_lj, formed from zero in ascending l, then c - t).  Products and
// sums are rounded as in the GEMM (fp64_mac_arith).  Returns 0 with the block written
// back, or -1 with the block untouched where a value left scales -108 .. 107, a NaR
// appeared or a pivot is not positive: the pattern-level recursion then decides.
constexpr int UPPER_PACKED = BLOCK_MAX * (BLOCK_MAX + 1) / 2;

template <int T>
static __device__ int potrf2_upper_decoded(int n, uint32_t* a, int lda, double* S) {
  const int tid = threadIdx.x;
  auto at = [&](int r, int c) -> double& { return S[c * (c + 1) / 2 + r]; };
  for (int e = tid; e < n * n; e += T) {
    const int r = e % n, c = e / n;
    if (r <= c) at(r, c) = to_double(a[r + (size_t)c * lda]);
  }
  __syncthreads();
  __shared__ int ok;
  for (int j = 0; j < n; ++j) {
    if (tid == 0) {  // the leaf: test and root (Rpotrf2.cpp:90-97)
      const double v = at(j, j);
      ok = v > 0.0;  // zero, negative or NaR: the pattern-level form decides
      if (ok) at(j, j) = fp64_posit_sqrt(v);
    }
    __syncthreads();
    if (!ok) return -1;  // uniform (every leaf but the last is followed by its node's
                         // barriers before lane 0 writes ok again)
    int s, mid, e;
    if (node_after(j, n, s, mid, e)) {
      const int n2 = e - mid;
      // A12 := U11^-T A12, rows s .. mid-1: row k is final once it has its terms from rows
      // s .. k-1 and its quotient, so the lane that gives element (k+1, c) its last term
      // forms that quotient at once (one barrier a row)
      for (int c = mid + tid; c < e; c += T) at(s, c) = fp64_posit_div(at(s, c), at(s, s));
      __syncthreads();
      for (int k = s; k + 1 < mid; ++k) {
        const int cnt = mid - 1 - k;
        for (int q = tid; q < cnt * n2; q += T) {
          const int i = k + 1 + q % cnt, c = mid + q / cnt;
          const double x = fp64_mac_arith(at(i, c), -at(k, i), at(k, c));
          at(i, c) = i == k + 1 ? fp64_posit_div(x, at(i, i)) : x;
        }
        __syncthreads();
      }
      for (int q = tid; q < n2 * n2; q += T) {  // A22 -= A12^T A12, upper triangle
        const int i = mid + q % n2, jj = mid + q / n2;
        if (i <= jj) {
          double t = 0.0;
          for (int l = s; l < mid; ++l) t = fp64_mac_arith(t, at(l, i), at(l, jj));
          at(i, jj) = fp64_mac_arith(at(i, jj), -1.0, t);
        }
      }
      __syncthreads();
    }
  }
  bool bad = false;
  for (int e = tid; e < n * n; e += T) {
    const int r = e % n, c = e / n;
    if (r <= c) bad |= at(r, c) != at(r, c);  // NaN: a scale left the range, or NaR
  }
  if (__syncthreads_or(bad)) return -1;
  for (int e = tid; e < n * n; e += T) {
    const int r = e % n, c = e / n;
    if (r <= c) a[r + (size_t)c * lda] = from_double(at(r, c));
  }
  return 0;
}

template <bool UPPER>
static __device__ int potrf2(int n, uint32_t* a, int lda, Shared& sh) {
  if (n <= BLOCK_MAX) {  // the FP64 form first (the GEMM tile buffer holds its block)
    const int f = UPPER ? potrf2_upper_decoded<NT>(n, a, lda, reinterpret_cast<double*>(sh.tile))
                        : potrf2_lower_decoded<NT>(n, a, lda);
    if (f == 0) return 0;
    __syncthreads();
  }
  for (int j = 0; j < n; ++j) {
    const uint32_t v = a[j + (size_t)j * lda];
    if (le(v, 0)) return j + 1;  // Rpotrf2.cpp:90 (NaR is below zero in pattern order)
    __syncthreads();
    if (threadIdx.x == 0) a[j + (size_t)j * lda] = pga::sqrt(v);
    __syncthreads();
    int s, mid, e;
    if (node_after(j, n, s, mid, e)) {
      const int n1 = mid - s, n2 = e - mid;
      if (UPPER) {  // Rpotrf2.cpp:115-120
        uint32_t* a12 = at(a, lda, s, mid);
        trsm_lutn(n1, n2, at(a, lda, s, s), lda, a12, lda);
        gemm<true, false, false>(n2, n2, n1, a12, lda, a12, lda, at(a, lda, mid, mid), lda, 2, sh);
      } else {  // Rpotrf2.cpp:132-137
        uint32_t* a21 = at(a, lda, mid, s);
        trsm_rltn(n2, n1, at(a, lda, s, s), lda, a21, lda);
        gemm<false, true, true>(n2, n2, n1, a21, lda, a21, lda, at(a, lda, mid, mid), lda, 1, sh);
      }
    }
  }
  return 0;
}

template <bool UPPER>
static __device__ int potrf(int n, uint32_t* a, int lda, int nb, Shared& sh) {
  if (nb <= 1 || nb >= n) return potrf2<UPPER>(n, a, lda, sh);
  for (int j = 0; j < n; j += nb) {
    const int jb = min(nb, n - j);
    if (UPPER) {  // Rpotrf_opencl.cpp:100-123
      gemm<true, false, false>(jb, jb, j, at(a, lda, 0, j), lda, at(a, lda, 0, j), lda, at(a, lda, j, j), lda, 2,
                               sh);
      const int f = potrf2<true>(jb, at(a, lda, j, j), lda, sh);
      if (f) return f + j;
      if (j + jb < n) {
        gemm<true, false, false>(jb, n - j - jb, j, at(a, lda, 0, j), lda, at(a, lda, 0, j + jb), lda,
                                 at(a, lda, j, j + jb), lda, 0, sh);
        trsm_lutn(jb, n - j - jb, at(a, lda, j, j), lda, at(a, lda, j, j + jb), lda);
      }
    } else {  // Rpotrf_opencl.cpp:130-153
      gemm<false, true, true>(jb, jb, j, at(a, lda, j, 0), lda, at(a, lda, j, 0), lda, at(a, lda, j, j), lda, 1,
                              sh);
      const int f = potrf2<false>(jb, at(a, lda, j, j), lda, sh);
      if (f) return f + j;
      if (j + jb < n) {
        gemm<false, true, false>(n - j - jb, jb, j, at(a, lda, j + jb, 0), lda, at(a, lda, j, 0), lda,
                                 at(a, lda, j + jb, j), lda, 0, sh);
        trsm_rltn(n - j - jb, jb, at(a, lda, j, j), lda, at(a, lda, j + jb, j), lda);
      }
    }
  }
  return 0;
}

// Rpotrs (Rpotrs.cpp:86-104)
static __device__ void potrs(bool upper, int n, int nrhs, const uint32_t* a, int lda, uint32_t* b, int ldb) {
  if (upper) {
    trsm_lutn(n, nrhs, a, lda, b, ldb);
    trsm_lunn(n, nrhs, a, lda, b, ldb);
  } else {
    trsm_llnn(n, nrhs, a, lda, b, ldb);
    trsm_lltn(n, nrhs, a, lda, b, ldb);
  }
}

}  // namespace wg
}  // namespace pga
