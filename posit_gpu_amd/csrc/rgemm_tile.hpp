// rgemm_tile.hpp -- the posit32 GEMM tile body shared by every kernel that updates C
// in the reference's per-element order (plain, batched and per-problem task kernels).
#pragma once

#include <type_traits>

#include "rgemm_kernels.hpp"

namespace pga {
namespace gemm {

// ---------------------------------------------------------------------------------
// The multiply-add in FP64.  Posit values with regimes up to 27 bits long have at most
// 28 significant bits, so every operand and accumulator is exact in FP64.  A value x is
// rounded to the posit grid of its binade by y = (x + C) - C with C = 1.5 * 2^(e + 52 - nf)
// (nf = 27 - |k|' fraction bits at scale e, |k|' = k for k >= 0 and -k - 1 below), whose
// FP64 ulp is the posit quantum: round-to-nearest-even rounds to a multiple of it, ties
// to the even one, which is the even posit pattern (see fp64_posit_add); with the factor
// 1.5, x + C stays in C's binade for either sign of x (fp64_grid_chi).  C depends only on
// x's exponent field, so it comes from a table in LDS indexed by those 11 bits: scales
// -108 .. 107 get their C; zero gets 1.5 (zero stays zero); every other exponent gets an
// infinite C, which turns the result into NaN -- the slice is then redone exactly.
//   product: t*a (up to 56 bits) is rounded once, by fma(t, a, C) - C, with C taken from
//     the exponent of t*a rounded to FP64 (fp64_posit_mul).  The FMA keeps the exact
//     product, so no double rounding arises.
//   sum: c + y rounded to FP64 and then to the grid (fp64_posit_add: a sum that lands on
//     a midpoint was exact, so ties round as the posit addition does).
// ---------------------------------------------------------------------------------
// The table holds C itself (8 bytes, low word zero), so that one 64-bit LDS read yields
// the FP64 constant with no register moves; values of opposite signs read the same entry,
// so they never meet in one LDS bank at different addresses.
constexpr int ROUND_TABLE_N = 2048;
static __shared__ double round_table[ROUND_TABLE_N];

// Fills the table (every lane of the block takes part; a barrier must follow before use).
// Refilling while other waves read it is harmless: the entries never change.
PGA_D void round_table_fill(int tid, int nt) {
  for (int idx = tid; idx < ROUND_TABLE_N; idx += nt)
    round_table[idx] = __hiloint2double((int)fp64_grid_chi((uint32_t)idx << 20), 0);
}

PGA_D uint32_t hi_word(double x) { return (uint32_t)(__double_as_longlong(x) >> 32); }

// C for x: the entry of x's exponent field
PGA_D double grid_c(double x) { return round_table[(hi_word(x) >> 20) & 0x7FFu]; }

// y = x rounded to the posit grid (NaN outside the table's scales); chi is C's high word
PGA_D double grid_round(double x, uint32_t& chi) {
  const double c = grid_c(x);
  chi = hi_word(c);
  return (x + c) - c;
}

// round(t * a) on the posit grid, rounded once: fma(t, a, C) - C with C chosen from the
// binade of t*a rounded to FP64 (if that rounding carried into the next binade, t*a lies
// within 2^-52 of its power of two and both quanta round it there)
PGA_D double fp64_posit_mul(double t, double a) {
  const double c = grid_c(t * a);
  return __builtin_fma(t, a, c) - c;
}

// c <- round(c + round(t * a))
PGA_D double fp64_mac(double c, double t, double a) {
  uint32_t chi;
  return grid_round(c + fp64_posit_mul(t, a), chi);
}

// binary exponent of a finite nonzero double
PGA_D int fp64_scale(double x) { return (int)((hi_word(x) >> 20) & 0x7FFu) - 1023; }

// LDS needed by gemm_tile<BM, BN, BK, ...>, in 32-bit words, for either form of the tile:
// the integer form keeps three decoded planes (mantissa, scale, sign) of the BK x BM
// A-slice and of the BK x BN B-side slice, the FP64 form each element as one 8-byte
// value; both add the per-wave slice bounds.  The buffer must be 8-byte aligned.
template <int BM, int BN, int BK>
struct TileLds {
  static constexpr int words = 3 * BK * BM + 3 * BK * BN + 16 * 4;
};

// Tiles of several elements per lane always take the FP64 form (gemm_tile_impl), so a
// kernel that runs only such a tile needs only the FP64 slices: 2 words per element.
template <int TM, int TN>
constexpr bool tile_fp64_only() { return TM * TN >= 4; }
template <int BM, int BN, int BK, int TM, int TN>
struct TileLdsFor {
  static constexpr int words = tile_fp64_only<TM, TN>() ? 2 * BK * BM + 2 * BK * BN : TileLds<BM, BN, BK>::words;
};

// LDS of the kernels that cover a T x T region per workgroup with gemm_tile_in: T = 64 as
// two 64 x 32 tiles, T = 32 and 16 as one T x T tile (T / 16 elements per lane each way).
template <int T>
constexpr int square_tile_lds_words() {
  return T == 64 ? TileLdsFor<64, 32, 16, 4, 2>::words : TileLdsFor<T, T, 16, T / 16, T / 16>::words;
}
// Workgroups per CU such kernels leave room for: five with 64 x 32 tiles (28 KB of LDS
// with the rounding table, at most 96 VGPRs), smaller tiles fit more anyway.
template <int T>
constexpr int square_tile_wgs_per_cu() { return T == 64 ? 5 : 3; }

// The tile body with the integer decoded multiply-add (gemm::mac): used for one element
// per lane and short K, where the FP64 form would wait on two LDS table reads per step
// with nothing else to issue and fill its table for little work.
template <int BM, int BN, int BK, int TM, int TN, bool TA, bool TB, bool SKIP0>
__device__ __forceinline__ void gemm_tile_impl_int(int m, int n, int k, uint32_t alpha,
                                               const uint32_t* __restrict__ A, int lda,
                                               const uint32_t* __restrict__ B, int ldb, uint32_t beta,
                                               uint32_t* __restrict__ C, int ldc, int i0, int j0, int tri,
                                               uint32_t (*__restrict__ sAm)[BM], int (*__restrict__ sAs)[BM],
                                               uint32_t (*__restrict__ sAg)[BM], uint32_t (*__restrict__ sTm)[BN],
                                               int (*__restrict__ sTs)[BN], uint32_t (*__restrict__ sTg)[BN],
                                               int (*__restrict__ sBound)[4]) {
  constexpr int NX = BM / TM;
  constexpr int NY = BN / TN;
  constexpr int NT = NX * NY;
  constexpr int NW = NT / 64;
  constexpr bool SUM_FIRST = TA;  // TN and TT sum a*b from zero, then scale
  constexpr int U = dec::FAST_MAX - growth_bits(BK);
  static_assert((BM * BK) % NT == 0 && (BN * BK) % NT == 0, "tile loads");
  static_assert(NT % 64 == 0, "whole waves");
  static_assert(NW <= 16, "slice-bound slots in TileLds");

  const int tid = threadIdx.x;
  const int tx = tid % NX;
  const int ty = tid / NX;
  const int wave = tid / 64;

  // ---- accumulators ----
  Val acc[TM][TN];
#pragma unroll
  for (int r = 0; r < TM; ++r) {
#pragma unroll
    for (int c = 0; c < TN; ++c) {
      int i = i0 + tx + r * NX, j = j0 + ty + c * NY;
      uint32_t init = 0;
      if (!SUM_FIRST && beta != 0 && i < m && j < n) {
        uint32_t cij = C[i + (size_t)j * ldc];
        init = (beta == P_ONE) ? cij : pga::mul(beta, cij);
      }
      acc[r][c] = dec::decode(init);
    }
  }

  const bool alpha_one = (alpha == P_ONE);
  const bool alpha_mone = (alpha == P_MONE);

  for (int kt = 0; kt < k; kt += BK) {
    // ---- stage the A slice (BM x BK) and the B-side slice (BK x BN), decoded ----
    int amax = INT32_MIN, amin = INT32_MAX, tmax = INT32_MIN, tmin = INT32_MAX;
#pragma unroll
    for (int q = 0; q < (BM * BK) / NT; ++q) {
      int e = tid + q * NT;
      int ii, ll;
      if (!TA) {  // A is m x k: walk rows fastest (coalesced)
        ii = e % BM;
        ll = e / BM;
      } else {  // A is k x m: walk l fastest
        ll = e % BK;
        ii = e / BK;
      }
      int i = i0 + ii, l = kt + ll;
      uint32_t v = 0;
      if (i < m && l < k) v = TA ? A[l + (size_t)i * lda] : A[i + (size_t)l * lda];
      Val d = dec::decode(v);
      sAm[ll][ii] = d.m;
      sAs[ll][ii] = d.sc;
      sAg[ll][ii] = d.sg;
      amax = max(amax, d.sc);
      if (v != 0) amin = min(amin, d.sc);
    }
#pragma unroll
    for (int q = 0; q < (BN * BK) / NT; ++q) {
      int e = tid + q * NT;
      int jj, ll;
      if (!TB) {  // B is k x n: walk l fastest
        ll = e % BK;
        jj = e / BK;
      } else {  // B is n x k: walk columns fastest
        jj = e % BN;
        ll = e / BN;
      }
      int j = j0 + jj, l = kt + ll;
      uint32_t v = 0;
      if (j < n && l < k) v = TB ? B[j + (size_t)l * ldb] : B[l + (size_t)j * ldb];
      // alpha*b; with SKIP0 a zero b stays zero (it marks a skipped step even if alpha is NaR)
      if (!SUM_FIRST && !(SKIP0 && v == 0))
        v = alpha_one ? v : (alpha_mone ? pga::neg(v) : pga::mul(alpha, v));
      Val d = dec::decode(v);
      sTm[ll][jj] = d.m;
      sTs[ll][jj] = d.sc;
      sTg[ll][jj] = d.sg;
      tmax = max(tmax, d.sc);
      if (v != 0) tmin = min(tmin, d.sc);
    }
    // wave reductions of the slice bounds
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      amax = max(amax, __shfl_xor(amax, off));
      amin = min(amin, __shfl_xor(amin, off));
      tmax = max(tmax, __shfl_xor(tmax, off));
      tmin = min(tmin, __shfl_xor(tmin, off));
    }
    if ((tid & 63) == 0) {
      sBound[wave][0] = amax;
      sBound[wave][1] = amin;
      sBound[wave][2] = tmax;
      sBound[wave][3] = tmin;
    }
    __syncthreads();
#pragma unroll
    for (int w = 0; w < NW; ++w) {
      amax = max(amax, sBound[w][0]);
      amin = min(amin, sBound[w][1]);
      tmax = max(tmax, sBound[w][2]);
      tmin = min(tmin, sBound[w][3]);
    }
    // products of nonzero operands lie in [amin + tmin, amax + tmax + 1]
    bool fast = (amax + tmax + 1 <= U);
    if (amin != INT32_MAX && tmin != INT32_MAX) fast = fast && (amin + tmin >= dec::FAST_MIN);
    // accumulators of this lane (zeros have scales far below every bound)
    int cmax = INT32_MIN, cmin = INT32_MAX;
#pragma unroll
    for (int r = 0; r < TM; ++r)
#pragma unroll
      for (int c = 0; c < TN; ++c) {
        cmax = max(cmax, acc[r][c].sc + 1);
        if (acc[r][c].m != 0) cmin = min(cmin, acc[r][c].sc);
      }
    fast = fast && (cmax <= U);
    const bool wave_fast = __all(fast);
    // Every operand of the slice is a multiple of 2^g, g = (smallest nonzero scale) - 27
    // (a posit has at most 27 fraction bits), and sums and roundings keep that, so a
    // nonzero sum is at least 2^g: with g >= FAST_MIN no step can leave the fast range
    // downwards and the per-step check is dropped.
    const int pmin = (amin != INT32_MAX && tmin != INT32_MAX) ? amin + tmin : INT32_MAX;
    const bool wave_nocheck = __all(min(pmin, cmin) >= dec::FAST_MIN + 27);

    const int lmax = min(BK, k - kt);
    bool redo = !wave_fast;
    if (wave_fast) {
      Val saved[TM][TN];
#pragma unroll
      for (int r = 0; r < TM; ++r)
#pragma unroll
        for (int c = 0; c < TN; ++c) saved[r][c] = acc[r][c];
      bool low = false;
      auto step = [&](int l, auto check) {
        uint32_t am[TM], ag[TM], tm[TN], tg[TN];
        int as_[TM], ts[TN];
#pragma unroll
        for (int r = 0; r < TM; ++r) {
          am[r] = sAm[l][tx + r * NX];
          as_[r] = sAs[l][tx + r * NX];
          ag[r] = sAg[l][tx + r * NX];
        }
#pragma unroll
        for (int c = 0; c < TN; ++c) {
          tm[c] = sTm[l][ty + c * NY];
          ts[c] = sTs[l][ty + c * NY];
          tg[c] = sTg[l][ty + c * NY];
        }
#pragma unroll
        for (int r = 0; r < TM; ++r)
#pragma unroll
          for (int c = 0; c < TN; ++c)
            low |= mac<decltype(check)::value>(acc[r][c], tm[c], ts[c], tg[c], am[r], as_[r], ag[r]);
      };
      // two steps per iteration: the loop then carries no register rotation
      if (wave_nocheck) {
        int l = 0;
        for (; l + 1 < lmax; l += 2) {
          step(l, std::false_type());
          step(l + 1, std::false_type());
        }
        if (l < lmax) step(l, std::false_type());
      } else {
        int l = 0;
        for (; l + 1 < lmax; l += 2) {
          step(l, std::true_type());
          step(l + 1, std::true_type());
        }
        if (l < lmax) step(l, std::true_type());
      }
      if (__any(low)) {
        redo = true;
#pragma unroll
        for (int r = 0; r < TM; ++r)
#pragma unroll
          for (int c = 0; c < TN; ++c) acc[r][c] = saved[r][c];
      }
    }
    if (redo) {
      // exact path for this slice: pattern-level SoftPosit operations
#pragma unroll
      for (int r = 0; r < TM; ++r) {
#pragma unroll
        for (int c = 0; c < TN; ++c) {
          uint32_t pc = dec::encode(acc[r][c]);
          for (int l = 0; l < lmax; ++l) {
            Val a{sAm[l][tx + r * NX], sAs[l][tx + r * NX], sAg[l][tx + r * NX]};
            Val t{sTm[l][ty + c * NY], sTs[l][ty + c * NY], sTg[l][ty + c * NY]};
            uint32_t tp = dec::encode(t);
            if (SKIP0 && tp == 0) continue;  // Rsyrk.cpp:149,169 skip zero multipliers
            pc = pga::add(pc, pga::mul(tp, dec::encode(a)));
          }
          acc[r][c] = dec::decode(pc);
        }
      }
    }
    // keep zero accumulators at the canonical zero scale (sums of zeros drift downwards)
#pragma unroll
    for (int r = 0; r < TM; ++r)
#pragma unroll
      for (int c = 0; c < TN; ++c)
        if (acc[r][c].sc < dec::ZERO_LIMIT) acc[r][c].sc = dec::ZERO_SC;
    __syncthreads();
  }

  // ---- epilogue and store ----
#pragma unroll
  for (int r = 0; r < TM; ++r) {
#pragma unroll
    for (int c = 0; c < TN; ++c) {
      int i = i0 + tx + r * NX, j = j0 + ty + c * NY;
      bool keep = (tri == 0) || (tri == 1 ? i >= j : i <= j);
      if (i < m && j < n && keep) {
        uint32_t v = dec::encode(acc[r][c]);
        if (SUM_FIRST) {
          uint32_t at = pga::mul(alpha, v);
          v = (beta == 0) ? at : pga::add(at, pga::mul(beta, C[i + (size_t)j * ldc]));
        }
        C[i + (size_t)j * ldc] = v;
      }
    }
  }
}

// Operand panels (PANEL form): A and the B side (alpha*B for NN / NT, B for TN / TT) held
// as FP64 values in HBM, cut into the slices the tile stages, each slice contiguous in
// the tile's LDS order ([l][i] for A, [l][j] for the B side) and zero-padded at the edges:
// A's slice (i-block ib, k-slice s) starts at PA + (ib * ks + s) * BK * BM, the B side's
// (j-block jb, s) at PT + (jb * ks + s) * BK * BN, with ks = ceil(k / BK).  A GEMM whose A
// slice serves N / BN tiles and whose B slice serves M / BM tiles then converts each
// operand once instead of once per tile.
struct Panels {
  const double* a;
  const double* t;
  int ks;
};

// The tile body with the FP64 multiply-add (fp64_mac), for several elements per lane.
// No range checks run before a slice: a product or sum whose scale leaves -108 .. 107,
// and any NaR, reads an infinite rounding constant and turns into NaN, which stays NaN
// to the end of the slice; a wave that finds a NaN redoes the slice from its saved
// accumulators with the pattern-level operations.  With PANEL the slices are copied from
// pre-converted operand panels (pn) instead of decoded from A and B.
template <int BM, int BN, int BK, int TM, int TN, bool TA, bool TB, bool SKIP0, bool PANEL = false>
__device__ __forceinline__ void gemm_tile_impl_fp64(int m, int n, int k, uint32_t alpha,
                                               const uint32_t* __restrict__ A, int lda,
                                               const uint32_t* __restrict__ B, int ldb, uint32_t beta,
                                               uint32_t* __restrict__ C, int ldc, int i0, int j0, int tri,
                                               double (*__restrict__ sA)[BM], double (*__restrict__ sT)[BN],
                                               Panels pn = Panels{}) {
  constexpr int NX = BM / TM;
  constexpr int NY = BN / TN;
  constexpr int NT = NX * NY;
  constexpr bool SUM_FIRST = TA;  // TN and TT sum a*b from zero, then scale
  static_assert((BM * BK) % NT == 0 && (BN * BK) % NT == 0, "tile loads");
  static_assert(NT % 64 == 0, "whole waves");

  const int tid = threadIdx.x;
  const int tx = tid % NX;
  const int ty = tid / NX;

  round_table_fill(tid, NT);  // visible after the first slice's barrier

  // ---- accumulators: posit values in FP64 (NaR as NaN) ----
  double acc[TM][TN];
#pragma unroll
  for (int r = 0; r < TM; ++r) {
#pragma unroll
    for (int c = 0; c < TN; ++c) {
      int i = i0 + tx + r * NX, j = j0 + ty + c * NY;
      uint32_t init = 0;
      if (!SUM_FIRST && beta != 0 && i < m && j < n) {
        uint32_t cij = C[i + (size_t)j * ldc];
        init = (beta == P_ONE) ? cij : pga::mul(beta, cij);
      }
      acc[r][c] = pga::to_double(init);
    }
  }

  const bool alpha_one = (alpha == P_ONE);
  const bool alpha_mone = (alpha == P_MONE);

  // the slices' patterns: read from A and B (zero outside the matrices), then converted
  // into LDS with alpha applied to the B side
  // (measured: the 32 x 32 tile, one workgroup per CU on small GEMMs, gains; one element
  // per lane, LDS-bound, does not; larger tiles hide the loads behind other workgroups)
  constexpr bool PREFETCH = !PANEL && TM * TN == 4;
  constexpr int QA = (BM * BK) / NT, QT = (BN * BK) / NT;
  uint32_t ra[QA], rt[QT];
  auto a_pos = [&](int q, int& ii, int& ll) {
    const int e = tid + q * NT;
    if (!TA) {  // A is m x k: walk rows fastest (coalesced)
      ii = e % BM;
      ll = e / BM;
    } else {  // A is k x m: walk l fastest
      ll = e % BK;
      ii = e / BK;
    }
  };
  auto t_pos = [&](int q, int& jj, int& ll) {
    const int e = tid + q * NT;
    if (!TB) {  // B is k x n: walk l fastest
      ll = e % BK;
      jj = e / BK;
    } else {  // B is n x k: walk columns fastest
      jj = e % BN;
      ll = e / BN;
    }
  };
  auto slice_load = [&](int kt0) {
#pragma unroll
    for (int q = 0; q < QA; ++q) {
      int ii, ll;
      a_pos(q, ii, ll);
      const int i = i0 + ii, l = kt0 + ll;
      ra[q] = (i < m && l < k) ? (TA ? A[l + (size_t)i * lda] : A[i + (size_t)l * lda]) : 0u;
    }
#pragma unroll
    for (int q = 0; q < QT; ++q) {
      int jj, ll;
      t_pos(q, jj, ll);
      const int j = j0 + jj, l = kt0 + ll;
      rt[q] = (j < n && l < k) ? (TB ? B[j + (size_t)l * ldb] : B[l + (size_t)j * ldb]) : 0u;
    }
  };
  auto slice_store = [&]() {
#pragma unroll
    for (int q = 0; q < QA; ++q) {
      int ii, ll;
      a_pos(q, ii, ll);
      sA[ll][ii] = pga::to_double(ra[q]);  // exact; NaR -> NaN
    }
#pragma unroll
    for (int q = 0; q < QT; ++q) {
      int jj, ll;
      t_pos(q, jj, ll);
      uint32_t v = rt[q];
      // alpha*b; with SKIP0 a zero b stays zero (it marks a skipped step even if alpha is NaR)
      if (!SUM_FIRST && !(SKIP0 && v == 0))
        v = alpha_one ? v : (alpha_mone ? pga::neg(v) : pga::mul(alpha, v));
      sT[ll][jj] = pga::to_double(v);
    }
  };

  for (int kt = 0; kt < k; kt += BK) {
    if constexpr (PANEL) {
      // ---- copy the slices from the panels: 16-byte pieces, in the LDS order ----
      static_assert((BM * BK) % (2 * NT) == 0 && (BN * BK) % (2 * NT) == 0, "panel copies");
      const double2* pa = reinterpret_cast<const double2*>(pn.a + ((size_t)(i0 / BM) * pn.ks + kt / BK) * BK * BM);
      const double2* pt = reinterpret_cast<const double2*>(pn.t + ((size_t)(j0 / BN) * pn.ks + kt / BK) * BK * BN);
      double2* la = reinterpret_cast<double2*>(&sA[0][0]);
      double2* lt = reinterpret_cast<double2*>(&sT[0][0]);
      double2 va[(BM * BK) / (2 * NT)], vt[(BN * BK) / (2 * NT)];
#pragma unroll
      for (int q = 0; q < (BM * BK) / (2 * NT); ++q) va[q] = pa[tid + q * NT];
#pragma unroll
      for (int q = 0; q < (BN * BK) / (2 * NT); ++q) vt[q] = pt[tid + q * NT];
#pragma unroll
      for (int q = 0; q < (BM * BK) / (2 * NT); ++q) la[tid + q * NT] = va[q];
#pragma unroll
      for (int q = 0; q < (BN * BK) / (2 * NT); ++q) lt[tid + q * NT] = vt[q];
      __syncthreads();
    } else {
      // ---- stage the A slice (BM x BK) and the B-side slice (BK x BN) as FP64 values ----
      // With PREFETCH the next slice's patterns are read before this one is computed, so
      // that its loads are in flight meanwhile.
      if (!PREFETCH || kt == 0) slice_load(kt);
      slice_store();
      __syncthreads();
      if (PREFETCH && kt + BK < k) slice_load(kt + BK);
    }

    const int lmax = min(BK, k - kt);
    double saved[TM][TN];
#pragma unroll
    for (int r = 0; r < TM; ++r)
#pragma unroll
      for (int c = 0; c < TN; ++c) saved[r][c] = acc[r][c];
    // fp64_mac on every element, in phases: all products and their table reads, then
    // all roundings of the products and the sums, then all sums' table reads, then the
    // sums' roundings -- every table read of a phase is in flight before the first
    // result is used (the scheduling barriers keep the compiler from interleaving them
    // one multiply-add at a time)
    auto step = [&](int l) {
      double ad[TM], td[TN], y[TM][TN], cc[TM][TN];
#pragma unroll
      for (int r = 0; r < TM; ++r) ad[r] = sA[l][tx + r * NX];
#pragma unroll
      for (int c = 0; c < TN; ++c) td[c] = sT[l][ty + c * NY];
#pragma unroll
      for (int r = 0; r < TM; ++r)
#pragma unroll
        for (int c = 0; c < TN; ++c) cc[r][c] = grid_c(td[c] * ad[r]);
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int r = 0; r < TM; ++r)
#pragma unroll
        for (int c = 0; c < TN; ++c) {
          y[r][c] = acc[r][c] + (__builtin_fma(td[c], ad[r], cc[r][c]) - cc[r][c]);
          cc[r][c] = grid_c(y[r][c]);
        }
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int r = 0; r < TM; ++r)
#pragma unroll
        for (int c = 0; c < TN; ++c) acc[r][c] = (y[r][c] + cc[r][c]) - cc[r][c];
    };
    // two steps per iteration: the loop then carries no register rotation
    int l = 0;
    for (; l + 1 < lmax; l += 2) {
      step(l);
      step(l + 1);
    }
    if (l < lmax) step(l);
    bool bad = false;
#pragma unroll
    for (int r = 0; r < TM; ++r)
#pragma unroll
      for (int c = 0; c < TN; ++c) bad |= acc[r][c] != acc[r][c];  // NaN: a scale left the table, or NaR
    if (__any(bad)) {
      // exact path for this slice: pattern-level SoftPosit operations
#pragma unroll
      for (int r = 0; r < TM; ++r) {
#pragma unroll
        for (int c = 0; c < TN; ++c) {
          uint32_t pc = pga::from_double(saved[r][c]);  // exact; NaN -> NaR
          for (int l2 = 0; l2 < lmax; ++l2) {
            const uint32_t ap = pga::from_double(sA[l2][tx + r * NX]);
            const uint32_t tp = pga::from_double(sT[l2][ty + c * NY]);
            if (SKIP0 && tp == 0) continue;  // Rsyrk.cpp:149,169 skip zero multipliers
            pc = pga::add(pc, pga::mul(tp, ap));
          }
          acc[r][c] = pga::to_double(pc);
        }
      }
    }
    __syncthreads();
  }

  // ---- epilogue and store ----
#pragma unroll
  for (int r = 0; r < TM; ++r) {
#pragma unroll
    for (int c = 0; c < TN; ++c) {
      int i = i0 + tx + r * NX, j = j0 + ty + c * NY;
      bool keep = (tri == 0) || (tri == 1 ? i >= j : i <= j);
      if (i < m && j < n && keep) {
        uint32_t v = pga::from_double(acc[r][c]);
        if (SUM_FIRST) {
          uint32_t at = pga::mul(alpha, v);
          v = (beta == 0) ? at : pga::add(at, pga::mul(beta, C[i + (size_t)j * ldc]));
        }
        C[i + (size_t)j * ldc] = v;
      }
    }
  }
}

// One element of the panels: A's slice element (ll, ii) of i-block ib, k-slice s, or the
// B side's (ll, jj) of j-block jb -- zero outside the matrix, alpha applied as the tile
// applies it.
template <int BM, int BK, bool TA>
__device__ __forceinline__ double panel_a_elem(int m, int k, const uint32_t* __restrict__ A, int lda, int ib, int s,
                                               int ll, int ii) {
  const int i = ib * BM + ii, l = s * BK + ll;
  uint32_t v = 0;
  if (i < m && l < k) v = TA ? A[l + (size_t)i * lda] : A[i + (size_t)l * lda];
  return pga::to_double(v);
}
template <int BN, int BK, bool TB, bool SUM_FIRST, bool SKIP0>
__device__ __forceinline__ double panel_t_elem(int n, int k, uint32_t alpha, const uint32_t* __restrict__ B, int ldb,
                                               int jb, int s, int ll, int jj) {
  const int j = jb * BN + jj, l = s * BK + ll;
  uint32_t v = 0;
  if (j < n && l < k) v = TB ? B[j + (size_t)l * ldb] : B[l + (size_t)j * ldb];
  if (!SUM_FIRST && !(SKIP0 && v == 0))
    v = (alpha == P_ONE) ? v : (alpha == P_MONE ? pga::neg(v) : pga::mul(alpha, v));
  return pga::to_double(v);
}

// One BM x BN tile of C at (i0, j0): the body shared by the plain and the batched kernel.
// Both forms give every element the reference's bits.  `lds` holds TileLds<BM, BN, BK>
// words, 8-byte aligned; each form lays out its slices there.
template <int BM, int BN, int BK, int TM, int TN, bool TA, bool TB, bool SKIP0>
__device__ __forceinline__ void gemm_tile_impl(int m, int n, int k, uint32_t alpha,
                                               const uint32_t* __restrict__ A, int lda,
                                               const uint32_t* __restrict__ B, int ldb, uint32_t beta,
                                               uint32_t* __restrict__ C, int ldc, int i0, int j0, int tri,
                                               uint32_t* lds) {
  // one element per lane: the FP64 form pays for its table and its LDS round trips only
  // over long K loops (a 512^3 GEMM: 1.45x the integer form; the LU panel's node updates,
  // K <= 128: 0.35x)
  if (tile_fp64_only<TM, TN>() || k >= 256) {
    double* d = reinterpret_cast<double*>(lds);
    gemm_tile_impl_fp64<BM, BN, BK, TM, TN, TA, TB, SKIP0>(
        m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, i0, j0, tri, reinterpret_cast<double(*)[BM]>(d),
        reinterpret_cast<double(*)[BN]>(d + BK * BM));
  } else if constexpr (!tile_fp64_only<TM, TN>()) {  // (not even instantiated for such tiles)
    gemm_tile_impl_int<BM, BN, BK, TM, TN, TA, TB, SKIP0>(
        m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, i0, j0, tri, reinterpret_cast<uint32_t(*)[BM]>(lds),
        reinterpret_cast<int(*)[BM]>(lds + BK * BM), reinterpret_cast<uint32_t(*)[BM]>(lds + 2 * BK * BM),
        reinterpret_cast<uint32_t(*)[BN]>(lds + 3 * BK * BM), reinterpret_cast<int(*)[BN]>(lds + 3 * BK * BM + BK * BN),
        reinterpret_cast<uint32_t(*)[BN]>(lds + 3 * BK * BM + 2 * BK * BN),
        reinterpret_cast<int(*)[4]>(lds + 3 * BK * BM + 3 * BK * BN));
  }
}

// The tile with its own LDS buffer, for kernels that run one tile shape.
template <int BM, int BN, int BK, int TM, int TN, bool TA, bool TB, bool SKIP0>
__device__ __forceinline__ void gemm_tile(int m, int n, int k, uint32_t alpha, const uint32_t* __restrict__ A,
                                          int lda, const uint32_t* __restrict__ B, int ldb, uint32_t beta,
                                          uint32_t* __restrict__ C, int ldc, int i0, int j0, int tri) {
  __shared__ __align__(16) uint32_t lds[TileLdsFor<BM, BN, BK, TM, TN>::words];
  gemm_tile_impl<BM, BN, BK, TM, TN, TA, TB, SKIP0>(m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, i0, j0, tri, lds);
}

// The tile in a caller's buffer of TileLds<BM, BN, BK>::words words (8-byte aligned), for
// kernels that run several tile shapes out of one LDS allocation.
template <int BM, int BN, int BK, int TM, int TN, bool TA, bool TB, bool SKIP0>
__device__ __forceinline__ void gemm_tile_in(int m, int n, int k, uint32_t alpha, const uint32_t* __restrict__ A,
                                             int lda, const uint32_t* __restrict__ B, int ldb, uint32_t beta,
                                             uint32_t* __restrict__ C, int ldc, int i0, int j0, int tri,
                                             uint32_t* lds) {
  gemm_tile_impl<BM, BN, BK, TM, TN, TA, TB, SKIP0>(m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, i0, j0, tri, lds);
}

// FP64 operand panels of op(A) and alpha * op(B) for the 64 x 32 tile, from the library's
// memory pool on stream s (rgemm.hip); *ws receives the workspace to hand back with
// hipFreeAsync on s once the kernels that read the panels are enqueued.  hipSuccess with
// *ws == nullptr: no pool or no room (the caller decodes its slices itself).
hipError_t panels_build(bool ta, bool tb, bool skip0, int m, int n, int k, uint32_t alpha, const uint32_t* A, int lda,
                        const uint32_t* B, int ldb, hipStream_t s, Panels* pn, void** ws);

}  // namespace gemm
}  // namespace pga
