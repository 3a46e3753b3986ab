// rgemm_kernels.hpp -- posit32 GEMM kernels for gfx950 (integer VALU, no MFMA).
//
// Semantics (per C element, matching libRpack/Rgemm.cpp:62-251):
//   NN / NT (Rgemm.cpp:163-185, 206-227):
//     c = beta==0 ? 0 : (beta==1 ? C : beta*C);  for l ascending: c = c + (alpha*b_lj)*a_il
//   TN / TT (Rgemm.cpp:186-204, 228-246):
//     t = 0;  for l ascending: t = t + a_li*b_lj;  C = beta==0 ? alpha*t : alpha*t + beta*C
// with each posit operation rounded.  The K loop of every element runs in ascending
// order in one thread (no split-K), so any tiling gives the reference's bits.
//
// Tiling: a workgroup owns a BM x BN tile of C; each lane owns TM x TN elements
// (rows tx + r*BM/TM, columns ty + c*BN/TN).  For each BK-slice of K the A-slice and
// the B-side slice (alpha*B for NN/NT, B for TN/TT) are decoded once into LDS in the
// (mantissa, scale, sign) form, then every lane runs TM*TN independent multiply-add
// chains out of LDS.  The decoded multiply-add is exact for results whose regime is at
// most 27 bits long; a slice whose operands or accumulators could leave that range,
// and a wave in which some sum did leave it, is computed with the pattern-level
// SoftPosit operations instead.
#pragma once

#include "posit32_core.hpp"

namespace pga {
namespace gemm {

using dec::Val;

// Single-instruction helpers.  The compiler would otherwise expand min(x, 1) into a
// compare and a select, (1 << n) - 1 into a shift and a not, and a zero test around the
// leading-zero count.
PGA_D uint32_t ffbh_raw(uint32_t x) {  // leading zeros; all ones for x == 0
  uint32_t r;
  asm("v_ffbh_u32 %0, %1" : "=v"(r) : "v"(x));
  return r;
}
PGA_D uint32_t nonzero_bit(uint32_t x) {  // x != 0 ? 1 : 0
  uint32_t r;
  asm("v_min_u32 %0, 1, %1" : "=v"(r) : "v"(x));
  return r;
}
PGA_D uint32_t low_mask(uint32_t n) {  // (1 << n) - 1 for n in [0, 31]
  uint32_t r;
  asm("v_bfm_b32 %0, %1, 0" : "=v"(r) : "v"(n));
  return r;
}
PGA_D uint32_t clear_mask(uint32_t mask, uint32_t x) {  // x & ~mask
  uint32_t r;
  asm("v_bfi_b32 %0, %1, 0, %2" : "=v"(r) : "v"(mask), "v"(x));
  return r;
}

// Rounds y (hidden bit at 30; bits >= 2 exact, bits 1..0 sticky) to the precision of
// scale `sc`: 28 - reg fraction bits, round to nearest, ties to even.  Valid for
// FAST_MIN <= sc <= FAST_MAX; for other scales the result is unspecified (but
// computed without undefined behaviour).  May return exactly 2^31 on a carry.
PGA_HD uint32_t shift_of_scale(int sc) {  // dropped bits: reg + 2 (< 32 in range)
  int k = sc >> 2;
  return (uint32_t)((k ^ (k >> 31)) + 3);
}
PGA_D uint32_t round_shift(uint32_t y, uint32_t sh) {
  uint32_t low = low_mask(sh);                     // hardware uses sh mod 32
  uint32_t lsb = __builtin_amdgcn_ubfe(y, sh, 1);  // last kept bit (tie breaker)
  return clear_mask(low, y + (low >> 1) + lsb);
}
PGA_D uint32_t round_to_scale(uint32_t y, int sc) { return round_shift(y, shift_of_scale(sc)); }

// Full 64-bit product of two 32-bit values in one instruction (the compiler otherwise
// issues v_mul_hi_u32 and v_mul_lo_u32).
PGA_D uint64_t mul_wide(uint32_t a, uint32_t b) {
  uint64_t r, carry;
  asm("v_mad_u64_u32 %0, %1, %2, %3, 0" : "=v"(r), "=s"(carry) : "v"(a), "v"(b));
  return r;
}

// c <- round(c + round(t * a)) in the decoded domain.  Returns true when the sum fell
// below the fast range (a nonzero result with scale < FAST_MIN); the caller must then
// recompute.  Preconditions (guaranteed by the slice checks): t and a are finite,
// every nonzero product has scale in [FAST_MIN, FAST_MAX - 1] and the accumulator
// stays at or below FAST_MAX.
// With CHECK false the caller has proved that no nonzero sum can fall below the fast
// range (every operand of the slice is a multiple of 2^g with g >= FAST_MIN) and the
// flag is not computed.
template <bool CHECK = true>
PGA_D bool mac(Val& c, uint32_t tm, int ts, uint32_t tg, uint32_t am, int as, uint32_t ag) {
  // ---- product t * a ----
  const uint64_t prod = mul_wide(tm, am);  // mantissas in [2^30, 2^31): product in [2^60, 2^62)
  uint32_t hi = (uint32_t)(prod >> 32);
  uint32_t lo = (uint32_t)prod;
  uint32_t cy = hi >> 29;
  int ps = ts + as + (int)cy;
  uint32_t pm = (hi << (2u - cy)) | nonzero_bit(lo);  // hidden at 30, sticky in bit 0
  uint32_t pg = tg ^ ag;
  pm = round_to_scale(pm, ps);
  uint32_t ov = pm >> 31;  // renormalise a rounding carry
  pm >>= ov;
  ps += (int)ov;

  // ---- sum c + p ----
  // order by magnitude with one signed 64-bit compare of (scale, mantissa)
  int64_t kp = (int64_t)(((uint64_t)(uint32_t)ps << 32) | pm);
  int64_t kc = (int64_t)(((uint64_t)(uint32_t)c.sc << 32) | c.m);
  bool pbig = kp > kc;
  uint32_t mb = pbig ? pm : c.m;
  uint32_t ms = pbig ? c.m : pm;
  int sb = pbig ? ps : c.sc;
  int ss = pbig ? c.sc : ps;
  uint32_t gb = pbig ? pg : c.sg;
  uint32_t d = min((uint32_t)(sb - ss), 31u);
  uint32_t lost = __builtin_amdgcn_ubfe(ms, 0, d);
  uint32_t al = (ms >> d) | nonzero_bit(lost);
  const uint32_t sd = pg ^ c.sg;       // all ones when the signs differ
  uint32_t x = ((al ^ sd) + mb) - sd;  // mb - al or mb + al
  // leading zeros; a zero sum gets a huge count so its scale lands far below zero
  uint32_t lz = min(ffbh_raw(x), 1u << 21);
  int sn = sb + 1 - (int)lz;
  uint32_t z = x << (lz & 31u);
  uint32_t y = (z >> 1) | (z & 1u);
  c.m = round_to_scale(y, sn);
  c.sc = sn;
  c.sg = gb;
  // a nonzero result below the fast range (zeros are below ZERO_LIMIT)
  if (!CHECK) return false;
  return (uint32_t)(sn - dec::ZERO_LIMIT) < (uint32_t)(dec::FAST_MIN - dec::ZERO_LIMIT);
}

// c <- round(c + p) for decoded values, exact in every case: the decoded sum when both
// operands are finite and the result stays in the fast range, the pattern-level
// SoftPosit addition otherwise.
PGA_D void add_checked(Val& c, const Val& p) {
  if (max(c.sc, p.sc) < dec::FAST_MAX) {
    Val t = c;
    uint32_t mb, ms, gb;
    int64_t kp = (int64_t)(((uint64_t)(uint32_t)p.sc << 32) | p.m);
    int64_t kc = (int64_t)(((uint64_t)(uint32_t)c.sc << 32) | c.m);
    bool pbig = kp > kc;
    mb = pbig ? p.m : c.m;
    ms = pbig ? c.m : p.m;
    int sb = pbig ? p.sc : c.sc, ss = pbig ? c.sc : p.sc;
    gb = pbig ? p.sg : c.sg;
    uint32_t d = min((uint32_t)(sb - ss), 31u);
    uint32_t al = (ms >> d) | nonzero_bit(__builtin_amdgcn_ubfe(ms, 0, d));
    const uint32_t sd = p.sg ^ c.sg;
    uint32_t x = ((al ^ sd) + mb) - sd;
    uint32_t lz = min(ffbh_raw(x), 1u << 21);
    int sn = sb + 1 - (int)lz;
    uint32_t z = x << (lz & 31u);
    t.m = round_to_scale((z >> 1) | (z & 1u), sn);
    t.sc = sn;
    t.sg = gb;
    if ((uint32_t)(sn - dec::ZERO_LIMIT) >= (uint32_t)(dec::FAST_MIN - dec::ZERO_LIMIT)) {
      if (sn < dec::ZERO_LIMIT) t.sc = dec::ZERO_SC;
      c = t;
      return;
    }
  }
  c = dec::decode(pga::add(dec::encode(c), dec::encode(p)));
}

// c <- round(c + p) on the fast decoded path only: returns false unless the sum's scale
// lies in [FAST_MIN, FAST_MAX], where the decoded arithmetic rounds exactly.  Every other
// case -- a NaR operand (its scale carries the sum far above the range), a zero sum (far
// below), a result too large or too small to round here -- returns false with a
// meaningless c, and the caller redoes the step exactly.  A zero operand needs no case of
// its own: the other operand comes back unchanged.
PGA_D bool add_fast(Val& c, uint32_t pm, int ps, uint32_t pg) {
  int64_t kp = (int64_t)(((uint64_t)(uint32_t)ps << 32) | pm);
  int64_t kc = (int64_t)(((uint64_t)(uint32_t)c.sc << 32) | c.m);
  bool pbig = kp > kc;
  uint32_t mb = pbig ? pm : c.m;
  uint32_t ms = pbig ? c.m : pm;
  int sb = pbig ? ps : c.sc, ss = pbig ? c.sc : ps;
  uint32_t gb = pbig ? pg : c.sg;
  uint32_t d = min((uint32_t)(sb - ss), 31u);
  uint32_t al = (ms >> d) | nonzero_bit(__builtin_amdgcn_ubfe(ms, 0, d));
  const uint32_t sd = pg ^ c.sg;
  uint32_t x = ((al ^ sd) + mb) - sd;
  uint32_t lz = min(ffbh_raw(x), 1u << 21);
  int sn = sb + 1 - (int)lz;
  uint32_t z = x << (lz & 31u);
  c.m = round_to_scale((z >> 1) | (z & 1u), sn);
  c.sc = sn;
  c.sg = gb;
  return (uint32_t)(sn - dec::FAST_MIN) <= (uint32_t)(dec::FAST_MAX - dec::FAST_MIN);
}

// Largest accumulator / product scale for which BK more steps cannot pass FAST_MAX:
// each step at most adds a product and rounds up by a factor <= 1.25 (a result in the
// fast range keeps at least one fraction bit), so after BK steps
// |c| < 1.25^BK * (BK + 1) * 2^(U+1).
constexpr int growth_bits(int bk) {
  // ceil(log2(1.25^bk * (bk + 1))) + 1, evaluated for the slice depths used here
  return bk <= 8 ? 7 : (bk <= 16 ? 11 : 18);
}

}  // namespace gemm
}  // namespace pga
