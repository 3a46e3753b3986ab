// posit32_core.hpp -- posit32 (es = 2) arithmetic for gfx950 kernels and host code.
//
// Two layers:
//
//  * Pattern-level operations (pga::add, sub, mul, div, sqrt, from_double, ...):
//    bit-exact with the SoftPosit routines the reference uses.  add and mul follow the
//    reference's OpenCL port of SoftPosit (libRpack/kernel_entry.cl:88-538): the same
//    64-bit fraction frames, the same truncating alignment shift and the same rounding
//    tail, including its treatment of regimes 29 and 30.  Regimes are decoded with a
//    count-leading-zeros instead of the reference's bit-by-bit loops.  div, sqrt and the
//    double conversion are correctly rounded (round to nearest, ties to the even
//    pattern, saturating at minpos / maxpos).
//
//  * Decoded-domain operations (pga::dec): the representation the GEMM inner loop works
//    in.  A value is (m, sc, sg): mantissa with its hidden bit at bit 30 (the fraction
//    below it, the three lowest bits clear), scale sc = 4k + e, and sign sg (0 or bit
//    31).  For results whose regime is at most 27 bits long the SoftPosit routines are
//    correctly rounded on the posit bit string, so in that range a fused
//    multiply-then-add can be computed with 32-bit integer arithmetic carrying a guard,
//    round and sticky bit.  Callers must route every other case to the pattern-level
//    operations (the GEMM does so with range flags).
#pragma once

#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define PGA_HD __host__ __device__ __forceinline__
#define PGA_D __device__ __forceinline__
#else
#define PGA_HD inline
#define PGA_D inline
#endif

namespace pga {

constexpr uint32_t P_ZERO = 0x00000000u;
constexpr uint32_t P_ONE = 0x40000000u;
constexpr uint32_t P_MONE = 0xC0000000u;
constexpr uint32_t P_NAR = 0x80000000u;
constexpr uint32_t P_MAXPOS = 0x7FFFFFFFu;
constexpr uint32_t P_MINPOS = 0x00000001u;

PGA_HD int clz32(uint32_t x) { return x ? __builtin_clz(x) : 32; }
PGA_HD int clz64(uint64_t x) { return x ? __builtin_clzll(x) : 64; }
PGA_HD uint32_t neg(uint32_t a) { return 0u - a; }
PGA_HD bool lt(uint32_t a, uint32_t b) { return (int32_t)a < (int32_t)b; }
PGA_HD bool le(uint32_t a, uint32_t b) { return (int32_t)a <= (int32_t)b; }
PGA_HD uint32_t abs_p(uint32_t a) { return ((int32_t)a < 0) ? 0u - a : a; }

// Regime, exponent and fraction of a positive, nonzero, non-NaR pattern.  `frac31` is
// the fraction with its hidden bit at bit 30, the layout SoftPosit uses.
struct Parts {
  int scale;        // 4k + e
  uint32_t frac31;  // hidden bit at 30, fraction bits from 29 downwards
};

PGA_HD Parts parts(uint32_t u) {
  uint32_t x = u << 1;                            // regime starts at bit 31
  uint32_t ones = (uint32_t)((int32_t)x >> 31);   // all ones for a run of ones
  int run = clz32(x ^ ones);                      // 1 .. 31
  uint32_t after = (x << run) << 1;               // drop the run and its terminator
  int k = ones ? run - 1 : -run;
  Parts p;
  p.scale = 4 * k + (int)(after >> 30);
  p.frac31 = 0x40000000u | (after & 0x3FFFFFFFu);
  return p;
}

// The rounding tail shared by add and mul (kernel_entry.cl:167-213 / 487-533): `f`
// carries the magnitude with its hidden bit at 62; returns the positive pattern.
PGA_HD uint32_t round_tail(int scale, uint64_t f) {
  int k = scale >> 2;  // arithmetic shift: floor(scale / 4)
  uint32_t e = (uint32_t)(scale & 3);
  int reg = (k < 0) ? -k : k + 1;
  if (reg > 30) return (k >= 0) ? P_MAXPOS : P_MINPOS;
  uint32_t regime = (k >= 0) ? (0x7FFFFFFFu - (0x7FFFFFFFu >> reg)) : (0x40000000u >> reg);
  uint64_t w = (f & 0x3FFFFFFFFFFFFFFFull) >> (reg + 2);
  uint32_t frac = (uint32_t)(w >> 32);
  uint32_t guard;
  if (reg <= 28) {
    guard = (uint32_t)(w >> 31) & 1u;
    e <<= (28 - reg);
  } else {
    // regime 29 keeps one exponent bit, regime 30 none; the fraction bits are all
    // below the last kept bit
    guard = (reg == 30) ? (e >> 1) : (e & 1u);
    e = (reg == 30) ? 0u : (e >> 1);
    frac = 0;
  }
  uint32_t ui = regime + e + frac;
  if (guard) {
    // as in the reference, the sticky bit is taken from the fraction window only
    uint32_t sticky = (w & 0x7FFFFFFFull) != 0;
    ui += (ui & 1u) | sticky;
  }
  return ui;
}

// posit32 addition, bit-exact with kernel_entry.cl:370-397.
PGA_HD uint32_t add(uint32_t a, uint32_t b) {
  if (a == 0 || b == 0) return a | b;
  if (a == P_NAR || b == P_NAR) return P_NAR;
  bool negz = (a >> 31) != 0;
  bool subtract = ((a ^ b) >> 31) != 0;
  uint32_t ua = negz ? 0u - a : a;
  uint32_t ub = (negz != subtract) ? 0u - b : b;  // |b|
  if (subtract && ua == ub) return 0;
  if ((int32_t)ua < (int32_t)ub) {
    uint32_t t = ua;
    ua = ub;
    ub = t;
    if (subtract) negz = !negz;
  }
  Parts pa = parts(ua), pb = parts(ub);
  int shift = pa.scale - pb.scale;  // == 4(kA-kB) + eA - eB >= 0
  uint64_t fa = (uint64_t)pa.frac31 << 32;
  uint64_t fb = (uint64_t)pb.frac31 << 32;
  int scale = pa.scale;
  uint64_t f;
  if (!subtract) {
    f = fa + ((shift > 63) ? 0ull : (fb >> shift));
    if (f >> 63) {
      f >>= 1;
      scale += 1;
    }
  } else {
    if (shift > 63) return negz ? 0u - ua : ua;
    f = fa - (fb >> shift);
    int z = clz64(f) - 1;  // f != 0 here
    f <<= z;
    scale -= z;
  }
  uint32_t ui = round_tail(scale, f);
  return negz ? 0u - ui : ui;
}

PGA_HD uint32_t sub(uint32_t a, uint32_t b) {
  if (a == P_NAR || b == P_NAR) return P_NAR;
  return add(a, 0u - b);
}

// posit32 multiplication, bit-exact with kernel_entry.cl:399-538.
PGA_HD uint32_t mul(uint32_t a, uint32_t b) {
  if (a == P_NAR || b == P_NAR) return P_NAR;
  if (a == 0 || b == 0) return 0;
  bool negz = ((a ^ b) >> 31) != 0;
  Parts pa = parts((a >> 31) ? 0u - a : a);
  Parts pb = parts((b >> 31) ? 0u - b : b);
  int scale = pa.scale + pb.scale;
  uint64_t p = (uint64_t)pa.frac31 * (uint64_t)pb.frac31;  // hidden at 60 or 61
  if (p >> 61) {
    p >>= 1;
    scale += 1;
  }
  uint32_t ui = round_tail(scale, p << 2);
  return negz ? 0u - ui : ui;
}

// ---- correctly rounded encoder (div, sqrt, conversions) ----

// Round (-1)^negz * m * 2^(scale - hb) to the nearest posit32 on the bit string (ties to
// the even pattern); m has its leading one at bit hb (29 <= hb <= 63) and `sticky` says
// whether nonzero bits lie below m.  Saturates at minpos / maxpos.
PGA_HD uint32_t encode_rne(bool negz, int scale, uint64_t m, int hb, bool sticky) {
  int k = scale >> 2;
  uint32_t e = (uint32_t)(scale & 3);
  int reg = (k < 0) ? -k : k + 1;
  uint32_t ui;
  if (reg > 30) {
    ui = (k >= 0) ? P_MAXPOS : P_MINPOS;
  } else {
    // normalise m to the leading one at bit 63, then the bit string after the sign is
    // [regime: reg+1 bits][exponent: 2][fraction: 63 bits]; keep 31 bits.
    uint64_t f = m << (63 - hb);
    uint64_t fr = f << 1;  // fraction bits, top-aligned
    int head = reg + 1 + 2;  // 4 .. 33 bits of regime and exponent
    uint32_t regime = (k >= 0) ? (0x7FFFFFFFu - (0x7FFFFFFFu >> reg)) : (0x40000000u >> reg);
    uint32_t kept, guard;
    bool rest;
    if (head <= 31) {
      int nf = 31 - head;  // fraction bits that fit
      uint32_t fkeep = nf ? (uint32_t)(fr >> (64 - nf)) : 0u;
      kept = regime + (e << nf) + fkeep;
      uint64_t below = nf ? (fr << nf) : fr;
      guard = (uint32_t)(below >> 63);
      rest = ((below << 1) != 0) || sticky;
    } else if (head == 32) {  // reg == 29: one exponent bit kept
      kept = regime + (e >> 1);
      guard = e & 1u;
      rest = (fr != 0) || sticky;
    } else {  // reg == 30: no exponent bit kept
      kept = regime;
      guard = e >> 1;
      rest = (e & 1u) || (fr != 0) || sticky;
    }
    if (guard && (rest || (kept & 1u))) kept++;
    ui = kept;
  }
  return negz ? 0u - ui : ui;
}

// Exact value of a finite nonzero pattern: (-1)^negz * m27 * 2^(scale - 27) with the
// leading one of m27 at bit 27.
PGA_HD void unpack27(uint32_t a, bool& negz, int& scale, uint32_t& m27) {
  negz = (a >> 31) != 0;
  Parts p = parts(negz ? 0u - a : a);
  scale = p.scale;
  m27 = p.frac31 >> 3;
}

PGA_HD uint32_t div(uint32_t a, uint32_t b) {
  if (a == P_NAR || b == P_NAR || b == 0) return P_NAR;
  if (a == 0) return 0;
  bool na, nb;
  int sa, sb;
  uint32_t ma, mb;
  unpack27(a, na, sa, ma);
  unpack27(b, nb, sb, mb);
  // q = floor(ma * 2^36 / mb): the quotient from the FP64 division (num has at most 28
  // significant bits, so both operands convert exactly, and the rounded quotient is
  // within one of the true one), corrected against the exact remainder.  A 64-bit
  // integer division costs several times as many instructions on the GPU.
  uint64_t num = (uint64_t)ma << 36;
  uint64_t q = (uint64_t)((double)num / (double)mb);
  uint64_t p = q * mb;  // q <= 2^37 + 1 and mb < 2^28: no overflow
  while (p > num) {
    --q;
    p -= mb;
  }
  while (num - p >= mb) {
    ++q;
    p += mb;
  }
  uint64_t r = num - p;
  int hb = (q >> 36) ? 36 : 35;
  return encode_rne(na != nb, sa - sb - 36 + hb, q, hb, r != 0);
}

// floor(sqrt(x)) for x < 2^62 with at most 53 significant bits, plus exactness: the
// FP64 square root (x converts exactly; the result is within one of the true root),
// corrected with exact integer squares.
PGA_HD uint64_t isqrt64(uint64_t x, bool& exact) {
  uint64_t res = (uint64_t)__builtin_sqrt((double)x);
  while (res * res > x) --res;
  while ((res + 1) * (res + 1) <= x) ++res;
  exact = (res * res == x);
  return res;
}

PGA_HD uint32_t sqrt(uint32_t a) {
  if (a == 0) return 0;
  if (a >> 31) return P_NAR;  // NaR and negative values
  bool n_;
  int s;
  uint32_t m;
  unpack27(a, n_, s, m);
  int ex = s - 27;
  uint64_t mm = m;
  if (ex & 1) {
    mm <<= 1;
    ex -= 1;
  }
  // sqrt(mm * 2^32) has 30-31 bits; take it with the remainder as the sticky bit
  bool exact;
  uint64_t r = isqrt64(mm << 32, exact);
  int hb = 63 - clz64(r);
  return encode_rne(false, ex / 2 - 16 + hb, r, hb, !exact);
}

PGA_HD uint32_t from_double(double x) {
  if (x == 0.0) return 0;
  uint64_t bits;
  __builtin_memcpy(&bits, &x, sizeof bits);
  int bexp = (int)((bits >> 52) & 0x7FF);
  bool negz = (bits >> 63) != 0;
  if (bexp == 0x7FF) return P_NAR;  // inf, nan
  if (bexp == 0) return negz ? 0u - 1u : 1u;  // subnormal: far below minpos
  uint64_t m = (bits & ((1ull << 52) - 1)) | (1ull << 52);
  return encode_rne(negz, bexp - 1023, m, 52, false);
}

PGA_HD double to_double(uint32_t a) {
  if (a == 0) return 0.0;
  if (a == P_NAR) return __builtin_nan("");
  bool negz;
  int s;
  uint32_t m;
  unpack27(a, negz, s, m);
  // m * 2^(s-27), exact in double for |s| <= 120
  uint64_t bits = ((uint64_t)(s + 1023) << 52) | (((uint64_t)m << 25) & ((1ull << 52) - 1));
  double v;
  __builtin_memcpy(&v, &bits, sizeof v);
  return negz ? -v : v;
}

// round(t + p) for posit values t and p held exactly in FP64 (the posit sum, as
// pga::add rounds it), for the long serial addition chains of the triangular solves.
// x = t + p is first rounded to FP64; then x is rounded to the posit grid of its binade
// -- nf = 27 - |k|' fraction bits at scale e = 4k + (e mod 4), |k|' = k for k >= 0 and
// -k - 1 below -- by adding and subtracting C = +-2^(e + 52 - nf), whose FP64 ulp is the
// posit quantum q = 2^(e - nf): round-to-nearest-even then takes x to a multiple of q,
// ties to the even one, which is the even posit pattern (nf >= 1).
// The first rounding cannot change the result: posit midpoints are FP64 numbers and
// rounding is monotone, so x lies strictly between the same two midpoints as the exact
// sum, or on one -- and then the FP64 sum was exact.  (An inexact sum needs a term with
// bits more than 52 places below the result's leading bit; a term that brings the sum
// within that distance of a midpoint lies within it of a nonzero multiple of half the
// result's quantum, and at every scale a posit that large has no fraction bits that
// far down.  A search over 3e7 constructed midpoint sums found no inexact one.)
// The grid above holds for scales -108 .. 107 (|k|' <= 26: both exponent bits and a
// fraction bit kept); `reach` is raised to |k|' of the result, and a caller whose
// `reach` ends above 26 (a scale outside that range, or an exact zero) must redo the
// additions exactly.
PGA_HD double fp64_posit_add(double t, double p, uint32_t& reach) {
  const double x = t + p;
  uint64_t xb;
  __builtin_memcpy(&xb, &x, sizeof xb);
  const uint32_t hi = (uint32_t)(xb >> 32);
  const int e = (int)((hi >> 20) & 0x7FFu) - 1023;  // binary exponent of x
  // |k|' for the regime value k = e >> 2 (the sign of k is that of e: one step shorter)
  const uint32_t ak = (uint32_t)((e >> 2) ^ (e >> 31));
  reach = reach > ak ? reach : ak;
  // C's high word: x's sign and exponent field raised by 25 + |k|' (no carry out of the
  // field for any finite x), formed so that only the last step waits for |k|'
  const uint32_t base = (hi & 0xFFF00000u) + (25u << 20);
#if defined(__HIP_DEVICE_COMPILE__)
  uint32_t chi;
  asm("v_lshl_add_u32 %0, %1, 20, %2" : "=v"(chi) : "v"(ak), "v"(base));
#else
  const uint32_t chi = (ak << 20) + base;
#endif
  const uint64_t cbits = (uint64_t)chi << 32;
  double c;
  __builtin_memcpy(&c, &cbits, sizeof c);
  return (x + c) - c;
}

// reach values for which fp64_posit_add's results are the posit sums
constexpr uint32_t FP64_ADD_REACH_MAX = 26;

// High word of the rounding constant C for a double x whose high word is `hi` (only x's
// sign and exponent field are used): C = +-2^(e + 52 - nf) for scales e in -108 .. 107,
// whose FP64 ulp is the posit quantum there; 1.0 for zero (a zero stays zero); an infinite
// C for every other exponent, which turns (x + C) - C into NaN.  The GEMM keeps these in
// an LDS table indexed by hi >> 20.
PGA_HD uint32_t fp64_round_chi(uint32_t hi) {
  const uint32_t sign = hi & 0x80000000u;
  const int eb = (int)((hi >> 20) & 0x7FFu);
  const int e = eb - 1023;
  if (eb == 0) return 0x3FF00000u;
  if (e < -108 || e > 107) return sign | 0x7FF00000u;
  const int ak = (e >> 2) ^ (e >> 31);
  return sign | ((uint32_t)(eb + 25 + ak) << 20);
}

// High word of the sign-free rounding constant C = 1.5 * 2^(e + 52 - nf) for a double x
// whose high word is `hi` (only x's exponent field is used).  Its FP64 ulp is the posit
// quantum q = 2^(e - nf) at x's binade, its last bit is zero, and x + C stays in C's
// binade for either sign of x (|x| < 2^(e+1) <= 2^-24 * C), so (x + C) - C rounds x to a
// multiple of q with ties to the even multiple -- the even posit pattern -- whatever x's
// sign.  Scales -108 .. 107 get their C, zero gets 1.5 (a zero stays zero), every other
// exponent +infinity, which turns (x + C) - C into NaN.  The GEMM keeps these in an LDS
// table indexed by the exponent field alone, so that values of opposite signs read the
// same entry.
PGA_HD uint32_t fp64_grid_chi(uint32_t hi) {
  const int eb = (int)((hi >> 20) & 0x7FFu);
  const int e = eb - 1023;
  if (eb == 0) return 0x3FF80000u;
  if (e < -108 || e > 107) return 0x7FF00000u;
  const int ak = (e >> 2) ^ (e >> 31);
  return ((uint32_t)(eb + 25 + ak) << 20) | 0x00080000u;
}

// The constant of fp64_grid_chi computed from x's binary exponent instead of read from a
// table: C = 1.5 * 2^(e + 25 + |k|') for scales e in -108 .. 107 (the same value),
// 1.5 * 2^24 for zero (a zero stays zero), +infinity for every other exponent (then
// (x + C) - C is NaN).  Costs a few VALU instructions and no LDS.
PGA_HD double fp64_grid_c_arith(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
  const int e = __builtin_amdgcn_frexp_exp(x) - 1;
#else
  int fe = 0;
  (void)__builtin_frexp(x, &fe);
  const int e = fe - 1;
#endif
  const int ak = (e >> 2) ^ (e >> 31);
  const bool in = (unsigned)(e + 108) <= 215u;
  return __builtin_ldexp(1.5, in ? e + 25 + ak : 2000);
}

// Half the grid's quantum for a constant c = 1.5 * 2^E from fp64_grid_c_arith: 2^(E - 53),
// which is c / 1.5 * 2^-53 exactly, from c's exponent field (for an infinite c, a finite
// value that no difference compares equal to, as the division would give infinity).
PGA_HD double fp64_grid_half(double c) {
  uint64_t b;
  __builtin_memcpy(&b, &c, sizeof b);
  const uint64_t hb = (((b >> 52) & 0x7FFu) - 53u) << 52;
  double h;
  __builtin_memcpy(&h, &hb, sizeof h);
  return h;
}

// The GEMM's FP64 product and multiply-add roundings (rgemm_tile.hpp: fp64_posit_mul,
// fp64_mac) with the constants computed (fp64_grid_c_arith) instead of read from a table:
// round(t * a) as fma(t, a, C) - C, and round(c + round(t * a)) as (x + C) - C.  NaN where
// a scale leaves -108 .. 107 or an operand is NaR.
PGA_HD double fp64_posit_mul_arith(double t, double a) {
  const double c = fp64_grid_c_arith(t * a);
  return __builtin_fma(t, a, c) - c;
}
PGA_HD double fp64_mac_arith(double c, double t, double a) {
  const double x = c + fp64_posit_mul_arith(t, a);
  const double k = fp64_grid_c_arith(x);
  return (x + k) - k;
}

// round(1 / a) on the posit grid for a nonzero posit value a held in FP64: the correctly
// rounded FP64 quotient, rounded to the grid of its binade ((q + C) - C).  Rounding twice
// differs from rounding once only where the FP64 quotient is exactly a posit midpoint
// (rounding to FP64 cannot carry a value across one, as midpoints are FP64 numbers);
// there the pattern-level division decides.  NaN where 1/a lies outside scales
// -108 .. 107.
PGA_HD double fp64_posit_recip(double a) {
  const double q = 1.0 / a;
  const double c = fp64_grid_c_arith(q);
  const double r = (q + c) - c;
  const double half = fp64_grid_half(c);  // C = 1.5 * 2^E: the grid's quantum is 2^(E - 52)
  if (__builtin_fabs(q - r) == half) return to_double(div(P_ONE, from_double(a)));
  return r;
}

// b / a on the posit grid for posit values held in FP64, as fp64_posit_recip forms 1 / a:
// the FP64 quotient (correctly rounded) rounded to the grid of its binade, which is the
// posit quotient unless it lands exactly on a posit midpoint (only there can the first
// rounding have decided the second); there the pattern-level quotient decides.  NaN for a
// zero or NaR divisor, a NaR dividend, and quotients outside scales -108 .. 107.
PGA_HD double fp64_posit_div(double b, double a) {
  const double q = b / a;
  const double c = fp64_grid_c_arith(q);
  const double r = (q + c) - c;
  const double half = fp64_grid_half(c);
  if (__builtin_fabs(q - r) == half) return to_double(div(from_double(b), from_double(a)));
  return r;
}

// sqrt(a) on the posit grid for a positive posit value a held in FP64: the FP64 root
// (within an ulp of the true one) rounded to the grid of its binade.  That gives the posit
// root unless the true root lies within a few FP64 ulps of a posit midpoint; there the
// pattern-level root decides.
PGA_HD double fp64_posit_sqrt(double a) {
  const double q = __builtin_sqrt(a);
  const double c = fp64_grid_c_arith(q);
  const double r = (q + c) - c;
  const double half = fp64_grid_half(c);
  const double slack = __builtin_ldexp(q, -49);  // more than 4 ulps of q
  if (__builtin_fabs(__builtin_fabs(q - r) - half) <= slack) return to_double(pga::sqrt(from_double(a)));
  return r;
}

// Serial sums of posit values on the grid of one binade, for the long addition chains of
// the triangular solves.  For t in binade e (2^e <= |t| < 2^(e+1)) at scales -108 .. 107
// the posits of the binade are the multiples of the quantum q = 2^(e - nf), nf = 27 - |k|'
// fraction bits (fp64_posit_add), so t = m * q for an integer m with 2^nf <= |m| <
// 2^(nf+1).  A sum t + p whose exact value lies inside the binade rounds to
// (m + d) * q, d = floor(p/q) plus one where p/q's fraction exceeds 1/2; where the
// fraction is exactly 1/2 (a tie) it rounds to the even one of m + floor(p/q) and that
// plus one, which is the even posit pattern.  A run of sums is thus integer arithmetic
// on m, and a sum whose result (m + d) lies strictly inside the binade's range had its
// exact value inside the binade (within half a unit of it), so it was the posit sum.
struct ChainGrid {
  int m0;      // t = m0 * 2^qe
  int qe;      // exponent of the quantum q
  int lo, hi;  // the binade on t's side of zero: lo < m < hi strictly inside it
  double qinv; // 1/q
};

// The grid of t's binade; false for zero and for scales outside -108 .. 107.
PGA_HD bool chain_grid_setup(double t, ChainGrid& g) {
  uint64_t tb;
  __builtin_memcpy(&tb, &t, sizeof tb);
  const uint32_t th = (uint32_t)(tb >> 32);
  const int eb = (int)((th >> 20) & 0x7FFu), e = eb - 1023;
  if (eb == 0 || e < -108 || e > 107) return false;
  const int ak = (e >> 2) ^ (e >> 31), nf = 27 - ak;
  g.qe = e - nf;
  const uint64_t qib = (uint64_t)(1023 - g.qe) << 52;
  __builtin_memcpy(&g.qinv, &qib, sizeof g.qinv);
  g.m0 = (int)(t * g.qinv);  // exact: t is a multiple of q, |m0| < 2^28
  const int a = 1 << nf, b = 2 << nf;
  g.lo = (int32_t)th < 0 ? -b : a;
  g.hi = (int32_t)th < 0 ? -a : b;
  return true;
}

// m * 2^qe, exact for |m| < 2^53
PGA_HD double chain_grid_value(const ChainGrid& g, int m) {
  const uint64_t qb = (uint64_t)(1023 + g.qe) << 52;
  double q;
  __builtin_memcpy(&q, &qb, sizeof q);
  return (double)m * q;
}

// A term's step on the grid: d = floor(p/q) (clamped to +-2^29, which any step leaving
// the binade exceeds by far), plus one where the fraction of p/q exceeds 1/2; `tie`
// where it is exactly 1/2 (the step is then d or d + 1, whichever makes m even).
// p/q, its floor and fraction are exact (q is a power of two).
PGA_HD void chain_grid_term(double p, double qinv, int& d, bool& tie) {
  const double y = p * qinv, fl = __builtin_floor(y), f = y - fl;
  const double c = fl < -536870912.0 ? -536870912.0 : (fl > 536870912.0 ? 536870912.0 : fl);
  d = (int)c + (f > 0.5 ? 1 : 0);
  tie = f == 0.5;
}

// ---------------------------------------------------------------------------------
// Decoded-domain representation used by the GEMM inner loop.
// ---------------------------------------------------------------------------------
namespace dec {

// scale of zero: far below every finite scale, so that zero loses every magnitude
// comparison and a product with a zero factor stays far below every finite product
constexpr int ZERO_SC = -(1 << 20);
// scale of NaR: far above every finite scale; range checks send it to the exact path
constexpr int NAR_SC = (1 << 24);
// any scale below this is a zero (a zero's scale drifts downwards by design)
constexpr int ZERO_LIMIT = -(1 << 19);
// scales for which a result is computed here: regime at most 27 bits long
constexpr int FAST_MIN = -108;
constexpr int FAST_MAX = 107;

struct Val {
  uint32_t m;  // hidden bit at 30 (or exactly 2^31 after a rounding carry); 0 for zero
  int sc;      // 4k + e
  uint32_t sg; // sign mask: 0 or 0xFFFFFFFF
};

PGA_HD Val decode(uint32_t a) {
  Val v;
  if (a == 0) {
    v.m = 0;
    v.sc = ZERO_SC;
    v.sg = 0;
  } else if (a == P_NAR) {
    v.m = 0x40000000u;
    v.sc = NAR_SC;
    v.sg = 0;
  } else {
    v.sg = (uint32_t)((int32_t)a >> 31);
    Parts p = parts(v.sg ? 0u - a : a);
    v.m = p.frac31;
    v.sc = p.scale;
  }
  return v;
}

// Exact pattern of a decoded value (the value must be a posit: no rounding happens).
PGA_HD uint32_t encode(Val v) {
  if (v.sc < ZERO_LIMIT) return 0;
  if (v.sc >= NAR_SC / 2) return P_NAR;
  uint32_t m = v.m;
  int sc = v.sc;
  if (m >> 31) {  // 2^31: a rounding carry that was not renormalised
    m >>= 1;
    sc += 1;
  }
  uint32_t ui = round_tail(sc, (uint64_t)m << 32);
  return v.sg ? 0u - ui : ui;
}

}  // namespace dec
}  // namespace pga
