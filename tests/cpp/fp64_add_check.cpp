// Host build of pga::fp64_posit_add and of the solve chain's grid steps (pga::chain_grid_*,
// posit32_core.hpp) and of the GEMM's FMA product and
// sum roundings (with pga::fp64_grid_chi, the sign-free C = 1.5 * 2^E of its LDS table),
// checked against the oracle's pattern-level posit addition and multiplication: wherever
// they do not ask for a redo, their results must be the posit sum and product.  Pairs come from several families: values of uniform[-1,1) and of wider
// ranges, short fractions (many exact midpoints), sums that cancel, terms far below the
// accumulator, constructed midpoints and FP64 neighbours of midpoints, and the running
// sums of long chains.  Prints the counts; exit status 0 when no result differs.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>

#include "posit32_core.hpp"

extern "C" uint32_t orc_p32_add(uint32_t, uint32_t);
extern "C" uint32_t orc_p32_mul(uint32_t, uint32_t);
extern "C" uint32_t orc_p32_div(uint32_t, uint32_t);
extern "C" uint32_t orc_p32_sqrt(uint32_t);

static uint64_t splitmix(uint64_t& s) {
  uint64_t z = (s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
static double unit(uint64_t& s) { return (double)(splitmix(s) >> 11) * 0x1.0p-53; }

static long checked = 0, redone = 0, bad = 0, mul_checked = 0, mul_redone = 0, gsum_checked = 0, gsum_redone = 0;

// the GEMM's rounding constant for x (rgemm_tile.hpp: grid_c) computed directly
static double grid_c_host(double x) {
  uint64_t b;
  __builtin_memcpy(&b, &x, sizeof b);
  const uint64_t cb = (uint64_t)pga::fp64_grid_chi((uint32_t)(b >> 32)) << 32;
  double c;
  __builtin_memcpy(&c, &cb, sizeof c);
  return c;
}

// the GEMM's product rounding (rgemm_tile.hpp: fp64_posit_mul): fma(t, a, C) - C, C from
// the exponent of t*a rounded to FP64
static double fp64_posit_mul_host(double t, double a) {
  const double c = grid_c_host(t * a);
  return std::fma(t, a, c) - c;
}

// the GEMM's sum rounding (rgemm_tile.hpp: grid_round of c + y): (x + C) - C
static void check_gemm_sum(uint32_t a, uint32_t b) {
  if (a == pga::P_NAR || b == pga::P_NAR) return;
  const double x = pga::to_double(a) + pga::to_double(b);
  const double c = grid_c_host(x);
  const double y = (x + c) - c;
  ++gsum_checked;
  if (y != y) {  // a scale outside the table: the caller redoes the slice
    ++gsum_redone;
    return;
  }
  const uint32_t want = orc_p32_add(a, b);
  if (y != pga::to_double(want)) {
    if (bad < 8) printf("gemm sum mismatch %08x + %08x: got %.17g want %08x\n", a, b, y, want);
    ++bad;
  }
}

static void check_mul(uint32_t a, uint32_t b) {
  if (a == pga::P_NAR || b == pga::P_NAR) return;
  const double y = fp64_posit_mul_host(pga::to_double(a), pga::to_double(b));
  ++mul_checked;
  if (y != y) {  // a scale outside the table: the caller redoes the step
    ++mul_redone;
    return;
  }
  const uint32_t want = orc_p32_mul(a, b);
  if (y != pga::to_double(want)) {
    if (bad < 8) printf("mul mismatch %08x * %08x: got %.17g want %08x\n", a, b, y, want);
    ++bad;
  }
}

// The solve chain's grid steps (pga::chain_grid_setup / chain_grid_term /
// chain_grid_value), applied the way the device applies them to a block of up to 64
// terms: steps d summed in term order, ties settled in order (each rounds to the even
// value), every running value checked strictly inside the binade.  Each running value
// before the first one outside must be the posit sum of the oracle; returns the number
// of terms consumed (that first term is formed with the oracle's addition), and the sum.
static long grid_terms = 0, grid_exits = 0, grid_ties = 0;
static int check_grid_block(uint32_t& t, const uint32_t* p, int cnt) {
  pga::ChainGrid g;
  if (!pga::chain_grid_setup(pga::to_double(t), g)) {
    t = orc_p32_add(t, p[0]);
    return 1;
  }
  if (pga::chain_grid_value(g, g.m0) != pga::to_double(t)) {
    if (bad < 8) printf("grid value of %08x differs\n", t);
    ++bad;
  }
  int D = 0, cs = 0, used = cnt;
  uint32_t want = t;
  for (int u = 0; u < cnt; ++u) {
    int d;
    bool tie;
    pga::chain_grid_term(pga::to_double(p[u]), g.qinv, d, tie);
    D += d;
    if (tie) {  // the device settles ties in lane order after the prefix sum
      const int up = (g.m0 + D + cs) & 1;
      cs += up;
      ++grid_ties;
    }
    const long m = (long)g.m0 + D + cs;
    want = orc_p32_add(want, p[u]);
    ++grid_terms;
    if (!(m > g.lo && m < g.hi)) {
      ++grid_exits;
      used = u + 1;
      break;
    }
    if (pga::chain_grid_value(g, (int)m) != pga::to_double(want)) {
      if (bad < 8) printf("grid step mismatch t=%08x term %d (%08x): m %ld want %08x\n", t, u, p[u], m, want);
      ++bad;
    }
  }
  t = want;
  return used;
}

static void check_grid_chain(uint32_t t, const uint32_t* p, int n) {
  for (int k = 0; k < n;) {
    if (t == pga::P_NAR) return;
    k += check_grid_block(t, p + k, n - k < 64 ? n - k : 64);
  }
}

static void check(uint32_t a, uint32_t b) {
  check_gemm_sum(a, b);
  if (a == pga::P_NAR || b == pga::P_NAR) return;
  uint32_t reach = 0;
  const double y = pga::fp64_posit_add(pga::to_double(a), pga::to_double(b), reach);
  ++checked;
  if (reach > pga::FP64_ADD_REACH_MAX) {
    ++redone;
    return;
  }
  const uint32_t want = orc_p32_add(a, b);
  if (y != pga::to_double(want) || pga::from_double(y) != want) {
    if (bad < 8) printf("mismatch %08x + %08x: got %.17g want %08x\n", a, b, y, want);
    ++bad;
  }
}

int main(int argc, char** argv) {
  const long n = argc > 1 ? atol(argv[1]) : 1000000;
  uint64_t s = 2024;
  for (long i = 0; i < n; ++i) {
    // uniform[-1,1) values and their products (the solve's terms)
    const uint32_t a = pga::from_double(2 * unit(s) - 1), b = pga::from_double(2 * unit(s) - 1);
    check(a, b);
    check(a, orc_p32_mul(b, pga::from_double(2 * unit(s) - 1)));
    // any two patterns, and patterns at a spread of scales
    check((uint32_t)splitmix(s), (uint32_t)splitmix(s));
    check(pga::from_double(ldexp(2 * unit(s) - 1, (int)(splitmix(s) % 200) - 100)),
          pga::from_double(ldexp(2 * unit(s) - 1, (int)(splitmix(s) % 200) - 100)));
    // the ends of the fast range and beyond (regimes of 28 .. 30 bits)
    {
      const int sgn = (splitmix(s) & 1) ? 1 : -1;
      const int e0 = sgn * (96 + (int)(splitmix(s) % 25));
      check(pga::from_double(ldexp(2 * unit(s) - 1, e0)), pga::from_double(ldexp(2 * unit(s) - 1, e0 - (int)(splitmix(s) % 8))));
    }
    // short fractions: exact midpoints are common
    check(pga::from_double((double)((int)(splitmix(s) % 4001) - 2000) / 64),
          pga::from_double((double)((int)(splitmix(s) % 4001) - 2000) / 2048));
    check(pga::from_double(ldexp((double)(splitmix(s) % 1024), (int)(splitmix(s) % 40) - 20)),
          pga::from_double(ldexp((double)(splitmix(s) % 1024), (int)(splitmix(s) % 60) - 30)));
    // products: uniform values (56-bit exact products), any patterns, short fractions, and
    // factors whose product lands near a power of two
    check_mul(a, b);
    check_mul((uint32_t)splitmix(s), (uint32_t)splitmix(s));
    check_mul(pga::from_double(ldexp(2 * unit(s) - 1, (int)(splitmix(s) % 120) - 60)),
              pga::from_double(ldexp(2 * unit(s) - 1, (int)(splitmix(s) % 120) - 60)));
    check_mul(pga::from_double((double)((int)(splitmix(s) % 4001) - 2000) / 64),
              pga::from_double((double)((int)(splitmix(s) % 4001) - 2000) / 2048));
    {
      const double x = 1 + unit(s);
      const uint32_t px = pga::from_double(x), py = pga::from_double(2.0 / x);
      check_mul(px, py);
      check_mul(px, py + 1u);
      check_mul(px, py - 1u);
    }
    // cancellation: b close to -a
    check(a, (0u - a) + (uint32_t)(splitmix(s) % 9) - 4u);
    // a term far below the accumulator
    check(a, pga::from_double(ldexp(2 * unit(s) - 1, -(int)(splitmix(s) % 60))));
    // constructed midpoints: t + p = a posit midpoint exactly, and its FP64 neighbours
    {
      const uint32_t t = a;
      const uint32_t up = t + 1u;  // next pattern up (same sign region)
      if (t != 0 && up != pga::P_NAR && (t >> 31) == (up >> 31)) {
        const double mid = (pga::to_double(t) + pga::to_double(up)) / 2;  // exact
        const double p0 = mid - pga::to_double(t);
        const uint32_t pp = pga::from_double(p0);
        check(t, pp);
        check(t, pp + 1u);
        check(t, pp - 1u);
      }
    }
  }
  // products whose FP64 rounding lands exactly on a posit midpoint while the exact
  // product does not: mantissas m_t, m_a in [2^27, 2^28) with m_t * m_a = 2^27 + r
  // (mod 2^28), r in 1..3, and a 56-bit product (the posit keeps its top 28 bits)
  for (int i = 0; i < 200000; ++i) {
    const uint64_t mt = (splitmix(s) & 0x7FFFFFFull) | 0x8000001ull;  // odd, top bit set
    uint64_t inv = mt;  // inverse of mt modulo 2^28 by Newton's iteration
    for (int it = 0; it < 5; ++it) inv = (inv * (2 - mt * inv)) & 0xFFFFFFFull;
    const uint64_t r = 1 + splitmix(s) % 3;
    const uint64_t ma = (((1ull << 27) + r) * inv) & 0xFFFFFFFull;
    if (ma < (1ull << 27) || mt * ma < (1ull << 55)) continue;
    check_mul(pga::from_double(ldexp((double)mt, -27)), pga::from_double(ldexp((double)ma, -27)));
  }
  // running sums of chains like the solve's: t_{k+1} = round(t_k - a_k * b_k)
  for (int c = 0; c < 200; ++c) {
    uint32_t t = pga::from_double(2 * unit(s) - 1);
    for (int k = 0; k < 5000; ++k) {
      const uint32_t p = 0u - orc_p32_mul(pga::from_double(2 * unit(s) - 1), pga::from_double(2 * unit(s) - 1));
      check(t, p);
      t = orc_p32_add(t, p);
    }
  }
  // grid chains: the solve's (products of uniform values), chains that start on or next
  // to a power of two and step across it either way, exact midpoints (ties), sign
  // changes, terms far below or far above the sum, arbitrary patterns, the range's ends
  {
    static uint32_t p[4096];
    for (int c = 0; c < 100; ++c) {
      for (int k = 0; k < 4096; ++k)
        p[k] = 0u - orc_p32_mul(pga::from_double((2 * unit(s) - 1) * 0.02), pga::from_double(2 * unit(s) - 1));
      check_grid_chain(pga::from_double(2 * unit(s) - 1), p, 4096);
    }
    for (long i = 0; i < n / 8; ++i) {
      const int e = (int)(splitmix(s) % 200) - 100;
      const double q = ldexp(1.0, e - 27 + ((e >> 2) ^ (e >> 31)));  // posit quantum at scale e
      const double sg = (splitmix(s) & 1) ? 1.0 : -1.0;
      const int off = (int)(splitmix(s) % 9) - 4;
      const uint32_t t = pga::from_double(sg * (ldexp(1.0, e + (int)(splitmix(s) & 1)) + off * q));
      const int fam = (int)(splitmix(s) % 6);
      for (int u = 0; u < 64; ++u) {
        if (fam == 0)  // a few (half-)quanta either way
          p[u] = pga::from_double(((int)(splitmix(s) % 17) - 8) * q * 0.5 * ((splitmix(s) & 3) ? 1 : 0.25));
        else if (fam == 1)  // any size at nearby scales
          p[u] = pga::from_double(ldexp(2 * unit(s) - 1, e - (int)(splitmix(s) % 30)));
        else if (fam == 2)  // odd multiples of half a quantum: ties
          p[u] = pga::from_double(q * 0.5 * (double)(2 * (splitmix(s) % 5) + 1) * ((splitmix(s) & 1) ? 1 : -1));
        else if (fam == 3)  // across zero, or over the binade
          p[u] = pga::from_double(-sg * ldexp(unit(s), e + (int)(splitmix(s) % 3) - 1));
        else if (fam == 4)
          p[u] = (uint32_t)splitmix(s);
        else  // far beyond the sum, now and then
          p[u] = pga::from_double(ldexp(2 * unit(s) - 1, e + ((splitmix(s) % 16) ? -3 : 40)));
      }
      check_grid_chain(fam == 4 ? (uint32_t)splitmix(s) : t, p, 64);
    }
  }
  // the arithmetic constant (pga::fp64_grid_c_arith) against the table's for every
  // exponent field, both signs and several fractions: the same constant where the table
  // has a finite one, infinity where it has infinity; zero stays zero
  long arith_checked = 0;
  for (uint32_t eb = 0; eb < 2048; ++eb)
    for (int sg = 0; sg < 2; ++sg)
      for (int f = 0; f < 4; ++f) {
        const uint64_t bits = ((uint64_t)sg << 63) | ((uint64_t)eb << 52) | (splitmix(s) >> 12) * (f != 0);
        double x;
        __builtin_memcpy(&x, &bits, sizeof x);
        if (eb == 2047) continue;  // inf / NaN: not a posit value
        const double ct = grid_c_host(x), ca = pga::fp64_grid_c_arith(x);
        ++arith_checked;
        bool ok;
        if (x == 0.0)
          ok = (x + ca) - ca == 0.0;
        else if (eb == 0)
          ok = std::isinf(ca);  // subnormal (no posit value or product is one): redo
        else
          ok = std::isinf(ct) ? std::isinf(ca) : ca == ct;
        if (!ok) {
          ++bad;
          if (bad < 8) printf("arith constant differs for exponent field %u: %.17g vs %.17g\n", eb, ca, ct);
        }
      }
  printf("checked %ld redone %ld mismatches %ld products %ld redone %ld gemm-sums %ld redone %ld grid-terms %ld exits %ld ties %ld\n",
         checked, redone, bad, mul_checked, mul_redone, gsum_checked, gsum_redone, grid_terms, grid_exits, grid_ties);
  // the reciprocal of the LU panel (pga::fp64_posit_recip) against the oracle's division:
  // random patterns, patterns with short fractions (quotients near midpoints), powers of
  // two and their neighbours; NaN only where 1/a is outside scales -108 .. 107
  long recip_checked = 0, recip_nan = 0;
  auto check_recip = [&](uint32_t a) {
    if (a == 0 || a == pga::P_NAR) return;
    const double r = pga::fp64_posit_recip(pga::to_double(a));
    const uint32_t want = orc_p32_div(pga::P_ONE, a);
    ++recip_checked;
    if (r != r) {
      ++recip_nan;
      const double w = std::fabs(pga::to_double(want));
      if (w >= 0x1p-108 && w < 0x1p108) {
        ++bad;
        if (bad < 8) printf("recip of %08x: NaN inside the range\n", a);
      }
      return;
    }
    if (pga::from_double(r) != want) {
      ++bad;
      if (bad < 8) printf("recip of %08x: got %08x want %08x\n", a, pga::from_double(r), want);
    }
  };
  for (long i = 0; i < 4 * n; ++i) check_recip((uint32_t)splitmix(s));
  for (uint32_t hi = 0; hi < (1u << 16); ++hi) {  // short fractions at every regime
    check_recip(hi << 16);
    check_recip((hi << 16) | 1u);
    check_recip((hi << 16) - 1u);
  }
  for (int e = -120; e <= 120; ++e) {
    const uint32_t p2 = pga::from_double(std::ldexp(1.0, e));
    for (int d = -3; d <= 3; ++d) check_recip(p2 + (uint32_t)d);
  }
  // the square root of the Cholesky diagonal blocks (pga::fp64_posit_sqrt) against the
  // oracle's: random positive patterns, short fractions at every regime, perfect squares
  // and their neighbours
  long sqrt_checked = 0;
  auto check_sqrt = [&](uint32_t a) {
    if (a == 0 || (a >> 31)) return;
    const double r = pga::fp64_posit_sqrt(pga::to_double(a));
    const uint32_t want = orc_p32_sqrt(a);
    ++sqrt_checked;
    if (pga::from_double(r) != want) {
      ++bad;
      if (bad < 8) printf("sqrt of %08x: got %08x want %08x\n", a, pga::from_double(r), want);
    }
  };
  for (long i = 0; i < 4 * n; ++i) check_sqrt((uint32_t)splitmix(s) >> 1);
  for (uint32_t hi = 0; hi < (1u << 15); ++hi) {
    check_sqrt(hi << 16);
    check_sqrt((hi << 16) | 1u);
    check_sqrt((hi << 16) - 1u);
  }
  for (uint32_t k = 1; k < 200000; ++k) {  // perfect squares of short values
    const uint32_t sq = pga::from_double((double)k * (double)k * 0x1p-20);
    check_sqrt(sq);
    check_sqrt(sq + 1u);
    check_sqrt(sq - 1u);
  }
  // the quotients of the upper Cholesky's solves (pga::fp64_posit_div) against the
  // oracle's division: random pairs, short fractions, and dividends over powers of two
  // (exact quotients whose dropped bits are often exactly half a quantum: ties, which the
  // FP64 rounding settles to the even value by itself).  The pattern-level fallback is for
  // an inexact quotient whose FP64 rounding lands exactly on a midpoint; posit midpoints
  // are FP64 values and rounding is monotonic, so any other FP64 quotient lies on the same
  // side of every midpoint as the exact one.  Such quotients are too rare to sample here.
  long div_checked = 0, div_nan = 0, div_ties = 0;
  auto check_div = [&](uint32_t b, uint32_t a) {
    if (a == 0 || a == pga::P_NAR || b == pga::P_NAR) return;
    const double r = pga::fp64_posit_div(pga::to_double(b), pga::to_double(a));
    const uint32_t want = orc_p32_div(b, a);
    ++div_checked;
    if (r != r) {
      ++div_nan;
      const double w = std::fabs(pga::to_double(want));
      if (w >= 0x1p-108 && w < 0x1p108) {
        ++bad;
        if (bad < 8) printf("%08x / %08x: NaN inside the range\n", b, a);
      }
      return;
    }
    if (pga::from_double(r) != want) {
      ++bad;
      if (bad < 8) printf("%08x / %08x: got %08x want %08x\n", b, a, pga::from_double(r), want);
    }
  };
  for (long i = 0; i < 4 * n; ++i) check_div((uint32_t)splitmix(s), (uint32_t)splitmix(s));
  for (long i = 0; i < 4 * n; ++i)
    check_div((uint32_t)(splitmix(s) >> 32) & 0xFFFF0000u, (uint32_t)(splitmix(s) >> 32) & 0xFFF00000u);
  for (long i = 0; i < 4 * n; ++i) {
    const uint32_t b = (uint32_t)splitmix(s);
    const int e = (int)(splitmix(s) % 81) - 40;
    const uint32_t a = pga::from_double(std::ldexp((splitmix(s) & 1) ? -1.0 : 1.0, e));
    const double q = pga::to_double(b) / pga::to_double(a);  // exact: a is a power of two
    if (q == q && pga::to_double(orc_p32_div(b, a)) != q) ++div_ties;  // the posit quotient rounded
    check_div(b, a);
  }
  printf("arithmetic constants checked %ld, reciprocals checked %ld (NaN %ld), roots checked %ld, quotients checked "
         "%ld (NaN %ld, rounded exact quotients %ld)\n",
         arith_checked, recip_checked, recip_nan, sqrt_checked, div_checked, div_nan, div_ties);
  return bad ? 1 : 0;
}
