// Host build of pga::fp64_posit_add (posit32_core.hpp), checked against the oracle's
// pattern-level posit addition: wherever it does not ask for a redo, its result must be
// the posit sum.  Pairs come from several families: values of uniform[-1,1) and of wider
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

static uint64_t splitmix(uint64_t& s) {
  uint64_t z = (s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
static double unit(uint64_t& s) { return (double)(splitmix(s) >> 11) * 0x1.0p-53; }

static long checked = 0, redone = 0, bad = 0;

static void check(uint32_t a, uint32_t b) {
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
  // running sums of chains like the solve's: t_{k+1} = round(t_k - a_k * b_k)
  for (int c = 0; c < 200; ++c) {
    uint32_t t = pga::from_double(2 * unit(s) - 1);
    for (int k = 0; k < 5000; ++k) {
      const uint32_t p = 0u - orc_p32_mul(pga::from_double(2 * unit(s) - 1), pga::from_double(2 * unit(s) - 1));
      check(t, p);
      t = orc_p32_add(t, p);
    }
  }
  printf("checked %ld redone %ld mismatches %ld\n", checked, redone, bad);
  return bad ? 1 : 0;
}
