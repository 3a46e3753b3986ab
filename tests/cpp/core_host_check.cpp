// Host build of the kernels' pattern-level posit32 core (posit32_core.hpp), checked
// against the oracle's independent restatement: sqrt on every pattern of a stride, div on
// seeded random pairs and on all pairs of mantissa / regime edge patterns.
// Usage: core_host_check <sqrt stride> <random div pairs>; exit status 0 when all agree.
#include <cstdint>
#include <cstdio>
#include <cstdlib>

#include "posit32_core.hpp"

extern "C" uint32_t orc_p32_div(uint32_t, uint32_t);
extern "C" uint32_t orc_p32_sqrt(uint32_t);

static uint64_t splitmix(uint64_t& s) {
  uint64_t z = (s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

int main(int argc, char** argv) {
  const long stride = argc > 1 ? atol(argv[1]) : 4099;
  const long pairs = argc > 2 ? atol(argv[2]) : 1000000;
  long bad = 0;
  for (uint64_t i = 0; i < (1ull << 32); i += (uint64_t)stride) {
    const uint32_t a = (uint32_t)i;
    if (pga::sqrt(a) != orc_p32_sqrt(a) && bad++ < 5) printf("sqrt %08x\n", a);
  }
  uint64_t s = 12345;
  for (long i = 0; i < pairs; ++i) {
    const uint64_t r = splitmix(s);
    uint32_t a = (uint32_t)r, b = (uint32_t)(r >> 32);
    if (i & 1) b &= 0xFFFF0000u;  // short fractions: exact and near-exact quotients
    if ((i & 3) == 2) a = b * 3u;
    if (pga::div(a, b) != orc_p32_div(a, b) && bad++ < 5) printf("div %08x %08x\n", a, b);
  }
  uint32_t e[512];
  int ne = 0;
  for (int k = 0; k < 31; ++k)
    for (uint32_t d = 0; d < 4; ++d) {
      const uint32_t base = 0x40000000u >> k;
      e[ne++] = base + d;
      e[ne++] = base - 1 - d;
      e[ne++] = 0u - (base + d);
      e[ne++] = 0x7FFFFFFFu - d - (uint32_t)k;
    }
  for (int i = 0; i < ne; ++i)
    for (int j = 0; j < ne; ++j)
      if (pga::div(e[i], e[j]) != orc_p32_div(e[i], e[j]) && bad++ < 5) printf("div %08x %08x\n", e[i], e[j]);
  printf("mismatches %ld\n", bad);
  return bad ? 1 : 0;
}
