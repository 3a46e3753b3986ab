// posit_ops_driver.cpp -- every operator of the drop-in posit32 type (include/softposit_cpp.h),
// the SoftPosit posit32 operators the reference code uses (e.g. Rtrsm.cpp:130 `-` and `/`,
// Rgetrf2.cpp:122 `one / a`, main_gemm.cpp:99-104 construction from double and toDouble,
// util.cpp:154 `+=`), applied to records read from a file.  Host code only: the type's
// arithmetic calls the library's scalar entry points, so no GPU is needed.
//
//   posit_ops_driver <in.bin> <count> <out.bin>
// in:  count records {uint32 a, uint32 b, double d}
// out: count records {uint32 r[NOPS], double todouble_a}
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "softposit_cpp.h"

struct In {
  uint32_t a, b;
  double d;
};

constexpr int NOPS = 21;
struct Out {
  uint32_t r[NOPS];
  double todouble_a;
};

int main(int argc, char** argv) {
  if (argc < 4) {
    fprintf(stderr, "usage: %s in.bin count out.bin\n", argv[0]);
    return 2;
  }
  const size_t count = strtoull(argv[2], nullptr, 10);
  std::vector<In> in(count);
  FILE* f = fopen(argv[1], "rb");
  if (!f || fread(in.data(), sizeof(In), count, f) != count) {
    fprintf(stderr, "cannot read %s\n", argv[1]);
    return 2;
  }
  fclose(f);
  std::vector<Out> out(count);
  for (size_t i = 0; i < count; ++i) {
    const posit32 a = posit32::from_bits(in[i].a), b = posit32::from_bits(in[i].b);
    const double d = in[i].d;
    Out& o = out[i];
    int t = 0;
    o.r[t++] = (a + b).value;
    o.r[t++] = (a - b).value;
    o.r[t++] = (a * b).value;
    o.r[t++] = (a / b).value;
    o.r[t++] = (-a).value;
    posit32 x = a;
    x += b;
    o.r[t++] = x.value;
    x = a;
    x -= b;
    o.r[t++] = x.value;
    x = a;
    x *= b;
    o.r[t++] = x.value;
    x = a;
    x /= b;
    o.r[t++] = x.value;
    o.r[t++] = sqrt(a).value;
    x = a;
    x.sqrt();
    o.r[t++] = x.value;
    o.r[t++] = posit32(d).value;
    x = a;
    x = d;  // operator=(double)
    o.r[t++] = x.value;
    o.r[t++] = (d + b).value;
    o.r[t++] = (d - b).value;
    o.r[t++] = (d * b).value;
    o.r[t++] = (d / b).value;
    const double di = (d == d && d > -1e6 && d < 1e6) ? d : 0.0;  // int conversions defined
    o.r[t++] = posit32((int)(di * 1000.0)).value;  // construction from int
    x = (int)(di * 7.0);                            // operator=(int)
    o.r[t++] = x.value;
    o.r[t++] = posit32((float)d).value;            // construction from float
    o.r[t++] = (uint32_t)((a == b) | (a != b) << 1 | (a < b) << 2 | (a <= b) << 3 | (a > b) << 4 |
                          (a >= b) << 5 | a.isNaR() << 6);
    o.todouble_a = a.toDouble();
    if (t != NOPS) return 3;
  }
  f = fopen(argv[3], "wb");
  if (!f || fwrite(out.data(), sizeof(Out), count, f) != count) return 2;
  fclose(f);
  return 0;
}
