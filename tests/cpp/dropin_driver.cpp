// dropin_driver.cpp -- a program written against the reference's C++ API (the calls of
// main_lu.cpp / main_cholesky.cpp / main_gemm.cpp), linked to the drop-in library.
// It reads a posit32 matrix (raw patterns) from argv[2], runs one routine and writes the
// outputs as raw binary files so that the GPU tests can compare them with the oracle.
//
//   dropin_driver lu  <A.bin> <n> <nb> <out_prefix>   Rgetrf_opencl + Rgemv(b=A*1) + Rgetrs
//   dropin_driver cho <A.bin> <n> <uplo> <out_prefix> Rpotrf_opencl + Rgemv + Rpotrs
//   dropin_driver gemm <A.bin> <n> <k> <out_prefix>   RgemmG NN C = -A(:,0:k)*A(0:k,:) + C
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "mplapack_posit.h"

static std::vector<posit32> read_matrix(const char* path, size_t count) {
  std::vector<posit32> v(count);
  FILE* f = fopen(path, "rb");
  if (!f || fread(v.data(), 4, count, f) != count) {
    fprintf(stderr, "cannot read %s\n", path);
    exit(2);
  }
  fclose(f);
  return v;
}

static void write_raw(const std::string& path, const void* p, size_t bytes) {
  FILE* f = fopen(path.c_str(), "wb");
  fwrite(p, 1, bytes, f);
  fclose(f);
}

int main(int argc, char** argv) {
  if (argc < 6) {
    fprintf(stderr, "usage: %s lu|cho|gemm A.bin n arg out_prefix\n", argv[0]);
    return 2;
  }
  initOpenCL();
  std::string mode = argv[1], out = argv[5];
  int n = atoi(argv[3]);
  std::vector<posit32> a = read_matrix(argv[2], (size_t)n * n);
  std::vector<posit32> a0 = a;
  std::vector<posit32> one(n, posit32(1.0)), b(n, posit32(0.0));
  const posit32 p_one = 1.0, p_zero = 0.0;
  if (mode == "lu") {
    int nb = atoi(argv[4]);
    std::vector<int> ipiv(n);
    int info = 0;
    Rgemv("N", n, n, p_one, a0.data(), n, one.data(), 1, p_zero, b.data(), 1);
    Rgetrf_opencl(n, n, a.data(), n, ipiv.data(), info, nb);
    int info2 = 0;
    Rgetrs("N", n, 1, a.data(), n, ipiv.data(), b.data(), n, info2);
    write_raw(out + "_factor.bin", a.data(), (size_t)n * n * 4);
    write_raw(out + "_ipiv.bin", ipiv.data(), (size_t)n * 4);
    write_raw(out + "_x.bin", b.data(), (size_t)n * 4);
    printf("info %d %d\n", info, info2);
  } else if (mode == "cho") {
    const char* uplo = argv[4];
    int info = 0, info2 = 0;
    Rgemv("N", n, n, p_one, a0.data(), n, one.data(), 1, p_zero, b.data(), 1);
    Rpotrf_opencl(uplo, n, a.data(), n, info);
    Rpotrs(uplo, n, 1, a.data(), n, b.data(), n, info2);
    write_raw(out + "_factor.bin", a.data(), (size_t)n * n * 4);
    write_raw(out + "_x.bin", b.data(), (size_t)n * 4);
    printf("info %d %d\n", info, info2);
  } else {
    int k = atoi(argv[4]);
    std::vector<posit32> c = a0;
    const posit32 m_one = -1.0;
    RgemmG("N", "N", n, n, k, m_one, a0.data(), n, a0.data(), n, p_one, c.data(), n);
    write_raw(out + "_c.bin", c.data(), (size_t)n * n * 4);
    // host scalar operators of the posit32 type
    posit32 x = 0.75, y = -2.5;
    posit32 z = (x * y + posit32(1.0)) / y;
    printf("scalar %08x %08x %.17g\n", (unsigned)z.value, (unsigned)sqrt(posit32(2.0)).value, z.toDouble());
  }
  return 0;
}
