// rgemm.hip -- posit32 GEMM kernels and their launchers (gfx950).
//
// Replaces the OpenCL kernels gemm_{NN,TN,NT,TT}_X of libRpack/kernel_entry.cl:598-792,
// with the per-element rounding order of the CPU Rgemm (libRpack/Rgemm.cpp:62-251),
// which is what every reference caller of RgemmG expects bit for bit.
#include <hip/hip_runtime.h>

#include <map>
#include <mutex>
#include <tuple>

#include "pga_internal.hpp"
#include "rgemm_tile.hpp"

namespace pga {
namespace gemm {

// ---------------------------------------------------------------------------------
// Tiled kernel
// ---------------------------------------------------------------------------------
// Workgroup -> tile of C.  The hardware hands consecutive workgroups to the 8 XCDs in
// turn, and each XCD has its own L2.  The workgroups of one XCD (bid % 8) are given one
// contiguous part of an order that walks bands of GROUP_M tile rows column by column, so
// the tiles that run at once on an XCD share a few slices of A (and their slices of B)
// in its L2; in plain column-major order an XCD would get every eighth tile row and
// reuse no slice of A.  Workgroups past the last multiple of 8 keep their place.
__device__ __forceinline__ void xcd_tile(int bid, int tiles_m, int tiles_n, int& ti, int& tj) {
  constexpr int XCDS = 8, GROUP_M = 8;
  const int total = tiles_m * tiles_n, per = total / XCDS;
  const int o = bid < per * XCDS ? (bid % XCDS) * per + bid / XCDS : bid;
  const int band = GROUP_M * tiles_n;
  const int g = o / band, r = o % band;
  const int gm = min(GROUP_M, tiles_m - g * GROUP_M);  // tile rows in this band
  tj = r / gm;
  ti = g * GROUP_M + r % gm;
}

// Workgroups per CU the compiler must leave room for: five 64 x 32 tiles (their FP64
// slices and table, 28 KB of LDS each, and at most 96 VGPRs) fill a CU; smaller tiles
// fit more anyway.
template <int TM, int TN>
constexpr int tile_wgs_per_cu() { return TM * TN >= 8 ? 5 : 3; }

template <int BM, int BN, int BK, int TM, int TN, bool TA, bool TB, bool SKIP0>
__global__ __launch_bounds__((BM / TM) * (BN / TN), (tile_wgs_per_cu<TM, TN>())) void rgemm_tiled(
    int m, int n, int k, uint32_t alpha, const uint32_t* __restrict__ A, int lda,
    const uint32_t* __restrict__ B, int ldb, uint32_t beta, uint32_t* __restrict__ C, int ldc,
    int tiles_m, int tri, int urgent, const int* __restrict__ stop) {
  PGA_KERNEL_VGPR_FLOOR();
  if (urgent) PGA_URGENT();
  int ti, tj;
  xcd_tile(blockIdx.x, tiles_m, (n + BN - 1) / BN, ti, tj);
  const int i0 = ti * BM, j0 = tj * BN;
  if (stop && *stop) return;  // an earlier step of a factorization failed
  // triangular update (Rsyrk): skip tiles with no element in the stored triangle
  if (tri == 1 && i0 + BM - 1 < j0) return;
  if (tri == 2 && j0 + BN - 1 < i0) return;
  gemm_tile<BM, BN, BK, TM, TN, TA, TB, SKIP0>(m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, i0, j0, tri);
}

// Batched form: one persistent launch works through the tiles of many independent
// problems, taking them from a global counter (problems are listed in the order the
// caller wants them scheduled, e.g. most expensive first).  Each tile is computed by the
// same gemm_tile as the plain kernel, so every element gets the same bits as a plain call.
template <int BM, int BN, int BK, int TM, int TN, bool TA, bool TB>
__global__ __launch_bounds__((BM / TM) * (BN / TN), (tile_wgs_per_cu<TM, TN>())) void rgemm_batched_persistent(
    const BatchProblem* __restrict__ probs, const int* __restrict__ tile_start, int nprob, int total_tiles,
    int* __restrict__ counter) {
  PGA_KERNEL_VGPR_FLOOR();
  __shared__ int s_tile;
  for (;;) {
    if (threadIdx.x == 0) s_tile = atomicAdd(counter, 1);
    __syncthreads();
    const int t = s_tile;
    __syncthreads();
    if (t >= total_tiles) return;  // every workgroup ends here once the queue is empty
    // problem owning tile t: last p with tile_start[p] <= t
    int lo = 0, hi = nprob - 1;
    while (lo < hi) {
      int mid = (lo + hi + 1) >> 1;
      if (tile_start[mid] <= t) lo = mid;
      else hi = mid - 1;
    }
    const BatchProblem& p = probs[lo];
    const int tl = t - tile_start[lo];
    const int tm = (p.m + BM - 1) / BM;
    gemm_tile<BM, BN, BK, TM, TN, TA, TB, false>(p.m, p.n, p.k, p.alpha, p.a, p.lda, p.b, p.ldb, p.beta, p.c,
                                                 p.ldc, (tl % tm) * BM, (tl / tm) * BN, 0);
  }
}

// Panel form of the 64 x 32 tile: the slices come from FP64 operand panels (Panels).
template <bool TA, bool TB, bool SKIP0>
__global__ __launch_bounds__(256, 3) void rgemm_tiled_panel(int m, int n, int k, uint32_t alpha,
                                                         const uint32_t* __restrict__ A, int lda,
                                                         const uint32_t* __restrict__ B, int ldb, uint32_t beta,
                                                         uint32_t* __restrict__ C, int ldc, int tiles_m, int tri,
                                                         int urgent, const int* __restrict__ stop, Panels pn) {
  PGA_KERNEL_VGPR_FLOOR();
  if (urgent) PGA_URGENT();
  int ti, tj;
  xcd_tile(blockIdx.x, tiles_m, (n + 31) / 32, ti, tj);
  const int i0 = ti * 64, j0 = tj * 32;
  if (stop && *stop) return;
  if (tri == 1 && i0 + 63 < j0) return;
  if (tri == 2 && j0 + 31 < i0) return;
  __shared__ __align__(16) double sA[16][64];
  __shared__ __align__(16) double sT[16][32];
  gemm_tile_impl_fp64<64, 32, 16, 4, 2, TA, TB, SKIP0, true>(m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, i0, j0,
                                                              tri, sA, sT, pn);
}

// The panels: one workgroup per slice (i-block or j-block, k-slice), lanes walking the
// source matrix's leading dimension so that the reads coalesce.
template <bool TA>
__global__ __launch_bounds__(256) void k_panel_a(int m, int k, const uint32_t* __restrict__ A, int lda,
                                                 double* __restrict__ pa, int ks) {
  PGA_KERNEL_VGPR_FLOOR();
  const int ib = blockIdx.x / ks, s = blockIdx.x % ks;
  double* out = pa + (size_t)blockIdx.x * (16 * 64);
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int e = threadIdx.x + q * 256;
    const int ll = TA ? e % 16 : e / 64, ii = TA ? e / 16 : e % 64;
    out[ll * 64 + ii] = panel_a_elem<64, 16, TA>(m, k, A, lda, ib, s, ll, ii);
  }
}
template <bool TB, bool SUM_FIRST, bool SKIP0>
__global__ __launch_bounds__(256) void k_panel_t(int n, int k, uint32_t alpha, const uint32_t* __restrict__ B,
                                                 int ldb, double* __restrict__ pt, int ks) {
  PGA_KERNEL_VGPR_FLOOR();
  const int jb = blockIdx.x / ks, s = blockIdx.x % ks;
  double* out = pt + (size_t)blockIdx.x * (16 * 32);
#pragma unroll
  for (int q = 0; q < 2; ++q) {
    const int e = threadIdx.x + q * 256;
    const int ll = TB ? e / 32 : e % 16, jj = TB ? e % 32 : e / 16;
    out[ll * 32 + jj] = panel_t_elem<32, 16, TB, SUM_FIRST, SKIP0>(n, k, alpha, B, ldb, jb, s, ll, jj);
  }
}

// FP64 operand panels (declared in rgemm_tile.hpp)
hipError_t panels_build(bool ta, bool tb, bool skip0, int m, int n, int k, uint32_t alpha, const uint32_t* A, int lda,
                        const uint32_t* B, int ldb, hipStream_t s, Panels* pn, void** ws) {
  *ws = nullptr;
  hipMemPool_t pool = runtime().pool;
  if (!pool || m <= 0 || n <= 0 || k <= 0) return hipSuccess;
  const int tm = (m + 63) / 64, tn = (n + 31) / 32, ks = (k + 15) / 16;
  if ((long long)tm * ks > 0x7FFFFFFF || (long long)tn * ks > 0x7FFFFFFF) return hipSuccess;
  const size_t a_elems = (size_t)tm * ks * 16 * 64, t_elems = (size_t)tn * ks * 16 * 32;
  if (hipMallocFromPoolAsync(ws, (a_elems + t_elems) * sizeof(double), pool, s) != hipSuccess) {
    (void)hipGetLastError();
    *ws = nullptr;
    return hipSuccess;
  }
  double* pa = static_cast<double*>(*ws);
  double* pt = pa + a_elems;
  *pn = Panels{pa, pt, ks};
  if (ta) hipLaunchKernelGGL((k_panel_a<true>), dim3(tm * ks), dim3(256), 0, s, m, k, A, lda, pa, ks);
  else hipLaunchKernelGGL((k_panel_a<false>), dim3(tm * ks), dim3(256), 0, s, m, k, A, lda, pa, ks);
  // the B side as the tile reads it: TN / TT sum first (no alpha), NN / NT carry alpha
  if (tb && ta) hipLaunchKernelGGL((k_panel_t<true, true, false>), dim3(tn * ks), dim3(256), 0, s, n, k, alpha, B, ldb, pt, ks);
  else if (!tb && ta) hipLaunchKernelGGL((k_panel_t<false, true, false>), dim3(tn * ks), dim3(256), 0, s, n, k, alpha, B, ldb, pt, ks);
  else if (tb && skip0) hipLaunchKernelGGL((k_panel_t<true, false, true>), dim3(tn * ks), dim3(256), 0, s, n, k, alpha, B, ldb, pt, ks);
  else if (tb) hipLaunchKernelGGL((k_panel_t<true, false, false>), dim3(tn * ks), dim3(256), 0, s, n, k, alpha, B, ldb, pt, ks);
  else if (skip0) hipLaunchKernelGGL((k_panel_t<false, false, true>), dim3(tn * ks), dim3(256), 0, s, n, k, alpha, B, ldb, pt, ks);
  else hipLaunchKernelGGL((k_panel_t<false, false, false>), dim3(tn * ks), dim3(256), 0, s, n, k, alpha, B, ldb, pt, ks);
  return hipGetLastError();
}

__global__ void reset_counter(int* counter) {
  PGA_KERNEL_VGPR_FLOOR();
  if (threadIdx.x == 0) *counter = 0;
}

// ---------------------------------------------------------------------------------
// Exact kernel: one element per thread, pattern-level operations only.  Used for
// validation of the tiled kernel and selectable at run time.
// ---------------------------------------------------------------------------------
template <bool TA, bool TB, bool SKIP0>
__global__ __launch_bounds__(256) void rgemm_exact(int m, int n, int k, uint32_t alpha,
                                                   const uint32_t* __restrict__ A, int lda,
                                                   const uint32_t* __restrict__ B, int ldb,
                                                   uint32_t beta, uint32_t* __restrict__ C,
                                                   int ldc, int tri, const int* __restrict__ stop) {
  PGA_KERNEL_VGPR_FLOOR();
  int i = blockIdx.x * 16 + (threadIdx.x % 16);
  int j = blockIdx.y * 16 + (threadIdx.x / 16);
  if (i >= m || j >= n) return;
  if (stop && *stop) return;
  if ((tri == 1 && i < j) || (tri == 2 && i > j)) return;
  uint32_t* cij = C + i + (size_t)j * ldc;
  if (!TA) {
    uint32_t c = (beta == 0) ? 0u : (beta == P_ONE ? *cij : pga::mul(beta, *cij));
    for (int l = 0; l < k; ++l) {
      uint32_t b = TB ? B[j + (size_t)l * ldb] : B[l + (size_t)j * ldb];
      if (SKIP0 && b == 0) continue;
      c = pga::add(c, pga::mul(pga::mul(alpha, b), A[i + (size_t)l * lda]));
    }
    *cij = c;
  } else {
    uint32_t t = 0;
    for (int l = 0; l < k; ++l) {
      uint32_t b = TB ? B[j + (size_t)l * ldb] : B[l + (size_t)j * ldb];
      t = pga::add(t, pga::mul(A[l + (size_t)i * lda], b));
    }
    uint32_t at = pga::mul(alpha, t);
    *cij = (beta == 0) ? at : pga::add(at, pga::mul(beta, *cij));
  }
}

// alpha == 0 (Rgemm.cpp:142-157): C = beta==0 ? 0 : beta*C
__global__ __launch_bounds__(256) void scale_c(int m, int n, uint32_t beta, uint32_t* C, int ldc,
                                               int j0, int tri, const int* stop) {
  PGA_KERNEL_VGPR_FLOOR();
  int i = blockIdx.x * 256 + threadIdx.x;
  int j = blockIdx.y;
  if (i >= m) return;
  if (stop && *stop) return;
  if ((tri == 1 && i < j0 + j) || (tri == 2 && i > j0 + j)) return;
  uint32_t* p = C + i + (size_t)j * ldc;
  *p = (beta == 0) ? 0u : pga::mul(beta, *p);
}

// Tile shapes: 64x64 with 4x4 elements per lane for large C; 32x32 (2x2) and 16x16
// (1x1) keep enough waves in flight for skinny C (e.g. Cholesky's N = 64 updates),
// since K is never split.

template <int BM, int TM, bool TA, bool TB, bool SKIP0>
static hipError_t launch_tiled_cfg(int m, int n, int k, uint32_t alpha, const uint32_t* A, int lda,
                                   const uint32_t* B, int ldb, uint32_t beta, uint32_t* C, int ldc,
                                   hipStream_t s, int tri, int urgent, const int* stop, int cap) {
  constexpr int BN = BM == 64 ? 32 : BM, TN = BM == 64 ? 2 : TM, BK = 16;
  constexpr int NT = (BM / TM) * (BN / TN);
  int tm = (m + BM - 1) / BM, tn = (n + BN - 1) / BN;
  long long blocks = (long long)tm * tn;
  if (blocks > 0x7FFFFFFF) return hipErrorInvalidValue;
  const unsigned dyn = pga::lds_for_cap((const void*)rgemm_tiled<BM, BN, BK, TM, TN, TA, TB, SKIP0>, cap);
  hipLaunchKernelGGL((rgemm_tiled<BM, BN, BK, TM, TN, TA, TB, SKIP0>), dim3((unsigned)blocks), dim3(NT),
                     dyn, s, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, tm, tri, urgent, stop);
  return hipGetLastError();
}

// Enough work for every CU several times over at the given tile size
static int pick_tile(int m, int n, int cus) {
  long long t64 = (long long)((m + 63) / 64) * ((n + 63) / 64);
  long long t32 = (long long)((m + 31) / 32) * ((n + 31) / 32);
  if (t64 >= 2LL * cus) return 64;
  if (t32 >= 2LL * cus) return 32;
  return 16;
}

template <bool TA, bool TB, bool SKIP0>
static hipError_t launch_panel(int m, int n, int k, uint32_t alpha, const uint32_t* A, int lda, const uint32_t* B,
                               int ldb, uint32_t beta, uint32_t* C, int ldc, hipStream_t s, int tri, int urgent,
                               const int* stop, int cap, bool& done) {
  done = false;
  hipMemPool_t pool = runtime().pool;
  if (!pool) return hipSuccess;
  const int tm = (m + 63) / 64, tn = (n + 31) / 32, ks = (k + 15) / 16;
  const long long blocks = (long long)tm * tn;
  if (blocks > 0x7FFFFFFF || (long long)tm * ks > 0x7FFFFFFF || (long long)tn * ks > 0x7FFFFFFF) return hipSuccess;
  const size_t a_elems = (size_t)tm * ks * 16 * 64, t_elems = (size_t)tn * ks * 16 * 32;
  void* ws = nullptr;
  if (hipMallocFromPoolAsync(&ws, (a_elems + t_elems) * sizeof(double), pool, s) != hipSuccess) {
    (void)hipGetLastError();  // no workspace: the tile decodes its own slices
    return hipSuccess;
  }
  done = true;
  Panels pn{static_cast<double*>(ws), static_cast<double*>(ws) + a_elems, ks};
  hipLaunchKernelGGL((k_panel_a<TA>), dim3(tm * ks), dim3(256), 0, s, m, k, A, lda, const_cast<double*>(pn.a), ks);
  hipLaunchKernelGGL((k_panel_t<TB, TA, SKIP0>), dim3(tn * ks), dim3(256), 0, s, n, k, alpha, B, ldb,
                     const_cast<double*>(pn.t), ks);
  const unsigned dyn = pga::lds_for_cap((const void*)rgemm_tiled_panel<TA, TB, SKIP0>, cap);
  hipLaunchKernelGGL((rgemm_tiled_panel<TA, TB, SKIP0>), dim3((unsigned)blocks), dim3(256), dyn, s, m, n, k, alpha,
                     A, lda, B, ldb, beta, C, ldc, tm, tri, urgent, stop, pn);
  hipError_t e = hipGetLastError();
  hipError_t f = hipFreeAsync(ws, s);
  return e != hipSuccess ? e : f;
}

template <bool TA, bool TB, bool SKIP0>
static hipError_t launch_tiled(int m, int n, int k, uint32_t alpha, const uint32_t* A, int lda,
                               const uint32_t* B, int ldb, uint32_t beta, uint32_t* C, int ldc,
                               hipStream_t s, int tri, int urgent, const int* stop, int tile, int cap) {
  // panels from m*n*k = 2^30 on (runtime().panel_min_work): below it the two extra launches
  // and the workspace cost more than the staging they save
  if (tile == 64 && runtime().gemm_panels && (long long)m * n * k >= runtime().panel_min_work) {
    bool done;
    hipError_t e = launch_panel<TA, TB, SKIP0>(m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, s, tri, urgent, stop,
                                               cap, done);
    if (done || e != hipSuccess) return e;
  }
  if (tile == 64)
    return launch_tiled_cfg<64, 4, TA, TB, SKIP0>(m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, s, tri, urgent, stop,
                                                  cap);
  if (tile == 32)
    return launch_tiled_cfg<32, 2, TA, TB, SKIP0>(m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, s, tri, urgent, stop,
                                                  cap);
  return launch_tiled_cfg<16, 1, TA, TB, SKIP0>(m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, s, tri, urgent, stop,
                                                  cap);
}

template <bool TA, bool TB, bool SKIP0>
static hipError_t launch_exact(int m, int n, int k, uint32_t alpha, const uint32_t* A, int lda,
                               const uint32_t* B, int ldb, uint32_t beta, uint32_t* C, int ldc,
                               hipStream_t s, int tri, const int* stop) {
  dim3 grid((m + 15) / 16, (n + 15) / 16);
  if (grid.y > 65535) return hipErrorInvalidValue;
  hipLaunchKernelGGL((rgemm_exact<TA, TB, SKIP0>), grid, dim3(256), 0, s, m, n, k, alpha, A, lda, B, ldb,
                     beta, C, ldc, tri, stop);
  return hipGetLastError();
}

template <bool SKIP0>
static hipError_t dispatch(bool ta, bool tb, int m, int n, int k, uint32_t alpha, const uint32_t* A,
                           int lda, const uint32_t* B, int ldb, uint32_t beta, uint32_t* C, int ldc,
                           hipStream_t s, GemmImpl impl, int tri, int urgent, const int* stop, int tile,
                           int cap) {
  if (impl == GemmImpl::Exact) {
    if (!ta && !tb) return launch_exact<false, false, SKIP0>(m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, s, tri, stop);
    if (!ta && tb) return launch_exact<false, true, SKIP0>(m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, s, tri, stop);
    if (ta && !tb) return launch_exact<true, false, SKIP0>(m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, s, tri, stop);
    return launch_exact<true, true, SKIP0>(m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, s, tri, stop);
  }
  if (!ta && !tb) return launch_tiled<false, false, SKIP0>(m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, s, tri, urgent, stop, tile, cap);
  if (!ta && tb) return launch_tiled<false, true, SKIP0>(m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, s, tri, urgent, stop, tile, cap);
  if (ta && !tb) return launch_tiled<true, false, SKIP0>(m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, s, tri, urgent, stop, tile, cap);
  return launch_tiled<true, true, SKIP0>(m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, s, tri, urgent, stop, tile, cap);
}

template <int BM, int TM, bool TA, bool TB>
static hipError_t launch_persistent(int grid, const BatchProblem* probs, const int* tile_start, int nprob,
                                    int total_tiles, int* counter, hipStream_t s) {
  // the 64 tile is 64 x 32 (4 x 2 per lane), as in the plain kernel
  constexpr int BN = BM == 64 ? 32 : BM, TN = BM == 64 ? 2 : TM;
  constexpr int NT = (BM / TM) * (BN / TN);
  hipLaunchKernelGGL(reset_counter, dim3(1), dim3(64), 0, s, counter);
  hipLaunchKernelGGL((rgemm_batched_persistent<BM, BN, 16, TM, TN, TA, TB>), dim3(grid), dim3(NT), 0, s, probs,
                     tile_start, nprob, total_tiles, counter);
  return hipGetLastError();
}

template <bool TA, bool TB>
static hipError_t launch_persistent_tile(int tile, int grid, const BatchProblem* probs, const int* tile_start,
                                         int nprob, int total_tiles, int* counter, hipStream_t s) {
  if (tile == 64) return launch_persistent<64, 4, TA, TB>(grid, probs, tile_start, nprob, total_tiles, counter, s);
  if (tile == 32) return launch_persistent<32, 2, TA, TB>(grid, probs, tile_start, nprob, total_tiles, counter, s);
  return launch_persistent<16, 1, TA, TB>(grid, probs, tile_start, nprob, total_tiles, counter, s);
}

}  // namespace gemm

// Tile edges chosen by pga_rgemm_autotune, per exact GEMM shape
namespace {
std::mutex g_shape_mu;
std::map<std::tuple<bool, bool, int, int, int>, int> g_shape_tile;
}  // namespace

int tuned_tile(bool ta, bool tb, int m, int n, int k) {
  std::lock_guard<std::mutex> lk(g_shape_mu);
  if (g_shape_tile.empty()) return 0;
  auto it = g_shape_tile.find(std::make_tuple(ta, tb, m, n, k));
  return it == g_shape_tile.end() ? 0 : it->second;
}
void set_tuned_tile(bool ta, bool tb, int m, int n, int k, int tile) {
  std::lock_guard<std::mutex> lk(g_shape_mu);
  g_shape_tile[std::make_tuple(ta, tb, m, n, k)] = tile;
}

unsigned lds_for_cap(const void* kernel, int cap) {
  if (cap <= 0) return 0;
  hipFuncAttributes fa;
  if (hipFuncGetAttributes(&fa, kernel) != hipSuccess) {
    (void)hipGetLastError();
    return 0;
  }
  const long need = (long)runtime().lds_per_cu / (cap + 1) + 1 - (long)fa.sharedSizeBytes;
  return need > 0 ? (unsigned)need : 0u;
}

// tiles of one problem for the persistent kernel's tile size `tile` (64 means 64 x 32)
int batched_tiles(int m, int n, int tile) {
  const int tn = tile == 64 ? 32 : tile;
  return ((m + tile - 1) / tile) * ((n + tn - 1) / tn);
}

hipError_t batched_launch(bool ta, bool tb, int tile, int grid, const BatchProblem* probs, const int* tile_start,
                          int nprob, int total_tiles, int* counter, hipStream_t s) {
  using namespace gemm;
  if (grid < 1) grid = 1;
  if (!ta && !tb) return launch_persistent_tile<false, false>(tile, grid, probs, tile_start, nprob, total_tiles, counter, s);
  if (!ta && tb) return launch_persistent_tile<false, true>(tile, grid, probs, tile_start, nprob, total_tiles, counter, s);
  if (ta && !tb) return launch_persistent_tile<true, false>(tile, grid, probs, tile_start, nprob, total_tiles, counter, s);
  return launch_persistent_tile<true, true>(tile, grid, probs, tile_start, nprob, total_tiles, counter, s);
}

// Device-pointer GEMM with Rgemm's quick returns; the caller has validated arguments.
hipError_t rgemm_launch(bool ta, bool tb, int m, int n, int k, uint32_t alpha, const uint32_t* A,
                        int lda, const uint32_t* B, int ldb, uint32_t beta, uint32_t* C, int ldc,
                        hipStream_t s, GemmImpl impl, const GemmOpts& o) {
  using namespace gemm;
  if (m == 0 || n == 0 || ((alpha == 0 || k == 0) && beta == P_ONE)) return hipSuccess;
  if (alpha == 0) {
    for (int j0 = 0; j0 < n; j0 += 65535) {
      int nn = (n - j0 < 65535) ? n - j0 : 65535;
      hipLaunchKernelGGL(scale_c, dim3((m + 255) / 256, nn), dim3(256), 0, s, m, nn, beta,
                         C + (size_t)j0 * ldc, ldc, j0, o.tri, o.stop);
    }
    return hipGetLastError();
  }
  int tile = o.tile ? o.tile : runtime().gemm_tile;
  if (!tile) tile = tuned_tile(ta, tb, m, n, k);
  if (!tile) tile = pick_tile(m, n, runtime().cus);
  if (o.skip_zero_b)
    return dispatch<true>(ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, s, impl, o.tri, o.urgent, o.stop, tile,
                          o.cap_per_cu);
  return dispatch<false>(ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, s, impl, o.tri, o.urgent, o.stop, tile,
                         o.cap_per_cu);
}

}  // namespace pga
