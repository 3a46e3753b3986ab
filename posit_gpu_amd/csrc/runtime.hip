// runtime.hip -- HIP runtime layer and C ABI of the posit32 engine.
//
// Replaces the reference's OpenCL host side (libRpack/Rgemm_opencl.cpp and
// libRpack/OclHelper.hpp): device selection, one stream, device buffers that grow on
// demand, host<->device copies of only the logical sub-matrices, kernel dispatch.
// Kernels are compiled offline for gfx950; nothing is compiled at run time.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>

#include "pga_internal.hpp"
#include "posit32_core.hpp"

namespace pga {

static Runtime g_rt;
static thread_local int t_illegal = 0;

Runtime& runtime() { return g_rt; }
std::recursive_mutex& host_call_mutex() {
  static std::recursive_mutex m;
  return m;
}
void set_illegal_argument(int index) { t_illegal = index; }

pga_status from_hip(hipError_t e) {
  if (e == hipSuccess) return PGA_SUCCESS;
  if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) return PGA_ERROR_OUT_OF_MEMORY;
  if (e == hipErrorNoDevice || e == hipErrorInvalidDevice) return PGA_ERROR_NO_DEVICE;
  return PGA_ERROR_HIP;
}

struct ScratchSlot {
  void* ptr = nullptr;
  size_t bytes = 0;
};
static ScratchSlot g_scratch[8];

// Host <-> device copies of a rows x cols column-major block (leading dimension ld on
// the host, rows on the device) for the synchronous host-pointer entry points.  The
// columns pass through one pinned staging buffer in a single DMA: a pitched copy of
// pageable memory moves one column at a time.  Each copy has finished when the call
// returns, so the one buffer serves every call in turn.
static std::mutex g_stage_mu;
static void* g_stage = nullptr;
static size_t g_stage_bytes = 0;

static hipError_t stage_reserve(size_t bytes) {
  if (bytes <= g_stage_bytes) return hipSuccess;
  if (g_stage) (void)hipHostFree(g_stage);
  g_stage = nullptr;
  g_stage_bytes = 0;
  const size_t want = bytes + bytes / 2;
  hipError_t e = hipHostMalloc(&g_stage, want, hipHostMallocDefault);
  if (e != hipSuccess) {
    g_stage = nullptr;
    return e;
  }
  g_stage_bytes = want;
  return hipSuccess;
}

hipError_t copy_h2d_block(uint32_t* d, const uint32_t* h, int rows, int cols, int ld, hipStream_t s) {
  if (rows <= 0 || cols <= 0) return hipSuccess;
  const size_t colb = (size_t)rows * sizeof(uint32_t), bytes = colb * cols;
  std::lock_guard<std::mutex> lk(g_stage_mu);
  hipError_t e = stage_reserve(bytes);
  if (e != hipSuccess) return e;
  for (int c = 0; c < cols; ++c) memcpy((char*)g_stage + c * colb, h + (size_t)c * ld, colb);
  if ((e = hipMemcpyAsync(d, g_stage, bytes, hipMemcpyHostToDevice, s)) != hipSuccess) return e;
  return hipStreamSynchronize(s);
}

hipError_t copy_d2h_block(uint32_t* h, const uint32_t* d, int rows, int cols, int ld, hipStream_t s) {
  if (rows <= 0 || cols <= 0) return hipSuccess;
  const size_t colb = (size_t)rows * sizeof(uint32_t), bytes = colb * cols;
  std::lock_guard<std::mutex> lk(g_stage_mu);
  hipError_t e = stage_reserve(bytes);
  if (e != hipSuccess) return e;
  if ((e = hipMemcpyAsync(g_stage, d, bytes, hipMemcpyDeviceToHost, s)) != hipSuccess) return e;
  if ((e = hipStreamSynchronize(s)) != hipSuccess) return e;
  for (int c = 0; c < cols; ++c) memcpy(h + (size_t)c * ld, (const char*)g_stage + c * colb, colb);
  return hipSuccess;
}

void* scratch(int slot, size_t bytes, pga_status* st) {
  *st = PGA_SUCCESS;
  ScratchSlot& s = g_scratch[slot];
  if (bytes <= s.bytes) return s.ptr;
  if (s.ptr) {
    (void)hipStreamSynchronize(g_rt.stream);
    (void)hipFree(s.ptr);
    s.ptr = nullptr;
    s.bytes = 0;
  }
  size_t want = bytes + bytes / 2;
  if (hipMalloc(&s.ptr, want) != hipSuccess) {
    (void)hipGetLastError();
    if (hipMalloc(&s.ptr, bytes) != hipSuccess) {
      (void)hipGetLastError();
      s.ptr = nullptr;
      *st = PGA_ERROR_OUT_OF_MEMORY;
      return nullptr;
    }
    want = bytes;
  }
  s.bytes = want;
  return s.ptr;
}

// ---------------------------------------------------------------------------------
// elementwise kernels
// ---------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_from_f64(const double* x, uint32_t* p, size_t n) {
  PGA_KERNEL_VGPR_FLOOR();
  for (size_t i = blockIdx.x * (size_t)256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256)
    p[i] = pga::from_double(x[i]);
}
__global__ __launch_bounds__(256) void k_to_f64(const uint32_t* p, double* x, size_t n) {
  PGA_KERNEL_VGPR_FLOOR();
  for (size_t i = blockIdx.x * (size_t)256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256)
    x[i] = pga::to_double(p[i]);
}
__global__ __launch_bounds__(256) void k_elementwise(int op, const uint32_t* a, const uint32_t* b,
                                                     uint32_t* o, size_t n) {
  PGA_KERNEL_VGPR_FLOOR();
  for (size_t i = blockIdx.x * (size_t)256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) {
    uint32_t x = a[i];
    uint32_t y = (op == 4) ? 0u : b[i];
    uint32_t r;
    switch (op) {
      case 0: r = pga::add(x, y); break;
      case 1: r = pga::sub(x, y); break;
      case 2: r = pga::mul(x, y); break;
      case 3: r = pga::div(x, y); break;
      default: r = pga::sqrt(x); break;
    }
    o[i] = r;
  }
}

// SplitMix64 of (seed, index) -> uniform double in [0, 1) with 53 random bits; the
// tests restate it in numpy to rebuild the same inputs on the host.
__device__ __forceinline__ double uniform01(uint64_t seed, uint64_t i) {
  uint64_t z = seed + (i + 1) * 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  return (double)(z >> 11) * 0x1.0p-53;
}

__global__ __launch_bounds__(256) void k_fill_uniform(uint32_t* p, size_t n, uint64_t seed,
                                                      double lo, double hi) {
  PGA_KERNEL_VGPR_FLOOR();
  for (size_t i = blockIdx.x * (size_t)256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256)
    p[i] = pga::from_double(lo + (hi - lo) * uniform01(seed, i));
}

static unsigned grid_for(size_t n) {
  size_t g = (n + 255) / 256;
  if (g > 65536) g = 65536;
  if (g == 0) g = 1;
  return (unsigned)g;
}

static hipStream_t as_stream(void* s) { return s ? (hipStream_t)s : g_rt.stream; }

static bool lsame(char a, char b) {
  auto up = [](char c) { return (c >= 'a' && c <= 'z') ? (char)(c - 'a' + 'A') : c; };
  return up(a) == up(b);
}

// Rgemm's argument check (libRpack/Rgemm.cpp:90-128); returns the Mxerbla index or 0.
static int check_gemm_args(char transa, char transb, int m, int n, int k, int lda, int ldb, int ldc,
                           bool& ta, bool& tb) {
  bool nota = lsame(transa, 'N'), notb = lsame(transb, 'N');
  int nrowa = nota ? m : k, nrowb = notb ? k : n;
  ta = !nota;
  tb = !notb;
  if (!nota && !lsame(transa, 'C') && !lsame(transa, 'T')) return 1;
  if (!notb && !lsame(transb, 'C') && !lsame(transb, 'T')) return 2;
  if (m < 0) return 3;
  if (n < 0) return 4;
  if (k < 0) return 5;
  if (lda < (nrowa > 1 ? nrowa : 1)) return 8;
  if (ldb < (nrowb > 1 ? nrowb : 1)) return 10;
  if (ldc < (m > 1 ? m : 1)) return 13;
  return 0;
}

}  // namespace pga

using namespace pga;

extern "C" {

pga_status pga_init(int device) {
  std::lock_guard<std::recursive_mutex> lk(host_call_mutex());
  if (g_rt.initialized) return PGA_SUCCESS;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count == 0) {
    (void)hipGetLastError();
    return PGA_ERROR_NO_DEVICE;
  }
  if (device < 0) {
    const char* env = getenv("PGA_DEVICE");
    device = env ? atoi(env) : 0;
  }
  if (device >= count) return PGA_ERROR_NO_DEVICE;
  PGA_HIP(hipSetDevice(device));
  PGA_HIP(hipStreamCreateWithFlags(&g_rt.stream, hipStreamNonBlocking));
  g_rt.device = device;
  {
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) {
      g_rt.cus = prop.multiProcessorCount;
      if (prop.maxSharedMemoryPerMultiProcessor > 0) g_rt.lds_per_cu = (int)prop.maxSharedMemoryPerMultiProcessor;
    }
  }
  // the fixed device array of the no-allocation batched GEMM, for this device
  PGA_HIP(batched_noalloc_setup(device));
  // a memory pool for the GEMM's operand panels that keeps what it has reserved (workspace
  // taken and returned in stream order on every large GEMM); without it the GEMM decodes
  // its operands in every tile
  {
    hipMemPoolProps props = {};
    props.allocType = hipMemAllocationTypePinned;
    props.handleTypes = hipMemHandleTypeNone;
    props.location.type = hipMemLocationTypeDevice;
    props.location.id = device;
    if (hipMemPoolCreate(&g_rt.pool, &props) == hipSuccess) {
      uint64_t keep = ~0ull;
      if (hipMemPoolSetAttribute(g_rt.pool, hipMemPoolAttrReleaseThreshold, &keep) != hipSuccess) {
        (void)hipMemPoolDestroy(g_rt.pool);
        g_rt.pool = nullptr;
      }
    } else {
      g_rt.pool = nullptr;
    }
    (void)hipGetLastError();
  }
  if (const char* pn = getenv("PGA_GEMM_PANELS")) g_rt.gemm_panels = atoi(pn) != 0;
  if (const char* pm = getenv("PGA_GEMM_PANEL_MIN")) g_rt.panel_min_work = atoll(pm);
  const char* impl = getenv("PGA_GEMM_IMPL");
  if (impl && strcmp(impl, "exact") == 0) g_rt.gemm_impl = GemmImpl::Exact;
  // the tile edge of the plain GEMM (the role of OPENCL_GEMM_BLOCKSIZE in the reference):
  // 16, 32 or 64; anything else keeps the per-shape choice.  K is never split, so every
  // tile gives the same bits.
  const char* tile = getenv("PGA_GEMM_TILE");
  if (tile) {
    const int t = atoi(tile);
    g_rt.gemm_tile = (t == 16 || t == 32 || t == 64) ? t : 0;
  }
  g_rt.initialized = true;
  return PGA_SUCCESS;
}

pga_status pga_finalize(void) {
  std::lock_guard<std::recursive_mutex> lk(host_call_mutex());
  if (!g_rt.initialized) return PGA_SUCCESS;
  (void)hipStreamSynchronize(g_rt.stream);
  for (auto& s : g_scratch) {
    if (s.ptr) (void)hipFree(s.ptr);
    s.ptr = nullptr;
    s.bytes = 0;
  }
  if (g_rt.pool) {
    (void)hipDeviceSynchronize();
    (void)hipMemPoolDestroy(g_rt.pool);
    g_rt.pool = nullptr;
  }
  (void)hipStreamDestroy(g_rt.stream);
  g_rt.stream = nullptr;
  g_rt.initialized = false;
  return PGA_SUCCESS;
}

int pga_is_initialized(void) { return g_rt.initialized ? 1 : 0; }

const char* pga_status_string(pga_status s) {
  switch (s) {
    case PGA_SUCCESS: return "success";
    case PGA_ERROR_ILLEGAL_ARGUMENT: return "illegal argument";
    case PGA_ERROR_NOT_INITIALIZED: return "not initialized";
    case PGA_ERROR_HIP: return "HIP runtime error";
    case PGA_ERROR_OUT_OF_MEMORY: return "out of device memory";
    case PGA_ERROR_NO_DEVICE: return "no usable HIP device";
    case PGA_ERROR_UNSUPPORTED: return "unsupported";
    case PGA_ERROR_EXCEEDS_SCRATCH: return "exceeds the fixed scratch of a no-allocation entry point";
  }
  return "unknown status";
}

int pga_illegal_argument(void) { return t_illegal; }

pga_status pga_get_device_info(pga_device_info* info) {
  PGA_CHECK_INIT();
  if (!info) return PGA_ERROR_ILLEGAL_ARGUMENT;
  hipDeviceProp_t p;
  PGA_HIP(hipGetDeviceProperties(&p, g_rt.device));
  memset(info, 0, sizeof *info);
  strncpy(info->name, p.name, sizeof info->name - 1);
  strncpy(info->arch, p.gcnArchName, sizeof info->arch - 1);
  info->compute_units = p.multiProcessorCount;
  info->clock_khz = p.clockRate;
  info->simd_per_cu = 4;
  info->lanes_per_simd = 16;
  info->int32_valu_peak = (double)p.multiProcessorCount * 4.0 * 16.0 * (double)p.clockRate * 1e3;
  info->hbm_bytes = p.totalGlobalMem;
  return PGA_SUCCESS;
}

pga_status pga_set_gemm_impl(int impl) {
  PGA_CHECK_INIT();
  if (impl != 0 && impl != 1) return PGA_ERROR_ILLEGAL_ARGUMENT;
  g_rt.gemm_impl = impl ? GemmImpl::Exact : GemmImpl::Tiled;
  return PGA_SUCCESS;
}

pga_status pga_malloc(void** dptr, size_t bytes) {
  PGA_CHECK_INIT();
  if (!dptr) return PGA_ERROR_ILLEGAL_ARGUMENT;
  PGA_HIP(hipMalloc(dptr, bytes ? bytes : 4));
  return PGA_SUCCESS;
}
pga_status pga_free(void* dptr) {
  PGA_CHECK_INIT();
  if (dptr) PGA_HIP(hipFree(dptr));
  return PGA_SUCCESS;
}
pga_status pga_memcpy_h2d(void* dst, const void* src, size_t bytes) {
  PGA_CHECK_INIT();
  PGA_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, g_rt.stream));
  PGA_HIP(hipStreamSynchronize(g_rt.stream));
  return PGA_SUCCESS;
}
pga_status pga_memcpy_d2h(void* dst, const void* src, size_t bytes) {
  PGA_CHECK_INIT();
  PGA_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, g_rt.stream));
  PGA_HIP(hipStreamSynchronize(g_rt.stream));
  return PGA_SUCCESS;
}
pga_status pga_memcpy_d2d(void* dst, const void* src, size_t bytes) {
  PGA_CHECK_INIT();
  PGA_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, g_rt.stream));
  PGA_HIP(hipStreamSynchronize(g_rt.stream));
  return PGA_SUCCESS;
}
pga_status pga_memset(void* dptr, int value, size_t bytes) {
  PGA_CHECK_INIT();
  PGA_HIP(hipMemsetAsync(dptr, value, bytes, g_rt.stream));
  PGA_HIP(hipStreamSynchronize(g_rt.stream));
  return PGA_SUCCESS;
}
pga_status pga_stream_create(void** stream) {
  PGA_CHECK_INIT();
  hipStream_t s;
  PGA_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  *stream = (void*)s;
  return PGA_SUCCESS;
}
pga_status pga_stream_destroy(void* stream) {
  PGA_CHECK_INIT();
  if (stream) PGA_HIP(hipStreamDestroy((hipStream_t)stream));
  return PGA_SUCCESS;
}
pga_status pga_stream_synchronize(void* stream) {
  PGA_CHECK_INIT();
  PGA_HIP(hipStreamSynchronize(as_stream(stream)));
  return PGA_SUCCESS;
}
pga_status pga_device_synchronize(void) {
  PGA_CHECK_INIT();
  PGA_HIP(hipDeviceSynchronize());
  return PGA_SUCCESS;
}
pga_status pga_event_create(void** event) {
  PGA_CHECK_INIT();
  hipEvent_t e;
  PGA_HIP(hipEventCreate(&e));
  *event = (void*)e;
  return PGA_SUCCESS;
}
pga_status pga_event_destroy(void* event) {
  PGA_CHECK_INIT();
  if (event) PGA_HIP(hipEventDestroy((hipEvent_t)event));
  return PGA_SUCCESS;
}
pga_status pga_event_record(void* event, void* stream) {
  PGA_CHECK_INIT();
  PGA_HIP(hipEventRecord((hipEvent_t)event, as_stream(stream)));
  return PGA_SUCCESS;
}
pga_status pga_event_synchronize(void* event) {
  PGA_CHECK_INIT();
  PGA_HIP(hipEventSynchronize((hipEvent_t)event));
  return PGA_SUCCESS;
}
pga_status pga_event_elapsed_ms(float* ms, void* start, void* stop) {
  PGA_CHECK_INIT();
  PGA_HIP(hipEventElapsedTime(ms, (hipEvent_t)start, (hipEvent_t)stop));
  return PGA_SUCCESS;
}

pga_status pga_rgemm_device(char transa, char transb, int m, int n, int k, uint32_t alpha,
                            const uint32_t* a, int lda, const uint32_t* b, int ldb, uint32_t beta,
                            uint32_t* c, int ldc, void* stream) {
  PGA_CHECK_INIT();
  bool ta, tb;
  int bad = check_gemm_args(transa, transb, m, n, k, lda, ldb, ldc, ta, tb);
  if (bad) {
    set_illegal_argument(bad);
    return PGA_ERROR_ILLEGAL_ARGUMENT;
  }
  PGA_HIP(rgemm_launch(ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, as_stream(stream),
                       g_rt.gemm_impl));
  return PGA_SUCCESS;
}

pga_status pga_rgemm_autotune(char transa, char transb, int m, int n, int k, uint32_t alpha, uint32_t beta,
                              int* tile) {
  PGA_CHECK_INIT();
  bool ta, tb;
  const bool ta0 = !(transa == 'N' || transa == 'n'), tb0 = !(transb == 'N' || transb == 'n');
  const int lda = std::max(1, ta0 ? k : m), ldb = std::max(1, tb0 ? n : k), ldc = std::max(1, m);
  int bad = check_gemm_args(transa, transb, m, n, k, lda, ldb, ldc, ta, tb);
  if (bad) {
    set_illegal_argument(bad);
    return PGA_ERROR_ILLEGAL_ARGUMENT;
  }
  int best_tile = 0;
  if (m > 0 && n > 0 && k > 0) {
    hipStream_t s = g_rt.stream;
    uint32_t *A = nullptr, *B = nullptr, *C = nullptr;
    const size_t na = (size_t)lda * (ta ? m : k), nb = (size_t)ldb * (tb ? k : n), nc = (size_t)ldc * n;
    hipError_t e = hipMalloc((void**)&A, na * 4);
    if (e == hipSuccess) e = hipMalloc((void**)&B, nb * 4);
    if (e == hipSuccess) e = hipMalloc((void**)&C, nc * 4);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e == hipSuccess) {
      hipLaunchKernelGGL(k_fill_uniform, dim3(grid_for(na)), dim3(256), 0, s, A, na, 11ull, -1.0, 1.0);
      hipLaunchKernelGGL(k_fill_uniform, dim3(grid_for(nb)), dim3(256), 0, s, B, nb, 12ull, -1.0, 1.0);
      hipLaunchKernelGGL(k_fill_uniform, dim3(grid_for(nc)), dim3(256), 0, s, C, nc, 13ull, -1.0, 1.0);
      e = hipGetLastError();
    }
    float best = 1e30f;
    for (int t : {64, 32, 16}) {
      if (e != hipSuccess) break;
      GemmOpts o;
      o.tile = t;
      e = rgemm_launch(ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, s, g_rt.gemm_impl, o);  // warm-up
      if (e != hipSuccess) break;
      (void)hipEventRecord(e0, s);
      for (int r = 0; r < 3 && e == hipSuccess; ++r)
        e = rgemm_launch(ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, s, g_rt.gemm_impl, o);
      (void)hipEventRecord(e1, s);
      if (e == hipSuccess) e = hipEventSynchronize(e1);
      float ms = 0.f;
      (void)hipEventElapsedTime(&ms, e0, e1);
      if (e == hipSuccess && ms < best) {
        best = ms;
        best_tile = t;
      }
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipStreamSynchronize(s);
    if (A) (void)hipFree(A);
    if (B) (void)hipFree(B);
    if (C) (void)hipFree(C);
    PGA_HIP(e);
    set_tuned_tile(ta, tb, m, n, k, best_tile);
  }
  if (tile) *tile = best_tile;
  return PGA_SUCCESS;
}

pga_status pga_rgemm(char transa, char transb, int m, int n, int k, uint32_t alpha,
                     const uint32_t* a, int lda, const uint32_t* b, int ldb, uint32_t beta,
                     uint32_t* c, int ldc) {
  PGA_CHECK_INIT();
  std::lock_guard<std::recursive_mutex> lk(host_call_mutex());
  bool ta, tb;
  int bad = check_gemm_args(transa, transb, m, n, k, lda, ldb, ldc, ta, tb);
  if (bad) {
    set_illegal_argument(bad);
    return PGA_ERROR_ILLEGAL_ARGUMENT;
  }
  if (m == 0 || n == 0 || ((alpha == 0 || k == 0) && beta == P_ONE)) return PGA_SUCCESS;
  hipStream_t s = g_rt.stream;
  const size_t w = sizeof(uint32_t);
  int arows = ta ? k : m, acols = ta ? m : k;
  int brows = tb ? n : k, bcols = tb ? k : n;
  bool need_ab = (alpha != 0) && k > 0;
  bool read_c = (beta != 0);
  pga_status st;
  uint32_t* dA = nullptr;
  uint32_t* dB = nullptr;
  if (need_ab) {
    dA = (uint32_t*)scratch(0, (size_t)arows * acols * w, &st);
    if (st) return st;
    dB = (uint32_t*)scratch(1, (size_t)brows * bcols * w, &st);
    if (st) return st;
    PGA_HIP(copy_h2d_block(dA, a, arows, acols, lda, s));
    PGA_HIP(copy_h2d_block(dB, b, brows, bcols, ldb, s));
  }
  uint32_t* dC = (uint32_t*)scratch(2, (size_t)m * n * w, &st);
  if (st) return st;
  if (read_c)
    PGA_HIP(copy_h2d_block(dC, c, m, n, ldc, s));
  PGA_HIP(rgemm_launch(ta, tb, m, n, k, alpha, dA, arows > 0 ? arows : 1, dB,
                       brows > 0 ? brows : 1, beta, dC, m, s, g_rt.gemm_impl));
  PGA_HIP(copy_d2h_block(c, dC, m, n, ldc, s));
  return PGA_SUCCESS;
}

pga_status pga_p32_from_f64_device(const double* x, uint32_t* p, size_t n, void* stream) {
  PGA_CHECK_INIT();
  if (n == 0) return PGA_SUCCESS;
  hipLaunchKernelGGL(k_from_f64, dim3(grid_for(n)), dim3(256), 0, as_stream(stream), x, p, n);
  PGA_HIP(hipGetLastError());
  return PGA_SUCCESS;
}

pga_status pga_p32_to_f64_device(const uint32_t* p, double* x, size_t n, void* stream) {
  PGA_CHECK_INIT();
  if (n == 0) return PGA_SUCCESS;
  hipLaunchKernelGGL(k_to_f64, dim3(grid_for(n)), dim3(256), 0, as_stream(stream), p, x, n);
  PGA_HIP(hipGetLastError());
  return PGA_SUCCESS;
}

pga_status pga_p32_elementwise_device(int op, const uint32_t* a, const uint32_t* b, uint32_t* out,
                                      size_t n, void* stream) {
  PGA_CHECK_INIT();
  if (op < 0 || op > 4) {
    set_illegal_argument(1);
    return PGA_ERROR_ILLEGAL_ARGUMENT;
  }
  if (n == 0) return PGA_SUCCESS;
  hipLaunchKernelGGL(k_elementwise, dim3(grid_for(n)), dim3(256), 0, as_stream(stream), op, a, b,
                     out, n);
  PGA_HIP(hipGetLastError());
  return PGA_SUCCESS;
}

pga_status pga_fill_uniform_p32_device(uint32_t* p, size_t n, uint64_t seed, double lo, double hi,
                                       void* stream) {
  PGA_CHECK_INIT();
  if (n == 0) return PGA_SUCCESS;
  hipLaunchKernelGGL(k_fill_uniform, dim3(grid_for(n)), dim3(256), 0, as_stream(stream), p, n, seed,
                     lo, hi);
  PGA_HIP(hipGetLastError());
  return PGA_SUCCESS;
}

uint32_t pga_p32_add(uint32_t a, uint32_t b) { return pga::add(a, b); }
uint32_t pga_p32_sub(uint32_t a, uint32_t b) { return pga::sub(a, b); }
uint32_t pga_p32_mul(uint32_t a, uint32_t b) { return pga::mul(a, b); }
uint32_t pga_p32_div(uint32_t a, uint32_t b) { return pga::div(a, b); }
uint32_t pga_p32_sqrt(uint32_t a) { return pga::sqrt(a); }
uint32_t pga_p32_from_double(double x) { return pga::from_double(x); }
double pga_p32_to_double(uint32_t a) { return pga::to_double(a); }

}  // extern "C"
