// runtime.hip -- the HIP runtime behind include/vkdfft.h.
//
// One library, one device at a time, default-priority streams.  Kernels are
// compiled with hiprtc for the device's gfx arch and loaded as code objects;
// FFT plans bind a kernel to its launch geometry and to twiddle tables that
// are generated here in FP64 (sincospi in double, rounded once to the storage
// precision).
#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>
#include <hip/hiprtc.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/vkdfft.h"
#include "vkx_error.h"

namespace {

thread_local std::string g_err;

int fail(const char* where, const std::string& what) {
  g_err = std::string(where) + ": " + what;
  return VKX_ERR;
}

int hip_fail(const char* where, hipError_t e) {
  return fail(where, std::string(hipGetErrorName(e)) + " (" + hipGetErrorString(e) + ")");
}

#define VKX_HIP(call)                              \
  do {                                             \
    hipError_t e_ = (call);                        \
    if (e_ != hipSuccess) return hip_fail(#call, e_); \
  } while (0)

inline hipStream_t S(uint64_t s) { return reinterpret_cast<hipStream_t>(s); }

constexpr int kParamPtrs = 16;
constexpr int kParamBytes = 256;
struct alignas(16) KParams {
  uint64_t p[kParamPtrs];
  unsigned char c[kParamBytes];
};
constexpr int kMaxUserBufs = 12;
constexpr int kSlotRoots = 13;
constexpr int kSlotTwFwd = 14;
constexpr int kSlotTwInv = 15;

// ---------------------------------------------------------------------------
// FP64 twiddle / chirp generation
// ---------------------------------------------------------------------------
struct PassInfo {
  int npass;
  int radix[VKX_MAX_PASSES];
};

// Entry i of the table belongs to the pass p with NS_p <= i+1 < NS_p*R_p.
template <typename T>
__global__ void twiddle_kernel(T* out, int n, PassInfo info) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n - 1) return;
  long long ns = 1;
  for (int p = 0; p < info.npass; ++p) {
    const long long span = ns * info.radix[p];
    if (i + 1 < span) {
      const long long local = i - (ns - 1);
      const long long r = local / ns + 1;
      const long long k = local % ns;
      const long long m = (r * k) % span;
      double s, c;
      sincospi(-2.0 * (double)m / (double)span, &s, &c);
      out[2 * i] = (T)c;
      out[2 * i + 1] = (T)s;
      return;
    }
    ns = span;
  }
}

// roots[k] = exp(-2 pi i k / n), k in [0, n)
template <typename T>
__global__ void roots_kernel(T* out, int n) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  double s, c;
  sincospi(-2.0 * (double)k / (double)n, &s, &c);
  out[2 * k] = (T)c;
  out[2 * k + 1] = (T)s;
}

// dst[i] = sum_b src[b * count + i] (fixed summation order: deterministic)
template <typename T>
__global__ void batch_sum_kernel(T* __restrict__ dst, const T* __restrict__ src, int batch, long long count) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  T acc = src[i];
  for (int b = 1; b < batch; ++b) acc += src[(long long)b * count + i];
  dst[i] = acc;
}

// w[m] = exp(sign * i * pi * m^2 / n) with m^2 reduced mod 2n in integers.
template <typename T>
__global__ void chirp_kernel(T* out, int n, int count, int sign) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= count) return;
  const unsigned long long two_n = 2ull * (unsigned long long)n;
  const unsigned long long mm = ((unsigned long long)m * (unsigned long long)m) % two_n;
  double s, c;
  sincospi((double)sign * (double)mm / (double)n, &s, &c);
  out[2 * m] = (T)c;
  out[2 * m + 1] = (T)s;
}

// ---------------------------------------------------------------------------
// plans
// ---------------------------------------------------------------------------
struct TwKey {
  int n;
  int dbl;
  std::vector<int> radices;
  bool operator<(const TwKey& o) const {
    if (n != o.n) return n < o.n;
    if (dbl != o.dbl) return dbl < o.dbl;
    return radices < o.radices;
  }
};

struct Plan {
  hipFunction_t fn;
  uint32_t grid[3];
  uint32_t block[3];
  uint32_t shared;
  uint64_t tw_fwd;
  uint64_t tw_inv;
  uint64_t roots;
};

std::mutex g_mu;
std::map<TwKey, uint64_t> g_tw_cache;  // device-resident twiddle tables, kept for the process
std::map<std::pair<int, int>, uint64_t> g_roots_cache;
hipStream_t g_table_stream = nullptr;  // the library's own stream for table generation

// Runs gen(stream) on the library's private non-blocking stream and waits for that
// stream only.  Table generation therefore never touches the null stream or a
// caller's stream: a plan may be created while the caller's stream is being
// captured into a graph (torch.cuda.graph, hipStreamBeginCapture), and nothing of
// the generation is recorded into that graph.  Called with g_mu held.
template <class F> int fill_table(uint64_t /*dst*/, F&& gen) {
  if (!g_table_stream) VKX_HIP(hipStreamCreateWithFlags(&g_table_stream, hipStreamNonBlocking));
  if (const int rc = gen(g_table_stream)) return rc;
  VKX_HIP(hipStreamSynchronize(g_table_stream));
  return 0;
}

int get_roots(uint64_t* out, int n, int dbl) {
  *out = 0;
  if (n <= 0) return 0;
  std::lock_guard<std::mutex> lk(g_mu);
  auto key = std::make_pair(n, dbl);
  auto it = g_roots_cache.find(key);
  if (it != g_roots_cache.end()) {
    *out = it->second;
    return 0;
  }
  void* p = nullptr;
  VKX_HIP(hipMalloc(&p, (size_t)n * (dbl ? 16 : 8)));
  const uint64_t dptr = reinterpret_cast<uint64_t>(p);
  if (const int rc = fill_table(dptr, [&](hipStream_t s) { return vkx_roots(dptr, n, dbl, (uint64_t)s); })) {
    (void)hipFree(p);
    return rc;
  }
  g_roots_cache[key] = dptr;
  *out = dptr;
  return 0;
}

int get_twiddles(uint64_t* out, int n, const int* radices, int npass, int dbl) {
  *out = 0;
  if (n <= 1 || npass <= 0) return 0;
  TwKey key{n, dbl, std::vector<int>(radices, radices + npass)};
  std::lock_guard<std::mutex> lk(g_mu);
  auto it = g_tw_cache.find(key);
  if (it != g_tw_cache.end()) {
    *out = it->second;
    return 0;
  }
  const size_t bytes = (size_t)(n - 1) * (dbl ? 16 : 8);
  void* p = nullptr;
  VKX_HIP(hipMalloc(&p, bytes));
  const uint64_t dptr = reinterpret_cast<uint64_t>(p);
  if (const int rc = fill_table(dptr, [&](hipStream_t s) { return vkx_twiddles(dptr, n, radices, npass, dbl, (uint64_t)s); })) {
    (void)hipFree(p);
    return rc;
  }
  g_tw_cache[key] = dptr;
  *out = dptr;
  return 0;
}

}  // namespace

int vkx_set_error(const char* where, const char* what) { return fail(where, what); }
int vkx_set_hip_error(const char* where, hipError_t e) { return hip_fail(where, e); }

extern "C" {

const char* vkx_last_error(void) { return g_err.c_str(); }
void vkx_clear_error(void) { g_err.clear(); }
int vkx_version(void) { return 100; }

int vkx_init(int device) {
  int count = 0;
  VKX_HIP(hipGetDeviceCount(&count));
  if (device < 0 || device >= count) return fail("vkx_init", "device index out of range");
  VKX_HIP(hipSetDevice(device));
  VKX_HIP(hipFree(nullptr));
  return 0;
}

int vkx_device_count(int* count) {
  *count = 0;
  hipError_t e = hipGetDeviceCount(count);
  if (e == hipErrorNoDevice) { *count = 0; return 0; }
  if (e != hipSuccess) return hip_fail("hipGetDeviceCount", e);
  return 0;
}

int vkx_device_props_get(int device, vkx_device_props* out) {
  hipDeviceProp_t p;
  VKX_HIP(hipGetDeviceProperties(&p, device));
  std::memset(out, 0, sizeof(*out));
  std::snprintf(out->name, sizeof(out->name), "%s", p.name);
  std::snprintf(out->arch, sizeof(out->arch), "%s", p.gcnArchName);
  out->compute_units = p.multiProcessorCount;
  out->wave_size = p.warpSize;
  out->max_threads_per_block = p.maxThreadsPerBlock;
  out->max_shared_per_block = (int)p.sharedMemPerBlock;
  out->l2_bytes = p.l2CacheSize;
  out->clock_khz = p.clockRate;
  out->mem_clock_khz = p.memoryClockRate;
  out->mem_bus_bits = p.memoryBusWidth;
  out->total_mem = p.totalGlobalMem;
  out->supports_fp64 = 1;
  out->supports_fp16 = 1;
  return 0;
}

int vkx_malloc(uint64_t* dptr, size_t bytes) {
  void* p = nullptr;
  VKX_HIP(hipMalloc(&p, bytes ? bytes : 1));
  *dptr = reinterpret_cast<uint64_t>(p);
  return 0;
}

int vkx_free(uint64_t dptr) {
  if (dptr) VKX_HIP(hipFree(reinterpret_cast<void*>(dptr)));
  return 0;
}

int vkx_memcpy_htod(uint64_t dst, const void* src, size_t bytes, uint64_t stream) {
  VKX_HIP(hipMemcpyAsync(reinterpret_cast<void*>(dst), src, bytes, hipMemcpyHostToDevice, S(stream)));
  VKX_HIP(hipStreamSynchronize(S(stream)));
  return 0;
}

int vkx_memcpy_dtoh(void* dst, uint64_t src, size_t bytes, uint64_t stream) {
  VKX_HIP(hipMemcpyAsync(dst, reinterpret_cast<const void*>(src), bytes, hipMemcpyDeviceToHost, S(stream)));
  VKX_HIP(hipStreamSynchronize(S(stream)));
  return 0;
}

int vkx_memcpy_dtod(uint64_t dst, uint64_t src, size_t bytes, uint64_t stream) {
  VKX_HIP(hipMemcpyAsync(reinterpret_cast<void*>(dst), reinterpret_cast<const void*>(src), bytes,
                         hipMemcpyDeviceToDevice, S(stream)));
  return 0;
}

int vkx_memset(uint64_t dst, int value, size_t bytes, uint64_t stream) {
  VKX_HIP(hipMemsetAsync(reinterpret_cast<void*>(dst), value, bytes, S(stream)));
  return 0;
}

int vkx_stream_create(uint64_t* stream) {
  hipStream_t s;
  VKX_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  *stream = reinterpret_cast<uint64_t>(s);
  return 0;
}

int vkx_stream_destroy(uint64_t stream) {
  if (stream) VKX_HIP(hipStreamDestroy(S(stream)));
  return 0;
}

int vkx_stream_synchronize(uint64_t stream) {
  VKX_HIP(hipStreamSynchronize(S(stream)));
  return 0;
}

int vkx_device_synchronize(void) {
  VKX_HIP(hipDeviceSynchronize());
  return 0;
}

int vkx_event_create(uint64_t* event) {
  hipEvent_t e;
  VKX_HIP(hipEventCreate(&e));
  *event = reinterpret_cast<uint64_t>(e);
  return 0;
}

int vkx_event_destroy(uint64_t event) {
  if (event) VKX_HIP(hipEventDestroy(reinterpret_cast<hipEvent_t>(event)));
  return 0;
}

int vkx_event_record(uint64_t event, uint64_t stream) {
  VKX_HIP(hipEventRecord(reinterpret_cast<hipEvent_t>(event), S(stream)));
  return 0;
}

int vkx_stream_wait_event(uint64_t stream, uint64_t event) {
  VKX_HIP(hipStreamWaitEvent(S(stream), reinterpret_cast<hipEvent_t>(event), 0));
  return 0;
}

int vkx_event_synchronize(uint64_t event) {
  VKX_HIP(hipEventSynchronize(reinterpret_cast<hipEvent_t>(event)));
  return 0;
}

int vkx_event_elapsed_ms(float* ms, uint64_t start, uint64_t stop) {
  VKX_HIP(hipEventElapsedTime(ms, reinterpret_cast<hipEvent_t>(start), reinterpret_cast<hipEvent_t>(stop)));
  return 0;
}

int vkx_compile(const char* source, const char* name, const char* const* options, int num_options, void** code,
                size_t* code_size, char** log) {
  *code = nullptr;
  *code_size = 0;
  if (log) *log = nullptr;
  hiprtcProgram prog;
  hiprtcResult r = hiprtcCreateProgram(&prog, source, name, 0, nullptr, nullptr);
  if (r != HIPRTC_SUCCESS) return fail("hiprtcCreateProgram", hiprtcGetErrorString(r));
  r = hiprtcCompileProgram(prog, num_options, const_cast<const char**>(options));
  size_t log_size = 0;
  hiprtcGetProgramLogSize(prog, &log_size);
  if (log && log_size > 1) {
    *log = static_cast<char*>(std::malloc(log_size + 1));
    hiprtcGetProgramLog(prog, *log);
    (*log)[log_size] = 0;
  }
  if (r != HIPRTC_SUCCESS) {
    hiprtcDestroyProgram(&prog);
    return fail("hiprtcCompileProgram", hiprtcGetErrorString(r));
  }
  size_t n = 0;
  hiprtcGetCodeSize(prog, &n);
  void* buf = std::malloc(n);
  hiprtcGetCode(prog, static_cast<char*>(buf));
  hiprtcDestroyProgram(&prog);
  *code = buf;
  *code_size = n;
  return 0;
}

int vkx_compiler_version(int* version) {
  // hiprtc reports only its API level (the same across ROCm releases); the runtime it
  // ships with names the release, and both come from the same installation
  VKX_HIP(hipRuntimeGetVersion(version));
  return 0;
}

void vkx_free_host(void* p) { std::free(p); }

int vkx_module_load(uint64_t* module, const void* code, size_t code_size) {
  (void)code_size;
  hipModule_t m;
  VKX_HIP(hipModuleLoadData(&m, code));
  *module = reinterpret_cast<uint64_t>(m);
  return 0;
}

int vkx_module_unload(uint64_t module) {
  if (module) VKX_HIP(hipModuleUnload(reinterpret_cast<hipModule_t>(module)));
  return 0;
}

int vkx_module_function(uint64_t* function, uint64_t module, const char* name) {
  hipFunction_t f;
  VKX_HIP(hipModuleGetFunction(&f, reinterpret_cast<hipModule_t>(module), name));
  *function = reinterpret_cast<uint64_t>(f);
  return 0;
}

int vkx_function_attrs(uint64_t function, int* num_regs, int* shared_bytes, int* max_threads, int* local_bytes) {
  hipFunction_t f = reinterpret_cast<hipFunction_t>(function);
  VKX_HIP(hipFuncGetAttribute(num_regs, HIP_FUNC_ATTRIBUTE_NUM_REGS, f));
  VKX_HIP(hipFuncGetAttribute(shared_bytes, HIP_FUNC_ATTRIBUTE_SHARED_SIZE_BYTES, f));
  VKX_HIP(hipFuncGetAttribute(max_threads, HIP_FUNC_ATTRIBUTE_MAX_THREADS_PER_BLOCK, f));
  VKX_HIP(hipFuncGetAttribute(local_bytes, HIP_FUNC_ATTRIBUTE_LOCAL_SIZE_BYTES, f));
  return 0;
}

int vkx_launch(uint64_t function, const uint32_t grid[3], const uint32_t block[3], uint32_t shared_bytes,
               uint64_t stream, const void* params, size_t params_size) {
  size_t sz = params_size;
  void* cfg[] = {HIP_LAUNCH_PARAM_BUFFER_POINTER, const_cast<void*>(params), HIP_LAUNCH_PARAM_BUFFER_SIZE, &sz,
                 HIP_LAUNCH_PARAM_END};
  VKX_HIP(hipModuleLaunchKernel(reinterpret_cast<hipFunction_t>(function), grid[0], grid[1], grid[2], block[0],
                                block[1], block[2], shared_bytes, S(stream), nullptr, cfg));
  return 0;
}

int vkx_twiddles(uint64_t dst, int n, const int* radices, int num_passes, int double_precision, uint64_t stream) {
  if (num_passes > VKX_MAX_PASSES) return fail("vkx_twiddles", "too many passes");
  long long prod = 1;
  PassInfo info;
  info.npass = num_passes;
  for (int i = 0; i < num_passes; ++i) {
    if (radices[i] < 1) return fail("vkx_twiddles", "radix < 1");
    info.radix[i] = radices[i];
    prod *= radices[i];
  }
  if (prod != n) return fail("vkx_twiddles", "product of radices != n");
  if (n <= 1) return 0;
  const int count = n - 1;
  const int threads = 256;
  const int blocks = (count + threads - 1) / threads;
  if (double_precision)
    hipLaunchKernelGGL(twiddle_kernel<double>, dim3(blocks), dim3(threads), 0, S(stream),
                       reinterpret_cast<double*>(dst), n, info);
  else
    hipLaunchKernelGGL(twiddle_kernel<float>, dim3(blocks), dim3(threads), 0, S(stream),
                       reinterpret_cast<float*>(dst), n, info);
  VKX_HIP(hipGetLastError());
  return 0;
}

int vkx_roots(uint64_t dst, int n, int double_precision, uint64_t stream) {
  if (n < 1) return fail("vkx_roots", "bad size");
  const int threads = 256;
  const int blocks = (n + threads - 1) / threads;
  if (double_precision)
    hipLaunchKernelGGL(roots_kernel<double>, dim3(blocks), dim3(threads), 0, S(stream),
                       reinterpret_cast<double*>(dst), n);
  else
    hipLaunchKernelGGL(roots_kernel<float>, dim3(blocks), dim3(threads), 0, S(stream),
                       reinterpret_cast<float*>(dst), n);
  VKX_HIP(hipGetLastError());
  return 0;
}

int vkx_batch_sum(uint64_t dst, uint64_t src, int batch, long long count, int double_precision, uint64_t stream) {
  if (batch < 1 || count < 0) return fail("vkx_batch_sum", "bad size");
  if (count == 0) return 0;
  const int threads = 256;
  const long long blocks = (count + threads - 1) / threads;
  if (double_precision)
    hipLaunchKernelGGL(batch_sum_kernel<double>, dim3((unsigned)blocks), dim3(threads), 0, S(stream),
                       reinterpret_cast<double*>(dst), reinterpret_cast<const double*>(src), batch, count);
  else
    hipLaunchKernelGGL(batch_sum_kernel<float>, dim3((unsigned)blocks), dim3(threads), 0, S(stream),
                       reinterpret_cast<float*>(dst), reinterpret_cast<const float*>(src), batch, count);
  VKX_HIP(hipGetLastError());
  return 0;
}

int vkx_chirp(uint64_t dst, int n, int count, int sign, int double_precision, uint64_t stream) {
  if (n < 1 || count < 1) return fail("vkx_chirp", "bad size");
  const int threads = 256;
  const int blocks = (count + threads - 1) / threads;
  if (double_precision)
    hipLaunchKernelGGL(chirp_kernel<double>, dim3(blocks), dim3(threads), 0, S(stream),
                       reinterpret_cast<double*>(dst), n, count, sign);
  else
    hipLaunchKernelGGL(chirp_kernel<float>, dim3(blocks), dim3(threads), 0, S(stream),
                       reinterpret_cast<float*>(dst), n, count, sign);
  VKX_HIP(hipGetLastError());
  return 0;
}

int vkx_plan_create(uint64_t* plan, const vkx_plan_desc* d) {
  *plan = 0;
  if (!d->function) return fail("vkx_plan_create", "null function");
  Plan* p = new Plan();
  p->fn = reinterpret_cast<hipFunction_t>(d->function);
  for (int i = 0; i < 3; ++i) { p->grid[i] = d->grid[i]; p->block[i] = d->block[i]; }
  p->shared = d->shared_bytes;
  p->tw_fwd = p->tw_inv = p->roots = 0;
  if (d->roots_n > 0) {
    if (const int rc = get_roots(&p->roots, d->roots_n, d->double_precision)) { delete p; return rc; }
  }
  if (d->n > 1 && d->num_passes_fwd > 0) {
    if (const int rc = get_twiddles(&p->tw_fwd, d->n, d->radices_fwd, d->num_passes_fwd, d->double_precision)) {
      delete p;
      return rc;
    }
  }
  if (d->n > 1 && d->num_passes_inv > 0) {
    if (const int rc = get_twiddles(&p->tw_inv, d->n, d->radices_inv, d->num_passes_inv, d->double_precision)) {
      delete p;
      return rc;
    }
  }
  *plan = reinterpret_cast<uint64_t>(p);
  return 0;
}

int vkx_plan_destroy(uint64_t plan) {
  delete reinterpret_cast<Plan*>(plan);
  return 0;
}

}  // extern "C"

namespace {
// One launch of a plan; with start / stop events the dispatch itself carries the
// timestamps (hipExtModuleLaunchKernel), so no marker sits between launches.
int plan_launch(const char* where, uint64_t plan, const uint64_t* bufs, int num_bufs, const void* consts,
                size_t consts_size, uint64_t stream, hipEvent_t start, hipEvent_t stop) {
  const Plan* p = reinterpret_cast<const Plan*>(plan);
  if (!p) return fail(where, "null plan");
  if (num_bufs < 0 || num_bufs > kMaxUserBufs) return fail(where, "too many buffers (max 12)");
  if (consts_size > (size_t)kParamBytes) return fail(where, "constants exceed 256 bytes");
  KParams prm;
  std::memset(&prm, 0, sizeof(prm));
  for (int i = 0; i < num_bufs; ++i) prm.p[i] = bufs[i];
  prm.p[kSlotTwFwd] = p->tw_fwd;
  prm.p[kSlotTwInv] = p->tw_inv;
  prm.p[kSlotRoots] = p->roots;
  if (consts_size) std::memcpy(prm.c, consts, consts_size);
  size_t sz = sizeof(prm);
  void* cfg[] = {HIP_LAUNCH_PARAM_BUFFER_POINTER, &prm, HIP_LAUNCH_PARAM_BUFFER_SIZE, &sz, HIP_LAUNCH_PARAM_END};
  hipError_t e;
  if (start || stop) {
    e = hipExtModuleLaunchKernel(p->fn, p->grid[0] * p->block[0], p->grid[1] * p->block[1], p->grid[2] * p->block[2],
                                 p->block[0], p->block[1], p->block[2], p->shared, S(stream), nullptr, cfg, start,
                                 stop, 0);
  } else {
    e = hipModuleLaunchKernel(p->fn, p->grid[0], p->grid[1], p->grid[2], p->block[0], p->block[1], p->block[2],
                              p->shared, S(stream), nullptr, cfg);
  }
  return e == hipSuccess ? 0 : hip_fail(where, e);
}
}  // namespace

extern "C" {

int vkx_plan_exec(uint64_t plan, const uint64_t* bufs, int num_bufs, const void* consts, size_t consts_size,
                  uint64_t stream) {
  return plan_launch("vkx_plan_exec", plan, bufs, num_bufs, consts, consts_size, stream, nullptr, nullptr);
}

int vkx_plan_exec_timed(uint64_t plan, const uint64_t* bufs, int num_bufs, const void* consts, size_t consts_size,
                        uint64_t stream, uint64_t start_event, uint64_t stop_event) {
  if (!start_event || !stop_event) return fail("vkx_plan_exec_timed", "null event");
  return plan_launch("vkx_plan_exec_timed", plan, bufs, num_bufs, consts, consts_size, stream,
                     reinterpret_cast<hipEvent_t>(start_event), reinterpret_cast<hipEvent_t>(stop_event));
}

int vkx_stream_is_capturing(uint64_t stream, int* capturing) {
  *capturing = 0;
  hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
  VKX_HIP(hipStreamIsCapturing(S(stream), &st));
  *capturing = st == hipStreamCaptureStatusActive ? 1 : (st == hipStreamCaptureStatusInvalidated ? 2 : 0);
  return 0;
}

int vkx_graph_begin(uint64_t stream) {
  VKX_HIP(hipStreamBeginCapture(S(stream), hipStreamCaptureModeThreadLocal));
  return 0;
}

int vkx_graph_end(uint64_t stream, uint64_t* graph_exec) {
  hipGraph_t g;
  VKX_HIP(hipStreamEndCapture(S(stream), &g));
  hipGraphExec_t ge;
  hipError_t e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
  hipGraphDestroy(g);
  if (e != hipSuccess) return hip_fail("hipGraphInstantiate", e);
  *graph_exec = reinterpret_cast<uint64_t>(ge);
  return 0;
}

int vkx_graph_launch(uint64_t graph_exec, uint64_t stream) {
  VKX_HIP(hipGraphLaunch(reinterpret_cast<hipGraphExec_t>(graph_exec), S(stream)));
  return 0;
}

int vkx_graph_destroy(uint64_t graph_exec) {
  if (graph_exec) VKX_HIP(hipGraphExecDestroy(reinterpret_cast<hipGraphExec_t>(graph_exec)));
  return 0;
}

}  // extern "C"
