/* vkdfft.h -- C-ABI of the vkdispatch_amd FFT runtime (libvkdfft.so).
 *
 * This is the drop-in boundary under the vd.fft Python API.  Every entry point
 * replaces one function of the reference's backend "native" table that the
 * vd.fft path binds (the table is listed in
 * vkdispatch/backends/opencl_backend.py:2028-2080 and served in C++ by
 * vkdispatch_native/wrapper.pyx:1-7).  Conventions follow that table:
 * plain integers / pointers, 64-bit integer handles with 0 meaning "none",
 * and a last-error string the caller turns into an exception
 * (vkdispatch/base/errors.py:15-44).
 *
 * All functions return 0 on success and a nonzero code on failure; on failure
 * vkx_last_error() describes the problem.  The error string is thread-local
 * (the reference's is process-global and sticky,
 * vkdispatch_native/context/errors.cpp:5-39; a thread-local string keeps
 * unrelated threads from seeing each other's failures).
 *
 * Streams, events, modules, functions, plans and graph executables are opaque
 * uint64_t handles.  Device pointers are uint64_t.
 */
#ifndef VKDFFT_H
#define VKDFFT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- errors -------------------------------------------------------------- */
/* replaces native.get_error_string (vkdispatch/base/errors.py:15-44) */
const char* vkx_last_error(void);
void vkx_clear_error(void);

/* ---- device / context ------------------------------------------------------ */
/* replaces native.init + native.context_create
 * (vkdispatch/base/init.py:520-620, vkdispatch/base/context.py:360-429) */
int vkx_init(int device);
int vkx_device_count(int* count);

/* Device limits the FFT planner reads (the reference reads max_shared_memory,
 * max_workgroup_size, max_workgroup_invocations and subgroup_size,
 * vkdispatch/base/context.py:199-254). */
typedef struct vkx_device_props {
  char name[256];
  char arch[64];            /* e.g. "gfx950" */
  int compute_units;
  int wave_size;            /* 64 on CDNA */
  int max_threads_per_block;
  int max_shared_per_block; /* bytes of LDS one workgroup may allocate */
  int l2_bytes;
  int clock_khz;
  int mem_clock_khz;
  int mem_bus_bits;
  size_t total_mem;
  int supports_fp64;
  int supports_fp16;
} vkx_device_props;
int vkx_device_props_get(int device, vkx_device_props* out);

/* ---- memory ---------------------------------------------------------------- */
/* replaces native.buffer_create / buffer_destroy (vkdispatch/base/buffer.py:101-128,140-147) */
int vkx_malloc(uint64_t* dptr, size_t bytes);
int vkx_free(uint64_t dptr);
/* replace native.buffer_write_staging + buffer_write and buffer_read +
 * buffer_read_staging (vkdispatch/base/buffer.py:154-258) */
int vkx_memcpy_htod(uint64_t dst, const void* src, size_t bytes, uint64_t stream);
int vkx_memcpy_dtoh(void* dst, uint64_t src, size_t bytes, uint64_t stream);
int vkx_memcpy_dtod(uint64_t dst, uint64_t src, size_t bytes, uint64_t stream);
int vkx_memset(uint64_t dst, int value, size_t bytes, uint64_t stream);

/* ---- streams and events ----------------------------------------------------- */
/* replace the per-queue worker threads and signals
 * (vkdispatch_native/queue/queue.cpp:81-183, vkdispatch/base/buffer.py:119-126).
 * Streams are created at the default priority. */
int vkx_stream_create(uint64_t* stream);
int vkx_stream_destroy(uint64_t stream);
int vkx_stream_synchronize(uint64_t stream);
int vkx_device_synchronize(void);
int vkx_event_create(uint64_t* event);
int vkx_event_destroy(uint64_t event);
int vkx_event_record(uint64_t event, uint64_t stream);
/* work queued on `stream` after this call waits for `event` (replaces a queue waiting
 * on another queue's signal, native.signal_wait, vkdispatch/backends/opencl_backend.py:1372);
 * also joins a second stream into a graph capture of the stream that recorded `event` */
int vkx_stream_wait_event(uint64_t stream, uint64_t event);
int vkx_event_synchronize(uint64_t event);
int vkx_event_elapsed_ms(float* ms, uint64_t start, uint64_t stop);

/* ---- kernel compilation and loading ----------------------------------------- */
/* replaces native.stage_compute_plan_create's compile step: glslang->SPIR-V
 * (vkdispatch_native/stages/stage_compute.cpp:13-82) / cl.Program().build()
 * (vkdispatch/backends/opencl_backend.py:1823-1862).  Compiles HIP source for
 * the given offload arch with hiprtc.  On success *code (malloc'd, free with
 * vkx_free_host) holds the code object; *log always receives the compiler log
 * (may be NULL when empty; free with vkx_free_host). */
int vkx_compile(const char* source, const char* name, const char* const* options, int num_options,
                void** code, size_t* code_size, char** log);
/* release of the HIP installation whose hiprtc vkx_compile runs (hipRuntimeGetVersion;
 * part of the on-disk code-object cache key, so code built by one compiler release is
 * never loaded as another's); no counterpart in the reference, which keeps no on-disk
 * code-object cache */
int vkx_compiler_version(int* version);
void vkx_free_host(void* p);
/* pipeline creation half of stage_compute_plan_create (stage_compute.cpp:84-250) */
int vkx_module_load(uint64_t* module, const void* code, size_t code_size);
int vkx_module_unload(uint64_t module);
int vkx_module_function(uint64_t* function, uint64_t module, const char* name);
int vkx_function_attrs(uint64_t function, int* num_regs, int* shared_bytes, int* max_threads, int* local_bytes);

/* ---- launch ---------------------------------------------------------------- */
/* replaces native.stage_compute_record + command_list_submit
 * (vkdispatch_native/stages/stage_compute.cpp:281-342,
 *  vkdispatch/base/command_list.py:64-71,130-132).  params/params_size is the
 * kernel's single by-value argument (a packed struct). */
int vkx_launch(uint64_t function, const uint32_t grid[3], const uint32_t block[3], uint32_t shared_bytes,
               uint64_t stream, const void* params, size_t params_size);

/* ---- FFT plans -------------------------------------------------------------- */
/* A plan binds one compiled FFT / convolution / transpose kernel to its launch
 * geometry and its FP64-generated twiddle tables; it is the C-side half of the
 * reference's cached ShaderFunction (vkdispatch/fft/shader_factories.py:11-60,
 * vkdispatch/shader/shader_function.py:189-253). */
#define VKX_MAX_PASSES 24
typedef struct vkx_plan_desc {
  uint64_t function;              /* from vkx_module_function */
  uint32_t grid[3];
  uint32_t block[3];
  uint32_t shared_bytes;          /* dynamic LDS (kernels normally use static LDS: 0) */
  int n;                          /* transform length (0: no twiddles) */
  int double_precision;           /* twiddles as complex128 instead of complex64 */
  int num_passes_fwd;             /* radix list of the forward pass sequence */
  int radices_fwd[VKX_MAX_PASSES];
  int num_passes_inv;             /* radix list of the inverse pass sequence (0: none) */
  int radices_inv[VKX_MAX_PASSES];
  int roots_n;                    /* >0: bind a table exp(-2 pi i k / roots_n), k < roots_n */
} vkx_plan_desc;
int vkx_plan_create(uint64_t* plan, const vkx_plan_desc* desc);
int vkx_plan_destroy(uint64_t plan);
/* replaces descriptor_set_write_buffer + stage_compute_record + command_list_submit
 * for one FFT dispatch (vkdispatch/shader/shader_function.py:277-380).
 * bufs: up to 12 device pointers in the reference's declared argument order
 * (vkdispatch/fft/io_manager.py:86-102); consts: packed Var/Const values. */
int vkx_plan_exec(uint64_t plan, const uint64_t* bufs, int num_bufs, const void* consts, size_t consts_size,
                  uint64_t stream);
/* vkx_plan_exec with the dispatch's own start / stop timestamps in two events
 * (hipExtModuleLaunchKernel): per-kernel times inside a stream of back-to-back
 * launches, without a marker between the launches.  No counterpart in the
 * reference. */
int vkx_plan_exec_timed(uint64_t plan, const uint64_t* bufs, int num_bufs, const void* consts, size_t consts_size,
                        uint64_t stream, uint64_t start_event, uint64_t stop_event);

/* ---- twiddles and chirps (generated in FP64 on the GPU) ----------------------- */
/* Stockham twiddle table for a pass list: entry (NS_p - 1) + (r-1)*NS_p + k holds
 * exp(-2 pi i r k / (NS_p R_p)); N-1 entries in total. */
int vkx_twiddles(uint64_t dst, int n, const int* radices, int num_passes, int double_precision, uint64_t stream);
/* Roots of unity r[k] = exp(-2 pi i k / n), k in [0, n) (real-FFT pre/post-processing). */
int vkx_roots(uint64_t dst, int n, int double_precision, uint64_t stream);
/* dst[i] = sum over b < batch of src[b*count + i] for float (or double) arrays, summed
 * in a fixed order (deterministic).  No counterpart in the reference: used by the
 * hand-written convolution adjoint to reduce the kernel gradient over the batch. */
int vkx_batch_sum(uint64_t dst, uint64_t src, int batch, long long count, int double_precision, uint64_t stream);
/* Bluestein chirp w[m] = exp(sign * i pi m^2 / n), m in [0, count); m^2 mod 2n in exact integers. */
int vkx_chirp(uint64_t dst, int n, int count, int sign, int double_precision, uint64_t stream);

/* ---- ragged batches ------------------------------------------------------------ */
/* No counterpart in the reference (its FFTs are one kernel per shape, one length
 * per launch).  Runs `ntasks` work items of 1-D C2C transforms of mixed lengths in
 * one persistent launch: `tasks` is a device array of 128-byte item records (layout
 * in vkdispatch_amd/fft/ragged.py, RTASK_DTYPE) in queue order, where an item only
 * waits for items before it; `counters` holds `ncounters` >= 2 device ints, all
 * zeroed by the call (queue position, abort flag, per-transform done counters);
 * `workspace` holds `scratch_per_wg` complex64 elements per workgroup of `grid`;
 * `block` in {64, 128, ..., 512} (512-thread workgroups take the whole 160 KB of
 * LDS and keep whole transforms up to 10232 points there; smaller ones 66 KB and
 * 4224 points; longer whole transforms alternate between the output and their
 * workspace, which must hold 2 n elements per workgroup for a whole transform of
 * n > 4096 points).  A waiting item polls its counter at most
 * `spin_limit` times (<= 0: 2^20), then sets the abort flag (counters[1] = 1) and
 * every waiting workgroup stops waiting, so the launch always ends; a result with
 * the flag set is invalid. */
int vkx_ragged_exec(uint64_t tasks, int ntasks, uint64_t counters, int ncounters, uint64_t workspace,
                    long long scratch_per_wg, int grid, int block, int spin_limit, uint64_t stream);
/* Bluestein kernel b[j] = exp(+i pi j^2 / n) for |j| < n (indices mod m), 0 elsewhere,
 * complex64, phases in FP64 with j^2 reduced mod 2n in integers; m >= 2n-1. */
int vkx_bluestein_b(uint64_t dst, int n, int m, uint64_t stream);

/* ---- graphs ---------------------------------------------------------------- */
/* replace vd.CommandGraph recording/submission
 * (vkdispatch/execution_pipeline/command_graph.py:193-316) and its CUDA-graph
 * capture hook (vkdispatch/execution_pipeline/cuda_graph_capture.py:9-52). */
int vkx_graph_begin(uint64_t stream);
/* *capturing = 1 while `stream` is being captured into a graph (by vkx_graph_begin or by
 * a caller such as torch.cuda.graph), 2 if that capture was invalidated, 0 otherwise.
 * The capture-state query behind vkdispatch's cuda_graph_capture scope
 * (vkdispatch/execution_pipeline/cuda_graph_capture.py:22-52): plans refuse to allocate
 * or initialise device memory while their stream is capturing. */
int vkx_stream_is_capturing(uint64_t stream, int* capturing);
int vkx_graph_end(uint64_t stream, uint64_t* graph_exec);
int vkx_graph_launch(uint64_t graph_exec, uint64_t stream);
int vkx_graph_destroy(uint64_t graph_exec);

/* library version (major*10000 + minor*100 + patch) */
int vkx_version(void);

#ifdef __cplusplus
}
#endif
#endif /* VKDFFT_H */
