// ragged.hip -- persistent executor for batches of mixed-size 1-D C2C FFTs.
//
// One launch processes a task list of transforms whose lengths are runtime
// values.  Workgroups stay resident and pull tasks from an atomic counter
// (tasks are sorted longest-first on the host, so the queue is an LPT
// schedule); every workgroup leaves the loop when the counter passes the task
// count.  Each task runs a Stockham pass list whose radices (2..13, 16) are
// dispatched at run time to the same compile-time butterflies the specialised
// kernels use (vkx_fft.h).  Buffers:
//   * length <= 4096: passes ping-pong between two LDS buffers;
//   * longer: passes ping-pong through this workgroup's slice of a global
//     scratch arena (L2-resident for the sizes at hand);
//   * lengths with a prime factor > 13: Bluestein (chirp-z) with an FP64
//     chirp and a precomputed power-of-two chirp spectrum.
#include <hip/hip_runtime.h>

#include "../../include/vkdfft.h"
#include "device/vkx_fft.h"

namespace {

using vkx::cpx;
typedef cpx<float> c64;

// Layout shared with vkdispatch_amd/fft/ragged.py (112 bytes).
struct RTask {
  unsigned long long in, out, tw, chirp, bspec;
  int n, m, npass, inverse;
  int radix[12];
  float scale;
  int count;  // transforms in this task (stride n, contiguous)
};
static_assert(sizeof(RTask) == 112, "RTask layout");

constexpr int kLdsPingPong = 4096;  // complex64 elements per LDS buffer

template <int R, int DIR>
__device__ __forceinline__ void bfly(const c64* __restrict__ src, c64* __restrict__ dst, const c64* __restrict__ tw,
                                     int n, int count, int ns, float scale) {
  const int W = n / R;
  const int total = W * count;
  for (int jj = threadIdx.x; jj < total; jj += blockDim.x) {
    const int c = jj / W;
    const int j = jj - c * W;
    const int off = c * n;
    const int k = j % ns;
    c64 v[R];
#pragma unroll
    for (int r = 0; r < R; ++r) v[r] = src[off + j + r * W];
    if (ns > 1) {
#pragma unroll
      for (int r = 1; r < R; ++r) {
        c64 w = tw[(ns - 1) + (r - 1) * ns + k];
        if (DIR > 0) w.y = -w.y;
        v[r] = vkx::cmul(v[r], w);
      }
    }
    vkx::DFT<R, DIR>::run(v);
    const int base = off + (j / ns) * ns * R + k;
#pragma unroll
    for (int r = 0; r < R; ++r) dst[base + r * ns] = v[r] * scale;
  }
}

template <int DIR>
__device__ void pass_dispatch(int R, const c64* src, c64* dst, const c64* tw, int n, int count, int ns, float scale) {
  switch (R) {
    case 2: bfly<2, DIR>(src, dst, tw, n, count, ns, scale); break;
    case 3: bfly<3, DIR>(src, dst, tw, n, count, ns, scale); break;
    case 4: bfly<4, DIR>(src, dst, tw, n, count, ns, scale); break;
    case 5: bfly<5, DIR>(src, dst, tw, n, count, ns, scale); break;
    case 6: bfly<6, DIR>(src, dst, tw, n, count, ns, scale); break;
    case 7: bfly<7, DIR>(src, dst, tw, n, count, ns, scale); break;
    case 8: bfly<8, DIR>(src, dst, tw, n, count, ns, scale); break;
    case 9: bfly<9, DIR>(src, dst, tw, n, count, ns, scale); break;
    case 10: bfly<10, DIR>(src, dst, tw, n, count, ns, scale); break;
    case 11: bfly<11, DIR>(src, dst, tw, n, count, ns, scale); break;
    case 12: bfly<12, DIR>(src, dst, tw, n, count, ns, scale); break;
    case 13: bfly<13, DIR>(src, dst, tw, n, count, ns, scale); break;
    case 16: bfly<16, DIR>(src, dst, tw, n, count, ns, scale); break;
    default: break;  // the host planner emits only the radices above
  }
}

__device__ void copy_elems(const c64* src, c64* dst, int total) {
  for (int i = threadIdx.x; i < total; i += blockDim.x) dst[i] = src[i];
}

// Stockham pass list from src to dst; T0/T1 are ping-pong buffers distinct from
// dst (T1 may alias src: src is only read by pass 0).  Ends with a barrier.
__device__ void run_fft(const c64* src, c64* dst, c64* T0, c64* T1, const RTask& t, int n, const int* radix,
                        int npass, int dir, float scale) {
  const c64* tw = reinterpret_cast<const c64*>(t.tw);
  if (npass == 1 && src == dst) {  // single pass in place: stage through T0
    copy_elems(src, T0, n * t.count);
    __syncthreads();
    src = T0;
    T0 = T1;
  }
  int ns = 1;
  const c64* in = src;
  for (int p = 0; p < npass; ++p) {
    const bool last = p == npass - 1;
    c64* out = last ? dst : ((p & 1) ? T1 : T0);
    if (dir > 0) pass_dispatch<1>(radix[p], in, out, tw, n, t.count, ns, last ? scale : 1.0f);
    else pass_dispatch<-1>(radix[p], in, out, tw, n, t.count, ns, last ? scale : 1.0f);
    __syncthreads();
    ns *= radix[p];
    in = out;
  }
}

__device__ void run_task(const RTask& t, c64* lds0, c64* lds1, c64* scr) {
  const c64* in = reinterpret_cast<const c64*>(t.in);
  c64* out = reinterpret_cast<c64*>(t.out);
  const int dir = t.inverse ? 1 : -1;
  if (t.m == 0) {
    const int total = t.n * t.count;
    if (total <= kLdsPingPong) {
      run_fft(in, out, lds0, lds1, t, t.n, t.radix, t.npass, dir, t.scale);
    } else {
      run_fft(in, out, scr, scr + total, t, t.n, t.radix, t.npass, dir, t.scale);
    }
    return;
  }
  // Bluestein: X[k] = w[k] * sum_j (x[j] w[j]) conj(w[k-j]),  w[j] = exp(-i pi j^2 / n)
  const int n = t.n, m = t.m;
  const c64* w = reinterpret_cast<const c64*>(t.chirp);
  const c64* bs = reinterpret_cast<const c64*>(t.bspec);
  c64* S0 = scr;
  c64* S1 = scr + m;
  c64* S2 = scr + 2 * m;
  for (int c = 0; c < t.count; ++c) {
    const c64* xin = in + (long long)c * n;
    c64* xout = out + (long long)c * n;
    for (int j = threadIdx.x; j < m; j += blockDim.x) {
      c64 v{0.f, 0.f};
      if (j < n) {
        c64 ww = w[j];
        if (t.inverse) ww.y = -ww.y;
        v = vkx::cmul(xin[j], ww);
      }
      S0[j] = v;
    }
    __syncthreads();
    RTask sub = t;
    sub.count = 1;
    run_fft(S0, S1, S2, S0, sub, m, t.radix, t.npass, -1, 1.0f);
    for (int j = threadIdx.x; j < m; j += blockDim.x) {
      c64 b = bs[j];
      if (t.inverse) {  // spectrum of conj(b) for the inverse chirp: B_inv[k] = conj(B[-k])
        b = bs[(m - j) & (m - 1)];
        b.y = -b.y;
      }
      S1[j] = vkx::cmul(S1[j], b);
    }
    __syncthreads();
    run_fft(S1, S0, S2, S1, sub, m, t.radix, t.npass, 1, 1.0f / (float)m);
    for (int k = threadIdx.x; k < n; k += blockDim.x) {
      c64 ww = w[k];
      if (t.inverse) ww.y = -ww.y;
      xout[k] = vkx::cmul(S0[k], ww) * t.scale;
    }
    __syncthreads();
  }
}

__global__ void ragged_reset_kernel(int* counter) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *counter = 0;
}

__global__ void __launch_bounds__(512) ragged_kernel(const RTask* __restrict__ tasks, int ntasks, int* counter,
                                                     c64* scratch, long long scratch_per_wg) {
  __shared__ c64 lds[2 * kLdsPingPong];
  __shared__ int task_id;
  c64* scr = scratch + (long long)blockIdx.x * scratch_per_wg;
  for (;;) {
    if (threadIdx.x == 0) task_id = atomicAdd(counter, 1);
    __syncthreads();
    const int id = task_id;
    __syncthreads();
    if (id >= ntasks) break;
    const RTask t = tasks[id];
    run_task(t, lds, lds + kLdsPingPong, scr);
  }
}

// b[j] = conj(w[j]) for |j| < n (j taken mod m), 0 elsewhere; w[j] = exp(-i pi j^2/n) in FP64.
__global__ void bluestein_b_kernel(c64* b, int n, int m) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= m) return;
  long long jj = -1;
  if (j < n) jj = j;
  else if (j > m - n) jj = m - j;
  c64 v{0.f, 0.f};
  if (jj >= 0) {
    const unsigned long long ph = ((unsigned long long)jj * (unsigned long long)jj) % (2ull * (unsigned long long)n);
    double s, c;
    sincospi((double)ph / (double)n, &s, &c);  // conj(exp(-i pi ph / n)) = exp(+i pi ph / n)
    v = c64{(float)c, (float)s};
  }
  b[j] = v;
}

}  // namespace

extern "C" {

int vkx_ragged_exec(uint64_t tasks, int ntasks, uint64_t counter, uint64_t scratch, long long scratch_per_wg,
                    int grid, int block, uint64_t stream) {
  if (ntasks <= 0) return 0;
  if (block < 64 || block > 512 || (block % 64) != 0 || grid < 1) {
    return 1;
  }
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  hipLaunchKernelGGL(ragged_reset_kernel, dim3(1), dim3(64), 0, s, reinterpret_cast<int*>(counter));
  hipLaunchKernelGGL(ragged_kernel, dim3(grid), dim3(block), 0, s, reinterpret_cast<const RTask*>(tasks), ntasks,
                     reinterpret_cast<int*>(counter), reinterpret_cast<c64*>(scratch), scratch_per_wg);
  return hipGetLastError() == hipSuccess ? 0 : 1;
}

int vkx_bluestein_b(uint64_t dst, int n, int m, uint64_t stream) {
  if (n < 1 || m < 2 * n - 1) return 1;
  const int threads = 256;
  hipLaunchKernelGGL(bluestein_b_kernel, dim3((m + threads - 1) / threads), dim3(threads), 0,
                     reinterpret_cast<hipStream_t>(stream), reinterpret_cast<c64*>(dst), n, m);
  return hipGetLastError() == hipSuccess ? 0 : 1;
}

}  // extern "C"
