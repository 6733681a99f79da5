// ragged.hip -- persistent executor for batches of mixed-size 1-D C2C FFTs.
//
// One launch processes a queue of work items for transforms whose lengths are
// runtime values.  Workgroups stay resident and pull items from an atomic counter
// (the host orders the queue); every workgroup leaves the loop when the counter
// passes the item count.  Stockham radices (2..13, 16) are dispatched at run time
// to the same compile-time butterflies the specialised kernels use (vkx_fft.h).
//   * whole-transform items (lengths up to one LDS buffer -- 4224 points, 10232 in
//     workgroups of 512 threads -- ping-pong between two LDS buffers; longer lengths
//     without a split alternate between the output and this workgroup's slice of a
//     global scratch arena, and Bluestein lengths whose power-of-two size has no
//     split run in that slice);
//   * longer lengths, split L = n1 * n2 by the host: column-tile and row-tile
//     items of the four-step algorithm, spread over all workgroups, so one long
//     transform no longer runs on one workgroup.  The host either launches the
//     levels one after another, or queues them in one launch where a row tile
//     waits on its transform's done counter for the column tiles;
//   * lengths with a prime factor > 13: Bluestein (chirp-z) with an FP64 chirp
//     and a precomputed power-of-two chirp spectrum, as four tile levels:
//     chirp x input -> columns, rows x spectrum, inverse columns, inverse rows x
//     chirp.
#include <hip/hip_runtime.h>

#include "../../include/vkdfft.h"
#include "device/vkx_fft.h"
#include "vkx_error.h"

namespace {

using vkx::cpx;
typedef cpx<float> c64;

// Layout shared with vkdispatch_amd/fft/ragged.py (128 bytes).
struct RTask {
  unsigned long long in, out, tw, chirp, bspec;
  int n, m, npass, inverse;  // npass: passes (bits 0-7), n1's passes (8-15), item kind (16-19), conjugate (20),
                             // log2 of the planner's tile width (21-24, not read here: `count` is the width)
  int radix[12];
  float scale;
  int count;   // whole transforms: transforms in the task (stride n); tiles: columns / rows in the tile
  int tile0;   // tiles: first column / row
  int dep;     // tiles: done counter to wait for (-1: none) ...
  int target;  // ... until it reaches this value
  int done;    // tiles: done counter to increment when finished (-1: none)
};
static_assert(sizeof(RTask) == 128, "RTask layout");

// One LDS buffer (complex64 elements; a workgroup has two): room for a padded tile of
// 4096 points.  Workgroups of 512 threads, one per CU at this kernel's register use,
// take the whole 160 KB instead, so whole transforms up to kLdsBig points stay in LDS.
constexpr int kLdsBuf = 4096 + 128;
constexpr int kLdsBig = 10232;  // 2 x 10232 x 8 B + the queue slot fit in 160 KB

template <int R, int DIR>
__device__ __forceinline__ void bfly(const c64* __restrict__ src, c64* __restrict__ dst, const c64* __restrict__ tw,
                                     int n, int count, int ns, float scale, int pitch) {
  const int W = n / R;
  const int total = W * count;
  for (int jj = threadIdx.x; jj < total; jj += blockDim.x) {
    const int c = jj / W;
    const int j = jj - c * W;
    const int off = c * pitch;
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
__device__ void pass_dispatch(int R, const c64* src, c64* dst, const c64* tw, int n, int count, int ns, float scale,
                              int pitch) {
  switch (R) {
    case 2: bfly<2, DIR>(src, dst, tw, n, count, ns, scale, pitch); break;
    case 3: bfly<3, DIR>(src, dst, tw, n, count, ns, scale, pitch); break;
    case 4: bfly<4, DIR>(src, dst, tw, n, count, ns, scale, pitch); break;
    case 5: bfly<5, DIR>(src, dst, tw, n, count, ns, scale, pitch); break;
    case 6: bfly<6, DIR>(src, dst, tw, n, count, ns, scale, pitch); break;
    case 7: bfly<7, DIR>(src, dst, tw, n, count, ns, scale, pitch); break;
    case 8: bfly<8, DIR>(src, dst, tw, n, count, ns, scale, pitch); break;
    case 9: bfly<9, DIR>(src, dst, tw, n, count, ns, scale, pitch); break;
    case 10: bfly<10, DIR>(src, dst, tw, n, count, ns, scale, pitch); break;
    case 11: bfly<11, DIR>(src, dst, tw, n, count, ns, scale, pitch); break;
    case 12: bfly<12, DIR>(src, dst, tw, n, count, ns, scale, pitch); break;
    case 13: bfly<13, DIR>(src, dst, tw, n, count, ns, scale, pitch); break;
    case 16: bfly<16, DIR>(src, dst, tw, n, count, ns, scale, pitch); break;
    default: break;  // the host planner emits only the radices above
  }
}

// One pass of a single length-n transform from global to global memory whose output
// runs (ns consecutive elements) are shorter than a cache line: each round takes whole
// groups of ns butterflies, so its R outputs per butterfly fill one contiguous range,
// which goes through LDS (`stage`, padded one element per 16) and out in full lines.
// Every thread runs the same rounds (the barriers are uniform).
template <int R, int DIR>
__device__ __forceinline__ void bfly_staged(const c64* __restrict__ src, c64* __restrict__ dst,
                                            const c64* __restrict__ tw, int n, int ns, float scale, c64* stage) {
  const int W = n / R;  // a multiple of ns
  const int C = (int)(blockDim.x / ns) * ns;
  for (int j0 = 0; j0 < W; j0 += C) {
    const int j = j0 + (int)threadIdx.x;
    if ((int)threadIdx.x < C && j < W) {
      const int k = j % ns;
      c64 v[R];
#pragma unroll
      for (int r = 0; r < R; ++r) v[r] = src[j + r * W];
      if (ns > 1) {
#pragma unroll
        for (int r = 1; r < R; ++r) {
          c64 w = tw[(ns - 1) + (r - 1) * ns + k];
          if (DIR > 0) w.y = -w.y;
          v[r] = vkx::cmul(v[r], w);
        }
      }
      vkx::DFT<R, DIR>::run(v);
      const int q0 = (j - j0 - k) * R + k;  // output position relative to j0 R
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const int q = q0 + r * ns;
        stage[q + (q >> 4)] = v[r] * scale;
      }
    }
    __syncthreads();
    const int cnt = min(C, W - j0) * R;
    for (int i = threadIdx.x; i < cnt; i += blockDim.x) dst[j0 * R + i] = stage[i + (i >> 4)];
    __syncthreads();
  }
}

template <int DIR>
__device__ void pass_dispatch_staged(int R, const c64* src, c64* dst, const c64* tw, int n, int ns, float scale,
                                     c64* stage) {
  switch (R) {
    case 2: bfly_staged<2, DIR>(src, dst, tw, n, ns, scale, stage); break;
    case 3: bfly_staged<3, DIR>(src, dst, tw, n, ns, scale, stage); break;
    case 4: bfly_staged<4, DIR>(src, dst, tw, n, ns, scale, stage); break;
    case 5: bfly_staged<5, DIR>(src, dst, tw, n, ns, scale, stage); break;
    case 6: bfly_staged<6, DIR>(src, dst, tw, n, ns, scale, stage); break;
    case 7: bfly_staged<7, DIR>(src, dst, tw, n, ns, scale, stage); break;
    case 8: bfly_staged<8, DIR>(src, dst, tw, n, ns, scale, stage); break;
    case 9: bfly_staged<9, DIR>(src, dst, tw, n, ns, scale, stage); break;
    case 10: bfly_staged<10, DIR>(src, dst, tw, n, ns, scale, stage); break;
    case 11: bfly_staged<11, DIR>(src, dst, tw, n, ns, scale, stage); break;
    case 12: bfly_staged<12, DIR>(src, dst, tw, n, ns, scale, stage); break;
    case 13: bfly_staged<13, DIR>(src, dst, tw, n, ns, scale, stage); break;
    case 16: bfly_staged<16, DIR>(src, dst, tw, n, ns, scale, stage); break;
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
    if (dir > 0) pass_dispatch<1>(radix[p], in, out, tw, n, t.count, ns, last ? scale : 1.0f, n);
    else pass_dispatch<-1>(radix[p], in, out, tw, n, t.count, ns, last ? scale : 1.0f, n);
    __syncthreads();
    ns *= radix[p];
    in = out;
  }
}

// The same pass list for one transform (count 1) through global memory: the passes
// alternate between T0 and dst, so that the last one lands in dst and a transform
// touches one scratch buffer of its length instead of two.  In place (src == dst), a
// first pass that would write dst goes to T1.  Passes whose output runs are shorter
// than 16 elements (128 bytes) are written through `stage` (LDS, room for
// blockDim.x * 17 elements).  Needs npass >= 2.  Ends with a barrier.
__device__ void run_fft_global(const c64* src, c64* dst, c64* T0, c64* T1, c64* stage, const RTask& t, int n,
                               const int* radix, int npass, int dir, float scale) {
  const c64* tw = reinterpret_cast<const c64*>(t.tw);
  int ns = 1;
  const c64* in = src;
  for (int p = 0; p < npass; ++p) {
    const bool last = p == npass - 1;
    c64* out = ((npass - 1 - p) & 1) ? T0 : dst;
    if (p == 0 && out == src) out = T1;
    const float sc = last ? scale : 1.0f;
    if (ns < 16) {
      if (dir > 0) pass_dispatch_staged<1>(radix[p], in, out, tw, n, ns, sc, stage);
      else pass_dispatch_staged<-1>(radix[p], in, out, tw, n, ns, sc, stage);
    } else {
      if (dir > 0) pass_dispatch<1>(radix[p], in, out, tw, n, 1, ns, sc, n);
      else pass_dispatch<-1>(radix[p], in, out, tw, n, 1, ns, sc, n);
    }
    __syncthreads();
    ns *= radix[p];
    in = out;
  }
}

// exp(-2 pi i p / L) for 0 <= p < 2^31 (the host keeps tiled lengths <= 2^18)
__device__ __forceinline__ c64 twiddle_L(int p, int L) {
  const int q = p % L;
  float s, c;
  sincospif(2.0f * (float)q / (float)L, &s, &c);
  return c64{c, -s};
}

// ---- tiles of a long transform ---------------------------------------------------
// A transform of length L = n1 n2 (x[j1 n2 + j2]) runs as work items of two kinds:
// column tiles (length-n1 transforms over j1 for a run of columns j2, times
// W_L^(k1 j2), stored row-major y[k1 n2 + j2]) and row tiles (length-n2 transforms
// over j2 for a run of rows k1, stored to X[k1 + n1 k2]).  Every pass is a forward
// Stockham pass; inverse transforms conjugate on the way in and out.  Pass 0 reads
// global memory and the last pass writes it, the passes between ping-pong in LDS at
// pitch n + 1, so a tile crosses HBM once each way.  Threads take consecutive columns
// (rows) where that makes the global access a run of the tile's width.

// Store functors with kRun take a butterfly's R outputs at once (run<R>), the others
// one element at a time (st).
struct LdsIO {  // element (c, j) of an LDS tile
  static constexpr bool kRun = false;
  c64* p;
  int pitch;
  __device__ __forceinline__ c64 ld(int c, int j) const { return p[c * pitch + j]; }
  __device__ __forceinline__ void st(int c, int j, c64 v) const { p[c * pitch + j] = v; }
};

struct ColLoad {  // element j1 of column c0 + c: x[j1 n2 + c0 + c]
  const c64* src;
  const c64* w;  // Bluestein chirp (mode 2)
  int n2, c0, nvalid, mode, conj;  // mode 0: x (conjugated if conj), 2: x w' for index < nvalid, else 0
  __device__ __forceinline__ c64 ld(int c, int j1) const {
    const int idx = j1 * n2 + c0 + c;
    if (mode == 0) {
      c64 v = src[idx];
      if (conj) v.y = -v.y;
      return v;
    }
    if (idx >= nvalid) return c64{0.f, 0.f};
    c64 ww = w[idx];
    if (conj) ww.y = -ww.y;
    return vkx::cmul(src[idx], ww);
  }
};

struct ColStore {  // y[k1 n2 + j2] = v W_L^(k1 j2), j2 = c0 + c
  static constexpr bool kRun = true;
  c64* dst;
  int n2, c0, L;
  // outputs k1 = base + r ns of one butterfly (k1 j2 < n1 n2 = L): the twiddle is
  // evaluated for every fourth r and stepped by W_L^(ns j2) in between
  template <int R>
  __device__ __forceinline__ void run(int c, int base, int ns, const c64 (&v)[R]) const {
    const int j2 = c0 + c;
    const c64 step = R > 1 ? twiddle_L(ns * j2, L) : c64{1.f, 0.f};
    c64 w{1.f, 0.f};
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int k1 = base + r * ns;
      w = (r & 3) == 0 ? twiddle_L(k1 * j2, L) : vkx::cmul(w, step);
      dst[k1 * n2 + j2] = vkx::cmul(v[r], w);
    }
  }
};

struct RowLoad {  // element j2 of row r0 + t: y[(r0 + t) n2 + j2]
  const c64* src;
  int n2, r0;
  __device__ __forceinline__ c64 ld(int t, int j2) const { return src[(r0 + t) * n2 + j2]; }
};

struct RowStore {  // X[k] with k = r0 + t + n1 k2
  static constexpr bool kRun = false;
  c64* dst;
  const c64* aux;  // mode 2: Bluestein kernel spectrum; mode 3: chirp
  int n1, r0, m, nvalid, mode, inv;
  float scale;
  __device__ __forceinline__ void st(int t, int k2, c64 v) const {
    const int k = r0 + t + n1 * k2;
    if (mode == 0) {  // v (conjugated for an inverse transform) times scale
      if (inv) v.y = -v.y;
      dst[k] = v * scale;
    } else if (mode == 2) {  // conj(v b'): the inverse transform that follows runs forward
      c64 b = aux[k];
      if (inv) {
        b = aux[(m - k) & (m - 1)];
        b.y = -b.y;
      }
      c64 r = vkx::cmul(v, b);
      r.y = -r.y;
      dst[k] = r;
    } else if (k < nvalid) {  // mode 3: conj(v) w'[k] scale for the first nvalid outputs
      c64 ww = aux[k];
      if (inv) ww.y = -ww.y;
      v.y = -v.y;
      dst[k] = vkx::cmul(v, ww) * scale;
    }
  }
};

// One forward Stockham pass of radix R over `count` length-n transforms (W = n / R
// butterflies each), reading through `ld` and writing through `st`.  cfast: thread
// index runs over the transforms first.
template <int R, class Ld, class St>
__device__ __forceinline__ void bfly_io(const c64* __restrict__ tw, int W, int count, int ns, bool cfast,
                                        const Ld& ld, const St& st) {
  const int total = W * count;
  for (int jj = threadIdx.x; jj < total; jj += blockDim.x) {
    int c, j;
    if (cfast) {
      j = jj / count;
      c = jj - j * count;
    } else {
      c = jj / W;
      j = jj - c * W;
    }
    const int k = j % ns;
    c64 v[R];
#pragma unroll
    for (int r = 0; r < R; ++r) v[r] = ld.ld(c, j + r * W);
    if (ns > 1) {
#pragma unroll
      for (int r = 1; r < R; ++r) v[r] = vkx::cmul(v[r], tw[(ns - 1) + (r - 1) * ns + k]);
    }
    vkx::DFT<R, -1>::run(v);
    const int base = (j / ns) * ns * R + k;
    if constexpr (St::kRun) {
      st.template run<R>(c, base, ns, v);
    } else {
#pragma unroll
      for (int r = 0; r < R; ++r) st.st(c, base + r * ns, v[r]);
    }
  }
}

template <class Ld, class St>
__device__ void pass_io(int R, const c64* tw, int n, int count, int ns, bool cfast, const Ld& ld, const St& st) {
  const int W = n / R;
  switch (R) {
    case 2: bfly_io<2>(tw, W, count, ns, cfast, ld, st); break;
    case 3: bfly_io<3>(tw, W, count, ns, cfast, ld, st); break;
    case 4: bfly_io<4>(tw, W, count, ns, cfast, ld, st); break;
    case 5: bfly_io<5>(tw, W, count, ns, cfast, ld, st); break;
    case 6: bfly_io<6>(tw, W, count, ns, cfast, ld, st); break;
    case 7: bfly_io<7>(tw, W, count, ns, cfast, ld, st); break;
    case 8: bfly_io<8>(tw, W, count, ns, cfast, ld, st); break;
    case 9: bfly_io<9>(tw, W, count, ns, cfast, ld, st); break;
    case 10: bfly_io<10>(tw, W, count, ns, cfast, ld, st); break;
    case 11: bfly_io<11>(tw, W, count, ns, cfast, ld, st); break;
    case 12: bfly_io<12>(tw, W, count, ns, cfast, ld, st); break;
    case 13: bfly_io<13>(tw, W, count, ns, cfast, ld, st); break;
    case 16: bfly_io<16>(tw, W, count, ns, cfast, ld, st); break;
    default: break;  // the host planner emits only the radices above
  }
}

// `count` length-n transforms, global -> (LDS ping-pong at `pitch`) -> global.  A tile
// of up to two passes uses only lds0 (and may extend over lds1); longer pass lists
// alternate lds0 / lds1.  cfast0 / cfast: thread order of pass 0 / the other passes.
template <class GLd, class GSt>
__device__ void tile_fft(const c64* tw, int n, int count, int pitch, const int* radix, int npass, c64* lds0,
                         c64* lds1, const GLd& gld, const GSt& gst, bool cfast0, bool cfast) {
  int ns = 1;
  for (int p = 0; p < npass; ++p) {
    const bool first = p == 0, last = p == npass - 1;
    const LdsIO src{(p & 1) ? lds0 : lds1, pitch};  // written by pass p - 1
    const LdsIO dst{(p & 1) ? lds1 : lds0, pitch};
    const bool cf = first ? cfast0 : cfast;
    if (first && last) pass_io(radix[p], tw, n, count, ns, cf, gld, gst);
    else if (first) pass_io(radix[p], tw, n, count, ns, cf, gld, dst);
    else if (last) pass_io(radix[p], tw, n, count, ns, cf, src, gst);
    else pass_io(radix[p], tw, n, count, ns, cf, src, dst);
    if (!last) __syncthreads();
    ns *= radix[p];
  }
}

// Item kinds (task bits 16-19); bit 20: conjugate (inverse direct transform).
enum : int { kWhole = 0, kCol = 1, kRow = 2, kBlueCol = 3, kBlueRow = 4, kBlueOut = 5 };

__device__ void run_tile(const RTask& t, int kind, c64* lds0, c64* lds1) {
  const int L = t.m ? t.m : t.n;
  const int npass = t.npass & 0xff, np1 = (t.npass >> 8) & 0xff;
  const int cj = (t.npass >> 20) & 1;
  int n1 = 1;
  for (int p = 0; p < np1; ++p) n1 *= t.radix[p];
  const int n2 = L / n1;
  const c64* tw = reinterpret_cast<const c64*>(t.tw);
  const c64* in = reinterpret_cast<const c64*>(t.in);
  c64* out = reinterpret_cast<c64*>(t.out);
  if (kind == kCol || kind == kBlueCol) {
    // direct columns conjugate the data of an inverse transform (bit 20); Bluestein
    // columns conjugate the chirp of an inverse transform
    const bool blue = kind == kBlueCol;
    const ColLoad ld{in, reinterpret_cast<const c64*>(t.chirp), n2, t.tile0, t.n, blue ? 2 : 0,
                     blue ? t.inverse : cj};
    const ColStore st{out, n2, t.tile0, L};
    tile_fft(tw, n1, t.count, n1 + 1, t.radix, np1, lds0, lds1, ld, st, true, true);
  } else {
    const int np2 = npass - np1;
    const RowLoad ld{in, n2, t.tile0};
    const int mode = kind == kRow ? 0 : (kind == kBlueRow ? 2 : 3);
    const c64* aux = reinterpret_cast<const c64*>(kind == kBlueRow ? t.bspec : t.chirp);
    const RowStore st{out, aux, n1, t.tile0, L, t.n, mode, kind == kRow ? cj : t.inverse, t.scale};
    tile_fft(tw + (n1 - 1), n2, t.count, n2 + 1, t.radix + np1, np2, lds0, lds1, ld, st, np2 == 1, true);
  }
}

// lds_cap: elements of each of the two LDS buffers
__device__ void run_task(const RTask& t, c64* lds0, c64* lds1, int lds_cap, c64* scr) {
  const c64* in = reinterpret_cast<const c64*>(t.in);
  c64* out = reinterpret_cast<c64*>(t.out);
  const int dir = t.inverse ? 1 : -1;
  const int npass = t.npass & 0xff;
  if (t.m == 0) {
    const int total = t.n * t.count;
    if (total <= lds_cap) {
      run_fft(in, out, lds0, lds1, t, t.n, t.radix, npass, dir, t.scale);
    } else if (npass < 2 || t.count != 1) {
      run_fft(in, out, scr, scr + total, t, t.n, t.radix, npass, dir, t.scale);
    } else {
      run_fft_global(in, out, scr, scr + total, lds0, t, t.n, t.radix, npass, dir, t.scale);
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
    run_fft(S0, S1, S2, S0, sub, m, t.radix, npass, -1, 1.0f);
    for (int j = threadIdx.x; j < m; j += blockDim.x) {
      c64 b = bs[j];
      if (t.inverse) {  // spectrum of conj(b) for the inverse chirp: B_inv[k] = conj(B[-k])
        b = bs[(m - j) & (m - 1)];
        b.y = -b.y;
      }
      S1[j] = vkx::cmul(S1[j], b);
    }
    __syncthreads();
    run_fft(S1, S0, S2, S1, sub, m, t.radix, npass, 1, 1.0f / (float)m);
    for (int k = threadIdx.x; k < n; k += blockDim.x) {
      c64 ww = w[k];
      if (t.inverse) ww.y = -ww.y;
      xout[k] = vkx::cmul(S0[k], ww) * t.scale;
    }
    __syncthreads();
  }
}

__global__ void ragged_reset_kernel(int* ctr, int n) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) ctr[i] = 0;
}

// Wait (thread 0) until counter ctr[2 + idx] reaches target.  Bounded: after
// spin_limit polls, or once another workgroup has given up, it sets the abort flag
// ctr[1] and returns, so every wave leaves the kernel whatever happens.
__device__ void wait_counter(int* ctr, int idx, int target, int spin_limit) {
  int spins = 0;
  while (__hip_atomic_load(&ctr[2 + idx], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
    if (__hip_atomic_load(&ctr[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) return;
    if (++spins > spin_limit) {
      __hip_atomic_store(&ctr[1], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      return;
    }
    __builtin_amdgcn_s_sleep(4);
  }
}

// ctr[0]: queue position, ctr[1]: abort flag, ctr[2 ...]: per-transform done counters.
// Items are fetched in queue order and an item only waits for items queued before it,
// so each wait is for work already held by a running workgroup.
// CAP: elements of each of the workgroup's two LDS buffers.
template <int CAP>
__global__ void __launch_bounds__(512) ragged_kernel(const RTask* __restrict__ tasks, int ntasks, int* ctr,
                                                     c64* scratch, long long scratch_per_wg, int spin_limit) {
  __shared__ c64 lds[2 * CAP];
  __shared__ int task_id;
  c64* scr = scratch + (long long)blockIdx.x * scratch_per_wg;
  for (;;) {
    if (threadIdx.x == 0) task_id = atomicAdd(&ctr[0], 1);
    __syncthreads();
    const int id = task_id;
    __syncthreads();
    if (id >= ntasks) break;
    const RTask t = tasks[id];
    const int kind = (t.npass >> 16) & 0xf;
    if (kind == kWhole) {
      run_task(t, lds, lds + CAP, CAP, scr);
      continue;
    }
    if (t.dep >= 0) {
      if (threadIdx.x == 0) wait_counter(ctr, t.dep, t.target, spin_limit);
      __syncthreads();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    }
    run_tile(t, kind, lds, lds + CAP);
    if (t.done >= 0) {
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
      __syncthreads();
      if (threadIdx.x == 0) __hip_atomic_fetch_add(&ctr[2 + t.done], 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    }
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

int vkx_ragged_exec(uint64_t tasks, int ntasks, uint64_t counters, int ncounters, uint64_t scratch,
                    long long scratch_per_wg, int grid, int block, int spin_limit, uint64_t stream) {
  if (ntasks <= 0) return 0;
  if (block < 64 || block > 512 || (block % 64) != 0)
    return vkx_set_error("vkx_ragged_exec", "block must be a multiple of 64 in [64, 512]");
  if (grid < 1) return vkx_set_error("vkx_ragged_exec", "grid must be >= 1");
  if (ncounters < 2) return vkx_set_error("vkx_ragged_exec", "need at least 2 counters (queue position, abort flag)");
  if (!tasks || !counters) return vkx_set_error("vkx_ragged_exec", "null task or counter array");
  if (spin_limit <= 0) spin_limit = 1 << 20;
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  hipLaunchKernelGGL(ragged_reset_kernel, dim3(1), dim3(256), 0, s, reinterpret_cast<int*>(counters), ncounters);
  // workgroups of 512 threads take the large LDS buffers (whole transforms up to kLdsBig
  // points in LDS); the host sizes the scratch for every whole transform above 4096
  // points either way
  if (block == 512)
    hipLaunchKernelGGL(ragged_kernel<kLdsBig>, dim3(grid), dim3(block), 0, s, reinterpret_cast<const RTask*>(tasks),
                       ntasks, reinterpret_cast<int*>(counters), reinterpret_cast<c64*>(scratch), scratch_per_wg,
                       spin_limit);
  else
    hipLaunchKernelGGL(ragged_kernel<kLdsBuf>, dim3(grid), dim3(block), 0, s, reinterpret_cast<const RTask*>(tasks),
                       ntasks, reinterpret_cast<int*>(counters), reinterpret_cast<c64*>(scratch), scratch_per_wg,
                       spin_limit);
  const hipError_t e = hipGetLastError();
  return e == hipSuccess ? 0 : vkx_set_hip_error("vkx_ragged_exec: launch", e);
}

int vkx_bluestein_b(uint64_t dst, int n, int m, uint64_t stream) {
  if (n < 1) return vkx_set_error("vkx_bluestein_b", "n must be >= 1");
  if (m < 2 * n - 1) return vkx_set_error("vkx_bluestein_b", "m must be >= 2n - 1");
  if (!dst) return vkx_set_error("vkx_bluestein_b", "null destination");
  const int threads = 256;
  hipLaunchKernelGGL(bluestein_b_kernel, dim3((m + threads - 1) / threads), dim3(threads), 0,
                     reinterpret_cast<hipStream_t>(stream), reinterpret_cast<c64*>(dst), n, m);
  const hipError_t e = hipGetLastError();
  return e == hipSuccess ? 0 : vkx_set_hip_error("vkx_bluestein_b: launch", e);
}

}  // extern "C"
