// vkx_conv2d.h -- 2-D FFT convolution of a batch of real images with one
// broadcast kernel spectrum:  y_b = irfft2(rfft2(x_b) . conj(K)).
//
// Two real images x_a, x_b are transformed as one complex image z = x_a + i x_b:
// the 2-D convolution with a real kernel is linear, so ifft2(fft2(z) . conj(Kf))
// = y_a + i y_b when Kf is the full (Hermitian) spectrum of a real kernel.  Kf is
// taken from the half spectrum K (H x (W/2+1)):
//   kx in 1..W/2-1   : K[ky][kx]
//   kx in W/2+1..W-1 : conj(K[-ky][W-kx])
//   kx in {0, W/2}   : (K[ky][kx] + conj(K[-ky][kx])) / 2
// The last rule reproduces exactly what the three-dispatch composition does with
// a half spectrum that is not Hermitian on those two columns (the c2r row
// transform keeps only the real part there).
//
// Data flow per image pair, through a complex scratch image S stored tile-major:
// 16-column tiles, each H rows x 16 columns (128-byte rows) contiguous, element
// (ky, kx) at (kx / 16) * 16H + ky * 16 + kx % 16.  Row work then reads and writes
// whole 128-byte lines from registers, and a column tile is one contiguous block.
//   K  (kernel prep): Kt[kx * H + ky] = conj(Kf[ky][kx]) * SCALE (column-major)
//   A  (rows [y0, y0+RA)): row FFTs of z, written to S
//   B  (one tile): staged in LDS column-major; each wave: column FFT, multiply by
//      its Kt column, column IFFT; tile written back
//   C  (rows [y0, y0+RA)): row IFFTs from S, real part -> image a, imaginary part
//      -> image b
//
// The phases run either as ordinary kernels in stream order (conv2d_phase_body)
// or as one cooperative launch (conv2d_pair_body).  In the cooperative form the
// L2 is per XCD and not coherent across XCDs, so all tasks of one pair run on
// workgroups of one XCD: there is one task queue per XCD (with its own Kt copy)
// and a workgroup takes tasks only from the queue of the XCD it runs on (XCC_ID
// register).  Producers complete their stores (s_waitcnt) before bumping a
// counter with a device-scope atomic, and consumers read with L1-bypassing
// (device-scope) loads after seeing the counter; the data stays in the XCD's L2
// unless evicted.  No agent-scope fence is used: it would write back /
// invalidate the whole L2.  Workgroups take tickets from their queue's counter
// (the next ticket is requested while the current task runs); a task only waits
// for tasks with smaller tickets of the same queue, held by running workgroups,
// so the grid cannot deadlock; every wait is bounded and a timeout sets an error
// bit.  The last workgroup to exit runs the tasks of any queue that no workgroup
// took (an XCD that received no workgroup), then resets the control block.
#pragma once
#include "vkx_fft.h"

namespace vkx {

// 64-lane, one-transform-per-wave configuration of the Stockham machinery
template <int N_, int NP_> struct WaveCfg {
  using real = float;
  static constexpr int N = N_, TT = 64, B = 1, PAD_SHIFT = 5, NP = NP_;
  static constexpr bool COLFAST = false, WAVE_LOCAL = true;
};

// Register-resident twiddles (loaded once per task, see TwSet) when a lane's set is
// small: 14 values for 512 = 8.8.8 and 9 for 256 = 4.4.4.4.  The 27 of 1024 =
// 16.16.4 would push the cooperative 1024 kernel into spills, so those kernels keep
// reading the table.
template <class WC, class L> constexpr bool conv_twreg() { return list_tw_regs<WC, L>::value <= 16; }

// control block layout (u32 words)
constexpr u32 CTL_TICKET = 0;      // + 32 q
constexpr u32 CTL_EXIT = 256;      // workgroups that have left
constexpr u32 CTL_ERROR = 512;     // bit 0: a wait timed out; bit 2: the last workgroup served a queue
constexpr u32 CTL_KDONE = 768;     // + 32 q: kernel-prep tasks done for queue q
constexpr u32 CTL_PAIRS = 1024;    // + 64 p: A-done at +0, B-done at +32
constexpr int CONV_QUEUES = 8;
constexpr u32 CONV_SPIN_LIMIT = 1u << 18;
constexpr u32 PHASE_A = 0, PHASE_B = 1, PHASE_C = 2, PHASE_K = 3;

__device__ __forceinline__ u32 ld_agent(const u32* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_agent(u32* p, u32 v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ u32 add_agent(u32* p, u32 v) {
  return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// complex64 load that bypasses the (non-coherent) L1 and reads the XCD's L2
__device__ __forceinline__ cpx<float> ld_c64_l2(const cpx<float>* p) {
  const unsigned long long u =
      __hip_atomic_load(reinterpret_cast<const unsigned long long*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  cpx<float> r;
  r.x = __uint_as_float((u32)u);
  r.y = __uint_as_float((u32)(u >> 32));
  return r;
}

// scratch / Kt load: L1-bypassing inside the cooperative kernel (data written by
// other workgroups of this launch), an ordinary cached load in the phase kernels
template <bool L2ONLY> __device__ __forceinline__ cpx<float> ld_c64(const cpx<float>* p) {
  if constexpr (L2ONLY) return ld_c64_l2(p);
  else return *p;
}

// thread 0 waits until *c >= target (bounded); the workgroup then synchronises
__device__ __forceinline__ void wait_count(u32* c, u32 target, u32* err) {
  if (threadIdx.x == 0) {
    u32 it = 0;
    while (ld_agent(c) < target) {
      __builtin_amdgcn_s_sleep(2);
      if (++it > CONV_SPIN_LIMIT) {
        __hip_atomic_fetch_or(err, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        break;
      }
    }
  }
  __syncthreads();
}

// all stores of the workgroup have reached L2, then one counter increment
__device__ __forceinline__ void signal_done(u32* c) {
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  if (threadIdx.x == 0) add_agent(c, 1u);
}

// C: H, W (= N of the wave transforms), PITCH (floats per real row), IMG (floats
// per image), NIMG, NPAIRS, NW (waves per workgroup), RA (rows per row task), NP
// (padded LDS region of one transform; odd so tile transposes are conflict-free),
// KW = W/2+1, SCALE, fwd / inv (Stockham pass lists, inv = reversed fwd), LAYOUT
// (0: tile-major scratch, 1: column-major), CB (columns per wave, column-major),
// PP (pairs per column task, column-major), TP (LDS stride per column of a
// row-block transpose tile, > RA).
template <class C> struct ConvArgs {
  float* x;
  const cpx<float>* K;
  cpx<float>* scr;
  cpx<float>* kt;
  const cpx<float>* twf;
  const cpx<float>* twi;
};

template <class C> __device__ __forceinline__ ConvArgs<C> conv_args(const KParams& prm) {
  return {bufp<float>(prm, 0), bufp<const cpx<float>>(prm, 1), bufp<cpx<float>>(prm, 2), bufp<cpx<float>>(prm, 5),
          bufp<const cpx<float>>(prm, SLOT_TW_FWD), bufp<const cpx<float>>(prm, SLOT_TW_INV)};
}

template <class C> struct ConvShape {
  static constexpr int NV = C::W / 64;               // points per lane
  static constexpr int RPW = C::RA / C::NW;          // rows per wave in a row task
  static constexpr int NT = C::NW * 64;              // threads per workgroup
  static constexpr int CT = 16;                      // columns per tile (128-byte rows)
  static constexpr int CB = CT / C::NW;              // columns per wave in a column task
  static constexpr u32 NA = C::H / C::RA;            // row tasks per pair
  static constexpr u32 NB = C::W / CT;               // column tasks (tiles) per pair
  static constexpr u32 KB = 64;                      // kernel-prep block (KB x KB)
  static constexpr u32 NK = (C::H / KB) * (C::W / KB);  // kernel-prep tasks
  static constexpr int XCH = C::NW * C::NP;          // per-wave exchange regions
  static constexpr int TILE = CT * C::NP;            // one column tile, column-major
  static constexpr int LDS_ALL = TILE > XCH ? TILE : XCH;
  static_assert(C::H == C::W, "square images");
  static_assert(C::W % 64 == 0 && C::RA % C::NW == 0 && CT % C::NW == 0, "tiling");
  static_assert(C::NP >= C::H + C::H / 32, "LDS region");
};

// element (ky, kx) of a tile-major pair spectrum
template <class C> __device__ __forceinline__ u32 tile_index(u32 ky, u32 kx) {
  return (kx >> 4) * (u32)(16 * C::H) + ky * 16u + (kx & 15u);
}

// ---- K: block `slice` (KB x KB) of Kt = conj(Kf) * SCALE, column-major.  Writes
// are coalesced along ky; the reads of K are strided but cached (1 MB). ----
template <class C>
__device__ __forceinline__ void conv_kprep(const cpx<float>* K, cpx<float>* kt, u32 slice, int tid) {
  using SH = ConvShape<C>;
  const u32 ky0 = (slice / (C::W / SH::KB)) * SH::KB, kx0 = (slice % (C::W / SH::KB)) * SH::KB;
  const float sc = (float)C::SCALE;
  static_assert(SH::KB * SH::KB % SH::NT == 0, "kernel-prep split");
#pragma unroll
  for (int i = 0; i < (int)(SH::KB * SH::KB) / SH::NT; ++i) {
    const int e = tid + i * SH::NT;
    const u32 ky = ky0 + (u32)e % SH::KB, kx = kx0 + (u32)e / SH::KB;
    const u32 kym = (C::H - ky) & (C::H - 1);
    const bool lo = kx <= (u32)(C::W / 2);
    cpx<float> kf = K[lo ? ky * C::KW + kx : kym * C::KW + (C::W - kx)];
    if (!lo) kf.y = -kf.y;
    if (kx == 0 || kx == (u32)(C::W / 2)) {
      const cpx<float> k1 = K[kym * C::KW + kx];
      kf = cpx<float>{0.5f * (kf.x + k1.x), 0.5f * (kf.y - k1.y)};
    }
    kt[kx * C::H + ky] = cpx<float>{kf.x * sc, -kf.y * sc};
  }
}

// ---- A: row FFTs of z = x_a + i x_b for rows [y0, y0+RA) into S ----
template <class C>
__device__ __forceinline__ void conv_rows_fwd(const ConvArgs<C>& a, cpx<float>* lds, u32 pair, u32 slice, int wave,
                                              int lane) {
  using SH = ConvShape<C>;
  using WC = WaveCfg<C::W, C::NP>;
  using L = typename C::fwd;
  using F0 = typename pl_first<L>::type;
  using FL = typename pl_last<L>::type;
  const u32 ia = 2 * pair, ib = 2 * pair + 1;
  const bool has_b = ib < C::NIMG;
  const u32 y0 = slice * C::RA;
  cpx<float>* S = a.scr + (u64)pair * (u64)(C::H * C::W);
  cpx<float> v[SH::RPW][SH::NV];
#pragma unroll
  for (int i = 0; i < SH::RPW; ++i) {
    const u32 y = y0 + (u32)(wave + C::NW * i);
    const float* ra = a.x + (u64)ia * C::IMG + y * C::PITCH;
    const float* rb = a.x + (u64)ib * C::IMG + y * C::PITCH;
    for_layout<WC, F0::R>(lane, [&](auto slot, u32 n) {
      v[i][decltype(slot)::value] = cpx<float>{ra[n], has_b ? rb[n] : 0.f};
    });
  }
#pragma unroll
  for (int i = 0; i < SH::RPW; ++i) {
    const u32 y = y0 + (u32)(wave + C::NW * i);
    run_passes<WC, -1>(L{}, v[i], lds + wave * C::NP, lane, 0, a.twf);
    for_layout<WC, FL::R>(lane, [&](auto slot, u32 kx) { S[tile_index<C>(y, kx)] = v[i][decltype(slot)::value]; });
  }
}

// ---- B: tile `slice` (16 columns) of S: staged in LDS column-major, each wave
// transforms CB columns in place (FFT, multiply by the Kt column, IFFT); the tile
// is written back.  Uses the whole workgroup and TILE elements of lds. ----
template <class C, bool L2ONLY>
__device__ __forceinline__ void conv_cols(const ConvArgs<C>& a, const cpx<float>* kt, cpx<float>* lds, u32 pair,
                                          u32 slice, int tid, int wave, int lane) {
  using SH = ConvShape<C>;
  using WC = WaveCfg<C::W, C::NP>;
  using L = typename C::fwd;
  using LI = typename C::inv;
  using F0 = typename pl_first<L>::type;
  using FL = typename pl_last<L>::type;
  using IL = typename pl_last<LI>::type;
  cpx<float>* T = a.scr + (u64)pair * (u64)(C::H * C::W) + slice * (u32)(16 * C::H);
  constexpr int IT = 16 * C::H / SH::NT;  // tile elements per thread
  static_assert(16 * C::H % SH::NT == 0, "tile split");
  {
    cpx<float> t[IT];  // all loads in flight before the first LDS write
#pragma unroll
    for (int i = 0; i < IT; ++i) t[i] = ld_c64<L2ONLY>(T + tid + i * SH::NT);
#pragma unroll
    for (int i = 0; i < IT; ++i) {
      const int e = tid + i * SH::NT, y = e >> 4, c = e & 15;
      lds[c * C::NP + y + (y >> 5)] = t[i];
    }
  }
  __syncthreads();
#pragma nounroll
  for (int cc = 0; cc < SH::CB; ++cc) {
    const int c = wave * SH::CB + cc;
    const u32 kx = slice * 16u + (u32)c;
    cpx<float>* col = lds + c * C::NP;
    cpx<float> kv[SH::NV];
    for_layout<WC, FL::R>(lane, [&](auto slot, u32 ky) { kv[decltype(slot)::value] = ld_c64<L2ONLY>(kt + kx * C::H + ky); });
    cpx<float> v[SH::NV];
    for_layout<WC, F0::R>(lane, [&](auto slot, u32 n) { v[decltype(slot)::value] = col[n + (n >> 5)]; });
    run_passes<WC, -1>(L{}, v, col, lane, 0, a.twf);
#pragma unroll
    for (int i = 0; i < SH::NV; ++i) v[i] = cmul(v[i], kv[i]);
    block_sync<WC>();
    run_passes<WC, 1>(LI{}, v, col, lane, 0, a.twi);
    block_sync<WC>();
    for_layout<WC, IL::R>(lane, [&](auto slot, u32 n) { col[n + (n >> 5)] = v[decltype(slot)::value]; });
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < IT; ++i) {
    const int e = tid + i * SH::NT, y = e >> 4, c = e & 15;
    T[e] = lds[c * C::NP + y + (y >> 5)];
  }
}

// ---- C: row IFFTs of rows [y0, y0+RA) from S; real part -> image a, imaginary
// part -> image b ----
template <class C, bool L2ONLY>
__device__ __forceinline__ void conv_rows_inv(const ConvArgs<C>& a, cpx<float>* lds, u32 pair, u32 slice, int wave,
                                              int lane) {
  using SH = ConvShape<C>;
  using WC = WaveCfg<C::W, C::NP>;
  using LI = typename C::inv;
  using I0 = typename pl_first<LI>::type;
  using IL = typename pl_last<LI>::type;
  const u32 ia = 2 * pair, ib = 2 * pair + 1;
  const bool has_b = ib < C::NIMG;
  const u32 y0 = slice * C::RA;
  const cpx<float>* S = a.scr + (u64)pair * (u64)(C::H * C::W);
  cpx<float> v[SH::RPW][SH::NV];
#pragma unroll
  for (int i = 0; i < SH::RPW; ++i) {
    const u32 y = y0 + (u32)(wave + C::NW * i);
    for_layout<WC, I0::R>(lane, [&](auto slot, u32 kx) {
      v[i][decltype(slot)::value] = ld_c64<L2ONLY>(S + tile_index<C>(y, kx));
    });
  }
#pragma unroll
  for (int i = 0; i < SH::RPW; ++i) {
    const u32 y = y0 + (u32)(wave + C::NW * i);
    run_passes<WC, 1>(LI{}, v[i], lds + wave * C::NP, lane, 0, a.twi);
    float* ra = a.x + (u64)ia * C::IMG + y * C::PITCH;
    float* rb = a.x + (u64)ib * C::IMG + y * C::PITCH;
    for_layout<WC, IL::R>(lane, [&](auto slot, u32 n) {
      const cpx<float> z = v[i][decltype(slot)::value];
      ra[n] = z.x;
      if (has_b) rb[n] = z.y;
    });
  }
}

// ---- column-major layout (S[kx * H + ky]): the transposes move into the row
// kernels (LDS tile of W x TP, RA-element runs per column) and the column kernel
// streams whole contiguous columns, each wave on its own CB columns. ----
template <class C>
__device__ __forceinline__ void conv_rows_fwd_cm(const ConvArgs<C>& a, cpx<float>* lds, u32 pair, u32 slice,
                                                 int tid, int wave, int lane) {
  using SH = ConvShape<C>;
  using WC = WaveCfg<C::W, C::NP>;
  using L = typename C::fwd;
  using F0 = typename pl_first<L>::type;
  using FL = typename pl_last<L>::type;
  constexpr int IT = C::W * C::RA / SH::NT;
  static_assert(C::W * C::RA % SH::NT == 0 && C::TP > C::RA, "transpose tile");
  const u32 ia = 2 * pair, ib = 2 * pair + 1;
  const bool has_b = ib < C::NIMG;
  const u32 y0 = slice * C::RA;
  cpx<float>* S = a.scr + (u64)pair * (u64)(C::H * C::W);
  TwSet<WC, L, conv_twreg<WC, L>()> ts;
  ts.load(lane, a.twf);  // in flight with the row loads
  cpx<float> v[SH::RPW][SH::NV];
#pragma unroll
  for (int i = 0; i < SH::RPW; ++i) {
    const u32 y = y0 + (u32)(wave + C::NW * i);
    const float* ra = a.x + (u64)ia * C::IMG + y * C::PITCH;
    const float* rb = a.x + (u64)ib * C::IMG + y * C::PITCH;
    for_layout<WC, F0::R>(lane, [&](auto slot, u32 n) {
      v[i][decltype(slot)::value] = cpx<float>{ra[n], has_b ? rb[n] : 0.f};
    });
  }
#pragma unroll
  for (int i = 0; i < SH::RPW; ++i) ts.template run<-1>(v[i], lds + wave * C::NP, lane, 0);
  __syncthreads();  // the transpose tile overlaps the exchange regions
#pragma unroll
  for (int i = 0; i < SH::RPW; ++i) {
    const u32 yl = (u32)(wave + C::NW * i);
    for_layout<WC, FL::R>(lane, [&](auto slot, u32 kx) { lds[kx * C::TP + yl] = v[i][decltype(slot)::value]; });
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < IT; ++i) {
    const u32 e = (u32)(tid + i * SH::NT), yl = e % C::RA, kx = e / C::RA;
    S[kx * C::H + y0 + yl] = lds[kx * C::TP + yl];
  }
}

// Column task: CB columns per wave for PP consecutive pairs; each wave loads its
// Kt columns once and reuses them for every pair of the task.
template <class C, bool L2ONLY>
__device__ __forceinline__ void conv_cols_cm(const ConvArgs<C>& a, const cpx<float>* kt, cpx<float>* lds, u32 pair0,
                                             u32 slice, int wave, int lane) {
  using SH = ConvShape<C>;
  using WC = WaveCfg<C::W, C::NP>;
  using L = typename C::fwd;
  using LI = typename C::inv;
  using F0 = typename pl_first<L>::type;
  using FL = typename pl_last<L>::type;
  using IL = typename pl_last<LI>::type;
  const u32 kx0 = slice * (u32)(C::NW * C::CB) + (u32)(wave * C::CB);
  cpx<float>* xch = lds + wave * C::NP;
  // every transform of the task uses the same per-lane twiddles: read them once.  A
  // twiddle table depends only on N and the radix list (the inverse passes conjugate),
  // so when the inverse list equals the forward one both directions share one set.
  constexpr bool SAME = __is_same(L, LI);
  TwSet<WC, L, conv_twreg<WC, L>()> tsf;
  TwSet<WC, LI, SAME ? false : conv_twreg<WC, LI>()> tsi_own;
  tsf.load(lane, a.twf);
  tsi_own.load(lane, a.twi);
#pragma unroll
  for (int cc = 0; cc < C::CB; ++cc) {
    const u32 kx = kx0 + (u32)cc;
    cpx<float> kv[SH::NV];
    for_layout<WC, FL::R>(lane, [&](auto slot, u32 ky) { kv[decltype(slot)::value] = ld_c64<L2ONLY>(kt + kx * C::H + ky); });
#pragma nounroll
    for (u32 pp = 0; pp < (u32)C::PP; ++pp) {
      const u32 pair = pair0 + pp;
      if (pair >= C::NPAIRS) break;
      cpx<float>* col = a.scr + (u64)pair * (u64)(C::H * C::W) + kx * C::H;
      cpx<float> v[SH::NV];
      for_layout<WC, F0::R>(lane, [&](auto slot, u32 n) { v[decltype(slot)::value] = ld_c64<L2ONLY>(col + n); });
      tsf.template run<-1>(v, xch, lane, 0);
#pragma unroll
      for (int i = 0; i < SH::NV; ++i) v[i] = cmul(v[i], kv[i]);
      block_sync<WC>();
      if constexpr (SAME) tsf.template run<1>(v, xch, lane, 0);
      else tsi_own.template run<1>(v, xch, lane, 0);
      for_layout<WC, IL::R>(lane, [&](auto slot, u32 n) { col[n] = v[decltype(slot)::value]; });
      block_sync<WC>();
    }
  }
}

template <class C, bool L2ONLY>
__device__ __forceinline__ void conv_rows_inv_cm(const ConvArgs<C>& a, cpx<float>* lds, u32 pair, u32 slice,
                                                 int tid, int wave, int lane) {
  using SH = ConvShape<C>;
  using WC = WaveCfg<C::W, C::NP>;
  using LI = typename C::inv;
  using I0 = typename pl_first<LI>::type;
  using IL = typename pl_last<LI>::type;
  constexpr int IT = C::W * C::RA / SH::NT;
  const u32 ia = 2 * pair, ib = 2 * pair + 1;
  const bool has_b = ib < C::NIMG;
  const u32 y0 = slice * C::RA;
  const cpx<float>* S = a.scr + (u64)pair * (u64)(C::H * C::W);
  TwSet<WC, LI, conv_twreg<WC, LI>()> ts;
  ts.load(lane, a.twi);  // in flight with the tile loads; reused by every row
  {
    cpx<float> t[IT];  // all loads in flight before the first LDS write
#pragma unroll
    for (int i = 0; i < IT; ++i) {
      const u32 e = (u32)(tid + i * SH::NT), yl = e % C::RA, kx = e / C::RA;
      t[i] = ld_c64<L2ONLY>(S + kx * C::H + y0 + yl);
    }
#pragma unroll
    for (int i = 0; i < IT; ++i) {
      const u32 e = (u32)(tid + i * SH::NT), yl = e % C::RA, kx = e / C::RA;
      lds[kx * C::TP + yl] = t[i];
    }
  }
  __syncthreads();
  cpx<float> v[SH::RPW][SH::NV];
#pragma unroll
  for (int i = 0; i < SH::RPW; ++i) {
    const u32 yl = (u32)(wave + C::NW * i);
    for_layout<WC, I0::R>(lane, [&](auto slot, u32 kx) { v[i][decltype(slot)::value] = lds[kx * C::TP + yl]; });
  }
  __syncthreads();  // the exchange regions overlap the transpose tile
#pragma unroll
  for (int i = 0; i < SH::RPW; ++i) {
    const u32 y = y0 + (u32)(wave + C::NW * i);
    ts.template run<1>(v[i], lds + wave * C::NP, lane, 0);
    float* ra = a.x + (u64)ia * C::IMG + y * C::PITCH;
    float* rb = a.x + (u64)ib * C::IMG + y * C::PITCH;
    for_layout<WC, IL::R>(lane, [&](auto slot, u32 n) {
      const cpx<float> z = v[i][decltype(slot)::value];
      ra[n] = z.x;
      if (has_b) rb[n] = z.y;
    });
  }
}

// One phase over all pairs as an ordinary (dependency-free) kernel.  PHASE 0: the
// row-FFT tasks of every pair (workgroup w < NPAIRS*NA), then the NK kernel-prep
// tasks; PHASE 1: column tiles, tile-major over the workgroups (workgroups that run
// together share one Kt tile in L2); PHASE 2: row-IFFT tasks.
template <class C, int PHASE>
__device__ __forceinline__ void conv2d_phase_body(const KParams& prm) {
  using SH = ConvShape<C>;
  constexpr bool CM = C::LAYOUT == 1;
  constexpr int ROWS_LDS = CM ? (C::W * C::TP > SH::XCH ? C::W * C::TP : SH::XCH) : SH::XCH;
  constexpr int LDS_ELEMS = PHASE == 1 ? (CM ? SH::XCH : SH::TILE) : ROWS_LDS;
  constexpr u32 NBT = CM ? (u32)(C::W / (C::NW * C::CB)) : SH::NB;  // column tasks per pair (group)
  constexpr u32 NPG = CM ? (C::NPAIRS + C::PP - 1) / C::PP : C::NPAIRS;  // pair groups
  __shared__ cpx<float> lds[LDS_ELEMS];
  const ConvArgs<C> a = conv_args<C>(prm);
  const int tid = (int)threadIdx.x;
  const int lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const u32 wg = blockIdx.x;
  if constexpr (PHASE == 0) {
    if (wg < C::NPAIRS * SH::NA) {
      if constexpr (CM) conv_rows_fwd_cm<C>(a, lds, wg / SH::NA, wg % SH::NA, tid, wave, lane);
      else conv_rows_fwd<C>(a, lds, wg / SH::NA, wg % SH::NA, wave, lane);
    } else if (wg < C::NPAIRS * SH::NA + SH::NK) {
      conv_kprep<C>(a.K, a.kt, wg - C::NPAIRS * SH::NA, tid);
    }
  } else if constexpr (PHASE == 1) {
    if (wg < NPG * NBT) {
      if constexpr (CM) conv_cols_cm<C, false>(a, a.kt, lds, (wg % NPG) * C::PP, wg / NPG, wave, lane);
      else conv_cols<C, false>(a, a.kt, lds, wg % C::NPAIRS, wg / C::NPAIRS, tid, wave, lane);
    }
  } else {
    if (wg < C::NPAIRS * SH::NA) {
      if constexpr (CM) conv_rows_inv_cm<C, false>(a, lds, wg / SH::NA, wg % SH::NA, tid, wave, lane);
      else conv_rows_inv<C, false>(a, lds, wg / SH::NA, wg % SH::NA, wave, lane);
    }
  }
}

// The phases in one cooperative launch (see the top of this file).  Slot 3: the
// control block; slot 4: the task table; slot 5: one Kt per queue.
template <class C>
__device__ __forceinline__ void conv2d_pair_body(const KParams& prm) {
  using SH = ConvShape<C>;
  static_assert(C::LAYOUT == 0, "the cooperative kernel uses the tile-major layout");
  __shared__ cpx<float> lds[SH::LDS_ALL];
  __shared__ u32 s_task;
  const ConvArgs<C> a = conv_args<C>(prm);
  u32* ctl = bufp<u32>(prm, 3);
  const u32* table = bufp<const u32>(prm, 4);

  const int tid = (int)threadIdx.x;
  const int lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  // queue = the XCD this workgroup actually runs on
  const u32 q = __builtin_amdgcn_readfirstlane(__builtin_amdgcn_s_getreg((3 << 11) | (0 << 6) | 20) & 7u);
  u32* err = ctl + CTL_ERROR;

  // qq: the queue the task belongs to (its Kt copy and kernel-prep counter)
  auto run_task = [&](u32 word, u32 qq) {
    const u32 phase = word >> 30;
    const u32 pair = (word >> 12) & 0x3FFFFu;
    const u32 slice = word & 0xFFFu;
    u32* cnt = ctl + CTL_PAIRS + 64 * pair;
    cpx<float>* ktq = a.kt + (u64)qq * (u64)(C::H * C::W);
    if (phase == PHASE_A) {
      conv_rows_fwd<C>(a, lds, pair, slice, wave, lane);
      signal_done(cnt);
    } else if (phase == PHASE_B) {
      wait_count(ctl + CTL_KDONE + 32 * qq, SH::NK, err);
      wait_count(cnt, SH::NA, err);
      conv_cols<C, true>(a, ktq, lds, pair, slice, tid, wave, lane);
      signal_done(cnt + 32);
    } else if (phase == PHASE_C) {
      wait_count(cnt + 32, SH::NB, err);
      conv_rows_inv<C, true>(a, lds, pair, slice, wave, lane);
      __syncthreads();  // LDS reuse by the next task
    } else {
      conv_kprep<C>(a.K, ktq, slice, tid);
      signal_done(ctl + CTL_KDONE + 32 * qq);
    }
  };

  // ---- take tickets from this XCD's queue ----
  {
    const u32 t_begin = table[q], ntask = table[q + 1] - table[q];
    u32* ticket = ctl + CTL_TICKET + 32 * q;
    if (tid == 0) s_task = add_agent(ticket, 1u);
    __syncthreads();
    u32 cur = __builtin_amdgcn_readfirstlane(s_task);
    __syncthreads();
    while (cur < ntask) {
      u32 nxt = 0;
      if (tid == 0) nxt = add_agent(ticket, 1u);  // next ticket, needed after this task
      run_task(table[9 + t_begin + cur], q);
      if (tid == 0) s_task = nxt;
      __syncthreads();
      cur = __builtin_amdgcn_readfirstlane(s_task);
      __syncthreads();
    }
  }

  // ---- the last workgroup to leave serves queues that no workgroup ran on (an
  // XCD that received none), then resets the control block for the next launch ----
  if (tid == 0) s_task = add_agent(ctl + CTL_EXIT, 1u);
  __syncthreads();
  const bool last = __builtin_amdgcn_readfirstlane(s_task) == gridDim.x - 1;
  __syncthreads();
  if (last) {
    for (u32 qq = 0; qq < (u32)CONV_QUEUES; ++qq) {
      if (table[qq + 1] == table[qq] || ld_agent(ctl + CTL_TICKET + 32 * qq) != 0) continue;
      if (tid == 0) __hip_atomic_fetch_or(err, 4u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      for (u32 t = table[qq]; t < table[qq + 1]; ++t) run_task(table[9 + t], qq);
    }
    if (tid == 0) {
      for (u32 p = 0; p < C::NPAIRS; ++p) {
        st_agent(ctl + CTL_PAIRS + 64 * p, 0u);
        st_agent(ctl + CTL_PAIRS + 64 * p + 32, 0u);
      }
      for (u32 qq = 0; qq < (u32)CONV_QUEUES; ++qq) {
        st_agent(ctl + CTL_TICKET + 32 * qq, 0u);
        st_agent(ctl + CTL_KDONE + 32 * qq, 0u);
      }
      st_agent(ctl + CTL_EXIT, 0u);
    }
  }
}

}  // namespace vkx
