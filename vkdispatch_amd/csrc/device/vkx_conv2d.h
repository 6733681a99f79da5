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
// Data flow per image pair: row FFTs of z, then column FFT . Kt . column IFFT,
// then row IFFTs (real part -> image a, imaginary part -> image b), with
//   Kt(ky, kx) = conj(Kf[ky][kx]) * SCALE.
// Two forms:
//  * conv2d_phase_body: three ordinary kernels in stream order over all pairs,
//    through a column-major scratch image per pair (S[kx * H + ky]) and a
//    column-major Kt prepared by the first kernel.
//  * conv2d_xcd_body: one launch; the three phases become tasks of per-XCD queues
//    and the transposes between them stay inside one XCD's L2 (description at the
//    body).  Its scratch is tile-major: 16-column tiles, each H rows x 16 columns
//    (128-byte rows) contiguous, element (ky, kx) at (kx / 16) * 16H + ky * 16 +
//    kx % 16, so row tasks read and write whole 128-byte lines from registers and a
//    column tile is one contiguous block.
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
// 16.16.4 keep reading the table (registers are better spent on data there).
template <class WC, class L> constexpr bool conv_twreg() { return list_tw_regs<WC, L>::value <= 16; }

constexpr u32 CONV_SPIN_LIMIT = 1u << 18;

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

// scratch / Kt load: L1-bypassing where other workgroups of the same launch wrote
// the data, an ordinary cached load in the phase kernels
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

// ---------------------------------------------------------------------------
// Single-launch form, XCD-resident exchange (conv2d_xcd_body).
//
// The three phases per image pair as tasks of one launch.  Each XCD has its own
// task queue and its own NSLOT pair spectra of scratch (the slot images), and a
// workgroup takes tasks only from the queue of the XCD it runs on (XCC_ID
// register), so every value a task reads from scratch was written on the same XCD
// and is found in that XCD's L2 (the transposes never leave the XCD unless the L2
// evicts them).  Pair i of queue q is pair q + 8 i and lives in slot i % NSLOT.
// Slot images are tile-major with CT-column tiles: element (ky, kx) at
// (kx / CT) CT H + ky CT + kx % CT, so a row task writes or reads CT-element runs
// from registers and a column tile is one contiguous block.
//
// Tasks (slice s = RA rows, tile c = CT columns):
//   A(i, s)     row FFTs of pair i, rows of slice s -> slot
//   B(i, c)     column FFT . Kt . column IFFT of tile c, in place in the slot; Kt is
//               the column-major conj(Kf) SCALE prepared by conv2d_kt_body
//   CA(i, s)    row IFFTs of pair i, slice s -> images, then A(i + NSLOT, s): the
//               next pair of the slot enters the rows pair i just left
//   C(i, s)     row IFFTs only (the last NSLOT pairs of a queue)
// A queue lists A(0..NSLOT-1), then per group of NSLOT pairs all their B tasks,
// then all their CA / C tasks.  B(i) waits for every row slice of pair i (the A or
// CA tasks that wrote it); CA(i) / C(i) wait for every tile of B(i).  A task waits
// only for tasks earlier in its queue, which running workgroups already hold, so
// the launch cannot deadlock; every wait is bounded (an expired wait sets error
// bit 0 and the task goes on; the host reports it).  Several small workgroups per
// CU run unrelated tasks side by side, so one's memory waits overlap another's
// arithmetic.
//
// Memory ordering inside one XCD: a task's stores complete to L2 (s_waitcnt) before
// its workgroup bumps the done counter; a waiting workgroup polls the counter, then
// reads slot data with device-scope loads, which miss the CU's L1 and read the
// XCD's L2 (or memory behind it).  No agent-scope release is used: on this part it
// writes back the whole L2, and the data never crosses XCDs.
//
// The last workgroup to leave runs, alone and in order, the tasks of any queue no
// workgroup took (an XCD that received none), then resets the control block, so
// every launch starts from zero without a memset (graph capture safe).
// ---------------------------------------------------------------------------
constexpr u32 XT_A = 0, XT_B = 1, XT_CA = 2, XT_C = 3;
constexpr u32 XCTL_TICKET = 0;  // + 32 q: next ticket of queue q
constexpr u32 XCTL_EXIT = 256;  // workgroups that have left
constexpr u32 XCTL_ERROR = 288;  // bit 0: a wait expired; bit 2: the last workgroup served a queue
constexpr u32 XCTL_CNT = 512;   // + 64 (q QMAX + i): rows of pair i done at +0, tiles done at +32

template <class C> struct Xcd {
  using WC = WaveCfg<C::W, C::NP>;
  static constexpr int CT = C::CT;           // columns per tile
  static constexpr int NV = C::W / 64;       // points per lane of one transform
  static constexpr int RPW = C::RA / C::NW;  // rows per wave in a row task
  static constexpr int NT = C::NW * 64;
  static constexpr int CB = CT / C::NW;      // columns per wave in a tile task
  static constexpr u32 NS = C::H / C::RA;    // row slices per pair
  static constexpr u32 NB = C::W / CT;       // column tiles per pair
  static constexpr u64 PAIR = (u64)C::H * (u64)C::W;
  static constexpr int IT = CT * C::H / NT;  // tile elements per thread
  static constexpr int LDS = (CT > C::NW ? CT : C::NW) * C::NP;
  static_assert(C::H == C::W && C::W % 64 == 0, "square images, rows a multiple of 64");
  static_assert(C::RA % C::NW == 0 && CT % C::NW == 0 && (CT * C::H) % NT == 0, "task split");
  static_assert(64 % CT == 0 && (CT & (CT - 1)) == 0, "tile width a power of two dividing a wave");
  static_assert(C::NP >= C::H + C::H / 32, "LDS column region");
};

// C::XNT (optional, default false): the streamed traffic (input rows, output rows, Kt)
// uses non-temporal accesses so that it does not push the slot images out of L2
template <class C, class = void> struct xcd_nt_of { static constexpr bool value = false; };
template <class C> struct xcd_nt_of<C, decltype((void)C::XNT)> { static constexpr bool value = C::XNT != 0; };
template <class C> constexpr bool xcd_nt() { return xcd_nt_of<C>::value; }

template <class C> __device__ __forceinline__ u32 xcd_tile_index(u32 ky, u32 kx) {
  return (kx / (u32)C::CT) * (u32)(C::CT * C::H) + ky * (u32)C::CT + (kx % (u32)C::CT);
}

template <class C> __device__ __forceinline__ u32* xcd_cnt(u32* ctl, u32 q, u32 i) {
  return ctl + XCTL_CNT + 64u * (q * (u32)C::QMAX + i);
}

// Kt for the single launch: block `slice` (64 x 64) of conj(Kf) SCALE, column-major
// (kt[kx H + ky]); one workgroup per block
template <class C> __device__ __forceinline__ void conv2d_kt_body(const KParams& prm) {
  conv_kprep<C>(bufp<const cpx<float>>(prm, 1), bufp<cpx<float>>(prm, 5), blockIdx.x, (int)threadIdx.x);
}

// the RPW rows of this wave (slice rows y0 + wave + NW j), in the first-pass layout
template <class C>
__device__ __forceinline__ void xcd_load_rows(const float* x, u32 pair, u32 slice, int wave, int lane,
                                              float (&ra)[Xcd<C>::RPW][Xcd<C>::NV],
                                              float (&rb)[Xcd<C>::RPW][Xcd<C>::NV]) {
  using X = Xcd<C>;
  using F0 = typename pl_first<typename C::fwd>::type;
  const bool has_b = 2 * pair + 1 < C::NIMG;
#pragma unroll
  for (int j = 0; j < X::RPW; ++j) {
    const u32 y = slice * C::RA + (u32)(wave + C::NW * j);
    const float* pa = x + (u64)(2 * pair) * C::IMG + y * C::PITCH;
    const float* pb = pa + C::IMG;
    for_layout<typename X::WC, F0::R>(lane, [&](auto slot, u32 n) {
      if constexpr (xcd_nt<C>()) {
        ra[j][decltype(slot)::value] = __builtin_nontemporal_load(pa + n);
        rb[j][decltype(slot)::value] = has_b ? __builtin_nontemporal_load(pb + n) : 0.f;
      } else {
        ra[j][decltype(slot)::value] = pa[n];
        rb[j][decltype(slot)::value] = has_b ? pb[n] : 0.f;
      }
    });
  }
}

// row FFTs of z = x_a + i x_b into the slot
template <class C, class TS>
__device__ __forceinline__ void xcd_rows_fwd(const TS& ts, cpx<float>* S, u32 slice, cpx<float>* xch, int wave,
                                             int lane, const float (&ra)[Xcd<C>::RPW][Xcd<C>::NV],
                                             const float (&rb)[Xcd<C>::RPW][Xcd<C>::NV]) {
  using X = Xcd<C>;
  using FL = typename pl_last<typename C::fwd>::type;
#pragma unroll
  for (int j = 0; j < X::RPW; ++j) {
    const u32 y = slice * C::RA + (u32)(wave + C::NW * j);
    cpx<float> v[X::NV];
#pragma unroll
    for (int k = 0; k < X::NV; ++k) v[k] = cpx<float>{ra[j][k], rb[j][k]};
    ts.template run<-1>(v, xch, lane, 0);
    for_layout<typename X::WC, FL::R>(lane, [&](auto slot, u32 kx) {
      S[xcd_tile_index<C>(y, kx)] = v[decltype(slot)::value];
    });
  }
}

// row IFFTs of the slot rows of this wave; real parts -> image a, imaginary -> image b
template <class C, class TS>
__device__ __forceinline__ void xcd_rows_inv(const TS& ts, const cpx<float>* S, float* x, u32 pair, u32 slice,
                                             cpx<float>* xch, int wave, int lane) {
  using X = Xcd<C>;
  using I0 = typename pl_first<typename C::inv>::type;
  using IL = typename pl_last<typename C::inv>::type;
  const bool has_b = 2 * pair + 1 < C::NIMG;
  cpx<float> v[X::RPW][X::NV];
#pragma unroll
  for (int j = 0; j < X::RPW; ++j) {  // every load of the task in flight first
    const u32 y = slice * C::RA + (u32)(wave + C::NW * j);
    for_layout<typename X::WC, I0::R>(lane, [&](auto slot, u32 kx) {
      v[j][decltype(slot)::value] = ld_c64_l2(S + xcd_tile_index<C>(y, kx));
    });
  }
#pragma unroll
  for (int j = 0; j < X::RPW; ++j) {
    const u32 y = slice * C::RA + (u32)(wave + C::NW * j);
    ts.template run<1>(v[j], xch, lane, 0);
    float* pa = x + (u64)(2 * pair) * C::IMG + y * C::PITCH;
    float* pb = pa + C::IMG;
    for_layout<typename X::WC, IL::R>(lane, [&](auto slot, u32 n) {
      const cpx<float> z = v[j][decltype(slot)::value];
      if constexpr (xcd_nt<C>()) {
        __builtin_nontemporal_store(z.x, pa + n);
        if (has_b) __builtin_nontemporal_store(z.y, pb + n);
      } else {
        pa[n] = z.x;
        if (has_b) pb[n] = z.y;
      }
    });
  }
}

// tile task: CT columns of the slot, in place
template <class C, class TF, class TI>
__device__ __forceinline__ void xcd_tile(const TF& tsf, const TI& tsi, const cpx<float>* kt, cpx<float>* S, u32 tile,
                                         cpx<float>* lds, int tid, int wave, int lane) {
  using X = Xcd<C>;
  using F0 = typename pl_first<typename C::fwd>::type;
  using FL = typename pl_last<typename C::fwd>::type;
  using IL = typename pl_last<typename C::inv>::type;
  constexpr int CT = X::CT;
  // 1. the tile (CT x H contiguous elements) into LDS, column-major
  cpx<float>* T = S + tile * (u32)(CT * C::H);
  {
    cpx<float> t[X::IT];
#pragma unroll
    for (int i = 0; i < X::IT; ++i) t[i] = ld_c64_l2(T + tid + i * X::NT);
#pragma unroll
    for (int i = 0; i < X::IT; ++i) {
      const int e = tid + i * X::NT, y = e / CT, c = e % CT;
      lds[c * C::NP + y + (y >> 5)] = t[i];
    }
  }
  __syncthreads();
  // 2. per wave: its columns, FFT, product with the Kt column (read while the FFT
  //    runs), IFFT, in their own LDS regions
#pragma unroll
  for (int cc = 0; cc < X::CB; ++cc) {
    const u32 kx = tile * (u32)CT + (u32)(wave * X::CB + cc);
    cpx<float>* col = lds + (wave * X::CB + cc) * C::NP;
    cpx<float> kv[X::NV];
    for_layout<typename X::WC, FL::R>(lane, [&](auto slot, u32 ky) {
      if constexpr (xcd_nt<C>()) {
        const float* kp = reinterpret_cast<const float*>(kt + kx * C::H + ky);
        kv[decltype(slot)::value] = cpx<float>{__builtin_nontemporal_load(kp), __builtin_nontemporal_load(kp + 1)};
      } else {
        kv[decltype(slot)::value] = kt[kx * C::H + ky];
      }
    });
    cpx<float> v[X::NV];
    for_layout<typename X::WC, F0::R>(lane, [&](auto slot, u32 n) { v[decltype(slot)::value] = col[n + (n >> 5)]; });
    tsf.template run<-1>(v, col, lane, 0);
#pragma unroll
    for (int k = 0; k < X::NV; ++k) v[k] = cmul(v[k], kv[k]);
    block_sync<typename X::WC>();
    tsi.template run<1>(v, col, lane, 0);
    block_sync<typename X::WC>();
    for_layout<typename X::WC, IL::R>(lane, [&](auto slot, u32 n) { col[n + (n >> 5)] = v[decltype(slot)::value]; });
  }
  __syncthreads();
  // 3. the tile back
#pragma unroll
  for (int i = 0; i < X::IT; ++i) {
    const int e = tid + i * X::NT, y = e / CT, c = e % CT;
    T[e] = lds[c * C::NP + y + (y >> 5)];
  }
}

// Slot 0: the images (in place), 1: the half spectrum K (unused here), 2: slot
// scratch (8 NSLOT pair images), 3: the control block, 4: the task table (9 queue
// offsets, then task words kind << 30 | local pair << 12 | slice), 5: Kt.
template <class C>
__device__ __forceinline__ void conv2d_xcd_body(const KParams& prm) {
  using X = Xcd<C>;
  using L = typename C::fwd;
  using LI = typename C::inv;
  __shared__ cpx<float> lds[X::LDS];
  __shared__ u32 s_word;
  float* x = bufp<float>(prm, 0);
  cpx<float>* scr = bufp<cpx<float>>(prm, 2);
  u32* ctl = bufp<u32>(prm, 3);
  const u32* table = bufp<const u32>(prm, 4);
  const cpx<float>* kt = bufp<const cpx<float>>(prm, 5);
  const cpx<float>* twf = bufp<const cpx<float>>(prm, SLOT_TW_FWD);
  const cpx<float>* twi = bufp<const cpx<float>>(prm, SLOT_TW_INV);

  const int tid0 = (int)threadIdx.x;
  const int wave = __builtin_amdgcn_readfirstlane(tid0 >> 6);
  const u32 xq = __builtin_amdgcn_readfirstlane(__builtin_amdgcn_s_getreg((3 << 11) | (0 << 6) | 20) & 7u);
  u32* err = ctl + XCTL_ERROR;

  // Twiddles come from the table (L1 / L2 hits): at four workgroups per CU the 28
  // registers of a register-resident set would push the tasks into spills, and the
  // table reads' latency is covered by the other workgroups of the CU.
  constexpr bool SAME = __is_same(L, LI);
  using TSF = TwSet<typename X::WC, L, false>;
  using TSI = TwSet<typename X::WC, LI, false>;

  auto run_task = [&](u32 word, u32 q) {
    // The thread index passes through an empty asm statement per task, so the
    // compiler derives every address from it inside the task: hoisted out of the
    // task loop, the per-thread LDS and global offsets of all task kinds would stay
    // live for the whole launch and crowd out the data registers.
    int tid = tid0;
    asm volatile("" : "+v"(tid));
    const int lane = tid & 63;
    cpx<float>* xch = lds + wave * C::NP;
    const u32 kind = word >> 30;
    const u32 i = (word >> 12) & 0x3FFFFu;
    const u32 s = word & 0xFFFu;
    const u32 pair = q + 8u * i;
    cpx<float>* S = scr + (u64)(q * (u32)C::NSLOT + i % (u32)C::NSLOT) * X::PAIR;
    u32* cnt = xcd_cnt<C>(ctl, q, i);
    if (kind == XT_A) {
      TSF tsf;
      tsf.load(lane, twf);
      float ra[X::RPW][X::NV], rb[X::RPW][X::NV];
      xcd_load_rows<C>(x, pair, s, wave, lane, ra, rb);
      xcd_rows_fwd<C>(tsf, S, s, xch, wave, lane, ra, rb);
      signal_done(cnt);
    } else if (kind == XT_B) {
      wait_count(cnt, X::NS, err);
      TSF tsf;
      tsf.load(lane, twf);
      if constexpr (SAME) {
        xcd_tile<C>(tsf, tsf, kt, S, s, lds, tid, wave, lane);
      } else {
        TSI tsi;
        tsi.load(lane, twi);
        xcd_tile<C>(tsf, tsi, kt, S, s, lds, tid, wave, lane);
      }
      signal_done(cnt + 32);
    } else {
      wait_count(cnt + 32, X::NB, err);
      TSI tsi;
      tsi.load(lane, twi);
      xcd_rows_inv<C>(tsi, S, x, pair, s, xch, wave, lane);
      if (kind == XT_CA) {
        // the next pair of this slot enters exactly the rows this wave just read
        TSF tsf;
        tsf.load(lane, twf);
        float ra[X::RPW][X::NV], rb[X::RPW][X::NV];
        xcd_load_rows<C>(x, pair + 8u * (u32)C::NSLOT, s, wave, lane, ra, rb);
        xcd_rows_fwd<C>(tsf, S, s, xch, wave, lane, ra, rb);
        signal_done(xcd_cnt<C>(ctl, q, i + (u32)C::NSLOT));
      } else {
        __syncthreads();  // LDS reuse by the next task
      }
    }
  };

  // ---- take tickets from this XCD's queue (the next one is requested while the
  // current task runs) ----
  {
    const u32 t_begin = table[xq], ntask = table[xq + 1] - table[xq];
    u32* ticket = ctl + XCTL_TICKET + 32 * xq;
    if (tid0 == 0) s_word = add_agent(ticket, 1u);
    __syncthreads();
    u32 cur = __builtin_amdgcn_readfirstlane(s_word);
    __syncthreads();
    while (cur < ntask) {
      u32 nxt = 0;
      if (tid0 == 0) nxt = add_agent(ticket, 1u);
      run_task(table[9 + t_begin + cur], xq);
      if (tid0 == 0) s_word = nxt;
      __syncthreads();
      cur = __builtin_amdgcn_readfirstlane(s_word);
      __syncthreads();
    }
  }

  // ---- the last workgroup to leave serves queues no workgroup ran on, then resets
  // the control block for the next launch ----
  if (tid0 == 0) s_word = add_agent(ctl + XCTL_EXIT, 1u);
  __syncthreads();
  const bool last = __builtin_amdgcn_readfirstlane(s_word) == gridDim.x - 1;
  __syncthreads();
  if (last) {
    for (u32 qq = 0; qq < 8u; ++qq) {
      if (table[qq + 1] == table[qq] || ld_agent(ctl + XCTL_TICKET + 32 * qq) != 0) continue;
      if (tid0 == 0) __hip_atomic_fetch_or(err, 4u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      for (u32 t = table[qq]; t < table[qq + 1]; ++t) run_task(table[9 + t], qq);
    }
    __syncthreads();
    if (tid0 == 0) {
      for (u32 w = 0; w < 8u * (u32)C::QMAX; ++w) {
        st_agent(ctl + XCTL_CNT + 64 * w, 0u);
        st_agent(ctl + XCTL_CNT + 64 * w + 32, 0u);
      }
      for (u32 qq = 0; qq < 8u; ++qq) st_agent(ctl + XCTL_TICKET + 32 * qq, 0u);
      st_agent(ctl + XCTL_EXIT, 0u);
    }
  }
}

}  // namespace vkx
