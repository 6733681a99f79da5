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
//   Kt(ky, kx) = conj(Kf[ky][kx]) * SCALE,
// as three ordinary kernels in stream order over all pairs (conv2d_phase_body),
// through a scratch image per pair and a Kt prepared by the first kernel.  The
// scratch is column-major (S[kx * H + ky], LAYOUT 1, the default) or tile-major
// (LAYOUT 0: 16-column tiles, each H rows x 16 columns contiguous, element (ky, kx)
// at (kx / 16) * 16H + ky * 16 + kx % 16).
#pragma once
#include "vkx_fft.h"

namespace vkx {

// One transform per TT_ threads (64: one wave, exchanges synchronise the wave only;
// 128: two waves, exchanges are workgroup barriers, half the points per lane)
template <int N_, int NP_, int TT_ = 64> struct WaveCfg {
  using real = float;
  static constexpr int N = N_, TT = TT_, B = 1, PAD_SHIFT = 5, NP = NP_;
  static constexpr bool COLFAST = false, WAVE_LOCAL = TT_ == 64;
};

// threads per transform of a convolution config (C::TT, default one wave)
template <class C, class = void> struct conv_tt_of { static constexpr int value = 64; };
template <class C> struct conv_tt_of<C, decltype((void)C::TT)> { static constexpr int value = C::TT; };
template <class C> using ConvWC = WaveCfg<C::W, C::NP, conv_tt_of<C>::value>;

// Register-resident twiddles (loaded once per task, see TwSet) when a lane's set is
// small: 14 values for 512 = 8.8.8 and 9 for 256 = 4.4.4.4.  The 27 of 1024 =
// 16.16.4 on one wave, or 19 of 1024 = 8.8.4.4 on two, keep reading the table
// (registers are better spent on data there).
template <class WC, class L> constexpr bool conv_twreg() { return list_tw_regs<WC, L>::value <= 16; }

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
  static constexpr int TT = conv_tt_of<C>::value;    // threads per transform
  static constexpr int NV = C::W / TT;               // points per lane
  static constexpr int NU = C::NW * 64 / TT;         // transform units per workgroup
  static constexpr int RPW = C::RA / NU;             // rows per unit in a row task
  static constexpr int NT = C::NW * 64;              // threads per workgroup
  static constexpr int CT = 16;                      // columns per tile (128-byte rows)
  static constexpr int CB = CT / C::NW;              // columns per wave in a column task (tile layout)
  static constexpr u32 NA = C::H / C::RA;            // row tasks per pair
  static constexpr u32 NB = C::W / CT;               // column tasks (tiles) per pair
  static constexpr u32 KB = 64;                      // kernel-prep block (KB x KB)
  static constexpr u32 NK = (C::H / KB) * (C::W / KB);  // kernel-prep tasks
  static constexpr int XCH = NU * C::NP;             // per-unit exchange regions
  static constexpr int TILE = CT * C::NP;            // one column tile, column-major
  static constexpr int LDS_ALL = TILE > XCH ? TILE : XCH;
  static_assert(C::H == C::W, "square images");
  static_assert(TT == 64 || TT == 128, "one or two waves per transform");
  static_assert(C::W % TT == 0 && C::RA % NU == 0 && CT % C::NW == 0, "tiling");
  static_assert(TT == 64 || C::LAYOUT == 1, "two-wave transforms take the column-major layout");
  static_assert(C::NP >= C::H + C::H / 32, "LDS region");
};

// transform unit of a thread and its index inside the unit (wave / lane for one wave)
template <class C> __device__ __forceinline__ int conv_unit(int tid) {
  return __builtin_amdgcn_readfirstlane(tid / ConvShape<C>::TT);
}
template <class C> __device__ __forceinline__ int conv_lane(int tid) { return tid % ConvShape<C>::TT; }

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
template <class C>
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
    for (int i = 0; i < IT; ++i) t[i] = T[tid + i * SH::NT];
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
    for_layout<WC, FL::R>(lane, [&](auto slot, u32 ky) { kv[decltype(slot)::value] = kt[kx * C::H + ky]; });
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
template <class C>
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
      v[i][decltype(slot)::value] = S[tile_index<C>(y, kx)];
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
                                                 int tid) {
  using SH = ConvShape<C>;
  using WC = ConvWC<C>;
  using L = typename C::fwd;
  using F0 = typename pl_first<L>::type;
  using FL = typename pl_last<L>::type;
  constexpr int IT = C::W * C::RA / SH::NT;
  static_assert(C::W * C::RA % SH::NT == 0 && C::TP > C::RA, "transpose tile");
  const int u = conv_unit<C>(tid), t = conv_lane<C>(tid);
  const u32 ia = 2 * pair, ib = 2 * pair + 1;
  const bool has_b = ib < C::NIMG;
  const u32 y0 = slice * C::RA;
  cpx<float>* S = a.scr + (u64)pair * (u64)(C::H * C::W);
  TwSet<WC, L, conv_twreg<WC, L>()> ts;
  ts.load(t, a.twf);  // in flight with the row loads
  cpx<float> v[SH::RPW][SH::NV];
#pragma unroll
  for (int i = 0; i < SH::RPW; ++i) {
    const u32 y = y0 + (u32)(u + SH::NU * i);
    const float* ra = a.x + (u64)ia * C::IMG + y * C::PITCH;
    const float* rb = a.x + (u64)ib * C::IMG + y * C::PITCH;
    for_layout<WC, F0::R>(t, [&](auto slot, u32 n) {
      v[i][decltype(slot)::value] = cpx<float>{ra[n], has_b ? rb[n] : 0.f};
    });
  }
#pragma unroll
  for (int i = 0; i < SH::RPW; ++i) ts.template run<-1>(v[i], lds + u * C::NP, t, 0);
  __syncthreads();  // the transpose tile overlaps the exchange regions
#pragma unroll
  for (int i = 0; i < SH::RPW; ++i) {
    const u32 yl = (u32)(u + SH::NU * i);
    for_layout<WC, FL::R>(t, [&](auto slot, u32 kx) { lds[kx * C::TP + yl] = v[i][decltype(slot)::value]; });
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < IT; ++i) {
    const u32 e = (u32)(tid + i * SH::NT), yl = e % C::RA, kx = e / C::RA;
    S[kx * C::H + y0 + yl] = lds[kx * C::TP + yl];
  }
}

// Column task: CB columns per unit for PP consecutive pairs; each unit loads its
// Kt columns once and reuses them for every pair of the task.
template <class C>
__device__ __forceinline__ void conv_cols_cm(const ConvArgs<C>& a, const cpx<float>* kt, cpx<float>* lds, u32 pair0,
                                             u32 slice, int tid) {
  using SH = ConvShape<C>;
  using WC = ConvWC<C>;
  using L = typename C::fwd;
  using LI = typename C::inv;
  using F0 = typename pl_first<L>::type;
  using FL = typename pl_last<L>::type;
  using IL = typename pl_last<LI>::type;
  const int u = conv_unit<C>(tid), t = conv_lane<C>(tid);
  const u32 kx0 = slice * (u32)(SH::NU * C::CB) + (u32)(u * C::CB);
  cpx<float>* xch = lds + u * C::NP;
  // every transform of the task uses the same per-lane twiddles: read them once.  A
  // twiddle table depends only on N and the radix list (the inverse passes conjugate),
  // so when the inverse list equals the forward one both directions share one set.
  constexpr bool SAME = __is_same(L, LI);
  TwSet<WC, L, conv_twreg<WC, L>()> tsf;
  TwSet<WC, LI, SAME ? false : conv_twreg<WC, LI>()> tsi_own;
  tsf.load(t, a.twf);
  tsi_own.load(t, a.twi);
#pragma unroll
  for (int cc = 0; cc < C::CB; ++cc) {
    const u32 kx = kx0 + (u32)cc;
    cpx<float> kv[SH::NV];
    for_layout<WC, FL::R>(t, [&](auto slot, u32 ky) { kv[decltype(slot)::value] = kt[kx * C::H + ky]; });
#pragma nounroll
    for (u32 pp = 0; pp < (u32)C::PP; ++pp) {
      const u32 pair = pair0 + pp;
      if (pair >= C::NPAIRS) break;  // uniform over the workgroup (pair0 is the task's)
      cpx<float>* col = a.scr + (u64)pair * (u64)(C::H * C::W) + kx * C::H;
      cpx<float> v[SH::NV];
      for_layout<WC, F0::R>(t, [&](auto slot, u32 n) { v[decltype(slot)::value] = col[n]; });
      tsf.template run<-1>(v, xch, t, 0);
#pragma unroll
      for (int i = 0; i < SH::NV; ++i) v[i] = cmul(v[i], kv[i]);
      block_sync<WC>();
      if constexpr (SAME) tsf.template run<1>(v, xch, t, 0);
      else tsi_own.template run<1>(v, xch, t, 0);
      for_layout<WC, IL::R>(t, [&](auto slot, u32 n) { col[n] = v[decltype(slot)::value]; });
      block_sync<WC>();
    }
  }
}

template <class C>
__device__ __forceinline__ void conv_rows_inv_cm(const ConvArgs<C>& a, cpx<float>* lds, u32 pair, u32 slice,
                                                 int tid) {
  using SH = ConvShape<C>;
  using WC = ConvWC<C>;
  using LI = typename C::inv;
  using I0 = typename pl_first<LI>::type;
  using IL = typename pl_last<LI>::type;
  constexpr int IT = C::W * C::RA / SH::NT;
  const int u = conv_unit<C>(tid), t = conv_lane<C>(tid);
  const u32 ia = 2 * pair, ib = 2 * pair + 1;
  const bool has_b = ib < C::NIMG;
  const u32 y0 = slice * C::RA;
  const cpx<float>* S = a.scr + (u64)pair * (u64)(C::H * C::W);
  TwSet<WC, LI, conv_twreg<WC, LI>()> ts;
  ts.load(t, a.twi);  // in flight with the tile loads; reused by every row
  {
    cpx<float> tl[IT];  // all loads in flight before the first LDS write
#pragma unroll
    for (int i = 0; i < IT; ++i) {
      const u32 e = (u32)(tid + i * SH::NT), yl = e % C::RA, kx = e / C::RA;
      tl[i] = S[kx * C::H + y0 + yl];
    }
#pragma unroll
    for (int i = 0; i < IT; ++i) {
      const u32 e = (u32)(tid + i * SH::NT), yl = e % C::RA, kx = e / C::RA;
      lds[kx * C::TP + yl] = tl[i];
    }
  }
  __syncthreads();
  cpx<float> v[SH::RPW][SH::NV];
#pragma unroll
  for (int i = 0; i < SH::RPW; ++i) {
    const u32 yl = (u32)(u + SH::NU * i);
    for_layout<WC, I0::R>(t, [&](auto slot, u32 kx) { v[i][decltype(slot)::value] = lds[kx * C::TP + yl]; });
  }
  __syncthreads();  // the exchange regions overlap the transpose tile
#pragma unroll
  for (int i = 0; i < SH::RPW; ++i) {
    const u32 y = y0 + (u32)(u + SH::NU * i);
    ts.template run<1>(v[i], lds + u * C::NP, t, 0);
    float* ra = a.x + (u64)ia * C::IMG + y * C::PITCH;
    float* rb = a.x + (u64)ib * C::IMG + y * C::PITCH;
    for_layout<WC, IL::R>(t, [&](auto slot, u32 n) {
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
  constexpr u32 NBT = CM ? (u32)(C::W / (SH::NU * C::CB)) : SH::NB;  // column tasks per pair (group)
  constexpr u32 NPG = CM ? (C::NPAIRS + C::PP - 1) / C::PP : C::NPAIRS;  // pair groups
  __shared__ cpx<float> lds[LDS_ELEMS];
  const ConvArgs<C> a = conv_args<C>(prm);
  const int tid = (int)threadIdx.x;
  const int lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const u32 wg = blockIdx.x;
  if constexpr (PHASE == 0) {
    if (wg < C::NPAIRS * SH::NA) {
      if constexpr (CM) conv_rows_fwd_cm<C>(a, lds, wg / SH::NA, wg % SH::NA, tid);
      else conv_rows_fwd<C>(a, lds, wg / SH::NA, wg % SH::NA, wave, lane);
    } else if (wg < C::NPAIRS * SH::NA + SH::NK) {
      conv_kprep<C>(a.K, a.kt, wg - C::NPAIRS * SH::NA, tid);
    }
  } else if constexpr (PHASE == 1) {
    // Task t = (column slice t / NPG, pair group t % NPG).  Workgroups are dealt to the
    // 8 XCDs round-robin; on a grid of exactly the 8k tasks, the workgroups of one XCD
    // take consecutive tasks, so each slice of Kt is read by one XCD (from its L2 for
    // the slice's other pair groups) instead of by all eight.
    constexpr u32 T = NPG * NBT;
    const u32 t = (T % 8u == 0u && gridDim.x == T) ? (wg & 7u) * (T / 8u) + (wg >> 3) : wg;
    if (t < T) {
      if constexpr (CM) conv_cols_cm<C>(a, a.kt, lds, (t % NPG) * C::PP, t / NPG, tid);
      else conv_cols<C>(a, a.kt, lds, t % C::NPAIRS, t / C::NPAIRS, tid, wave, lane);
    }
  } else {
    if (wg < C::NPAIRS * SH::NA) {
      if constexpr (CM) conv_rows_inv_cm<C>(a, lds, wg / SH::NA, wg % SH::NA, tid);
      else conv_rows_inv<C>(a, lds, wg / SH::NA, wg % SH::NA, wave, lane);
    }
  }
}

}  // namespace vkx
