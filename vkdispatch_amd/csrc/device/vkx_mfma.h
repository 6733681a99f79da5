// vkx_mfma.h -- 256-point FFT stages as complex GEMMs on the gfx950 matrix cores
// (v_mfma_f32_16x16x32_bf16 / _f16), for 16-bit (bf16 or fp16) storage.
//
// A length-256 transform is viewed as X[n1][n2] = x[n1 + 16 n2]:
//   stage 1  Y = X . F16            (DFT over n2, one 16x16x16 complex GEMM)
//   twiddle  Y[n1][k2] *= W256^(n1 k2)
//   stage 2  Z = F16 . Y            (DFT over n1)
//   output   X[k2 + 16 k1] = Z[k1][k2]
// Each complex GEMM is one real 16x32 by 32x16 product per output part, using the
// 2x2 real embedding ([Xr | Xi] . [[Fr, Fi], [-Fi, Fr]]).  The DFT matrix is held as
// hi + lo parts; FP32 data operands are split the same way and three products
// (hi.hi + lo.hi + hi.lo) accumulate in FP32 (~16 significant bits); data read from
// 16-bit storage is exact on the MFMA of its own type and needs two products.  F16
// is symmetric, so the same per-lane constant operands serve both stages.
//
// MFMA 16x16x32 lane layout (checked on the device): A[r][k] is element k%8 of
// lane r + 16*(k/8); B[k][c] is element k%8 of lane c + 16*(k/8); the result
// D[r][c] is element r%4 of lane c + 16*(r/4).
//
// Contents: the per-wave tile machinery (tile256x, tiles16, cross16_*), the
// transposed-write 256-point pass (fft256_mfma_tw_body, three passes per 256^3
// volume) and the two-pass 256^3 transform (fft3_p1_body, fft3_p2_body).
#pragma once
#include "vkx_fft.h"

namespace vkx {

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float f32x4_t __attribute__((ext_vector_type(4)));

__device__ __forceinline__ f32x4_t mma_bf16(bf16x8_t a, bf16x8_t b, f32x4_t c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

typedef _Float16 f16x8_t __attribute__((ext_vector_type(8)));
__device__ __forceinline__ f32x4_t mma_f16(f16x8_t a, f16x8_t b, f32x4_t c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ void split_bf16(float x, __bf16& hi, __bf16& lo) {
  hi = (__bf16)x;
  lo = (__bf16)(x - (float)hi);
}

typedef unsigned short u16x8_t __attribute__((ext_vector_type(8)));
typedef float f32x2_t __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2_t __attribute__((ext_vector_type(2)));

// Split two FP32 values into packed bf16 pairs with x ~= hi + lo: hi is x truncated to its
// upper 16 bits (one byte permute for both), lo = x - hi is exact in FP32 and rounded to
// bf16 (|lo| < 2^-7 |x|, so the pair keeps ~16 significant bits).  Element 0 sits in the
// low half of each word, as in the bf16x8 operand vectors.
__device__ __forceinline__ void split2_bf16(float x0, float x1, unsigned& hi, unsigned& lo) {
  const unsigned u0 = __builtin_bit_cast(unsigned, x0), u1 = __builtin_bit_cast(unsigned, x1);
  hi = __builtin_amdgcn_perm(u1, u0, 0x07060302u);
  const f32x2_t l = {x0 - __builtin_bit_cast(float, u0 & 0xffff0000u), x1 - __builtin_bit_cast(float, u1 & 0xffff0000u)};
  lo = __builtin_bit_cast(unsigned, __builtin_convertvector(l, bf16x2_t));
}

// Stage-1 operands straight from 16-bit storage: eight complex values -> the real (part 0)
// or imaginary (part 1) halves packed in order, one byte permute per two values.  The
// stored value is exact on the matching MFMA (bf16 data on the bf16 one, fp16 data on the
// fp16 one), so stage 1 needs only the constant's hi and lo products.
template <class S>
__device__ __forceinline__ void pack_part(const cpx<S> (&cur)[8], int part, unsigned (&w)[4]) {
  static_assert(sizeof(cpx<S>) == 4, "16-bit complex");
  const unsigned sel = part ? 0x07060302u : 0x05040100u;
#pragma unroll
  for (int i = 0; i < 4; ++i)
    w[i] = __builtin_amdgcn_perm(__builtin_bit_cast(unsigned, cur[2 * i + 1]), __builtin_bit_cast(unsigned, cur[2 * i]), sel);
}

// Per-lane constant operands: the DFT-16 matrix as hi/lo parts in the 2x2 real embedding
// (B operand when the data is A, A operand when the data is B), in bf16 and in fp16, and
// the 256-point twiddles W256^((4g + e) r) of the lane's stage-1 outputs.  ``colmul``
// scales output column r (lane r holds column r of the B operand), which folds a per-
// output twiddle or scale into the GEMM; the matrix is then no longer symmetric and
// serves only as B.
struct Dft16Ops {
  bf16x8_t fre_hi, fre_lo, fim_hi, fim_lo;
  f16x8_t hre_hi, hre_lo, him_hi, him_lo;
  cpx<float> tw[4];
};

template <int DIR>
__device__ __forceinline__ Dft16Ops dft16_ops(const cpx<float>* roots, int r, int g,
                                              cpx<float> colmul = cpx<float>{1.f, 0.f}) {
  const int h = g & 1, part = g >> 1;
  Dft16Ops o;
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    const int n = 8 * h + i;
    cpx<float> w = roots[(16 * ((n * r) & 15)) & 255];  // W16^(n r)
    if constexpr (DIR > 0) w.y = -w.y;
    w = cmul(w, colmul);
    const float bre = part == 0 ? w.x : -w.y;  // [Fr ; -Fi]
    const float bim = part == 0 ? w.y : w.x;   // [Fi ;  Fr]
    __bf16 hi, lo;
    split_bf16(bre, hi, lo);
    o.fre_hi[i] = hi;
    o.fre_lo[i] = lo;
    split_bf16(bim, hi, lo);
    o.fim_hi[i] = hi;
    o.fim_lo[i] = lo;
    const _Float16 rh = (_Float16)bre, ih = (_Float16)bim;
    o.hre_hi[i] = rh;
    o.hre_lo[i] = (_Float16)(bre - (float)rh);
    o.him_hi[i] = ih;
    o.him_lo[i] = (_Float16)(bim - (float)ih);
  }
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    cpx<float> w = roots[((4 * g + e) * r) & 255];  // W256^(n1 k2), n1 = 4g + e, k2 = r
    if constexpr (DIR > 0) w.y = -w.y;
    o.tw[e] = w;
  }
  return o;
}

__device__ __forceinline__ void wave_sync_lds() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// NT independent 256-point transforms per wave (their MFMA chains interleave): lane
// (r, g) holds cur[t][i] = x_t[r + 16 (8h + i)], i = 0..7, and receives
// z_t[e] = X_t[r + 16 (4g + e)] (FP32, unscaled).  ypw: NT x 2 KB of per-wave planes.
//   stage 1: A operand = the data (real part for g < 2, imaginary for g >= 2), exact on
//            the MFMA of its storage type, times the hi and lo parts of the DFT-16 matrix
//   twiddle: W256^(n1 k2)
//   stage 2: B operand from four bf16 planes (re/im x hi/lo), A operand = the DFT-16
//            matrix, three products.  Plane layout [n1 / 4][k2][n1 % 4]: the 16 lanes of
//            one g write 128 contiguous bytes, a B operand is two 8-byte reads.
template <int DIR, class SIn, int NT>
__device__ __forceinline__ void tile256x(const Dft16Ops& o, const cpx<SIn> (*cur)[8], unsigned short* ypw, int r,
                                         int g, f32x4_t (&zr)[NT], f32x4_t (&zi)[NT]) {
  constexpr int N = 256;
  const int h = g & 1, part = g >> 1;
  f32x4_t yr[NT], yi[NT];
#pragma unroll
  for (int t = 0; t < NT; ++t) {
    yr[t] = f32x4_t{0.f, 0.f, 0.f, 0.f};
    yi[t] = f32x4_t{0.f, 0.f, 0.f, 0.f};
    unsigned w[4];
    pack_part(cur[t], part, w);
    if constexpr (__is_same(SIn, bf16)) {
      const bf16x8_t a = __builtin_bit_cast(bf16x8_t, w);
      yr[t] = mma_bf16(a, o.fre_hi, yr[t]);
      yr[t] = mma_bf16(a, o.fre_lo, yr[t]);
      yi[t] = mma_bf16(a, o.fim_hi, yi[t]);
      yi[t] = mma_bf16(a, o.fim_lo, yi[t]);
    } else {
      static_assert(__is_same(SIn, _Float16), "16-bit storage");
      const f16x8_t a = __builtin_bit_cast(f16x8_t, w);
      yr[t] = mma_f16(a, o.hre_hi, yr[t]);
      yr[t] = mma_f16(a, o.hre_lo, yr[t]);
      yi[t] = mma_f16(a, o.him_hi, yi[t]);
      yi[t] = mma_f16(a, o.him_lo, yi[t]);
    }
  }
  typedef unsigned u32x2_t __attribute__((ext_vector_type(2)));
#pragma unroll
  for (int t = 0; t < NT; ++t) {
    cpx<float> y[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) y[e] = cmul(cpx<float>{yr[t][e], yi[t][e]}, o.tw[e]);
    unsigned rh[2], rl[2], ih[2], il[2];
#pragma unroll
    for (int e = 0; e < 2; ++e) {
      split2_bf16(y[2 * e].x, y[2 * e + 1].x, rh[e], rl[e]);
      split2_bf16(y[2 * e].y, y[2 * e + 1].y, ih[e], il[e]);
    }
    unsigned short* pl = ypw + t * 4 * N + g * 64 + r * 4;
    *(u32x2_t*)(pl + 0 * N) = u32x2_t{rh[0], rh[1]};
    *(u32x2_t*)(pl + 1 * N) = u32x2_t{rl[0], rl[1]};
    *(u32x2_t*)(pl + 2 * N) = u32x2_t{ih[0], ih[1]};
    *(u32x2_t*)(pl + 3 * N) = u32x2_t{il[0], il[1]};
  }
  wave_sync_lds();
#pragma unroll
  for (int t = 0; t < NT; ++t) {
    const unsigned short* pl = ypw + t * 4 * N + 2 * h * 64 + r * 4;
    const u32x2_t h0 = *(const u32x2_t*)(pl + (2 * part) * N), h1 = *(const u32x2_t*)(pl + (2 * part) * N + 64);
    const u32x2_t l0 = *(const u32x2_t*)(pl + (2 * part + 1) * N), l1 = *(const u32x2_t*)(pl + (2 * part + 1) * N + 64);
    typedef unsigned u32x4_t __attribute__((ext_vector_type(4)));
    const bf16x8_t bhi = __builtin_bit_cast(bf16x8_t, u32x4_t{h0[0], h0[1], h1[0], h1[1]});
    const bf16x8_t blo = __builtin_bit_cast(bf16x8_t, u32x4_t{l0[0], l0[1], l1[0], l1[1]});
    f32x4_t a = {0.f, 0.f, 0.f, 0.f}, b = {0.f, 0.f, 0.f, 0.f};
    a = mma_bf16(o.fre_hi, bhi, a);
    a = mma_bf16(o.fre_lo, bhi, a);
    a = mma_bf16(o.fre_hi, blo, a);
    b = mma_bf16(o.fim_hi, bhi, b);
    b = mma_bf16(o.fim_lo, bhi, b);
    b = mma_bf16(o.fim_hi, blo, b);
    zr[t] = a;
    zi[t] = b;
  }
  wave_sync_lds();
}

// The same stages split for software pipelining over 16 tiles (NT per group, planes
// double-buffered): the stage-1 GEMMs of group j + 1 and the stage-2 GEMMs of group j
// share one LDS interval, so their MFMA chains and the twiddle/split VALU work overlap.
template <class SIn, int NT>
__device__ __forceinline__ void tile_stage1(const Dft16Ops& o, const cpx<SIn> (&cur)[NT][8], int part, f32x4_t (&yr)[NT],
                                            f32x4_t (&yi)[NT]) {
#pragma unroll
  for (int t = 0; t < NT; ++t) {
    yr[t] = f32x4_t{0.f, 0.f, 0.f, 0.f};
    yi[t] = f32x4_t{0.f, 0.f, 0.f, 0.f};
    unsigned w[4];
    pack_part(cur[t], part, w);
    if constexpr (__is_same(SIn, bf16)) {
      const bf16x8_t a = __builtin_bit_cast(bf16x8_t, w);
      yr[t] = mma_bf16(a, o.fre_hi, yr[t]);
      yr[t] = mma_bf16(a, o.fre_lo, yr[t]);
      yi[t] = mma_bf16(a, o.fim_hi, yi[t]);
      yi[t] = mma_bf16(a, o.fim_lo, yi[t]);
    } else {
      static_assert(__is_same(SIn, _Float16), "16-bit storage");
      const f16x8_t a = __builtin_bit_cast(f16x8_t, w);
      yr[t] = mma_f16(a, o.hre_hi, yr[t]);
      yr[t] = mma_f16(a, o.hre_lo, yr[t]);
      yi[t] = mma_f16(a, o.him_hi, yi[t]);
      yi[t] = mma_f16(a, o.him_lo, yi[t]);
    }
  }
}

template <int NT>
__device__ __forceinline__ void tile_twiddle_write(const Dft16Ops& o, const f32x4_t (&yr)[NT], const f32x4_t (&yi)[NT],
                                                   unsigned short* ypw, int r, int g) {
  constexpr int N = 256;
  typedef unsigned u32x2_t __attribute__((ext_vector_type(2)));
#pragma unroll
  for (int t = 0; t < NT; ++t) {
    cpx<float> y[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) y[e] = cmul(cpx<float>{yr[t][e], yi[t][e]}, o.tw[e]);
    unsigned rh[2], rl[2], ih[2], il[2];
#pragma unroll
    for (int e = 0; e < 2; ++e) {
      split2_bf16(y[2 * e].x, y[2 * e + 1].x, rh[e], rl[e]);
      split2_bf16(y[2 * e].y, y[2 * e + 1].y, ih[e], il[e]);
    }
    unsigned short* pl = ypw + t * 4 * N + g * 64 + r * 4;
    *(u32x2_t*)(pl + 0 * N) = u32x2_t{rh[0], rh[1]};
    *(u32x2_t*)(pl + 1 * N) = u32x2_t{rl[0], rl[1]};
    *(u32x2_t*)(pl + 2 * N) = u32x2_t{ih[0], ih[1]};
    *(u32x2_t*)(pl + 3 * N) = u32x2_t{il[0], il[1]};
  }
}

template <int NT>
__device__ __forceinline__ void tile_stage2(const Dft16Ops& o, const unsigned short* ypw, int r, int g, f32x4_t (&zr)[NT],
                                            f32x4_t (&zi)[NT]) {
  constexpr int N = 256;
  const int h = g & 1, part = g >> 1;
  typedef unsigned u32x2_t __attribute__((ext_vector_type(2)));
  typedef unsigned u32x4_t __attribute__((ext_vector_type(4)));
#pragma unroll
  for (int t = 0; t < NT; ++t) {
    const unsigned short* pl = ypw + t * 4 * N + 2 * h * 64 + r * 4;
    const u32x2_t h0 = *(const u32x2_t*)(pl + (2 * part) * N), h1 = *(const u32x2_t*)(pl + (2 * part) * N + 64);
    const u32x2_t l0 = *(const u32x2_t*)(pl + (2 * part + 1) * N), l1 = *(const u32x2_t*)(pl + (2 * part + 1) * N + 64);
    const bf16x8_t bhi = __builtin_bit_cast(bf16x8_t, u32x4_t{h0[0], h0[1], h1[0], h1[1]});
    const bf16x8_t blo = __builtin_bit_cast(bf16x8_t, u32x4_t{l0[0], l0[1], l1[0], l1[1]});
    f32x4_t a = {0.f, 0.f, 0.f, 0.f}, b = {0.f, 0.f, 0.f, 0.f};
    a = mma_bf16(o.fre_hi, bhi, a);
    a = mma_bf16(o.fre_lo, bhi, a);
    a = mma_bf16(o.fre_hi, blo, a);
    b = mma_bf16(o.fim_hi, bhi, b);
    b = mma_bf16(o.fim_lo, bhi, b);
    b = mma_bf16(o.fim_hi, blo, b);
    zr[t] = a;
    zi[t] = b;
  }
}

// 16 tiles of a wave, pipelined; get(j, cur) fills the stage-1 inputs of tile j and
// put(j, zr, zi) takes its result, X_j[r + 16 (4g + e)] = (zr[e], zi[e]).
// ypw: 2 x NT x 2 KB.
template <int NT, class SIn, class Get, class Put>
__device__ __forceinline__ void tiles16(const Dft16Ops& o, Get&& get, Put&& put, unsigned short* ypw, int r, int g) {
  constexpr int G = 16 / NT, BUF = NT * 4 * 256;
  static_assert(16 % NT == 0, "tile groups");
  const int part = g >> 1;
  {
    cpx<SIn> cur[NT][8];
#pragma unroll
    for (int t = 0; t < NT; ++t) get(t, cur[t]);
    f32x4_t yr[NT], yi[NT];
    tile_stage1<SIn, NT>(o, cur, part, yr, yi);
    tile_twiddle_write<NT>(o, yr, yi, ypw, r, g);
  }
  wave_sync_lds();
#pragma unroll
  for (int j = 0; j < G; ++j) {
    f32x4_t yr[NT], yi[NT];
    if (j + 1 < G) {
      cpx<SIn> cur[NT][8];
#pragma unroll
      for (int t = 0; t < NT; ++t) get((j + 1) * NT + t, cur[t]);
      tile_stage1<SIn, NT>(o, cur, part, yr, yi);
    }
    f32x4_t zr[NT], zi[NT];
    tile_stage2<NT>(o, ypw + (j & 1) * BUF, r, g, zr, zi);
#pragma unroll
    for (int t = 0; t < NT; ++t) put(j * NT + t, zr[t], zi[t]);
    if (j + 1 < G) tile_twiddle_write<NT>(o, yr, yi, ypw + ((j + 1) & 1) * BUF, r, g);
    wave_sync_lds();
  }
}

// Tile results kept between the tile stages and the cross stage: FP32 (v32, 128 VGPRs for
// 16 tiles) or fp16 pairs (v16, 64 VGPRs; one 11-bit rounding, the values pre-scaled by
// an exact power of two so they stay in fp16 range).
__device__ __forceinline__ unsigned pack_f16(float re, float im) {
  const _Float16 a = (_Float16)re, b = (_Float16)im;
  return (unsigned)__builtin_bit_cast(unsigned short, a) | ((unsigned)__builtin_bit_cast(unsigned short, b) << 16);
}

// cross stage on fp16 tile results: planes [g0][part][j / 8][r'][j % 8] (4 KB per wave),
// the stored values are exact fp16 operands, so each output part is two products
__device__ __forceinline__ void cross16_write16(const unsigned (&v)[16][4], int e0, unsigned short* xw, int r, int g) {
  typedef unsigned u32x4_t __attribute__((ext_vector_type(4)));
#pragma unroll
  for (int pp = 0; pp < 2; ++pp) {
    const unsigned sel = pp ? 0x07060302u : 0x05040100u;
    u32x4_t w0, w1;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      w0[k] = __builtin_amdgcn_perm(v[2 * k + 1][e0], v[2 * k][e0], sel);
      w1[k] = __builtin_amdgcn_perm(v[2 * k + 9][e0], v[2 * k + 8][e0], sel);
    }
    unsigned short* d = xw + ((g * 2 + pp) * 2) * 128 + r * 8;
    *(u32x4_t*)(d) = w0;
    *(u32x4_t*)(d + 128) = w1;
  }
}

__device__ __forceinline__ void cross16_gemm16(const Dft16Ops& o, const unsigned short* xw, int g0, int r, int g,
                                               f32x4_t& wr, f32x4_t& wi) {
  const int h = g & 1, part = g >> 1;
  const unsigned short* sp = xw + ((g0 * 2 + part) * 2 + h) * 128 + r * 8;
  const f16x8_t a = __builtin_bit_cast(f16x8_t, *(const u16x8_t*)(sp));
  f32x4_t x = {0.f, 0.f, 0.f, 0.f}, y = {0.f, 0.f, 0.f, 0.f};
  x = mma_f16(a, o.hre_hi, x);
  x = mma_f16(a, o.hre_lo, x);
  y = mma_f16(a, o.him_hi, y);
  y = mma_f16(a, o.him_lo, y);
  wr = x;
  wi = y;
}

// ---------------------------------------------------------------------------
// Transposed-write pass: 256-point transforms along the contiguous last axis of
// a [NB][256] volume, result written frequency-major ([256][NB]).  Three such
// passes over a 256^3 volume return it to natural order ([i0][i1][i2] ->
// [k2][i0][i1] -> [k1][k2][i0] -> [k0][k1][k2]), so every pass reads each
// transform as 1 KB of contiguous 16-bit data.
//
// Per wave, per transform (TPW transforms per wave, the next one's loads issued
// before the current one is computed):
//   stage 1: A operand straight from global memory (lane r + 16g holds
//            x[r + 16 n2], n2 = 8(g&1)..+7, real part for g < 2, imaginary for
//            g >= 2), exact on the MFMA of its storage type (two products)
//   twiddle: per-lane constants W256^(n1 k2)
//   stage 2: B operand from four per-wave bf16 planes (re/im x hi/lo, [k2][n1],
//            one 16-byte LDS read each)
//   output:  Z[k1][k2] -> staging tile [k][t] (row stride T + 1), then the
//            workgroup writes each frequency's T consecutive transforms (128 B)
// ---------------------------------------------------------------------------
template <class C, class SIn, class SOut>
__device__ __forceinline__ void fft256_mfma_tw_body(const KParams& prm) {
  constexpr int N = 256;
  constexpr int T = C::T;
  constexpr int WAVES = C::BLOCK / 64;
  constexpr int TPW = T / WAVES;
  constexpr int TS = T + 2;  // even: pairs of transforms are 8/16-byte aligned in the staging tile
  static_assert(T % WAVES == 0 && T % 2 == 0, "transforms per wave");
  static_assert(N * T % (2 * C::BLOCK) == 0, "store split");
  __shared__ cpx<SOut> stage[N * TS];
  constexpr int NT = TPW % 2 == 0 ? 2 : 1;  // transforms per plane exchange
  __shared__ __attribute__((aligned(16))) unsigned short yp[WAVES][NT][4][N];  // [re hi, re lo, im hi, im lo][k2*16+n1]

  const int tid = (int)threadIdx.x;
  const int lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int r = lane & 15, g = lane >> 4, h = g & 1;
  const cpx<SIn>* in = bufp<const cpx<SIn>>(prm, 1);
  cpx<SOut>* out = bufp<cpx<SOut>>(prm, 0);
  const cpx<float>* roots = bufp<const cpx<float>>(prm, SLOT_ROOTS);  // exp(-2 pi i k / 256)
  const u32 t0 = blockIdx.x * (u32)T;
  const Dft16Ops ops = dft16_ops<C::DIR>(roots, r, g);
  const float scale = (float)C::OUT_SCALE;

  // x[r + 16 n2] for n2 = 8h .. 8h+7 of every transform of this wave, all loads
  // issued up front (they complete in order: transform 0 computes while the
  // others are still in flight)
  cpx<SIn> ld[TPW][8];
#pragma unroll
  for (int j = 0; j < TPW; ++j) {
    const cpx<SIn>* src = in + (u64)(t0 + (u32)(wave * TPW + j)) * N + r + 128 * h;
#pragma unroll
    for (int i = 0; i < 8; ++i) ld[j][i] = src[16 * i];
  }
  unsigned short* ypw = &yp[wave][0][0][0];
#pragma unroll
  for (int j = 0; j < TPW; j += NT) {
    f32x4_t zr[NT], zi[NT];
    tile256x<C::DIR, SIn, NT>(ops, &ld[j], ypw, r, g, zr, zi);
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      const int tl = wave * TPW + j + t;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int k = r + 16 * (4 * g + e);  // X[k2 + 16 k1], k1 = 4g + e
        stage[k * TS + tl] = ccvt<SOut>(cpx<float>{zr[t][e] * scale, zi[t][e] * scale});
      }
    }
  }
  __syncthreads();
  // ---- frequency-major store: T consecutive transforms per frequency, two per lane ----
  struct alignas(2 * sizeof(cpx<SOut>)) pair_t { cpx<SOut> a, b; };
#pragma unroll
  for (int i = 0; i < N * T / (2 * C::BLOCK); ++i) {
    const int idx = tid + i * C::BLOCK;
    const int k = idx / (T / 2), tl = 2 * (idx % (T / 2));
    *(pair_t*)(out + (u64)k * C::NB + t0 + (u32)tl) = *(const pair_t*)(stage + k * TS + tl);
  }
}

// ---------------------------------------------------------------------------
// DFT-16 across the 16 tiles a wave has transformed.  v[j][e] of lane (r, g) is
// point j of column r + 16 (4g + e) (the tile256 output layout, tile j).  Chunk e0
// handles the four tiles m = 4 g0 + e0 (g0 = 0..3) whose columns are r' + 16 m,
// r' = 0..15: the owner lane (r', g0) writes its column's 16 points as bf16 hi/lo
// planes [g0][part][hi, lo][j / 8][r'][j % 8] (8 KB per wave; 16 lanes write 256
// contiguous bytes), lane (r', g') reads points 8h'..8h'+7 of part g' >> 1 as its A operand, and the GEMM with the DFT-16 matrix as
// B leaves W[q][r' + 16 m] in lane (q, g), element e for r' = 4g + e:
//   (wr[g0][e], wi[g0][e]) = W[q][4g + e + 16 (4 g0 + e0)]
// ---------------------------------------------------------------------------
__device__ __forceinline__ void cross16_chunk(const Dft16Ops& o, const cpx<float> (&v)[16][4], int e0,
                                              unsigned short* xw, int r, int g, f32x4_t (&wr)[4],
                                              f32x4_t (&wi)[4]) {
  const int h = g & 1, part = g >> 1;
  typedef unsigned u32x4_t __attribute__((ext_vector_type(4)));
#pragma unroll
  for (int pp = 0; pp < 2; ++pp) {
    u32x4_t hi0, hi1, lo0, lo1;
#pragma unroll
    for (int jj = 0; jj < 8; ++jj) {
      const float x0 = pp ? v[2 * jj][e0].y : v[2 * jj][e0].x;
      const float x1 = pp ? v[2 * jj + 1][e0].y : v[2 * jj + 1][e0].x;
      unsigned hw, lw;
      split2_bf16(x0, x1, hw, lw);
      if (jj < 4) {
        hi0[jj] = hw;
        lo0[jj] = lw;
      } else {
        hi1[jj - 4] = hw;
        lo1[jj - 4] = lw;
      }
    }
    unsigned short* d = xw + ((g * 2 + pp) * 2) * 256 + r * 8;
    *(u32x4_t*)(d) = hi0;
    *(u32x4_t*)(d + 128) = hi1;
    *(u32x4_t*)(d + 256) = lo0;
    *(u32x4_t*)(d + 256 + 128) = lo1;
  }
  wave_sync_lds();
#pragma unroll
  for (int g0 = 0; g0 < 4; ++g0) {
    const unsigned short* sp = xw + ((g0 * 2 + part) * 2) * 256 + h * 128 + r * 8;
    const bf16x8_t ahi = __builtin_bit_cast(bf16x8_t, *(const u16x8_t*)(sp));
    const bf16x8_t alo = __builtin_bit_cast(bf16x8_t, *(const u16x8_t*)(sp + 256));
    f32x4_t a = {0.f, 0.f, 0.f, 0.f}, b = {0.f, 0.f, 0.f, 0.f};
    a = mma_bf16(ahi, o.fre_hi, a);
    a = mma_bf16(alo, o.fre_hi, a);
    a = mma_bf16(ahi, o.fre_lo, a);
    b = mma_bf16(ahi, o.fim_hi, b);
    b = mma_bf16(alo, o.fim_hi, b);
    b = mma_bf16(ahi, o.fim_lo, b);
    wr[g0] = a;
    wi[g0] = b;
  }
  wave_sync_lds();
}

// ---------------------------------------------------------------------------
// Two-pass 256^3 transform.  Each axis index is split n = a + 16 b, k = c + 16 d:
//   X[c + 16 d] = sum_a W16^(a d) W256^(a c) sum_b W16^(b c) x[a + 16 b]
// Pass 1 (workgroup = slab i0 and four values of a1, one per wave): the full axis-2
// transform of the 16 rows i1 = a1 + 16 b1 (tile256 per row), then the first axis-1
// stage across those rows (cross16: DFT over b1, twiddle W256^(a1 c1)); the result
// goes out as fp16 pairs scaled by 1/64 to the intermediate
//   M[c1][k2 / P][i0][a1][k2 % P]        (P = 4 or 8)
// through an LDS tile, in runs of 4 a1 x P k2 (16 P bytes).  Pass 2 (workgroup = c1
// and P consecutive k2, one wave each) loads its contiguous 16 P KB block of M into
// LDS, runs the full axis-0 transform of each (a1, k2) column and the second axis-1
// stage across a1 (cross16: DFT over a1), and writes X[k0][c1 + 16 d1][k2] as FP32
// through a staging tile, 8 P contiguous bytes (P consecutive k2) per row.
// ---------------------------------------------------------------------------
template <class C, class = void> struct p1_ntin_of { static constexpr bool value = false; };
template <class C> struct p1_ntin_of<C, decltype((void)C::NTIN)> { static constexpr bool value = C::NTIN != 0; };
template <class C> constexpr bool p1_ntin() { return p1_ntin_of<C>::value; }

template <class C, class SIn>
__device__ __forceinline__ void fft3_p1_body(const KParams& prm) {
  constexpr int N = 256, PF = C::PF, P = C::P;
  static_assert(C::BLOCK == 256, "four waves");
  static_assert(P == 4 || P == 8, "k2 per intermediate run");
  constexpr int NT = C::NT;
  struct alignas(16) quad_t { cpx<_Float16> q[4]; };
  // per-wave double-buffered planes; after the tiles the same bytes hold the output
  // staging of one cross16 chunk: row (g0, c1, g) of four waves' quads, 16 bytes of
  // padding after every four rows (conflict-free writes along c1)
  constexpr int YP = 4 * 2 * NT * 4 * N * 2, OST = 1088 * 16;
  __shared__ __attribute__((aligned(16))) unsigned char ypo[YP > OST ? YP : OST];
  constexpr bool VH = C::VH;  // tile results held as fp16 pairs
  __shared__ __attribute__((aligned(16))) unsigned short xw[4][VH ? 2048 : 4096];
  quad_t* ost = (quad_t*)ypo;
  const int tid = (int)threadIdx.x;
  const int lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int r = lane & 15, g = lane >> 4, h = g & 1;
  const cpx<SIn>* in = bufp<const cpx<SIn>>(prm, 1);
  cpx<_Float16>* mid = bufp<cpx<_Float16>>(prm, 0);
  const cpx<float>* roots = bufp<const cpx<float>>(prm, SLOT_ROOTS);
  // consecutive workgroups of one XCD take the four a1 groups of one slab
  const u32 b = blockIdx.x, xcd = b & 7, slot = b >> 3;
  const u32 ag = slot & 3, i0 = (slot >> 2) * 8 + xcd;
  const int a1 = (int)ag * 4 + wave;
  Dft16Ops ops = dft16_ops<C::DIR>(roots, r, g);
  // fp16 tile results are scaled by 1/16 (in the stage twiddles), undone in the cross GEMM
  constexpr float VS = VH ? 1.f / 16.f : 1.f;
  if constexpr (VH)
#pragma unroll
    for (int e = 0; e < 4; ++e) ops.tw[e] = cpx<float>{ops.tw[e].x * VS, ops.tw[e].y * VS};
  // the cross16 GEMM's output column c1 = r carries the twiddle W256^(a1 c1) and the scale
  cpx<float> t1 = roots[(a1 * r) & 255];
  if constexpr (C::DIR > 0) t1.y = -t1.y;
  const float os = (float)C::OUT_SCALE / VS;
  const Dft16Ops ops1 = dft16_ops<C::DIR>(roots, r, g, cpx<float>{t1.x * os, t1.y * os});

  const cpx<SIn>* rows = in + ((u64)i0 * N + (u32)a1) * N + r + 128 * h;  // row b1 at + 16 N b1
  // C::NTIN (optional): the input is read once per transform; non-temporal loads keep it
  // from displacing the intermediate and the output in the caches
  auto ldin = [&](u64 off) -> cpx<SIn> {
    if constexpr (p1_ntin<C>()) {
      const unsigned w = __builtin_nontemporal_load(reinterpret_cast<const unsigned*>(rows + off));
      return __builtin_bit_cast(cpx<SIn>, w);
    } else {
      return rows[off];
    }
  };
  cpx<SIn> ld[16][8];
#pragma unroll
  for (int j = 0; j < PF; ++j)
#pragma unroll
    for (int i = 0; i < 8; ++i) ld[j][i] = ldin((u64)(16 * j) * N + 16 * i);
  cpx<float> v[16][4];
  unsigned vh[16][4];
  tiles16<NT, SIn>(ops, [&](int j, cpx<SIn> (&cur)[8]) {
    if (j + PF < 16) {
#pragma unroll
      for (int i = 0; i < 8; ++i) ld[j + PF][i] = ldin((u64)(16 * (j + PF)) * N + 16 * i);
    }
#pragma unroll
    for (int i = 0; i < 8; ++i) cur[i] = ld[j][i];
  }, [&](int j, const f32x4_t& zr, const f32x4_t& zi) {
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      if constexpr (VH)
        vh[j][e] = pack_f16(zr[e], zi[e]);
      else
        v[j][e] = cpx<float>{zr[e], zi[e]};
    }
  }, (unsigned short*)ypo + wave * (2 * NT * 4 * N), r, g);
  __syncthreads();  // the planes become the output staging
  unsigned short* xww = &xw[wave][0];
#pragma unroll
  for (int e0 = 0; e0 < 4; ++e0) {
    f32x4_t wr[4], wi[4];
    if constexpr (VH) {
      cross16_write16(vh, e0, xww, r, g);
      wave_sync_lds();
#pragma unroll
      for (int g0 = 0; g0 < 4; ++g0) cross16_gemm16(ops1, xww, g0, r, g, wr[g0], wi[g0]);
      wave_sync_lds();
    } else {
      cross16_chunk(ops1, v, e0, xww, r, g, wr, wi);
    }
    // lane (c1 = r, g) holds k2 = 16 (4 g0 + e0) + 4g + e; its 16 bytes go next to the
    // other waves' (a1) in the staging tile, one 16 P-byte row per (g0, c1, k2 / P)
#pragma unroll
    for (int g0 = 0; g0 < 4; ++g0) {
      quad_t q;
#pragma unroll
      for (int e = 0; e < 4; ++e) q.q[e] = ccvt<_Float16>(cpx<float>{wr[g0][e], wi[g0][e]});
      if constexpr (P == 4) {
        const int row = (g0 * 16 + r) * 4 + g;
        ost[row * 4 + (row >> 2) + wave] = q;
      } else {
        const int row = (g0 * 16 + r) * 2 + (g >> 1);
        ost[row * 8 + (row >> 1) + 2 * wave + (g & 1)] = q;
      }
    }
    __syncthreads();
#pragma unroll
    for (int it = 0; it < 4; ++it) {  // 1024 quads: runs of 4 a1 x P k2 (16 P bytes)
      const int qi = it * 256 + tid;
      if constexpr (P == 4) {
        const int run = qi >> 2, w = qi & 3;
        const int g0 = run >> 6, c1 = (run >> 2) & 15, gg = run & 3;
        const u32 kb = (u32)(4 * (4 * g0 + e0) + gg);
        *(quad_t*)(mid + ((((u64)c1 * 64 + kb) * N + i0) * 16 + ag * 4 + (u32)w) * 4) = ost[qi + (qi >> 4)];
      } else {
        const int row = qi >> 3, pos = qi & 7;
        const int g0 = row >> 5, c1 = (row >> 1) & 15, gh = row & 1;
        const u32 kb = (u32)(2 * (4 * g0 + e0) + gh);
        *(quad_t*)(mid + ((((u64)c1 * 32 + kb) * N + i0) * 16 + ag * 4 + (u32)(pos >> 1)) * 8 + 4 * (pos & 1)) =
            ost[qi + (qi >> 4)];
      }
    }
    __syncthreads();
  }
}

template <class C>
__device__ __forceinline__ void fft3_p2_body(const KParams& prm) {
  constexpr int N = 256;
  constexpr int NT = C::NT, P = C::P, KB = 256 / P, NTH = 64 * P;
  static_assert(C::BLOCK == NTH && (P == 4 || P == 8), "one wave per k2");
  // block tile: dword (a0, b0, a1, p) at a0 * 256 P + b0 * 16 P + ((P a1 + p) ^ (a0 + 16 (b0 >> 3))),
  // i0 = a0 + 16 b0: the 32 addresses of one stage-1 operand read fall in 32 distinct banks
  __shared__ __attribute__((aligned(16))) unsigned blk[4096 * P];
  __shared__ __attribute__((aligned(16))) unsigned short yp[P][2 * NT * 4 * N];
  const int tid = (int)threadIdx.x;
  const int lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int r = lane & 15, g = lane >> 4, h = g & 1;
  const cpx<_Float16>* mid = bufp<const cpx<_Float16>>(prm, 1);
  cpx<float>* out = bufp<cpx<float>>(prm, 0);
  const cpx<float>* roots = bufp<const cpx<float>>(prm, SLOT_ROOTS);
  // The eight XCDs work on the same c1 at a time, so the output rows of one c1 are
  // written by all of them together: XCD x takes the groups of four consecutive k2
  // blocks q with q mod 8 = x, the four blocks of a group (which complete each other's
  // 128-byte output lines) on consecutive workgroups of that XCD.  (c1 per XCD and
  // consecutive k2 blocks of one c1 in turn: 0.9 us slower per transform.)
  static_assert(KB % 32 == 0, "k2 blocks per c1: groups of four on each of the eight XCDs");
  const u32 b = blockIdx.x, xcd = b & 7, slot = b >> 3;
  constexpr u32 PER = KB / 8;  // blocks of one c1 per XCD
  const u32 c1 = slot / PER, loc = slot % PER;
  const u32 kb = ((loc / 4) * 8 + xcd) * 4 + loc % 4;
  constexpr bool VH = C::VH;  // tile results held as fp16 pairs, scaled by 1/16
  constexpr float VS = VH ? 1.f / 16.f : 1.f;
  Dft16Ops ops = dft16_ops<C::DIR>(roots, r, g);
  if constexpr (VH)
#pragma unroll
    for (int e = 0; e < 4; ++e) ops.tw[e] = cpx<float>{ops.tw[e].x * VS, ops.tw[e].y * VS};
  const Dft16Ops ops1 = dft16_ops<C::DIR>(roots, r, g, cpx<float>{(float)C::OUT_SCALE / VS, 0.f});

  // ---- the block M[c1][kb][i0][a1][0..P) (4096 P fp16 pairs) into LDS ----
  {
    const uint4* src = (const uint4*)(mid + ((u64)c1 * KB + kb) * (4096 * P));
    uint4 q[16];
#pragma unroll
    for (int it = 0; it < 16; ++it) q[it] = src[it * NTH + tid];
#pragma unroll
    for (int it = 0; it < 16; ++it) {
      const int e4 = it * NTH + tid, qq = e4 % (P / 4), a1 = (e4 / (P / 4)) & 15, i0 = e4 / (4 * P);
      const int a0 = i0 & 15, b0 = i0 >> 4, f = a0 + 16 * (b0 >> 3);
      unsigned* d = blk + a0 * (256 * P) + b0 * (16 * P);
      const int o = P * a1 + 4 * qq;
      d[(o + 0) ^ f] = q[it].x;
      d[(o + 1) ^ f] = q[it].y;
      d[(o + 2) ^ f] = q[it].z;
      d[(o + 3) ^ f] = q[it].w;
    }
  }
  __syncthreads();
  // ---- axis 0: wave p = k2 % P, tile j = a1, lane (r, g) reads i0 = r + 16 (8h + i) ----
  const int p = wave;
  cpx<float> v[16][4];
  unsigned vh[16][4];
  const int f = r + 16 * h;
  tiles16<NT, _Float16>(ops, [&](int j, cpx<_Float16> (&cur)[8]) {
#pragma unroll
    for (int i = 0; i < 8; ++i)
      cur[i] = __builtin_bit_cast(cpx<_Float16>, blk[r * (256 * P) + (8 * h + i) * (16 * P) + ((P * j + p) ^ f)]);
  }, [&](int j, const f32x4_t& zr, const f32x4_t& zi) {
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      if constexpr (VH)
        vh[j][e] = pack_f16(zr[e], zi[e]);
      else
        v[j][e] = cpx<float>{zr[e], zi[e]};
    }
  }, &yp[wave][0], r, g);
  __syncthreads();  // blk is reused: cross16 planes (8 KB per wave) + output staging
  unsigned short* xww = (unsigned short*)blk + wave * 4096;
  // output staging [g0][k0 % 16][d1][p] in 8-byte slots, p rotated by d1 / 4 (conflict-free
  // writes along d1); one row (g0, k0 % 16, d1) = the P consecutive k2 of one output run
  float2* stg = (float2*)(blk + 2048 * P);
  const int rot = (r >> 2) * (P / 4);
#pragma unroll
  for (int e0 = 0; e0 < 4; ++e0) {
    f32x4_t wr[4], wi[4];
    if constexpr (VH) {
      cross16_write16(vh, e0, xww, r, g);
      wave_sync_lds();
#pragma unroll
      for (int g0 = 0; g0 < 4; ++g0) cross16_gemm16(ops1, xww, g0, r, g, wr[g0], wi[g0]);
      wave_sync_lds();
    } else {
      cross16_chunk(ops1, v, e0, xww, r, g, wr, wi);
    }
#pragma unroll
    for (int g0 = 0; g0 < 4; ++g0)
#pragma unroll
      for (int e = 0; e < 4; ++e)  // lane (d1 = r, g): k0 = 4g + e + 16 (4 g0 + e0)
        stg[((g0 * 16 + 4 * g + e) * 16 + r) * P + ((p + rot) & (P - 1))] = float2{wr[g0][e], wi[g0][e]};
    __syncthreads();
#pragma unroll
    for (int it = 0; it < 8; ++it) {  // 1024 runs of 8 P bytes, P / 2 lanes each
      const int qi = it * NTH + tid, row = qi / (P / 2), quarter = qi % (P / 2);  // row = (g0, k0 % 16, d1)
      const int g0 = row >> 8, k0 = ((row >> 4) & 15) + 16 * (4 * g0 + e0), d1 = row & 15;
      const float2* sr = stg + row * P;
      const int rr = (d1 >> 2) * (P / 4);
      const float2 v0 = sr[(2 * quarter + rr) & (P - 1)], v1 = sr[(2 * quarter + 1 + rr) & (P - 1)];
      *(float4*)(out + ((u64)k0 * 65536 + (c1 + 16 * (u32)d1) * 256 + P * kb + 2 * quarter)) = float4{v0.x, v0.y, v1.x, v1.y};
    }
    __syncthreads();
  }
}

}  // namespace vkx
