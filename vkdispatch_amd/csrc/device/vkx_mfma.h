// vkx_mfma.h -- 256-point FFT passes whose two radix-16 stages are complex
// GEMMs on the gfx950 matrix cores (v_mfma_f32_16x16x32_bf16).
//
// A length-256 transform is viewed as X[n1][n2] = x[n1 + 16 n2]:
//   stage 1  Y = X . F16            (DFT over n2, one 16x16x16 complex GEMM)
//   twiddle  Y[n1][k2] *= W256^(n1 k2)
//   stage 2  Z = F16 . Y            (DFT over n1)
//   output   X[k2 + 16 k1] = Z[k1][k2]
// Each complex GEMM is one real 16x32 by 32x16 product per output part, using
// the 2x2 real embedding ([Xr | Xi] . [[Fr, Fi], [-Fi, Fr]]).  Operands are
// split into bf16 hi + lo parts and three products (hi.hi + lo.hi + hi.lo) are
// accumulated in FP32, so the arithmetic keeps ~16 significant bits.  F16 is
// symmetric, so the same per-lane constant operands serve both stages.
//
// MFMA 16x16x32 lane layout (checked on the device): A[r][k] is element k%8 of
// lane r + 16*(k/8); B[k][c] is element k%8 of lane c + 16*(k/8); the result
// D[r][c] is element r%4 of lane c + 16*(r/4).
//
#pragma once
#include "vkx_fft.h"

namespace vkx {

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float f32x4_t __attribute__((ext_vector_type(4)));

__device__ __forceinline__ f32x4_t mma_bf16(bf16x8_t a, bf16x8_t b, f32x4_t c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ void split_bf16(float x, __bf16& hi, __bf16& lo) {
  hi = (__bf16)x;
  lo = (__bf16)(x - (float)hi);
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
//            g >= 2); bf16 data is its own hi part (lo = 0: two products)
//   twiddle: per-lane constants W256^(n1 k2)
//   stage 2: B operand from four per-wave bf16 planes (re/im x hi/lo, [k2][n1],
//            one 16-byte LDS read each)
//   output:  Z[k1][k2] -> staging tile [k][t] (row stride T + 1), then the
//            workgroup writes each frequency's T consecutive transforms (128 B)
// ---------------------------------------------------------------------------
typedef unsigned short u16x8_t __attribute__((ext_vector_type(8)));
typedef unsigned short u16x4_t __attribute__((ext_vector_type(4)));

__device__ __forceinline__ unsigned short bf16_bits(__bf16 v) { return __builtin_bit_cast(unsigned short, v); }

template <class S> struct PartLoad;
template <> struct PartLoad<bf16> {
  static constexpr bool EXACT = true;  // the stored value is its own bf16 hi part
  __device__ __forceinline__ static unsigned short bits(cpx<bf16> v, int part) { return part ? v.y.b : v.x.b; }
};
template <> struct PartLoad<_Float16> {
  static constexpr bool EXACT = false;
  __device__ __forceinline__ static float value(cpx<_Float16> v, int part) { return (float)(part ? v.y : v.x); }
};

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
  __shared__ __attribute__((aligned(16))) unsigned short yp[WAVES][4][N];  // [re hi, re lo, im hi, im lo][k2*16+n1]

  const int tid = (int)threadIdx.x;
  const int lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int r = lane & 15, g = lane >> 4, h = g & 1, part = g >> 1;
  const cpx<SIn>* in = bufp<const cpx<SIn>>(prm, 1);
  cpx<SOut>* out = bufp<cpx<SOut>>(prm, 0);
  const cpx<float>* roots = bufp<const cpx<float>>(prm, SLOT_ROOTS);  // exp(-2 pi i k / 256)
  const u32 t0 = blockIdx.x * (u32)T;

  // per-lane constants: DFT-16 operands (hi/lo, symmetric) and stage twiddles
  bf16x8_t fre_hi, fre_lo, fim_hi, fim_lo;
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    const int n = 8 * h + i;
    cpx<float> w = roots[(16 * ((n * r) & 15)) & 255];  // W16^(n r)
    if constexpr (C::DIR > 0) w.y = -w.y;
    const float bre = part == 0 ? w.x : -w.y;  // [Fr ; -Fi]
    const float bim = part == 0 ? w.y : w.x;   // [Fi ;  Fr]
    __bf16 hi, lo;
    split_bf16(bre, hi, lo);
    fre_hi[i] = hi;
    fre_lo[i] = lo;
    split_bf16(bim, hi, lo);
    fim_hi[i] = hi;
    fim_lo[i] = lo;
  }
  cpx<float> tw[4];
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    cpx<float> w = roots[((4 * g + e) * r) & 255];  // W256^(n1 k2), n1 = 4g + e, k2 = r
    if constexpr (C::DIR > 0) w.y = -w.y;
    tw[e] = w;
  }
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
  unsigned short* ypw = &yp[wave][0][0];
#pragma unroll
  for (int j = 0; j < TPW; ++j) {
    const cpx<SIn>* cur = ld[j];
    // ---- stage 1: Y[n1][k2] = sum_n2 X[n1][n2] F[n2][k2] ----
    f32x4_t yr = {0.f, 0.f, 0.f, 0.f}, yi = {0.f, 0.f, 0.f, 0.f};
    if constexpr (PartLoad<SIn>::EXACT) {
      u16x8_t bits;
#pragma unroll
      for (int i = 0; i < 8; ++i) bits[i] = PartLoad<SIn>::bits(cur[i], part);
      const bf16x8_t a = __builtin_bit_cast(bf16x8_t, bits);
      yr = mma_bf16(a, fre_hi, yr);
      yr = mma_bf16(a, fre_lo, yr);
      yi = mma_bf16(a, fim_hi, yi);
      yi = mma_bf16(a, fim_lo, yi);
    } else {
      bf16x8_t ahi, alo;
#pragma unroll
      for (int i = 0; i < 8; ++i) {
        __bf16 hi, lo;
        split_bf16(PartLoad<SIn>::value(cur[i], part), hi, lo);
        ahi[i] = hi;
        alo[i] = lo;
      }
      yr = mma_bf16(ahi, fre_hi, yr);
      yr = mma_bf16(alo, fre_hi, yr);
      yr = mma_bf16(ahi, fre_lo, yr);
      yi = mma_bf16(ahi, fim_hi, yi);
      yi = mma_bf16(alo, fim_hi, yi);
      yi = mma_bf16(ahi, fim_lo, yi);
    }
    // ---- twiddle, split, into the per-wave planes [k2][n1] (lane holds n1 = 4g..4g+3, k2 = r) ----
    {
      u16x4_t rh, rl, ih, il;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const cpx<float> y = cmul(cpx<float>{yr[e], yi[e]}, tw[e]);
        __bf16 hi, lo;
        split_bf16(y.x, hi, lo);
        rh[e] = bf16_bits(hi);
        rl[e] = bf16_bits(lo);
        split_bf16(y.y, hi, lo);
        ih[e] = bf16_bits(hi);
        il[e] = bf16_bits(lo);
      }
      const int o = r * 16 + 4 * g;
      *(u16x4_t*)(ypw + 0 * N + o) = rh;
      *(u16x4_t*)(ypw + 1 * N + o) = rl;
      *(u16x4_t*)(ypw + 2 * N + o) = ih;
      *(u16x4_t*)(ypw + 3 * N + o) = il;
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    // ---- stage 2: Z[k1][k2] = sum_n1 F[k1][n1] Y[n1][k2]; lane reads Y[8h..8h+7][r] ----
    {
      const int o = r * 16 + 8 * h;
      const bf16x8_t bhi = __builtin_bit_cast(bf16x8_t, *(const u16x8_t*)(ypw + (2 * part) * N + o));
      const bf16x8_t blo = __builtin_bit_cast(bf16x8_t, *(const u16x8_t*)(ypw + (2 * part + 1) * N + o));
      f32x4_t zr = {0.f, 0.f, 0.f, 0.f}, zi = {0.f, 0.f, 0.f, 0.f};
      zr = mma_bf16(fre_hi, bhi, zr);
      zr = mma_bf16(fre_lo, bhi, zr);
      zr = mma_bf16(fre_hi, blo, zr);
      zi = mma_bf16(fim_hi, bhi, zi);
      zi = mma_bf16(fim_lo, bhi, zi);
      zi = mma_bf16(fim_hi, blo, zi);
      const int tl = wave * TPW + j;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int k = r + 16 * (4 * g + e);  // X[k2 + 16 k1], k1 = 4g + e
        stage[k * TS + tl] = ccvt<SOut>(cpx<float>{zr[e] * scale, zi[e] * scale});
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
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

}  // namespace vkx
