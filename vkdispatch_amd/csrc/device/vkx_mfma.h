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
// One workgroup handles C::T transforms of one batch tile: a coalesced load of
// the tile (16-bit complex, or any storage type) into LDS as FP32, each wave
// transforms C::T / WAVES tiles' rows, and a coalesced store converts to the
// output storage type with a power-of-two scale (keeps fp16 intermediates in
// range without rounding).
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

template <class C, class SIn, class SOut>
__device__ __forceinline__ void fft256_mfma_body(const KParams& prm) {
  constexpr int N = 256;
  constexpr int T = C::T;
  constexpr int LP = N + 8;  // padded LDS row
  constexpr int WAVES = C::BLOCK / 64;
  __shared__ cpx<float> lds[T * LP];
  const int tid = (int)threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const cpx<SIn>* in = bufp<const cpx<SIn>>(prm, 1);
  cpx<SOut>* out = bufp<cpx<SOut>>(prm, 0);
  const cpx<float>* roots = bufp<const cpx<float>>(prm, SLOT_ROOTS);  // exp(-2 pi i k / 256)

  const u32 wg = blockIdx.x;
  u32 outer, first;
  if constexpr (C::S > 1) {
    outer = wg / C::NWG_INNER;
    first = (wg % C::NWG_INNER) * T;
  } else {
    outer = 0;
    first = wg * T;
  }
  const u32 count = C::S > 1 ? C::S : C::OUTER;

  // ---- load tile into LDS (FP32); every load is issued before the first LDS write ----
  constexpr int IT = T * N / C::BLOCK;
  static_assert(T * N % C::BLOCK == 0, "tile split");
  {
    cpx<SIn> raw[IT];
#pragma unroll
    for (int i = 0; i < IT; ++i) {
      const int idx = tid + i * C::BLOCK;
      int c, n;
      if constexpr (C::S > 1) { c = idx % T; n = idx / T; }
      else { c = idx / N; n = idx % N; }
      const u32 t = first + (u32)c;
      const u32 gi = C::S > 1 ? outer * C::IN_OS + t + (u32)n * C::S : t * C::IN_OS + (u32)n;
      raw[i] = t < count ? in[gi] : cpx<SIn>{};
    }
#pragma unroll
    for (int i = 0; i < IT; ++i) {
      const int idx = tid + i * C::BLOCK;
      int c, n;
      if constexpr (C::S > 1) { c = idx % T; n = idx / T; }
      else { c = idx / N; n = idx % N; }
      lds[c * LP + n] = ccvt<float>(raw[i]);
    }
  }

  // ---- per-lane constants: DFT-16 operands (hi/lo) and stage twiddles ----
  const int r = lane & 15;
  const int g = lane >> 4;
  bf16x8_t fre_hi, fre_lo, fim_hi, fim_lo;
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    const int n = 8 * (g & 1) + i;
    cpx<float> w = roots[(16 * ((n * r) & 15)) & 255];  // W16^(n r)
    if constexpr (C::DIR > 0) w.y = -w.y;
    const float fr = w.x, fi = w.y;
    const float bre = g < 2 ? fr : -fi;  // [Fr ; -Fi]
    const float bim = g < 2 ? fi : fr;   // [Fi ;  Fr]
    __bf16 h, l;
    split_bf16(bre, h, l);
    fre_hi[i] = h;
    fre_lo[i] = l;
    split_bf16(bim, h, l);
    fim_hi[i] = h;
    fim_lo[i] = l;
  }
  cpx<float> tw[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    cpx<float> w = roots[((4 * g + i) * r) & 255];  // W256^(n1 k2)
    if constexpr (C::DIR > 0) w.y = -w.y;
    tw[i] = w;
  }
  __syncthreads();

  // ---- transforms: two MFMA stages per 256-point row ----
  for (int c = wave; c < T; c += WAVES) {
    cpx<float>* row = lds + c * LP;
    bf16x8_t ahi, alo;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const int n2 = 8 * (g & 1) + i;
      const cpx<float> v = row[r + 16 * n2];
      __bf16 h, l;
      split_bf16(g < 2 ? v.x : v.y, h, l);
      ahi[i] = h;
      alo[i] = l;
    }
    f32x4_t yr = {0.f, 0.f, 0.f, 0.f}, yi = {0.f, 0.f, 0.f, 0.f};
    yr = mma_bf16(ahi, fre_hi, yr);
    yr = mma_bf16(alo, fre_hi, yr);
    yr = mma_bf16(ahi, fre_lo, yr);
    yi = mma_bf16(ahi, fim_hi, yi);
    yi = mma_bf16(alo, fim_hi, yi);
    yi = mma_bf16(ahi, fim_lo, yi);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const cpx<float> y = cmul(cpx<float>{yr[i], yi[i]}, tw[i]);
      row[(4 * g + i) * 16 + r] = y;  // Y[n1][k2]
    }
    bf16x8_t bhi, blo;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const int n1 = 8 * (g & 1) + i;
      const cpx<float> v = row[n1 * 16 + r];
      __bf16 h, l;
      split_bf16(g < 2 ? v.x : v.y, h, l);
      bhi[i] = h;
      blo[i] = l;
    }
    f32x4_t zr = {0.f, 0.f, 0.f, 0.f}, zi = {0.f, 0.f, 0.f, 0.f};
    zr = mma_bf16(fre_hi, bhi, zr);
    zr = mma_bf16(fre_lo, bhi, zr);
    zr = mma_bf16(fre_hi, blo, zr);
    zi = mma_bf16(fim_hi, bhi, zi);
    zi = mma_bf16(fim_lo, bhi, zi);
    zi = mma_bf16(fim_hi, blo, zi);
#pragma unroll
    for (int i = 0; i < 4; ++i) row[r + 16 * (4 * g + i)] = cpx<float>{zr[i], zi[i]};
  }
  __syncthreads();

  // ---- store tile ----
  const float scale = (float)C::OUT_SCALE;
#pragma unroll
  for (int i = 0; i < IT; ++i) {
    const int idx = tid + i * C::BLOCK;
    int c, n;
    if constexpr (C::S > 1) { c = idx % T; n = idx / T; }
    else { c = idx / N; n = idx % N; }
    const u32 t = first + (u32)c;
    if (t < count) {
      const u32 gi = C::S > 1 ? outer * C::OUT_OS + t + (u32)n * C::S : t * C::OUT_OS + (u32)n;
      out[gi] = ccvt<SOut>(lds[c * LP + n] * scale);
    }
  }
}

}  // namespace vkx
