// vkx_fft.h -- device-side building blocks of the gfx950 FFT kernels.
//
// Every FFT / convolution / transpose kernel of vkdispatch_amd is an
// instantiation of the templates below: the host planner (Python) picks the
// radix passes, threads-per-transform, transforms-per-workgroup and the I/O
// functors, writes a ~30 line instantiation, and compiles it for gfx950 with
// hiprtc.  Nothing here is generated: the butterflies, twiddles, LDS exchange
// and HBM access patterns are written by hand for wave64 / 160 KB LDS parts.
//
// Algorithm: Stockham autosort, decimation in time.  A transform of length N
// is a sequence of passes (R_p, NS_p) with NS_0 = 1, NS_{p+1} = NS_p * R_p and
// prod R_p = N.  Pass p processes N/R_p butterflies; butterfly j reads
// x[j + r*N/R_p], multiplies by W_{NS*R}^{r*(j mod NS)}, runs a radix-R DFT and
// writes y[(j/NS)*NS*R + j mod NS + r*NS].  With this form the first pass reads
// and the last pass writes "thread t handles n = t + T*m", so global loads and
// stores are always unit-stride across lanes (coalesced); the passes in
// between exchange through LDS.
//
// The semantics (sign convention exp(-2 pi i k n / N) forward, 1/N on the
// inverse when normalising, in-place, r2c = full complex FFT of (x, 0) with
// the N/2+1 packed layout) follow the reference vkdispatch FFT:
//   vkdispatch/fft/cooley_tukey.py:8-9, config.py:276-277 (sign),
//   vkdispatch/fft/shader_factories.py:52-53 (1/N),
//   vkdispatch/fft/global_memory_iterators.py:17-92,197-229 (r2c layout).
#pragma once

namespace vkx {

typedef unsigned int u32;
typedef int i32;
typedef unsigned long long u64;
typedef long long i64;

// ---------------------------------------------------------------------------
// compile-time loop helpers
// ---------------------------------------------------------------------------
template <int I> struct ic { static constexpr int value = I; };

template <int B, int E, class F>
__device__ __forceinline__ void sfor(F&& f) {
  if constexpr (B < E) {
    f(ic<B>{});
    sfor<B + 1, E>(f);
  }
}

template <class A, class B> struct same_t { static constexpr bool value = false; };
template <class A> struct same_t<A, A> { static constexpr bool value = true; };

// ---------------------------------------------------------------------------
// scalar storage types
// ---------------------------------------------------------------------------
struct bf16 { unsigned short b; };

__device__ __forceinline__ float bf16_to_f(bf16 v) { return __uint_as_float(((u32)v.b) << 16); }
__device__ __forceinline__ bf16 f_to_bf16(float f) {
  u32 u = __float_as_uint(f);
  if ((u & 0x7fffffffu) > 0x7f800000u) return bf16{(unsigned short)((u >> 16) | 0x40u)};  // quiet NaN
  u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest even
  return bf16{(unsigned short)(u >> 16)};
}

// scalar conversion To <- From
template <class To, class From> __device__ __forceinline__ To sc(From v) { return (To)v; }
template <> __device__ __forceinline__ float sc<float, bf16>(bf16 v) { return bf16_to_f(v); }
template <> __device__ __forceinline__ double sc<double, bf16>(bf16 v) { return (double)bf16_to_f(v); }
template <> __device__ __forceinline__ bf16 sc<bf16, float>(float v) { return f_to_bf16(v); }
template <> __device__ __forceinline__ bf16 sc<bf16, double>(double v) { return f_to_bf16((float)v); }
template <> __device__ __forceinline__ bf16 sc<bf16, bf16>(bf16 v) { return v; }
template <> __device__ __forceinline__ _Float16 sc<_Float16, bf16>(bf16 v) { return (_Float16)bf16_to_f(v); }
template <> __device__ __forceinline__ bf16 sc<bf16, _Float16>(_Float16 v) { return f_to_bf16((float)v); }

// ---------------------------------------------------------------------------
// complex type (registers and storage share the layout: {re, im})
// ---------------------------------------------------------------------------
template <class T> struct alignas(2 * sizeof(T)) cpx { T x, y; };

template <class T> __device__ __forceinline__ cpx<T> mk(T x, T y) { return cpx<T>{x, y}; }
template <class T> __device__ __forceinline__ cpx<T> operator+(cpx<T> a, cpx<T> b) { return {a.x + b.x, a.y + b.y}; }
template <class T> __device__ __forceinline__ cpx<T> operator-(cpx<T> a, cpx<T> b) { return {a.x - b.x, a.y - b.y}; }
template <class T> __device__ __forceinline__ cpx<T> operator-(cpx<T> a) { return {-a.x, -a.y}; }
// GLSL vec2 semantics for the hook language: complex*complex is component-wise.
// Complex multiplication is cmul().  Kernel code below never uses operator*.
template <class T> __device__ __forceinline__ cpx<T> operator*(cpx<T> a, cpx<T> b) { return {a.x * b.x, a.y * b.y}; }
template <class T> __device__ __forceinline__ cpx<T> operator/(cpx<T> a, cpx<T> b) { return {a.x / b.x, a.y / b.y}; }
template <class T> __device__ __forceinline__ cpx<T> operator*(cpx<T> a, T s) { return {a.x * s, a.y * s}; }
template <class T> __device__ __forceinline__ cpx<T> operator*(T s, cpx<T> a) { return {a.x * s, a.y * s}; }
template <class T> __device__ __forceinline__ cpx<T> operator/(cpx<T> a, T s) { return {a.x / s, a.y / s}; }
template <class T> __device__ __forceinline__ cpx<T> cmul(cpx<T> a, cpx<T> b) {
  return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <class T> __device__ __forceinline__ cpx<T> conj(cpx<T> a) { return {a.x, -a.y}; }
template <class T> __device__ __forceinline__ cpx<T> mul_i(cpx<T> a) { return {-a.y, a.x}; }
template <class T> __device__ __forceinline__ cpx<T> mul_negi(cpx<T> a) { return {a.y, -a.x}; }

// complex conversion between storage / compute precisions
template <class To, class From> __device__ __forceinline__ cpx<To> ccvt(cpx<From> v) {
  if constexpr (same_t<To, From>::value) return v;
  else return cpx<To>{sc<To>(v.x), sc<To>(v.y)};
}
template <class To, class From> __device__ __forceinline__ cpx<To> rcvt(From v) {
  return cpx<To>{sc<To>(v), sc<To>(0.0f)};
}

// ---------------------------------------------------------------------------
// compile-time trigonometry: cos/sin(2*pi*k/n) with exact octant reduction
// ---------------------------------------------------------------------------
constexpr double kPI = 3.141592653589793238462643383279502884;
struct ce_sc { double c, s; };
constexpr double ce_sin_poly(double x) {
  double x2 = x * x, term = x, s = x;
  for (int i = 1; i < 14; ++i) { term *= -x2 / (double)((2 * i) * (2 * i + 1)); s += term; }
  return s;
}
constexpr double ce_cos_poly(double x) {
  double x2 = x * x, term = 1.0, s = 1.0;
  for (int i = 1; i < 14; ++i) { term *= -x2 / (double)((2 * i - 1) * (2 * i)); s += term; }
  return s;
}
constexpr ce_sc ce_sincos2pi(i64 k, i64 n) {
  k %= n;
  if (k < 0) k += n;
  double sgn_s = 1.0, sgn_c = 1.0;
  i64 a = k, d = n;
  if (2 * a > d) { a = d - a; sgn_s = -1.0; }             // theta -> 2pi - theta
  if (4 * a > d) { a = d - 2 * a; d = 2 * d; sgn_c = -1.0; }  // theta -> pi - theta
  bool swap = false;
  if (8 * a > d) { a = d - 4 * a; d = 4 * d; swap = true; }  // theta -> pi/2 - theta
  double th = 2.0 * kPI * (double)a / (double)d;
  double c = ce_cos_poly(th), s = ce_sin_poly(th);
  if (swap) { double t = c; c = s; s = t; }
  return ce_sc{sgn_c * c, sgn_s * s};
}

// a * W_R^K with W_R = exp(DIR * 2 pi i / R); DIR = -1 forward, +1 inverse.
template <int R, int K, int DIR, class T>
__device__ __forceinline__ cpx<T> twc(cpx<T> a) {
  constexpr int k = ((K % R) + R) % R;
  if constexpr (k == 0) {
    return a;
  } else if constexpr (2 * k == R) {
    return cpx<T>{-a.x, -a.y};
  } else if constexpr (4 * k == R) {
    if constexpr (DIR < 0) return mul_negi(a); else return mul_i(a);
  } else if constexpr (4 * k == 3 * R) {
    if constexpr (DIR < 0) return mul_i(a); else return mul_negi(a);
  } else if constexpr ((8 * k) % R == 0) {
    // odd multiples of pi/4: (+-1 +- i)/sqrt(2): 2 adds + 2 muls
    constexpr ce_sc w = ce_sincos2pi(k, R);
    constexpr T h = (T)0.707106781186547524400844362104849;
    constexpr bool cpos = w.c > 0.0;
    constexpr bool spos = (DIR * w.s) > 0.0;
    const T ax = cpos ? a.x : -a.x, ay = cpos ? a.y : -a.y;  // c * a
    const T bx = spos ? a.y : -a.y, by = spos ? a.x : -a.x;  // s * (a.y, a.x)
    return cpx<T>{h * (ax - bx), h * (by + ay)};
  } else {
    constexpr ce_sc w = ce_sincos2pi(k, R);
    constexpr T c = (T)w.c;
    constexpr T s = (T)(DIR * w.s);
    return cpx<T>{a.x * c - a.y * s, a.x * s + a.y * c};
  }
}

// ---------------------------------------------------------------------------
// radix-R DFT in registers
// ---------------------------------------------------------------------------
constexpr int first_factor(int r) {
  if (r % 4 == 0 && r > 4) return 4;
  for (int f = 2; f * f <= r; ++f)
    if (r % f == 0) return f;
  return r;
}

// Prime-length DFT, symmetric form: pairs x[n], x[P-n] share cos/sin products.
template <int P, int DIR> struct PrimeDFT {
  template <class T> __device__ __forceinline__ static void run(cpx<T>* v) {
    if constexpr (P == 2) {
      const cpx<T> a = v[0] + v[1];
      v[1] = v[0] - v[1];
      v[0] = a;
    } else if constexpr (P == 3) {
      constexpr T c = (T)-0.5;
      constexpr T s = (T)(DIR * 0.866025403784438646763723170752936);
      const cpx<T> sm = v[1] + v[2], df = v[1] - v[2];
      const cpx<T> x0 = v[0];
      const cpx<T> A{x0.x + c * sm.x, x0.y + c * sm.y};
      v[0] = x0 + sm;
      v[1] = cpx<T>{A.x - s * df.y, A.y + s * df.x};
      v[2] = cpx<T>{A.x + s * df.y, A.y - s * df.x};
    } else {
      constexpr int H = (P - 1) / 2;
      cpx<T> s[H], d[H];
      const cpx<T> x0 = v[0];
      cpx<T> sum = v[0];
      sfor<0, H>([&](auto nc) {
        constexpr int n = decltype(nc)::value;
        s[n] = v[n + 1] + v[P - 1 - n];
        d[n] = v[n + 1] - v[P - 1 - n];
        sum = sum + s[n];
      });
      sfor<1, H + 1>([&](auto kc) {
        constexpr int k = decltype(kc)::value;
        cpx<T> A = x0;
        cpx<T> Bc{(T)0, (T)0};
        sfor<0, H>([&](auto nc) {
          constexpr int n = decltype(nc)::value;
          constexpr ce_sc w = ce_sincos2pi((i64)(n + 1) * k, P);
          constexpr T c = (T)w.c;
          constexpr T sn = (T)w.s;
          A.x += s[n].x * c;
          A.y += s[n].y * c;
          Bc.x += d[n].x * sn;
          Bc.y += d[n].y * sn;
        });
        // X[k] = A + i*DIR*Bc,  X[P-k] = A - i*DIR*Bc
        v[k] = cpx<T>{A.x - DIR * Bc.y, A.y + DIR * Bc.x};
        v[P - k] = cpx<T>{A.x + DIR * Bc.y, A.y - DIR * Bc.x};
      });
      v[0] = sum;
    }
  }
};

// Composite radix: R = A*B, in-register Cooley-Tukey with constant twiddles.
// x[n1 + A*n2] -> DFT_B over n2 -> * W_R^{n1 k2} -> DFT_A over n1 -> X[k2 + B*k1]
template <int R, int DIR> struct DFT {
  template <class T> __device__ __forceinline__ static void run(cpx<T>* v) {
    constexpr int A = first_factor(R);
    if constexpr (R == 1) {
      return;
    } else if constexpr (A == R) {
      PrimeDFT<R, DIR>::run(v);
    } else {
      constexpr int B = R / A;
      cpx<T> tmp[R];
      sfor<0, A>([&](auto n1c) {
        constexpr int n1 = decltype(n1c)::value;
        cpx<T> w[B];
        sfor<0, B>([&](auto n2c) { w[decltype(n2c)::value] = v[n1 + A * decltype(n2c)::value]; });
        DFT<B, DIR>::run(w);
        sfor<0, B>([&](auto k2c) {
          constexpr int k2 = decltype(k2c)::value;
          tmp[n1 * B + k2] = twc<R, n1 * k2, DIR>(w[k2]);
        });
      });
      sfor<0, B>([&](auto k2c) {
        constexpr int k2 = decltype(k2c)::value;
        cpx<T> w[A];
        sfor<0, A>([&](auto n1c) { w[decltype(n1c)::value] = tmp[decltype(n1c)::value * B + k2]; });
        DFT<A, DIR>::run(w);
        sfor<0, A>([&](auto k1c) { v[k2 + B * decltype(k1c)::value] = w[decltype(k1c)::value]; });
      });
    }
  }
};

// ---------------------------------------------------------------------------
// pass lists
// ---------------------------------------------------------------------------
template <int R_, int NS_> struct P { static constexpr int R = R_; static constexpr int NS = NS_; };
template <class... Ps> struct PL {};

template <class L> struct pl_first;
template <class F, class... Rs> struct pl_first<PL<F, Rs...>> { using type = F; };
template <class L> struct pl_last;
template <class F> struct pl_last<PL<F>> { using type = F; };
template <class F, class S, class... Rs> struct pl_last<PL<F, S, Rs...>> { using type = typename pl_last<PL<S, Rs...>>::type; };

// ---------------------------------------------------------------------------
// kernel parameters: buffer pointers in the reference's declared order
// (output-map args, input-map args / input buffer, kernel-map args), then
// twiddle tables in the last two slots; hook Var/Const values in c[].
// ---------------------------------------------------------------------------
struct KParams {
  u64 p[16];
  alignas(16) unsigned char c[256];
};
constexpr int SLOT_ROOTS = 13;   // exp(-2 pi i k / roots_n), bound by the plan
constexpr int SLOT_TW_FWD = 14;
constexpr int SLOT_TW_INV = 15;

// Buffer pointers are device (global) memory: the cast through the global address
// space lets the compiler emit global_* instead of flat_* memory instructions (flat
// accesses also count against the LDS wait counter, which couples LDS waits to
// outstanding global loads).
template <class T> __device__ __forceinline__ T* bufp(const KParams& prm, int slot) {
  return (T*)(__attribute__((address_space(1))) T*)prm.p[slot];
}
template <class T> __device__ __forceinline__ T cval(const KParams& prm, int off) { return *(const T*)(prm.c + off); }

// ---------------------------------------------------------------------------
// per-thread geometry
// ---------------------------------------------------------------------------
struct Geo {
  int t;       // thread index within the transform
  int b;       // transform slot within the workgroup
  u32 outer;   // outer batch index
  u32 inner;   // inner (strided) batch index
  bool valid;  // this slot maps to a real transform
};

// Cfg::XCDMAP (optional, with Cfg::NWG a multiple of 8): workgroups are dealt to the
// XCDs round-robin (blockIdx mod 8), and the logical index makes the workgroups of
// one XCD take consecutive tiles, so that neighbouring tiles share that XCD's L2 and
// a cache line straddling two of them (rows gathered from unaligned offsets) leaves
// memory once.
template <class C, class = void> struct xcdmap_of { static constexpr bool value = false; };
template <class C> struct xcdmap_of<C, decltype((void)C::XCDMAP)> { static constexpr bool value = C::XCDMAP; };
template <class C> __device__ __forceinline__ u32 logical_wg() {
  if constexpr (xcdmap_of<C>::value) {
    static_assert(C::NWG % 8u == 0u, "XCD mapping needs a grid of 8k workgroups");
    return (blockIdx.x & 7u) * (C::NWG / 8u) + (blockIdx.x >> 3);
  } else {
    return blockIdx.x;
  }
}

template <class C> __device__ __forceinline__ Geo make_geo() {
  Geo g;
  const int tid = (int)threadIdx.x;
  if constexpr (C::COLFAST) { g.b = tid % C::B; g.t = tid / C::B; }
  else { g.t = tid % C::TT; g.b = tid / C::TT; }
  const u32 wg = logical_wg<C>();
  if constexpr (C::S > 1) {
    g.outer = wg / C::NWG_INNER;
    g.inner = (wg % C::NWG_INNER) * C::B + (u32)g.b;
    g.valid = g.inner < C::S;
  } else {
    g.outer = wg * C::B + (u32)g.b;
    g.inner = 0;
    g.valid = g.outer < C::OUTER;
  }
  return g;
}

// ---------------------------------------------------------------------------
// hook operations (the reference's GlobalReadOp / GlobalWriteOp,
// vkdispatch/fft/global_memory_iterators.py:51-229)
// ---------------------------------------------------------------------------
enum { RD_INPUT = 0, RD_KERNEL = 1 };

// input_signal_range: sample i of a transform is read when SIG_LO <= i < SIG_HI (zero
// otherwise); SIG_LO > SIG_HI is a wrapped kept range [SIG_LO, N) + [0, SIG_HI), i.e. an
// interior block [SIG_HI, SIG_LO) read as zeros (VkFFT-style zero padding).
template <class C> constexpr bool sig_limited() { return C::SIG_LO > 0 || C::SIG_HI < C::N; }
template <class C> __device__ __forceinline__ bool sig_kept(u32 i) {
  if constexpr (C::SIG_LO <= C::SIG_HI) return i >= (u32)C::SIG_LO && i < (u32)C::SIG_HI;
  else return i >= (u32)C::SIG_LO || i < (u32)C::SIG_HI;
}

template <class C, int ROLE> struct ReadOp {
  using real = typename C::real;
  using reg_t = cpx<real>;
  static constexpr u32 fft_size = C::N;
  reg_t& reg;
  u32 io_index;
  u32 fft_index;
  u32 batch_offset;

  template <class S>
  __device__ __forceinline__ void read_from_buffer(const cpx<S>* buf, reg_t& dst, u32 io) const {
    if constexpr (ROLE == RD_INPUT && sig_limited<C>()) {
      if (!sig_kept<C>(fft_index)) { dst = reg_t{(real)0, (real)0}; return; }
    }
    if constexpr (ROLE == RD_INPUT && C::R2C && !C::INV) {
      // packed real rows: buffer[io // 2][io % 2]
      dst = reg_t{sc<real>(((const S*)buf)[io]), (real)0};
    } else if constexpr (ROLE == RD_INPUT && C::R2C && C::INV) {
      if (fft_index >= (u32)(C::N / 2 + 1)) {
        const u32 io2 = 2u * batch_offset + (u32)C::N * C::S - io;
        const reg_t v = ccvt<real>(buf[io2]);
        dst = reg_t{v.x, -v.y};
      } else {
        dst = ccvt<real>(buf[io]);
      }
    } else {
      dst = ccvt<real>(buf[io]);
    }
  }
  // real-valued buffer: element io becomes (x, 0)
  template <class S>
  __device__ __forceinline__ void read_from_buffer(const S* buf, reg_t& dst, u32 io) const {
    if constexpr (ROLE == RD_INPUT && sig_limited<C>()) {
      if (!sig_kept<C>(fft_index)) { dst = reg_t{(real)0, (real)0}; return; }
    }
    dst = reg_t{sc<real>(buf[io]), (real)0};
  }
  template <class Ptr> __device__ __forceinline__ void read_from_buffer(Ptr buf, reg_t& dst) const { read_from_buffer(buf, dst, io_index); }
  template <class Ptr> __device__ __forceinline__ void read_from_buffer(Ptr buf) const { read_from_buffer(buf, reg, io_index); }
};

template <class C> struct WriteOp {
  using real = typename C::real;
  using reg_t = cpx<real>;
  static constexpr u32 fft_size = C::N;
  reg_t& reg;
  u32 io_index;
  u32 fft_index;
  u32 batch_offset;

  template <class S>
  __device__ __forceinline__ void write_to_buffer(cpx<S>* buf, const reg_t& src, u32 io) const {
    if constexpr (C::R2C && !C::INV) {
      if (fft_index < (u32)(C::N / 2 + 1)) buf[io] = ccvt<S>(src);
    } else if constexpr (C::R2C && C::INV) {
      ((S*)buf)[io] = sc<S>(src.x);
    } else {
      buf[io] = ccvt<S>(src);
    }
  }
  // real-valued buffer: the real part is stored
  template <class S>
  __device__ __forceinline__ void write_to_buffer(S* buf, const reg_t& src, u32 io) const {
    if constexpr (C::R2C && !C::INV) {
      if (fft_index < (u32)(C::N / 2 + 1)) buf[io] = sc<S>(src.x);
    } else {
      buf[io] = sc<S>(src.x);
    }
  }
  template <class Ptr> __device__ __forceinline__ void write_to_buffer(Ptr buf, const reg_t& src) const { write_to_buffer(buf, src, io_index); }
  template <class Ptr> __device__ __forceinline__ void write_to_buffer(Ptr buf) const { write_to_buffer(buf, reg, io_index); }
};

// Default I/O functors (no mapping hook): read/write one buffer slot.
template <int SLOT, class S> struct DefaultIn {
  template <class Op> __device__ __forceinline__ static void apply(Op& op, const KParams& prm) {
    op.read_from_buffer(bufp<const cpx<S>>(prm, SLOT));
  }
};
template <int SLOT, class S> struct DefaultOut {
  template <int K, class Op> __device__ __forceinline__ static void apply(Op& op, const KParams& prm) {
    op.write_to_buffer(bufp<cpx<S>>(prm, SLOT));
  }
};
// The reference's default convolution kernel map: X <- X * conj(K[io])
// (vkdispatch/fft/shader_factories.py:136-145).
template <int SLOT, class S> struct DefaultKernel {
  template <int K, class Op> __device__ __forceinline__ static void apply(Op& op, const KParams& prm) {
    typename Op::reg_t kv;
    op.read_from_buffer(bufp<const cpx<S>>(prm, SLOT), kv);
    op.reg = cmul(op.reg, conj(kv));
  }
};

// ---------------------------------------------------------------------------
// LDS addressing.  COLFAST (strided axes, transform slot fastest across lanes):
// element (b, pos) at pos*B + b.  Otherwise one padded row per transform.
// ---------------------------------------------------------------------------
template <class C> __device__ __forceinline__ int lds_addr(int b, int pos) {
  if constexpr (C::COLFAST) return pos * C::B + b;
  else return b * C::NP + pos + (pos >> C::PAD_SHIFT);
}

template <class C> __device__ __forceinline__ void block_sync() {
  if constexpr (C::WAVE_LOCAL) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  } else {
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Stockham pass pieces
// ---------------------------------------------------------------------------
template <class C, int R, int NS, int DIR>
__device__ __forceinline__ void pass_compute(cpx<typename C::real>* v, int t, const cpx<typename C::real>* tw) {
  using real = typename C::real;
  constexpr int W = C::N / R;
  constexpr int Q = (W + C::TT - 1) / C::TT;
  sfor<0, Q>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    constexpr bool full = (q + 1) * C::TT <= W;
    const int j = t + q * C::TT;
    if (full || j < W) {
      if constexpr (NS > 1) {
        const int k = j % NS;
        sfor<1, R>([&](auto rc) {
          constexpr int r = decltype(rc)::value;
          cpx<real> w = tw[(NS - 1) + (r - 1) * NS + k];
          if constexpr (DIR > 0) w.y = -w.y;
          v[q * R + r] = cmul(v[q * R + r], w);
        });
      }
      DFT<R, DIR>::run(v + q * R);
    }
  });
}

template <class C, int R, int NS>
__device__ __forceinline__ void lds_put(cpx<typename C::real>* lds, const cpx<typename C::real>* v, int t, int b) {
  constexpr int W = C::N / R;
  constexpr int Q = (W + C::TT - 1) / C::TT;
  sfor<0, Q>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    constexpr bool full = (q + 1) * C::TT <= W;
    const int j = t + q * C::TT;
    if (full || j < W) {
      const int base = (j / NS) * NS * R + (j % NS);
      sfor<0, R>([&](auto rc) {
        constexpr int r = decltype(rc)::value;
        lds[lds_addr<C>(b, base + r * NS)] = v[q * R + r];
      });
    }
  });
}

template <class C, int R>
__device__ __forceinline__ void lds_get(const cpx<typename C::real>* lds, cpx<typename C::real>* v, int t, int b) {
  constexpr int W = C::N / R;
  constexpr int Q = (W + C::TT - 1) / C::TT;
  sfor<0, Q>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    constexpr bool full = (q + 1) * C::TT <= W;
    const int j = t + q * C::TT;
    if (full || j < W) {
      sfor<0, R>([&](auto rc) {
        constexpr int r = decltype(rc)::value;
        v[q * R + r] = lds[lds_addr<C>(b, j + r * W)];
      });
    }
  });
}

template <class C, int DIR, class Cur, class... Rest>
__device__ __forceinline__ void run_passes_(cpx<typename C::real>* v, cpx<typename C::real>* lds, int t, int b,
                                            const cpx<typename C::real>* tw) {
  pass_compute<C, Cur::R, Cur::NS, DIR>(v, t, tw);
  if constexpr (sizeof...(Rest) > 0) {
    using Nx = typename pl_first<PL<Rest...>>::type;
    lds_put<C, Cur::R, Cur::NS>(lds, v, t, b);
    block_sync<C>();
    lds_get<C, Nx::R>(lds, v, t, b);
    block_sync<C>();
    run_passes_<C, DIR, Rest...>(v, lds, t, b, tw);
  }
}
template <class C, int DIR, class... Ps>
__device__ __forceinline__ void run_passes(PL<Ps...>, cpx<typename C::real>* v, cpx<typename C::real>* lds, int t, int b,
                                           const cpx<typename C::real>* tw) {
  run_passes_<C, DIR, Ps...>(v, lds, t, b, tw);
}

// Register-resident twiddles.  A thread's twiddles depend only on its index t and
// the pass list, so a kernel that runs the same transform many times per thread can
// read them once (tw_load) and run every transform from registers (run_passes_reg).
// Inside a loop that also stores to global memory the compiler cannot hoist the
// table reads itself (it cannot prove the stores leave the table alone), and each
// pass would wait on a fresh table read.  Slot q (R - 1) + (r - 1) of a pass holds
// its leg-r twiddle for butterfly j = t + q TT, in the order pass_compute reads them.
template <class C, class Pa> struct pass_tw_regs {
  static constexpr int W = C::N / Pa::R;
  static constexpr int Q = (W + C::TT - 1) / C::TT;
  static constexpr int value = Pa::NS > 1 ? Q * (Pa::R - 1) : 0;
};
template <class C, class L> struct list_tw_regs;
template <class C, class... Ps> struct list_tw_regs<C, PL<Ps...>> {
  static constexpr int value = (0 + ... + pass_tw_regs<C, Ps>::value);
};

template <class C, int OFF, class Cur, class... Rest>
__device__ __forceinline__ void tw_load_(cpx<typename C::real>* w, int t, const cpx<typename C::real>* tw) {
  constexpr int R = Cur::R, NS = Cur::NS, W = C::N / R, Q = (W + C::TT - 1) / C::TT;
  if constexpr (NS > 1) {
    sfor<0, Q>([&](auto qc) {
      constexpr int q = decltype(qc)::value;
      constexpr bool full = (q + 1) * C::TT <= W;
      const int j = t + q * C::TT;
      const int k = j % NS;
      sfor<1, R>([&](auto rc) {
        constexpr int r = decltype(rc)::value;
        cpx<typename C::real> x{0, 0};
        if (full || j < W) x = tw[(NS - 1) + (r - 1) * NS + k];  // rows j >= W have no butterfly
        w[OFF + q * (R - 1) + (r - 1)] = x;
      });
    });
  }
  if constexpr (sizeof...(Rest) > 0) tw_load_<C, OFF + pass_tw_regs<C, Cur>::value, Rest...>(w, t, tw);
}
template <class C, class... Ps>
__device__ __forceinline__ void tw_load(PL<Ps...>, cpx<typename C::real>* w, int t, const cpx<typename C::real>* tw) {
  tw_load_<C, 0, Ps...>(w, t, tw);
}

template <class C, int R, int NS, int DIR>
__device__ __forceinline__ void pass_compute_reg(cpx<typename C::real>* v, int t, const cpx<typename C::real>* w) {
  constexpr int W = C::N / R;
  constexpr int Q = (W + C::TT - 1) / C::TT;
  sfor<0, Q>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    constexpr bool full = (q + 1) * C::TT <= W;
    const int j = t + q * C::TT;
    if (full || j < W) {
      if constexpr (NS > 1) {
        sfor<1, R>([&](auto rc) {
          constexpr int r = decltype(rc)::value;
          cpx<typename C::real> x = w[q * (R - 1) + (r - 1)];
          if constexpr (DIR > 0) x.y = -x.y;
          v[q * R + r] = cmul(v[q * R + r], x);
        });
      }
      DFT<R, DIR>::run(v + q * R);
    }
  });
}

template <class C, int DIR, int OFF, class Cur, class... Rest>
__device__ __forceinline__ void run_passes_reg_(cpx<typename C::real>* v, cpx<typename C::real>* lds, int t, int b,
                                                const cpx<typename C::real>* w) {
  pass_compute_reg<C, Cur::R, Cur::NS, DIR>(v, t, w + OFF);
  if constexpr (sizeof...(Rest) > 0) {
    using Nx = typename pl_first<PL<Rest...>>::type;
    lds_put<C, Cur::R, Cur::NS>(lds, v, t, b);
    block_sync<C>();
    lds_get<C, Nx::R>(lds, v, t, b);
    block_sync<C>();
    run_passes_reg_<C, DIR, OFF + pass_tw_regs<C, Cur>::value, Rest...>(v, lds, t, b, w);
  }
}
template <class C, int DIR, class... Ps>
__device__ __forceinline__ void run_passes_reg(PL<Ps...>, cpx<typename C::real>* v, cpx<typename C::real>* lds, int t,
                                               int b, const cpx<typename C::real>* w) {
  run_passes_reg_<C, DIR, 0, Ps...>(v, lds, t, b, w);
}

// C::TWREG (a plan option; false when the configuration does not name it): the
// plan's kernels take their twiddles from registers, loaded at the start of the
// kernel together with the data, instead of from the table after each exchange.
// More registers, fewer exposed waits; the autotuner decides per plan.
template <class C, class = void> struct twreg_of { static constexpr bool value = false; };
template <class C> struct twreg_of<C, decltype((void)C::TWREG)> { static constexpr bool value = C::TWREG; };

template <class C, class L, bool REG = twreg_of<C>::value> struct TwSet {
  const cpx<typename C::real>* tw;
  __device__ __forceinline__ void load(int, const cpx<typename C::real>* table) { tw = table; }
  template <int DIR>
  __device__ __forceinline__ void run(cpx<typename C::real>* v, cpx<typename C::real>* lds, int t, int b) const {
    run_passes<C, DIR>(L{}, v, lds, t, b, tw);
  }
};
template <class C, class L> struct TwSet<C, L, true> {
  static constexpr int NW = list_tw_regs<C, L>::value;
  cpx<typename C::real> w[NW > 0 ? NW : 1];
  __device__ __forceinline__ void load(int t, const cpx<typename C::real>* table) { tw_load<C>(L{}, w, t, table); }
  template <int DIR>
  __device__ __forceinline__ void run(cpx<typename C::real>* v, cpx<typename C::real>* lds, int t, int b) const {
    run_passes_reg<C, DIR>(L{}, v, lds, t, b, w);
  }
};

// Iterate the register slots of the layout produced by pass (R): slot q*R+r
// holds element n = j + r*N/R, j = t + q*TT (first pass input / last pass output).
template <class C, int R, class F>
__device__ __forceinline__ void for_layout(int t, F&& f) {
  constexpr int W = C::N / R;
  constexpr int Q = (W + C::TT - 1) / C::TT;
  sfor<0, Q>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    constexpr bool full = (q + 1) * C::TT <= W;
    const int j = t + q * C::TT;
    if (full || j < W) {
      sfor<0, R>([&](auto rc) {
        constexpr int r = decltype(rc)::value;
        f(ic<q * R + r>{}, (u32)(j + r * W));
      });
    }
  });
}

template <class C> __device__ __forceinline__ u32 in_boff(const Geo& g) { return g.outer * C::IN_OS + g.inner; }
template <class C> __device__ __forceinline__ u32 out_boff(const Geo& g) { return g.outer * C::OUT_OS + g.inner; }

// ---------------------------------------------------------------------------
// plain FFT kernel body: read (hook) -> passes -> [1/N] -> write (hook)
// ---------------------------------------------------------------------------
template <class C, class In, class Out>
__device__ __forceinline__ void fft_body(const KParams& prm) {
  using real = typename C::real;
  using L = typename C::fwd;
  using F0 = typename pl_first<L>::type;
  using FL = typename pl_last<L>::type;
  __shared__ cpx<real> lds[C::LDS_ELEMS];
  const Geo g = make_geo<C>();
  cpx<real> v[C::VMAX];
  const cpx<real>* tw = bufp<const cpx<real>>(prm, SLOT_TW_FWD);

  if (g.valid) {
    const u32 boff = in_boff<C>(g);
    for_layout<C, F0::R>(g.t, [&](auto slot, u32 n) {
      ReadOp<C, RD_INPUT> op{v[decltype(slot)::value], boff + n * C::S, n, boff};
      In::apply(op, prm);
    });
  } else {
    sfor<0, C::VMAX>([&](auto s) { v[decltype(s)::value] = cpx<real>{(real)0, (real)0}; });
  }
  TwSet<C, L> ts;
  ts.load(g.t, tw);

  ts.template run<(C::INV ? 1 : -1)>(v, lds, g.t, g.b);

  if (g.valid) {
    const u32 boff = out_boff<C>(g);
    const real scale = (real)C::SCALE;
    for_layout<C, FL::R>(g.t, [&](auto slot, u32 n) {
      constexpr int s = decltype(slot)::value;
      if constexpr (C::NORMALIZE) v[s] = v[s] * scale;
      WriteOp<C> op{v[s], boff + n * C::S, n, boff};
      Out::template apply<0>(op, prm);
    });
  }
}

// ---------------------------------------------------------------------------
// fused convolution kernel body (the reference's make_convolution_shader,
// vkdispatch/fft/shader_factories.py:107-195):
//   read -> forward FFT -> for each kernel: kernel_map -> inverse FFT -> [1/N] -> write
// The inverse pass list is the forward list reversed, so the forward output
// layout is exactly the inverse input layout: no register shuffle, no LDS trip.
// ---------------------------------------------------------------------------
template <class C> __device__ __forceinline__ u32 kernel_index_base(const Geo& g) {
  if constexpr (C::KERNEL_INNER_ONLY) return g.inner;
  else return g.outer * C::K_OS + g.inner;
}

// Transposed-kernel layout: slot-major per workgroup so that every kernel read
// of the convolution is a unit-stride 64-lane access.
template <class C> __device__ __forceinline__ u32 transposed_index(const Geo& g, int slot) {
  const u32 tid = threadIdx.x;
  if constexpr (C::KERNEL_INNER_ONLY) {
    if constexpr (C::S > 1) return ((logical_wg<C>() % C::NWG_INNER) * C::KSLOTS + (u32)slot) * C::BLOCK + tid;
    else return (u32)slot * C::TT + (u32)g.t;
  } else {
    return (logical_wg<C>() * C::KSLOTS + (u32)slot) * C::BLOCK + tid;
  }
}

template <class C, class In, class KMap, class Out>
__device__ __forceinline__ void conv_body(const KParams& prm) {
  using real = typename C::real;
  using L = typename C::fwd;
  using LI = typename C::inv;
  using F0 = typename pl_first<L>::type;
  using FL = typename pl_last<L>::type;
  using IL = typename pl_last<LI>::type;
  __shared__ cpx<real> lds[C::LDS_ELEMS];
  const Geo g = make_geo<C>();
  cpx<real> v[C::VMAX];
  const cpx<real>* twf = bufp<const cpx<real>>(prm, SLOT_TW_FWD);
  const cpx<real>* twi = bufp<const cpx<real>>(prm, SLOT_TW_INV);

  if (g.valid) {
    const u32 boff = in_boff<C>(g);
    for_layout<C, F0::R>(g.t, [&](auto slot, u32 n) {
      ReadOp<C, RD_INPUT> op{v[decltype(slot)::value], boff + n * C::S, n, boff};
      In::apply(op, prm);
    });
  } else {
    sfor<0, C::VMAX>([&](auto s) { v[decltype(s)::value] = cpx<real>{(real)0, (real)0}; });
  }
  // a twiddle table depends only on N and the radix list (the inverse passes
  // conjugate): a palindromic radix list shares one set between the directions
  constexpr bool SAME = __is_same(L, LI);
  TwSet<C, L> tsf;
  tsf.load(g.t, twf);
  TwSet<C, LI, SAME ? false : twreg_of<C>::value> tsi_own;
  tsi_own.load(g.t, twi);

  tsf.template run<-1>(v, lds, g.t, g.b);

  cpx<real> bak[C::KNUM > 1 ? C::VMAX : 1];
  if constexpr (C::KNUM > 1) {
    sfor<0, C::VMAX>([&](auto s) { bak[decltype(s)::value] = v[decltype(s)::value]; });
  }

  sfor<0, C::KNUM>([&](auto kc) {
    constexpr int kidx = decltype(kc)::value;
    if constexpr (kidx > 0) {
      sfor<0, C::VMAX>([&](auto s) { v[decltype(s)::value] = bak[decltype(s)::value]; });
    }
    if (g.valid) {
      const u32 kb = kernel_index_base<C>(g);
      for_layout<C, FL::R>(g.t, [&](auto slot, u32 n) {
        constexpr int s = decltype(slot)::value;
        const u32 io = C::KERNEL_TRANSPOSED ? transposed_index<C>(g, s) : kb + n * C::S;
        ReadOp<C, RD_KERNEL> op{v[s], io, n, kb};
        KMap::template apply<kidx>(op, prm);
      });
    }
    if constexpr (SAME) tsf.template run<1>(v, lds, g.t, g.b);
    else tsi_own.template run<1>(v, lds, g.t, g.b);
    if (g.valid) {
      const u32 boff = out_boff<C>(g);
      const real scale = (real)C::SCALE;
      for_layout<C, IL::R>(g.t, [&](auto slot, u32 n) {
        constexpr int s = decltype(slot)::value;
        if constexpr (C::NORMALIZE) v[s] = v[s] * scale;
        WriteOp<C> op{v[s], boff + n * C::S, n, boff};
        Out::template apply<kidx>(op, prm);
      });
    }
    if constexpr (kidx + 1 < C::KNUM) block_sync<C>();
  });
}

// ---------------------------------------------------------------------------
// Real-to-complex / complex-to-real rows of even length 2M through an M-point
// complex FFT (the "half-length" method).  Same results as the reference's
// full-length complex r2c/c2r (global_memory_iterators.py:85-92,214-227) at
// half the arithmetic.  C::N is M; rows are contiguous; IN_OS / OUT_OS are the
// RFFTBuffer row pitches (2M+2 reals, M+1 complex).  SLOT_ROOTS binds
// exp(-2 pi i k / 2M).
//
// forward: z[m] = x[2m] + i x[2m+1]  ->  Z = FFT_M(z)
//          E = (Z[k] + conj Z[M-k]) / 2,  O = -i (Z[k] - conj Z[M-k]) / 2
//          X[k] = E + W^k O,  X[M-k] = conj(E - W^k O)        (k = 0..M/2)
// inverse: Z'[k] = (X[k] + conj X[M-k]) + i conj(W^k) (X[k] - conj X[M-k])
//          z = IFFT_M(Z'),  x[2m] = Re z[m] / 2M,  x[2m+1] = Im z[m] / 2M
//          (imaginary parts of X[0] and X[M] do not contribute, as in a
//           complex inverse followed by taking the real part)
// ---------------------------------------------------------------------------
// OOP = true: real rows are read from slot 1 (pitch C::IN_OS reals) and the
// spectrum written to slot 0; otherwise the RFFTBuffer in slot 0 is used in place.
template <class C, class S, bool OOP = false>
__device__ __forceinline__ void rfft_body(const KParams& prm) {
  using real = typename C::real;
  using L = typename C::fwd;
  using F0 = typename pl_first<L>::type;
  using FL = typename pl_last<L>::type;
  constexpr int M = C::N;
  __shared__ cpx<real> lds[C::LDS_ELEMS];
  const Geo g = make_geo<C>();
  cpx<real> v[C::VMAX];
  const cpx<real>* tw = bufp<const cpx<real>>(prm, SLOT_TW_FWD);
  const cpx<real>* roots = bufp<const cpx<real>>(prm, SLOT_ROOTS);
  const S* xin = bufp<const S>(prm, OOP ? 1 : 0);
  cpx<S>* xout = bufp<cpx<S>>(prm, 0);

  if (g.valid) {
    const u32 base = g.outer * C::IN_OS;
    for_layout<C, F0::R>(g.t, [&](auto slot, u32 n) {
      const cpx<S> p = *(const cpx<S>*)(xin + base + 2u * n);
      real a = sc<real>(p.x), b = sc<real>(p.y);
      if constexpr (C::SIG_LO > 0 || C::SIG_HI < 2 * M) {
        const u32 e0 = 2u * n, e1 = 2u * n + 1u;
        if (!(e0 >= (u32)C::SIG_LO && e0 < (u32)C::SIG_HI)) a = (real)0;
        if (!(e1 >= (u32)C::SIG_LO && e1 < (u32)C::SIG_HI)) b = (real)0;
      }
      v[decltype(slot)::value] = cpx<real>{a, b};
    });
  } else {
    sfor<0, C::VMAX>([&](auto s) { v[decltype(s)::value] = cpx<real>{(real)0, (real)0}; });
  }
  TwSet<C, L> ts;
  ts.load(g.t, tw);

  ts.template run<-1>(v, lds, g.t, g.b);

  // natural-order spectrum into LDS, then Hermitian post-processing in pairs (k, M-k)
  for_layout<C, FL::R>(g.t, [&](auto slot, u32 n) { lds[lds_addr<C>(g.b, (int)n)] = v[decltype(slot)::value]; });
  block_sync<C>();
  if (g.valid) {
    const u32 ob = g.outer * C::OUT_OS;
    const real h = (real)0.5;
    for (int k = g.t; k <= M / 2; k += C::TT) {
      const cpx<real> zk = lds[lds_addr<C>(g.b, k)];
      const cpx<real> zm = lds[lds_addr<C>(g.b, k == 0 ? 0 : M - k)];
      const cpx<real> e{h * (zk.x + zm.x), h * (zk.y - zm.y)};
      const cpx<real> d{zk.x - zm.x, zk.y + zm.y};
      const cpx<real> o{h * d.y, -h * d.x};
      const cpx<real> t = cmul(roots[k], o);
      xout[ob + (u32)k] = ccvt<S>(e + t);
      xout[ob + (u32)(M - k)] = ccvt<S>(conj(e - t));
    }
  }
}

// OOP = true: the spectrum is read from slot 1 (pitch C::IN_OS complex) and real
// rows written to slot 0 (pitch C::OUT_OS reals).
template <class C, class S, bool OOP = false>
__device__ __forceinline__ void irfft_body(const KParams& prm) {
  using real = typename C::real;
  using L = typename C::fwd;
  using F0 = typename pl_first<L>::type;
  using FL = typename pl_last<L>::type;
  constexpr int M = C::N;
  __shared__ cpx<real> lds[C::LDS_ELEMS];
  const Geo g = make_geo<C>();
  cpx<real> v[C::VMAX];
  const cpx<real>* tw = bufp<const cpx<real>>(prm, SLOT_TW_FWD);
  const cpx<real>* roots = bufp<const cpx<real>>(prm, SLOT_ROOTS);
  const cpx<S>* Xin = bufp<const cpx<S>>(prm, OOP ? 1 : 0);
  S* xout = bufp<S>(prm, 0);
  TwSet<C, L> ts;
  ts.load(g.t, tw);

  if (g.valid) {
    const u32 ib = g.outer * C::IN_OS;
    for_layout<C, F0::R>(g.t, [&](auto slot, u32 n) {
      cpx<real> a = ccvt<real>(Xin[ib + n]);
      cpx<real> bm = ccvt<real>(Xin[ib + (u32)M - n]);
      if (n == 0u) { a.y = (real)0; bm.y = (real)0; }
      const cpx<real> e{a.x + bm.x, a.y - bm.y};
      const cpx<real> d{a.x - bm.x, a.y + bm.y};
      const cpx<real> o = cmul(d, conj(roots[n]));
      v[decltype(slot)::value] = e + mul_i(o);
    });
  } else {
    sfor<0, C::VMAX>([&](auto s) { v[decltype(s)::value] = cpx<real>{(real)0, (real)0}; });
  }

  ts.template run<1>(v, lds, g.t, g.b);
  block_sync<C>();  // every load of the row precedes every (in-place) store

  if (g.valid) {
    const u32 ob = g.outer * C::OUT_OS;
    const real scale = (real)C::SCALE;
    for_layout<C, FL::R>(g.t, [&](auto slot, u32 n) {
      cpx<real> z = v[decltype(slot)::value];
      if constexpr (C::NORMALIZE) z = z * scale;
      *(cpx<S>*)(xout + ob + 2u * n) = ccvt<S>(z);
    });
  }
}

// ---------------------------------------------------------------------------
// kernel-spectrum transpose into the transposed-kernel layout
// (the reference's make_transpose_shader, shader_factories.py:73-97)
// slot 0 = output (transposed), slot 1 = input spectrum
// ---------------------------------------------------------------------------
template <class C, class SIn, class SOut>
__device__ __forceinline__ void transpose_body(const KParams& prm) {
  using FL = typename pl_last<typename C::fwd>::type;
  const Geo g = make_geo<C>();
  if (!g.valid) return;
  if constexpr (C::KERNEL_INNER_ONLY) {
    if (g.outer != 0) return;
  }
  const cpx<SIn>* src = bufp<const cpx<SIn>>(prm, 1);
  cpx<SOut>* dst = bufp<cpx<SOut>>(prm, 0);
  const u32 kb = g.outer * C::K_OS + g.inner;
  for_layout<C, FL::R>(g.t, [&](auto slot, u32 n) {
    dst[transposed_index<C>(g, decltype(slot)::value)] = ccvt<SOut>(src[kb + n * C::S]);
  });
}

}  // namespace vkx
