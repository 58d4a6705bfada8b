// Device building blocks of the xSPDE ensemble path on gfx950:
// complex arithmetic, the Philox4x32-10 counter stream with Box-Muller
// normals, and fully unrolled in-register DFTs of length <= 64.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

namespace xsh {

template <class R>
struct cpx {
  R x, y;
};

template <class R>
__device__ __forceinline__ cpx<R> mk(R a, R b) { return cpx<R>{a, b}; }
template <class R>
__device__ __forceinline__ cpx<R> operator+(cpx<R> a, cpx<R> b) { return {a.x + b.x, a.y + b.y}; }
template <class R>
__device__ __forceinline__ cpx<R> operator-(cpx<R> a, cpx<R> b) { return {a.x - b.x, a.y - b.y}; }
template <class R>
__device__ __forceinline__ cpx<R> operator*(cpx<R> a, cpx<R> b) {
  return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <class R>
__device__ __forceinline__ cpx<R> operator*(R s, cpx<R> a) { return {s * a.x, s * a.y}; }
template <class R>
__device__ __forceinline__ cpx<R> conjc(cpx<R> a) { return {a.x, -a.y}; }
template <class R>
__device__ __forceinline__ R norm2(cpx<R> a) { return a.x * a.x + a.y * a.y; }
// multiply by -i (SIGN=-1) or +i (SIGN=+1)
template <int SIGN, class R>
__device__ __forceinline__ cpx<R> mul_i(cpx<R> a) {
  return SIGN > 0 ? cpx<R>{-a.y, a.x} : cpx<R>{a.y, -a.x};
}

// cos/sin(2 pi m / 64): compile-time twiddles of the register DFTs
#include "xsh_twiddle64.inc"

// ------------------------------------------------------------------ DFT
// Compile-time loop: f(integral_constant<int, I>) for I in [B, E).  Every
// register-array index below is a constant expression, so the arrays stay in
// VGPRs whatever the unroller decides.
template <int B, int E, class F>
__device__ __forceinline__ void static_for(F&& f) {
  if constexpr (B < E) {
    f(std::integral_constant<int, B>{});
    static_for<B + 1, E>(f);
  }
}

__host__ __device__ constexpr int ilog2(int n) { return n <= 1 ? 0 : 1 + ilog2(n / 2); }
__host__ __device__ constexpr int bit_reverse(int i, int bits) {
  int r = 0;
  for (int b = 0; b < bits; ++b)
    if (i & (1 << b)) r |= 1 << (bits - 1 - b);
  return r;
}

__device__ __forceinline__ float fmad(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fmad(double a, double b, double c) { return __builtin_fma(a, b, c); }

// (a + w b, a - w b) for the compile-time twiddle w = cos(2 pi m / 64) +
// SIGN i sin(2 pi m / 64) in six fused multiply-adds instead of a complex
// multiply and two complex additions (eight): w b = c (b + i (s/c) b) when
// |c| >= |s|, else s ((c/s) b + i b), and the outer factor goes into the
// multiply-adds with a (the ratio never exceeds 1 in magnitude).
template <class R, int SIGN, int M>
__device__ __forceinline__ void twiddle_butterfly(cpx<R> a, cpx<R> b, cpx<R>& out0, cpx<R>& out1) {
  constexpr double c = kCos64[M], s = SIGN * kSin64[M];
  constexpr bool cos_major = (c < 0 ? -c : c) >= (s < 0 ? -s : s);
  constexpr R f = R(cos_major ? c : s);
  constexpr R r = R(cos_major ? s / c : c / s);
  cpx<R> t;
  if constexpr (cos_major) {
    t = cpx<R>{fmad(-r, b.y, b.x), fmad(r, b.x, b.y)};   // w b = c t
  } else {
    t = cpx<R>{fmad(r, b.x, -b.y), fmad(r, b.y, b.x)};   // w b = s t
  }
  out0 = cpx<R>{fmad(f, t.x, a.x), fmad(f, t.y, a.y)};
  out1 = cpx<R>{fmad(-f, t.x, a.x), fmad(-f, t.y, a.y)};
}

// In-register DFT of length N (power of two, N <= 64), natural order in and
// out.  SIGN = -1 forward (exp(-2 pi i nk/N)), +1 inverse (no 1/N).
// Radix-2 decimation in time over the bit-reversed order; the twiddles are
// compile-time constants (k = 0 and k = len/4 need no multiply, the others
// run as twiddle_butterfly).
template <class R, int N, int SIGN>
__device__ __forceinline__ void dft(cpx<R> (&v)[N]) {
  if constexpr (N == 1) {
    return;
  } else if constexpr (N == 2) {
    const cpx<R> a = v[0], b = v[1];
    v[0] = a + b;
    v[1] = a - b;
  } else if constexpr (N == 4) {
    const cpx<R> a0 = v[0] + v[2], a1 = v[0] - v[2];
    const cpx<R> b0 = v[1] + v[3], b1 = mul_i<SIGN>(v[1] - v[3]);
    v[0] = a0 + b0;
    v[2] = a0 - b0;
    v[1] = a1 + b1;
    v[3] = a1 - b1;
  } else {
    constexpr int BITS = ilog2(N);
    static_for<0, N>([&](auto I) __attribute__((always_inline)) {
      constexpr int i = decltype(I)::value;
      constexpr int j = bit_reverse(i, BITS);
      if constexpr (j > i) {
        const cpx<R> t = v[i];
        v[i] = v[j];
        v[j] = t;
      }
    });
    static_for<1, BITS + 1>([&](auto S) __attribute__((always_inline)) {
      constexpr int len = 1 << decltype(S)::value;
      constexpr int half = len / 2;
      static_for<0, N / len>([&](auto G) __attribute__((always_inline)) {
        static_for<0, half>([&](auto K) __attribute__((always_inline)) {
          constexpr int i = decltype(G)::value * len;
          constexpr int k = decltype(K)::value;
          const cpx<R> a = v[i + k];
          cpx<R> b = v[i + k + half];
          if constexpr (k == 0 || 4 * k == len) {
            if constexpr (k != 0) b = mul_i<SIGN>(b);
            v[i + k] = a + b;
            v[i + k + half] = a - b;
          } else {
            twiddle_butterfly<R, SIGN, k * (64 / len)>(a, b, v[i + k], v[i + k + half]);
          }
        });
      });
    });
  }
}

// --------------------------------------------------------------- Philox
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11).  Stream layout:
//   key     = (nsp, tag | seq << 8)
//   counter = (group, draw, traj, inst),  group = L / 4, lane = L % 4
// where L = site * comps + comp.
struct u4 {
  uint32_t x, y, z, w;
};

// a ^ b ^ c in one instruction: gfx950's three-input bit operation with the
// truth table of the three-way XOR (0x96; symmetric, so operand order is free)
__device__ __forceinline__ uint32_t xor3(uint32_t a, uint32_t b, uint32_t c) {
  return __builtin_amdgcn_bitop3_b32(a, b, c, 0x96);
}

__device__ __forceinline__ u4 philox10(u4 c, uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    const uint64_t p0 = (uint64_t)0xD2511F53u * (uint64_t)c.x;
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * (uint64_t)c.z;
    c = u4{xor3((uint32_t)(p1 >> 32), c.y, k0), (uint32_t)p1, xor3((uint32_t)(p0 >> 32), c.w, k1),
           (uint32_t)p0};
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  return c;
}

enum : uint32_t { TAG_NOISE = 0, TAG_GAUSS_INIT = 1, TAG_UNIFORM_INIT = 2, TAG_KNOISE = 3, TAG_KRAND_INIT = 4 };

__device__ __forceinline__ uint32_t key1(uint32_t tag, uint32_t seq) {
  return (tag & 0xFFu) | (seq << 8);
}

// Box-Muller on (x0,x1): u = (x + 1/2) 2^-32 in (0,1).  FP32 on the hardware
// transcendentals directly: v_log_f32 is log2 (so -2 ln u = -2 ln2 log2 u),
// v_sin_f32 / v_cos_f32 take the angle in revolutions (sin(2 pi u1) =
// v_sin(u1), u1 < 1 is inside their +-256 input range), v_sqrt_f32 (1 ulp);
// u0 >= 2^-33 and -2 ln u0 <= 46 keep every operand normal, so none of the
// library's denormal scaling is needed.
__device__ __forceinline__ void box_muller(uint32_t x0, uint32_t x1, float& n0, float& n1) {
  const float u0 = (float)x0 * 2.3283064365386963e-10f + 1.1641532182693481e-10f;
  const float u1 = (float)x1 * 2.3283064365386963e-10f + 1.1641532182693481e-10f;
  const float r = __builtin_amdgcn_sqrtf(-1.3862943611198906f * __builtin_amdgcn_logf(u0));
  n0 = r * __builtin_amdgcn_cosf(u1);
  n1 = r * __builtin_amdgcn_sinf(u1);
}
__device__ __forceinline__ void box_muller(uint32_t x0, uint32_t x1, double& n0, double& n1) {
  const double u0 = ((double)x0 + 0.5) * 2.3283064365386963e-10;
  const double u1 = ((double)x1 + 0.5) * 2.3283064365386963e-10;
  const double r = sqrt(-2.0 * log(u0));
  double s, c;
  sincos(6.283185307179586476925 * u1, &s, &c);
  n0 = r * c;
  n1 = r * s;
}

template <class R>
__device__ __forceinline__ void normals4(u4 c, uint32_t k0, uint32_t k1, R (&n)[4]) {
  const u4 x = philox10(c, k0, k1);
  box_muller(x.x, x.y, n[0], n[1]);
  box_muller(x.z, x.w, n[2], n[3]);
}

template <class R>
__device__ __forceinline__ R uniform01(uint32_t x) {
  return (R)(((double)x + 0.5) * 2.3283064365386963e-10);
}

}  // namespace xsh
