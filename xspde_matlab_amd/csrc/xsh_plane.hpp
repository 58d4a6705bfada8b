// 2D periodic transform of one N x N plane per workgroup, arranged so that
// every wave owns whole lattice rows.
//
// Register layout (that of Lattice with TPB = 1): thread t = 64 w + l keeps
// sites x = 2 (t mod N/2) + v, y = r + RB i (v < 2, i < CPT), r = 2 t / N, with
// RB = 2 NT / N rows per chunk.  For N = 128, NT = 1024: r = w, so wave w owns
// the rows w, w + 16, ..., w + 112 completely.
//
// a -> T^-1 op(T a), N = CPT * RB along y and N = XN1 * XN2 along x:
//   1. y, first radix (CPT) in registers: the CPT values of one column that a
//      thread holds are the n1 samples of one n2 = r; twiddles W_N^{r k1} are
//      wave-uniform (scalar loads)
//   2. x forward on the wave's own rows in LDS, two radix passes (XN1, XN2):
//      wave-local, ordered by wavefront fences, no workgroup barrier
//   3. barrier; y second radix (RB) forward, op(X, kx, ky), inverse, per column
//      block (one LDS pass)
//   4. barrier; x inverse on the wave's own rows (two wave-local passes) and the
//      rows back to registers
//   5. y first radix inverse in registers
// Two workgroup barriers and six LDS sweeps per transform, where the generic
// Lattice::spectral needs nine barriers and eight sweeps.  op receives the
// natural frequency indices (kx, ky) as in Lattice::spectral.
#pragma once
#include "xsh_path.hpp"

namespace xsh {

template <class R, int N, int NT>
struct Plane {
  using C = cpx<R>;
  static constexpr int CPT = N * N / (2 * NT);  // register radix along y
  static constexpr int RB = 2 * NT / N;         // LDS radix along y (rows per chunk)
  static constexpr int XN1 = 16, XN2 = N / 16;  // x: first (stride XN2) and second radix
  // LDS layout: every XN2-element block of a row is followed by one pad
  // element (row length RL), and the rows of chunk i start 16 (i & 1) +
  // 8 ((i >> 1) & 1) + 32 (i >> 1) elements further on (non-decreasing in i, so
  // rows never overlap; 0, 16, 8, 24 modulo 32).  With it the strided (stage 1) and
  // the blocked (stage 2) x passes of the 64 lanes of a wave, and the row
  // stores and loads, spread over all 64 banks of 4 bytes.
  static constexpr int RL = N + N / XN2;
  static constexpr int LDS = N * RL + 16 + 8 + 32 * (CPT / 2 - 1);
  static constexpr int WROWS = CPT;             // rows owned by a wave
  static_assert(N == 128 && NT == 1024, "rows must be wave-owned: one row per wave and chunk");
  static_assert(CPT * RB == N && XN1 * XN2 == N, "radix splits must factor N");
  static_assert(WROWS * XN2 == 64, "x pass 1 maps one (row, n2) item to every lane");

  __device__ static __forceinline__ int wave() { return __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6)); }
  // lane through opaque_tid: per-lane addresses and twiddles are recomputed
  // where used instead of being hoisted out of the time loop into VGPRs
  __device__ static __forceinline__ int lane() { return opaque_tid() & 63; }
  // order this wave's LDS writes before its following LDS reads (no workgroup barrier)
  __device__ static __forceinline__ void wave_sync() { __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront"); }

  // 1./5. first radix along y in registers, both sites v of every chunk
  template <int SIGN>
  __device__ static __forceinline__ void yreg(C (&f)[2 * CPT], const C* __restrict__ twy) {
    const int r = wave();  // = 2 t / N for N = 128, NT = 1024
    static_for<0, 2>([&](auto V) __attribute__((always_inline)) {
      constexpr int v = decltype(V)::value;
      C a[CPT];
      static_for<0, CPT>([&](auto I) __attribute__((always_inline)) { a[decltype(I)::value] = f[2 * decltype(I)::value + v]; });
      if constexpr (SIGN < 0) {
        dft<R, CPT, -1>(a);
        static_for<1, CPT>([&](auto K) __attribute__((always_inline)) {
          constexpr int k = decltype(K)::value;
          a[k] = a[k] * gld(twy, r * k);
        });
      } else {
        static_for<1, CPT>([&](auto K) __attribute__((always_inline)) {
          constexpr int k = decltype(K)::value;
          a[k] = a[k] * conjc(gld(twy, r * k));
        });
        dft<R, CPT, 1>(a);
      }
      static_for<0, CPT>([&](auto I) __attribute__((always_inline)) { f[2 * decltype(I)::value + v] = a[decltype(I)::value]; });
    });
  }

  // LDS row of chunk i of wave w (position RB k1 + n2 with n2 = w, k1 = i)
  __device__ static __forceinline__ int row_of(int w, int i) { return w + RB * i; }
  // element offset of row y and of position x within a row
  __device__ static __forceinline__ int row_base(int y) {
    const int i = y / RB;
    return y * RL + 16 * (i & 1) + 8 * ((i >> 1) & 1) + 32 * (i >> 1);
  }
  __device__ static __forceinline__ constexpr int col(int x) { return x + x / XN2; }

  __device__ static __forceinline__ void store_rows(C* L, const C (&f)[2 * CPT]) {
    const int w = wave(), l = lane();
    static_for<0, CPT>([&](auto I) __attribute__((always_inline)) {
      constexpr int i = decltype(I)::value;
      C* p = L + row_base(row_of(w, i)) + col(2 * l);
      p[0] = f[2 * i];
      p[1] = f[2 * i + 1];
    });
  }
  __device__ static __forceinline__ void load_rows(const C* L, C (&f)[2 * CPT]) {
    const int w = wave(), l = lane();
    static_for<0, CPT>([&](auto I) __attribute__((always_inline)) {
      constexpr int i = decltype(I)::value;
      const C* p = L + row_base(row_of(w, i)) + col(2 * l);
      f[2 * i] = p[0];
      f[2 * i + 1] = p[1];
    });
  }

  // 2. forward x on the wave's rows: radix XN1 over stride XN2, then twiddles
  // and radix XN2 on each contiguous block; position XN2 k1 + k2 ends up
  // holding X[k1 + XN1 k2]
  __device__ static __forceinline__ void x_forward(C* L, const C* __restrict__ twx) {
    const int w = wave(), l = lane();
    {
      const int i = l / XN2, n2 = l % XN2;
      C* row = L + row_base(row_of(w, i)) + n2;
      C v[XN1];
#pragma unroll
      for (int n1 = 0; n1 < XN1; ++n1) v[n1] = row[(XN2 + 1) * n1];
      dft<R, XN1, -1>(v);
#pragma unroll
      for (int k1 = 0; k1 < XN1; ++k1) row[(XN2 + 1) * k1] = v[k1];
    }
    wave_sync();
    // the twiddles W_N^{n2 k1} between the stages are applied as the second
    // stage reads its block (XN2 - 1 of them live, not XN1 - 1)
#pragma unroll 1
    for (int u = 0; u < WROWS * XN1 / 64; ++u) {
      const int item = l + 64 * u, i = item / XN1, k1 = item % XN1;
      C* blk = L + row_base(row_of(w, i)) + (XN2 + 1) * k1;
      C v[XN2];
#pragma unroll
      for (int n = 0; n < XN2; ++n) v[n] = blk[n];
#pragma unroll
      for (int n2 = 1; n2 < XN2; ++n2) v[n2] = v[n2] * gld(twx, n2 * k1);
      dft<R, XN2, -1>(v);
#pragma unroll
      for (int n = 0; n < XN2; ++n) blk[n] = v[n];
    }
  }

  // 4. inverse x on the wave's rows (the two passes of x_forward reversed)
  __device__ static __forceinline__ void x_inverse(C* L, const C* __restrict__ twx) {
    const int w = wave(), l = lane();
#pragma unroll 1
    for (int u = 0; u < WROWS * XN1 / 64; ++u) {
      const int item = l + 64 * u, i = item / XN1, k1 = item % XN1;
      C* blk = L + row_base(row_of(w, i)) + (XN2 + 1) * k1;
      C v[XN2];
#pragma unroll
      for (int n = 0; n < XN2; ++n) v[n] = blk[n];
      dft<R, XN2, 1>(v);
#pragma unroll
      for (int n2 = 1; n2 < XN2; ++n2) v[n2] = v[n2] * conjc(gld(twx, n2 * k1));
#pragma unroll
      for (int n = 0; n < XN2; ++n) blk[n] = v[n];
    }
    wave_sync();
    {
      const int i = l / XN2, n2 = l % XN2;
      C* row = L + row_base(row_of(w, i)) + n2;
      C v[XN1];
#pragma unroll
      for (int k1 = 0; k1 < XN1; ++k1) v[k1] = row[(XN2 + 1) * k1];
      dft<R, XN1, 1>(v);
#pragma unroll
      for (int n1 = 0; n1 < XN1; ++n1) row[(XN2 + 1) * n1] = v[n1];
    }
  }

  // frequency index held at x position p after x_forward
  __device__ static __forceinline__ int kx_of_pos(int p) { return p / XN2 + XN1 * (p % XN2); }

  // 3. second radix along y, op, inverse: item = (x position p, k1 = first-radix
  // frequency); the RB rows RB k1 .. RB k1 + RB - 1 of column p hold n2 = 0..RB-1
  template <class Op>
  __device__ static __forceinline__ void y_mid(C* L, const Op& op) {
    static_assert(N * CPT == NT, "one (column, k1) item per thread");
    const int t = opaque_tid();
    const int p = t % N, k1 = __builtin_amdgcn_readfirstlane(t / N);
    // rows RB k1 + n all belong to chunk k1: one base, constant row stride RL
    C* cl = L + row_base(RB * k1) + col(p);
    C v[RB];
#pragma unroll
    for (int n = 0; n < RB; ++n) v[n] = cl[n * RL];
    dft<R, RB, -1>(v);
    const int kx = kx_of_pos(p);
#pragma unroll
    for (int k2 = 0; k2 < RB; ++k2) v[k2] = op(v[k2], kx, k1 + CPT * k2, 0);
    dft<R, RB, 1>(v);
#pragma unroll
    for (int n = 0; n < RB; ++n) cl[n * RL] = v[n];
  }

  // a -> T^-1 op(T a) of the plane held in f (every thread's 2 CPT sites)
  template <class Op>
  __device__ static __forceinline__ void spectral(C* L, C (&f)[2 * CPT], const C* twx, const C* twy, const Op& op) {
    yreg<-1>(f, twy);
    store_rows(L, f);
    wave_sync();
    x_forward(L, twx);
    __syncthreads();
    y_mid(L, op);
    __syncthreads();
    x_inverse(L, twx);
    wave_sync();
    load_rows(L, f);
    yreg<1>(f, twy);
  }
};

}  // namespace xsh
