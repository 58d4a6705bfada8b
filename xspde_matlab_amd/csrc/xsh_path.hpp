// Path kernels of the xSPDE ensemble integration path on gfx950.
//
// One launch runs whole paths (xpath.m:29-86): every workgroup owns a group
// of trajectories of one (sub-ensemble j, check nc) path and keeps their
// fields resident on chip - in VGPRs - from the initial state to the last
// surplus step.  LDS is the workspace of the spectral transforms only:
//
//   forward  : in-place decimation-in-frequency stages (natural -> digit
//              reversed order), one LDS read+write per stage
//   propagate: the last forward stage, the multiply by the propagator
//              P(kx,ky) = exp(L dtr/2) and the first inverse stage are one
//              register-resident step (xprop.m:33-35)
//   inverse  : in-place decimation-in-time stages (digit reversed ->
//              natural order), so no permutation pass is ever needed.
//
// Noise is generated in registers per step from a Philox4x32-10 counter
// (xnoise.m:31-34 semantics), held constant over the method's stages
// (RK4.m:13-37).  Observables are reduced in FP64 and accumulated with
// FP64 atomics into the per-path ensemble means (xdata.m:84).
#pragma once
#include <type_traits>
#include "xsh_common.hpp"
#include "../../include/xspde_hip.h"

namespace xsh {

struct PathArgs {
  int nx, ny, ns;
  double x0, dx, y0, dy, dv;
  int npts, steps, checks, method, iterations;
  double hdt[2];           // dtr/2 of the fine (0) and coarse (1) runs
  int E, e2, e3, npar, seed, seq, inst;
  int j_only, nc_only;     // > 0: one path (j, nc)
  int groups;              // trajectory groups per path
  int noises, inrandoms, urandoms;
  int knoises, krandoms;   // k-space filtered noise / random components (0..2)
  const void* kn_filter;   // complex R [knoises][ns]: noise filter F(k) at the k
                           // sites in FFT order, times sqrt(nscale knoisefactor) / ns
  const void* kr_filter;   // complex R [krandoms][ns]: random filter, times
                           // sqrt(knoisefactor) / ns
  void* kn_work;           // complex R [work items * TPB][ns]: x-space k-noise of the step
  double noise_scale;      // sqrt(nscale * noisefactor), xnoise.m:32
  double init_scale;       // sqrt(noisefactor), xrandom.m:22
  int deriv;
  double dp[6];
  int initial;
  double ip[6];
  int nobs;
  int obs[XSH_MAX_OBS];
  long long obs_off[XSH_MAX_OBS];
  long long avlen;
  double inv_E;
  int has_linear;
  const void* px[2];       // [nc] propagator factor along x (incl. exp(l0 dt), 1/nx)
  const void* py[2];       // [nc] propagator factor along y (incl. 1/ny)
  const void* twx;         // W_nx^m, m < nx
  const void* twy;
  int mat;                 // 1: dense per-dimension transforms (non-periodic lattice)
  int cplx[2];             // per dimension: 1 if its matrix is complex (periodic DFT)
  const void* tm[4];       // forward x, forward y, inverse x, inverse y (row-major N x N)
  const void* tmp[4];      // real matrices in MFMA operand order, else NULL: FP32 (N % 32 == 0)
                           // Tp[kb][kk/2][lane] = T[32 kb + lane%32][kk + lane/32],
                           // FP64 (N % 16 == 0) Tp[kb][kk/4][lane] = T[16 kb + lane%16][kk + lane/16]
  int tsym[4];             // per packed matrix: 0 the full N x N GEMM; 1 / 2 the
                           // even/odd split of a DCT-I / DST-I (T[k][N-1-n] =
                           // +-(-1)^k T[k][n]): rows [0, N/2) hold T[2r][n], rows
                           // [N/2, N) T[2r+1][n], n < N/2, applied to
                           // x_n + sgn x_{N-1-n} with sgn = +1 for the even rows
                           // of a DCT-I and the odd rows of a DST-I, else -1
  int dirichlet;           // any Dirichlet end (xsetbound.m:40-53)
  double* traj;            // optional [paths][E][npts][nobs] per-trajectory lattice means
  double* raw;             // optional [paths][npts][E][ns] complex double fields
  int nmodes;              // Fourier modes of the initial noise to keep
  const int* fmap;         // [nx*ny] mode index of each FFT position, -1 if not kept
  float* feat;             // [paths][E][2*nmodes]
  double feat_scale;       // sqrt(dv / nsites): unit-variance Re/Im
  double kscale;           // prod_d kfact_d^2, kfact = dx / sqrt(2 pi) (xpreferences.m:167)
  int dir_lo[2], dir_hi[2];// per dimension: lower / upper end is Dirichlet
  int has_patch;           // non-zero boundary values: patch around every propagation
  const void* bval;        // complex R [level][dim][lower,upper][nb]: boundary values
                           // along the other dimension, level m at t0 + m dtr_fine / 2
  int bval_level;          // elements per level (0: one level for the whole path)
  int bval_nb;             // values per boundary line
  const void* bshape;      // real R [dim][lower,upper][n]: patch shape functions,
                           // pa(x, y) = sum_dim,side b(side, other) * shape(side, x_dim)
  int bshape_n;            // entries per shape function (max(nx, ny))
  double lpf[2][2][2];     // [nc][dim][re,im]: Neumann correction per unit of
                           // b(upper) - b(lower), da / range (xpropfactor.m:40-51)
  void* work;              // streaming kernels: per-workgroup field planes in device memory
  int work_planes;         // planes per workgroup (2, or 3 with the corr observable)
  const void* twh;         // streaming kernels: W_{ny/2}^m, m < ny/2
  double* ft;              // SDE time-domain spectra: [paths][ft_len][E] complex double
                           // step fields (time-major, trajectories fastest)
  int ft_len;              // points * steps
  int tspec_mask;          // bit o: observable o is taken of the time transform
  int* nonfinite;          // optional: set to 1 when a field value at a stored point is
                           // not finite (divergence check; NULL: not checked)
  double* av;              // [paths][avlen] complex double
  const double* a0;        // optional [E][ns] complex double initial fields
  double* a_out;           // optional [E][ns] complex double final fields
  // coupled SDE components (fields > 1, ns = fields): da_i/dt = sum_j fm[i][j] a_j
  // + fg[i] |a_i|^2 a_i + sum_k fs[i][k] w_k; propagator fp[nc][i]; constants fc0[i]
  int fields;
  double fm[XSH_MAX_FIELDS][XSH_MAX_FIELDS][2];
  double fg[XSH_MAX_FIELDS][2];
  double fs[XSH_MAX_FIELDS][4];
  double fp[2][XSH_MAX_FIELDS][2];
  double fc0[XSH_MAX_FIELDS][2];
  int obs_field[XSH_MAX_OBS];
};

// ------------------------------------------------------------ drift forms
// Every built-in drift is D(a, xi) = (g |a|^2 + u + i c xi_1) a + s xi with the
// coefficients of the form (the others zero): Kubo.m:17 (i c w a), and the
// GP/NLS/CGL family (g |a|^2 a + u a + s (w1 + i w2)).  One branch-free
// expression: the stages run it for every site without control flow.
template <class R>
struct Drift {
  cpx<R> g, u;
  R s, c;
  __device__ __forceinline__ void init(const PathArgs& A) {
    const bool gp = A.deriv == XSH_DERIV_GP, kubo = A.deriv == XSH_DERIV_KUBO;
    g = gp ? cpx<R>{(R)A.dp[0], (R)A.dp[1]} : cpx<R>{R(0), R(0)};
    u = gp ? cpx<R>{(R)A.dp[2], (R)A.dp[3]} : cpx<R>{R(0), R(0)};
    s = gp ? (R)A.dp[4] : R(0);
    c = kubo ? (R)A.dp[0] : R(0);
  }
  __device__ __forceinline__ cpx<R> operator()(cpx<R> a, cpx<R> xi) const {
    const R n = norm2(a);
    const cpx<R> k = {g.x * n + u.x, g.y * n + u.y + c * xi.x};
    const cpx<R> r = k * a;
    return cpx<R>{r.x + s * xi.x, r.y + s * xi.y};
  }
};

// a complex value with a NaN or infinite part
template <class R>
__device__ __forceinline__ bool not_finite(cpx<R> a) {
  return !(__builtin_isfinite(a.x) && __builtin_isfinite(a.y));
}

// block-wide sum of two doubles; every thread gets the result
template <int NT>
__device__ __forceinline__ void block_sum2(double& a, double& b, double* red) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    a += __shfl_down(a, o, 64);
    b += __shfl_down(b, o, 64);
  }
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  __syncthreads();
  if (l == 0) {
    red[2 * w] = a;
    red[2 * w + 1] = b;
  }
  __syncthreads();
  a = 0.0;
  b = 0.0;
#pragma unroll
  for (int i = 0; i < NT / 64; ++i) {
    a += red[2 * i];
    b += red[2 * i + 1];
  }
}

__device__ __forceinline__ void atomic_add_c(double* p, double re, double im) {
  unsafeAtomicAdd(p, re);
  unsafeAtomicAdd(p + 1, im);
}

// ------------------------------------------------------- lattice engine
// Trajectory planes of NX x NY sites, TPB planes per workgroup of NT
// threads.  Field registers: chunk c = tid + NT*i (i < CPT) holds VEC=2
// consecutive sites of one row.  LDS: plane p, row y, column x at
// p*PS + y*RS + x with one pad element per row.
// Thread index through an empty asm: address arithmetic derived from it is
// recomputed where it is used instead of being hoisted out of the time loop
// and held in VGPRs for the whole path (that hoisting caused register spills).
// Element i of a read-only table in device memory (twiddles, propagator
// factors), reached through a pointer of the argument block that the compiler
// sees as generic: reading it through the global address space makes the load
// a global load (vector-memory counter only, 32-bit offsets) instead of a flat
// load, which also counts against the LDS counter and so delays the LDS waits
// of the transforms around it.
// (The host pass of the compiler, which also analyses device templates, keeps
// the plain read.)
template <class R>
__device__ __forceinline__ cpx<R> gld(const cpx<R>* p, int i) {
#if defined(__HIP_DEVICE_COMPILE__)
  const __attribute__((address_space(1))) R* q = (const __attribute__((address_space(1))) R*)p;
  return cpx<R>{q[2 * i], q[2 * i + 1]};
#else
  return p[i];
#endif
}

__device__ __forceinline__ int opaque_tid() {
  int t = threadIdx.x;
  asm volatile("" : "+v"(t));
  return t;
}

// N = N1 * N2, both <= 32, DIF radix order (N1 first); R1 > 0 chooses N1
template <int N, int R1 = 0>
struct Split {
  static constexpr int N1 = R1 > 0 ? R1
                          : N <= 4 ? (N >= 2 ? 2 : 1)
                          : N == 8 ? 4 : N == 16 ? 4 : N == 32 ? 4 : N == 64 ? 8
                          : N == 128 ? 16 : N == 256 ? 16 : N == 512 ? 16 : 32;
  static constexpr int N2 = N / N1;
  static_assert(N1 * N2 == N && N1 <= 32 && N2 <= 32, "radix split must factor N into two factors <= 32");
};

// RX > 0: first radix of the x transforms (an autotuning choice)
template <class R, int NX, int NY, int NT, int TPB, int RX = 0>
struct Lattice {
  static constexpr int NS = NX * NY;
  static constexpr int VEC = 2;
  static constexpr int NPTS = TPB * NS;
  static constexpr int CPT = NPTS / (VEC * NT);
  static constexpr int PPT = CPT * VEC;
  static constexpr int RS = NX + 1;
  static constexpr int PS = NY * RS;
  static constexpr int LDS = TPB * PS;
  static_assert(NPTS % (VEC * NT) == 0, "points per workgroup must tile the threads");
  static_assert(NX % 2 == 0, "rows hold an even number of sites");
  using C = cpx<R>;

  // point index of register slot (i, v)
  __device__ static __forceinline__ int point(int i, int v) { return (opaque_tid() + NT * i) * VEC + v; }
  __device__ static __forceinline__ int lds_of(int pt) {
    const int p = pt / NS, s = pt - p * NS;
    const int y = s / NX, x = s - y * NX;
    return p * PS + y * RS + x;
  }

  __device__ static __forceinline__ void store(C* L, const C (&f)[PPT]) {
#pragma unroll
    for (int i = 0; i < CPT; ++i) {
      const int a = lds_of(point(i, 0));
#pragma unroll
      for (int v = 0; v < VEC; ++v) L[a + v] = f[i * VEC + v];
    }
  }
  __device__ static __forceinline__ void load(const C* L, C (&f)[PPT]) {
#pragma unroll
    for (int i = 0; i < CPT; ++i) {
      const int a = lds_of(point(i, 0));
#pragma unroll
      for (int v = 0; v < VEC; ++v) f[i * VEC + v] = L[a + v];
    }
  }

  // line geometry along dim 0 (x, rows) or 1 (y, columns)
  template <int DIM>
  struct Line {
    static constexpr int N = DIM == 0 ? NX : NY;
    static constexpr int N1 = Split<N, DIM == 0 ? RX : 0>::N1, N2 = Split<N, DIM == 0 ? RX : 0>::N2;
    static constexpr int NL = DIM == 0 ? TPB * NY : TPB * NX;
    static constexpr int ES = DIM == 0 ? 1 : RS;
    __device__ static __forceinline__ int base(int l) {
      if (DIM == 0) {
        const int p = l / NY, y = l - p * NY;
        return p * PS + y * RS;
      } else {
        const int p = l / NX, x = l - p * NX;
        return p * PS + x;
      }
    }
  };

  // decimation-in-frequency stage 1: natural -> (k1, n2) with twiddles
  template <int DIM>
  __device__ static __forceinline__ void dif1(C* L, const C* tw) {
    using Ln = Line<DIM>;
    constexpr int N1 = Ln::N1, N2 = Ln::N2;
#pragma unroll 1
    for (int it = opaque_tid(); it < Ln::NL * N2; it += NT) {
      const int l = it % Ln::NL, n2 = it / Ln::NL;
      const int b = Ln::base(l);
      C v[N1];
#pragma unroll
      for (int n1 = 0; n1 < N1; ++n1) v[n1] = L[b + (n1 * N2 + n2) * Ln::ES];
      dft<R, N1, -1>(v);
#pragma unroll
      for (int k1 = 1; k1 < N1; ++k1) v[k1] = v[k1] * gld(tw, n2 * k1);
#pragma unroll
      for (int k1 = 0; k1 < N1; ++k1) L[b + (k1 * N2 + n2) * Ln::ES] = v[k1];
    }
  }
  // forward stage 2 (no middle op): block k1 -> X[k1 + N1 k2] at k1*N2 + k2
  template <int DIM>
  __device__ static __forceinline__ void dif2(C* L) {
    using Ln = Line<DIM>;
    constexpr int N1 = Ln::N1, N2 = Ln::N2;
#pragma unroll 1
    for (int it = opaque_tid(); it < Ln::NL * N1; it += NT) {
      const int l = it % Ln::NL, k1 = it / Ln::NL;
      const int b = Ln::base(l) + k1 * N2 * Ln::ES;
      C v[N2];
#pragma unroll
      for (int n = 0; n < N2; ++n) v[n] = L[b + n * Ln::ES];
      dft<R, N2, -1>(v);
#pragma unroll
      for (int n = 0; n < N2; ++n) L[b + n * Ln::ES] = v[n];
    }
  }
  // inverse stage A (no middle op): inverse DFT-N2 of block k1, twiddle conj
  template <int DIM>
  __device__ static __forceinline__ void dita(C* L, const C* tw) {
    using Ln = Line<DIM>;
    constexpr int N1 = Ln::N1, N2 = Ln::N2;
#pragma unroll 1
    for (int it = opaque_tid(); it < Ln::NL * N1; it += NT) {
      const int l = it % Ln::NL, k1 = it / Ln::NL;
      const int b = Ln::base(l) + k1 * N2 * Ln::ES;
      C v[N2];
#pragma unroll
      for (int n = 0; n < N2; ++n) v[n] = L[b + n * Ln::ES];
      dft<R, N2, 1>(v);
#pragma unroll
      for (int n2 = 1; n2 < N2; ++n2) v[n2] = v[n2] * conjc(gld(tw, n2 * k1));
#pragma unroll
      for (int n = 0; n < N2; ++n) L[b + n * Ln::ES] = v[n];
    }
  }
  // forward stage 2 + op(X, l, k) + inverse stage A in registers
  template <int DIM, class Op>
  __device__ static __forceinline__ void mid(C* L, const C* tw, const Op& op) {
    using Ln = Line<DIM>;
    constexpr int N1 = Ln::N1, N2 = Ln::N2;
#pragma unroll 1
    for (int it = opaque_tid(); it < Ln::NL * N1; it += NT) {
      const int l = it % Ln::NL, k1 = it / Ln::NL;
      const int b = Ln::base(l) + k1 * N2 * Ln::ES;
      C v[N2];
#pragma unroll
      for (int n = 0; n < N2; ++n) v[n] = L[b + n * Ln::ES];
      dft<R, N2, -1>(v);
#pragma unroll
      for (int k2 = 0; k2 < N2; ++k2) v[k2] = op(v[k2], l, k1 + N1 * k2);
      dft<R, N2, 1>(v);
#pragma unroll
      for (int n2 = 1; n2 < N2; ++n2) v[n2] = v[n2] * conjc(gld(tw, n2 * k1));
#pragma unroll
      for (int n = 0; n < N2; ++n) L[b + n * Ln::ES] = v[n];
    }
  }
  // decimation-in-time stage B: (k1, n2) -> natural order
  template <int DIM>
  __device__ static __forceinline__ void ditb(C* L) {
    using Ln = Line<DIM>;
    constexpr int N1 = Ln::N1, N2 = Ln::N2;
#pragma unroll 1
    for (int it = opaque_tid(); it < Ln::NL * N2; it += NT) {
      const int l = it % Ln::NL, n2 = it / Ln::NL;
      const int b = Ln::base(l);
      C v[N1];
#pragma unroll
      for (int k1 = 0; k1 < N1; ++k1) v[k1] = L[b + (k1 * N2 + n2) * Ln::ES];
      dft<R, N1, 1>(v);
#pragma unroll
      for (int n1 = 0; n1 < N1; ++n1) L[b + (n1 * N2 + n2) * Ln::ES] = v[n1];
    }
  }

  // Dense transform along DIM, in place in LDS: out[k] = sum_n T[k*N + n] in[n]
  // with T real (the DCT/DST matrices of the non-periodic boundaries) or
  // complex (the DFT of a periodic dimension on a mixed lattice).  Work items
  // are (line, block of KB outputs); lanes run over lines, so the block - and
  // the row of T - is wave-uniform.  Results are buffered in registers until
  // every thread has read its inputs, then written back through op(out, line, k).
  template <int DIM, bool CPLX, class Op>
  __device__ static __forceinline__ void mat(C* L, const R* __restrict__ T, const Op& op) {
    using Ln = Line<DIM>;
    constexpr int N = Ln::N, NL = Ln::NL, KB = 8;
    // lanes run over lines: with whole waves of lines the block index (and the
    // row of T) is wave-uniform; with 32-line blocks it differs between the two
    // half-waves and T is read per lane
    static_assert(NL % 32 == 0, "matrix transforms need half waves of lines");
    static_assert(N % KB == 0 && N % 8 == 0, "line length must be a multiple of 8");
    constexpr int ITEMS = NL * (N / KB);
    static_assert(ITEMS % NT == 0, "items must tile the threads");
    constexpr int IT = ITEMS / NT;
    constexpr int TW = CPLX ? 2 : 1;
    C out[IT * KB];
    static_for<0, IT>([&](auto I) __attribute__((always_inline)) {
      constexpr int it = decltype(I)::value;
      const int item = opaque_tid() + it * NT;
      const int l = item % NL;
      const int kb = NL % 64 == 0 ? __builtin_amdgcn_readfirstlane(item / NL) : item / NL;
      const int b = Ln::base(l);
      const R* __restrict__ Trow = T + (long long)kb * KB * N * TW;
      C acc[KB];
      static_for<0, KB>([&](auto J) __attribute__((always_inline)) {
        acc[decltype(J)::value] = C{R(0), R(0)};
      });
#pragma unroll 1
      for (int n0 = 0; n0 < N; n0 += 8) {
        C x[8];
        static_for<0, 8>([&](auto Q) __attribute__((always_inline)) {
          constexpr int q = decltype(Q)::value;
          x[q] = L[b + (n0 + q) * Ln::ES];
        });
        static_for<0, KB>([&](auto J) __attribute__((always_inline)) {
          constexpr int j = decltype(J)::value;
          static_for<0, 8>([&](auto Q) __attribute__((always_inline)) {
            constexpr int q = decltype(Q)::value;
            if constexpr (CPLX) {
              const C t = {Trow[2 * (j * N + n0 + q)], Trow[2 * (j * N + n0 + q) + 1]};
              acc[j] = acc[j] + t * x[q];
            } else {
              acc[j] = acc[j] + Trow[j * N + n0 + q] * x[q];
            }
          });
        });
      }
      static_for<0, KB>([&](auto J) __attribute__((always_inline)) {
        constexpr int j = decltype(J)::value;
        out[it * KB + j] = op(acc[j], l, kb * KB + j);
      });
    });
    __syncthreads();
    static_for<0, IT>([&](auto I) __attribute__((always_inline)) {
      constexpr int it = decltype(I)::value;
      const int item = opaque_tid() + it * NT;
      const int l = item % NL;
      const int kb = item / NL;
      const int b = Ln::base(l);
      static_for<0, KB>([&](auto J) __attribute__((always_inline)) {
        constexpr int j = decltype(J)::value;
        L[b + (kb * KB + j) * Ln::ES] = out[it * KB + j];
      });
    });
  }

  // FP32 real transforms of the lattice block on the matrix cores: along DIM,
  // OUT[k][line] = sum_n T[k][n] IN[n][line] is a GEMM with the lines' real and
  // imaginary parts as columns.  v_mfma_f32_32x32x2f32 (lane l supplies
  // A[l%32][l/32] and B[l/32][l%32] and receives D[8(r/4) + 4(l/32) + r%4][l%32]
  // in register r): each wave owns PW pairs of 32x32 output tiles (real and
  // imaginary columns of the same 32 lines, one 32-row block of k), reads T in
  // the packed operand order from memory and the plane from LDS, and writes
  // op(X, line, k) back after every wave has read its inputs.
  template <int DIM>
  struct MfmaPlan {
    static constexpr int N = Line<DIM>::N, NL = Line<DIM>::NL, NW = NT / 64;
    static constexpr int KB = N / 32, LB = NL / 32, P = KB * LB;
    static constexpr int PW = P >= NW && P % NW == 0 ? P / NW : 0;
    // 2D planes (lines = rows or columns of one trajectory) and 1D blocks
    // (lines = the TPB trajectories of the workgroup) alike
    static constexpr bool ok = std::is_same<R, float>::value && (NY > 1 ? TPB == 1 : DIM == 0) &&
                               N % 32 == 0 && NL % 32 == 0 && PW > 0 && LB % PW == 0;
  };
  template <int DIM, bool SYM, class Op>
  __device__ static __forceinline__ void mat_mfma(C* L, const float* __restrict__ Tp, int sym, const Op& op) {
    using Pl = MfmaPlan<DIM>;
    constexpr int N = Pl::N, PW = Pl::PW, LB = Pl::LB, KB = Pl::KB;
    // SYM: half the GEMM depth; row tiles [0, KB/2) are the even k, the rest
    // the odd k, and the input pairs (n, N-1-n) fold into x_n + sgn x_{N-1-n}
    constexpr int K = SYM ? N / 2 : N;
    static_assert(!SYM || KB % 2 == 0, "the even/odd split needs N % 64 == 0");
    typedef float v16f __attribute__((ext_vector_type(16)));
    const int lane = threadIdx.x & 63;
    const int p0 = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6)) * PW;
    const int kb = p0 / LB;
    const bool odd = SYM && kb >= KB / 2;
    const float sgn = odd == (sym == 1) ? -1.f : 1.f;
    v16f acc[2 * PW];
    static_for<0, 2 * PW>([&](auto T) __attribute__((always_inline)) {
      acc[decltype(T)::value] = v16f{0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    });
    const float* __restrict__ Ta = Tp + (size_t)kb * (K / 2) * 64 + lane;
#pragma unroll 4
    for (int kk = 0; kk < K; kk += 2) {
      const float a = Ta[(kk / 2) * 64];
      const int n = kk + (lane >> 5);
      static_for<0, PW>([&](auto T) __attribute__((always_inline)) {
        constexpr int t = decltype(T)::value;
        const int ln = ((p0 + t) % LB) * 32 + (lane & 31);
        C x = L[DIM == 0 ? ln * RS + n : n * RS + ln];
        if constexpr (SYM) {
          const C y = L[DIM == 0 ? ln * RS + (N - 1 - n) : (N - 1 - n) * RS + ln];
          x = C{fmad(sgn, y.x, x.x), fmad(sgn, y.y, x.y)};
        }
        acc[2 * t] = __builtin_amdgcn_mfma_f32_32x32x2f32(a, x.x, acc[2 * t], 0, 0, 0);
        acc[2 * t + 1] = __builtin_amdgcn_mfma_f32_32x32x2f32(a, x.y, acc[2 * t + 1], 0, 0, 0);
      });
    }
    __syncthreads();
    static_for<0, PW>([&](auto T) __attribute__((always_inline)) {
      constexpr int t = decltype(T)::value;
      const int ln = ((p0 + t) % LB) * 32 + (lane & 31);
      static_for<0, 16>([&](auto Rg) __attribute__((always_inline)) {
        constexpr int r = decltype(Rg)::value;
        const int kr = kb * 32 + 8 * (r / 4) + 4 * (lane >> 5) + (r % 4);
        const int k = SYM ? (kr < N / 2 ? 2 * kr : 2 * (kr - N / 2) + 1) : kr;
        const C X = op(C{acc[2 * t][r], acc[2 * t + 1][r]}, ln, k);
        L[DIM == 0 ? ln * RS + k : k * RS + ln] = X;
      });
    });
  }

  // FP64 real transforms on the matrix cores: the same GEMM on
  // v_mfma_f64_16x16x4f64 (lane l supplies A[l%16][l/16] and B[l/16][l%16] and
  // receives D[4 r + l/16][l%16] in register r), T in the packed operand
  // order Tp[kb][kk/4][lane] = T[16 kb + lane%16][kk + lane/16] (one coalesced
  // 512 B load per K step), each wave owning PW pairs of 16x16 tiles (real and
  // imaginary columns of the same 16 lines, one 16-row block of k).
  template <int DIM>
  struct MfmaPlan64 {
    static constexpr int N = Line<DIM>::N, NL = Line<DIM>::NL, NW = NT / 64;
    static constexpr int KB = N / 16, LB = NL / 16, P = KB * LB;
    static constexpr int PW = P >= NW && P % NW == 0 ? P / NW : 0;
    static constexpr bool ok = std::is_same<R, double>::value && (NY > 1 ? TPB == 1 : DIM == 0) &&
                               N % 16 == 0 && NL % 16 == 0 && PW > 0 && LB % PW == 0;
  };
  template <int DIM, bool SYM, class Op>
  __device__ static __forceinline__ void mat_mfma64(C* L, const double* __restrict__ Tp, int sym, const Op& op) {
    using Pl = MfmaPlan64<DIM>;
    constexpr int N = Pl::N, PW = Pl::PW, LB = Pl::LB, KB = Pl::KB;
    constexpr int K = SYM ? N / 2 : N;  // the even/odd split as in mat_mfma
    static_assert(!SYM || KB % 2 == 0, "the even/odd split needs N % 32 == 0");
    typedef double v4d __attribute__((ext_vector_type(4)));
    const int lane = threadIdx.x & 63;
    const int p0 = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6)) * PW;
    const int kb = p0 / LB;
    const bool odd = SYM && kb >= KB / 2;
    const double sgn = odd == (sym == 1) ? -1.0 : 1.0;
    v4d acc[2 * PW];
    static_for<0, 2 * PW>([&](auto T) __attribute__((always_inline)) {
      acc[decltype(T)::value] = v4d{0.0, 0.0, 0.0, 0.0};
    });
    const double* __restrict__ Ta = Tp + (size_t)kb * (K / 4) * 64 + lane;
#pragma unroll 4
    for (int kk = 0; kk < K; kk += 4) {
      const double a = Ta[(kk / 4) * 64];
      const int n = kk + (lane >> 4);
      static_for<0, PW>([&](auto T) __attribute__((always_inline)) {
        constexpr int t = decltype(T)::value;
        const int ln = ((p0 + t) % LB) * 16 + (lane & 15);
        C x = L[DIM == 0 ? ln * RS + n : n * RS + ln];
        if constexpr (SYM) {
          const C y = L[DIM == 0 ? ln * RS + (N - 1 - n) : (N - 1 - n) * RS + ln];
          x = C{fmad(sgn, y.x, x.x), fmad(sgn, y.y, x.y)};
        }
        acc[2 * t] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, x.x, acc[2 * t], 0, 0, 0);
        acc[2 * t + 1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, x.y, acc[2 * t + 1], 0, 0, 0);
      });
    }
    __syncthreads();
    static_for<0, PW>([&](auto T) __attribute__((always_inline)) {
      constexpr int t = decltype(T)::value;
      const int ln = ((p0 + t) % LB) * 16 + (lane & 15);
      static_for<0, 4>([&](auto Rg) __attribute__((always_inline)) {
        constexpr int r = decltype(Rg)::value;
        const int kr = kb * 16 + 4 * r + (lane >> 4);  // D[4 r + lane / 16][lane % 16]
        const int k = SYM ? (kr < N / 2 ? 2 * kr : 2 * (kr - N / 2) + 1) : kr;
        const C X = op(C{acc[2 * t][r], acc[2 * t + 1][r]}, ln, k);
        L[DIM == 0 ? ln * RS + k : k * RS + ln] = X;
      });
    });
  }

  template <int DIM, class Op>
  __device__ static __forceinline__ void mat_any(C* L, const void* T, const void* Tp, bool cplx, const Op& op,
                                                 int sym = 0) {
    if constexpr (MfmaPlan<DIM>::ok) {
      if (!cplx && Tp) {
        if constexpr (MfmaPlan<DIM>::KB % 2 == 0) {
          if (sym) {
            mat_mfma<DIM, true>(L, static_cast<const float*>(Tp), sym, op);
            return;
          }
        }
        mat_mfma<DIM, false>(L, static_cast<const float*>(Tp), 0, op);
        return;
      }
    }
    if constexpr (MfmaPlan64<DIM>::ok) {
      if (!cplx && Tp) {
        if constexpr (MfmaPlan64<DIM>::KB % 2 == 0) {
          if (sym) {
            mat_mfma64<DIM, true>(L, static_cast<const double*>(Tp), sym, op);
            return;
          }
        }
        mat_mfma64<DIM, false>(L, static_cast<const double*>(Tp), 0, op);
        return;
      }
    }
    if (cplx)
      mat<DIM, true>(L, static_cast<const R*>(T), op);
    else
      mat<DIM, false>(L, static_cast<const R*>(T), op);
  }

  // a -> T^-1 op(T a) with dense per-dimension transforms (boundary types of
  // xtransform.m:33-70); op receives natural-order (kx, ky).
  template <class Op>
  __device__ static __forceinline__ void spectral_mat(C* L, C (&f)[PPT], const void* const* Tm,
                                                      const void* const* Tp, int cplx_x, int cplx_y,
                                                      const Op& op, const int* tsym) {
    auto id = [](C X, int, int) __attribute__((always_inline)) { return X; };
    __syncthreads();
    store(L, f);
    __syncthreads();
    if constexpr (NY == 1) {
      mat_any<0>(L, Tm[0], Tp[0], cplx_x,
                 [&](C X, int l, int k) __attribute__((always_inline)) { return op(X, k, 0); }, tsym[0]);
      __syncthreads();
      mat_any<0>(L, Tm[2], Tp[2], cplx_x, id, tsym[2]);
    } else {
      mat_any<0>(L, Tm[0], Tp[0], cplx_x, id, tsym[0]);
      __syncthreads();
      mat_any<1>(L, Tm[1], Tp[1], cplx_y,
                 [&](C X, int l, int k) __attribute__((always_inline)) { return op(X, l % NX, k); }, tsym[1]);
      __syncthreads();
      mat_any<1>(L, Tm[3], Tp[3], cplx_y, id, tsym[3]);
      __syncthreads();
      mat_any<0>(L, Tm[2], Tp[2], cplx_x, id, tsym[2]);
    }
    __syncthreads();
    load(L, f);
  }

  // frequency index held at position pos along x after the forward stages
  __device__ static __forceinline__ int kx_of_pos(int pos) {
    constexpr int N1 = Split<NX, RX>::N1, N2 = Split<NX, RX>::N2;
    return pos / N2 + N1 * (pos % N2);
  }

  // frequency index held at position pos along y after the forward stages
  __device__ static __forceinline__ int ky_of_pos(int pos) {
    constexpr int N1 = Split<NY>::N1, N2 = Split<NY>::N2;
    return pos / N2 + N1 * (pos % N2);
  }
  // FFT-order k site (kx + NX ky) whose value sits at lattice site s after the
  // forward stages
  __device__ static __forceinline__ int k_site(int s) {
    const int y = s / NX, x = s - y * NX;
    return kx_of_pos(x) + (NY > 1 ? NX * ky_of_pos(y) : 0);
  }

  // L holds a k-space field in the forward stages' frequency order (lattice
  // position s holds k_site(s)): L -> its inverse DFT sum_k L_k exp(+2 pi i k.x / N),
  // no 1/N.  Called with L written and synchronised; returns synchronised.
  __device__ static __forceinline__ void inverse_lds(C* L, const C* twx, const C* twy) {
    if constexpr (NY > 1) {
      dita<1>(L, twy);
      __syncthreads();
      ditb<1>(L);
      __syncthreads();
    }
    dita<0>(L, twx);
    __syncthreads();
    ditb<0>(L);
    __syncthreads();
  }

  // a -> T^-1 op(T a): periodic transform of every plane held in f
  template <class Op>
  __device__ static __forceinline__ void spectral(C* L, C (&f)[PPT], const C* twx, const C* twy,
                                                  const Op& op) {
    __syncthreads();
    store(L, f);
    __syncthreads();
    if constexpr (NY == 1) {
      dif1<0>(L, twx);
      __syncthreads();
      mid<0>(L, twx, [&](C X, int l, int k) __attribute__((always_inline)) { return op(X, k, 0, l); });
      __syncthreads();
      ditb<0>(L);
    } else {
      dif1<0>(L, twx);
      __syncthreads();
      dif2<0>(L);
      __syncthreads();
      dif1<1>(L, twy);
      __syncthreads();
      mid<1>(L, twy, [&](C X, int l, int k) __attribute__((always_inline)) {
        return op(X, kx_of_pos(l % NX), k, l / NX);
      });
      __syncthreads();
      ditb<1>(L);
      __syncthreads();
      dita<0>(L, twx);
      __syncthreads();
      ditb<0>(L);
    }
    __syncthreads();
    load(L, f);
  }
};

// propagator op: X * Px[kx] (1D) or X * Px[kx] * Py[ky] (py set: 2D)
template <class R>
struct PropOp {
  const cpx<R>* px;
  const cpx<R>* py;
  // (kx, ky): frequency indices; p: trajectory within the workgroup
  __device__ __forceinline__ cpx<R> operator()(cpx<R> X, int kx, int ky, int p = 0) const {
    cpx<R> r = X * gld(px, kx);
    return py ? r * gld(py, ky) : r;
  }
};
// the same on a 2D lattice: no test of py per element
template <class R>
struct PropOp2 {
  const cpx<R>* px;
  const cpx<R>* py;
  __device__ __forceinline__ cpx<R> operator()(cpx<R> X, int kx, int ky, int p = 0) const {
    return (X * gld(px, kx)) * gld(py, ky);
  }
};
// observable ops, sharing one transform instance:
//   CORR : |X|^2 * scale (equal-time correlation, the observe form 'corr')
//   MODES: record selected modes X(kx,ky) * scale as features, pass X through
//   LOWK : |X|^2 * scale - 1/2 on the selected modes other than k = 0, else 0;
//          the unnormalised inverse transform then holds their sum at x = 0
//   ABS2K: add |X|^2 * kscale to the k-space site kx + nx ky of acc (trajectories
//          p of the workgroup with traj0 + p < E), pass X through
template <class R>
struct ObsOp {
  int mode;  // 0 CORR, 1 MODES, 2 LOWK, 3 ABS2K
  R scale;
  int nx;
  const int* fmap;
  float* feat;
  double* acc;
  double kscale;
  int traj0, E;
  __device__ __forceinline__ cpx<R> operator()(cpx<R> X, int kx, int ky, int p = 0) const {
    if (mode == 0) return cpx<R>{norm2(X) * scale, R(0)};
    if (mode == 3) {
      if (traj0 + p < E) unsafeAtomicAdd(acc + 2 * (long long)(kx + nx * ky), (double)norm2(X) * kscale);
      return X;
    }
    const int fi = fmap[kx + nx * ky];
    if (mode == 2) return cpx<R>{(fi >= 0 && (kx | ky)) ? norm2(X) * scale - R(0.5) : R(0), R(0)};
    if (fi >= 0) {
      feat[2 * fi] = (float)(X.x * scale);
      feat[2 * fi + 1] = (float)(X.y * scale);
    }
    return X;
  }
};

}  // namespace xsh
