// Streaming whole-path kernel for lattices whose state does not fit on chip
// (FP64 128x128: one field is 256 KB, the register file of a CU 512 KB and
// its LDS 160 KB).  The fields live in a per-workgroup workspace in device
// memory; the 2D FFT runs through LDS half a plane at a time:
//
//   x pass   rows of each half-plane (64 x 128), forward DIF in LDS
//   y pass   radix-2 split of every column across the halves in registers
//            (u = a[y] + a[y+H], w = (a[y] - a[y+H]) W^y), then the even and
//            odd half-length column transforms in LDS, each fused with the
//            propagator multiply and its inverse, then the radix-2 merge
//   x pass   inverse DIT of each half-plane
//
// Every pointwise stage of the method (drift evaluations, RK combinations)
// streams its sites between the workspace and registers with coalesced
// 2-site accesses; noise is regenerated from the counter stream at each
// stage, exactly as in the register-resident kernel, so both kernels follow
// the same arithmetic (xpath.m:29-86, RK4.m:7-43, MP.m:10-28).
#pragma once
#include "xsh_kernels.hpp"

namespace xsh {

template <class R, int NX, int NY, int NT>
struct Stream {
  static constexpr int NS = NX * NY;
  static constexpr int H = NY / 2;            // rows per half-plane
  static constexpr int NSH = NX * H;          // sites per half-plane
  static constexpr int PAIRS = NS / (2 * NT); // 2-site chunks per thread (plane)
  static constexpr int HPAIRS = NSH / (2 * NT);
  static_assert(NY % 2 == 0 && NSH % (2 * NT) == 0, "half-planes must tile the threads");
  using Lh = Lattice<R, NX, H, NT, 1, 0>;     // LDS geometry of a half-plane (rows)
  using Lv = Lattice<R, NX / 2, NY, NT, 1, 0>; // LDS geometry of a column half (dense y passes)
  using C = cpx<R>;
  static constexpr int RS = Lh::RS;
  static constexpr int LDS = Lh::LDS > Lv::LDS ? Lh::LDS : Lv::LDS;

  // first site of 2-site chunk i of a region of the plane (consecutive threads,
  // consecutive chunks: coalesced)
  __device__ static __forceinline__ int site(int i) { return (opaque_tid() + NT * i) * 2; }

  __device__ static __forceinline__ void to_lds(C* L, const C* F) {
#pragma unroll
    for (int i = 0; i < HPAIRS; ++i) {
      const int s = site(i), y = s / NX, x = s - y * NX;
      const C a = F[s], b = F[s + 1];
      L[y * RS + x] = a;
      L[y * RS + x + 1] = b;
    }
  }
  __device__ static __forceinline__ void from_lds(const C* L, C* F) {
#pragma unroll
    for (int i = 0; i < HPAIRS; ++i) {
      const int s = site(i), y = s / NX, x = s - y * NX;
      F[s] = L[y * RS + x];
      F[s + 1] = L[y * RS + x + 1];
    }
  }

  // columns [c NX/2, (c+1) NX/2) of the plane, every row, to and from LDS
  __device__ static __forceinline__ void cols_to_lds(C* L, const C* F, int c) {
#pragma unroll 4
    for (int i = 0; i < HPAIRS; ++i) {
      const int sl = site(i), y = sl / (NX / 2), xl = sl - y * (NX / 2);
      const int g = y * NX + c * (NX / 2) + xl;
      L[y * Lv::RS + xl] = F[g];
      L[y * Lv::RS + xl + 1] = F[g + 1];
    }
  }
  __device__ static __forceinline__ void cols_from_lds(const C* L, C* F, int c) {
#pragma unroll 4
    for (int i = 0; i < HPAIRS; ++i) {
      const int sl = site(i), y = sl / (NX / 2), xl = sl - y * (NX / 2);
      const int g = y * NX + c * (NX / 2) + xl;
      F[g] = L[y * Lv::RS + xl];
      F[g + 1] = L[y * Lv::RS + xl + 1];
    }
  }

  // Dense-transform lattices (non-periodic boundaries, xtransform.m:33-70):
  // F -> Tx^-1 Ty^-1 op(Ty Tx F) in place, the x passes on row halves and the y
  // passes on column halves in LDS (each half fits), matrices Tm[0..3] =
  // forward x, forward y, inverse x, inverse y, complex for periodic dimensions
  // of a mixed lattice.  op(X, kx, ky) in the transforms' natural k order.
  template <class Op>
  __device__ static __forceinline__ void propagate_mat(C* L, C* F, const void* const* Tm, const void* const* Tp,
                                                        int cplx_x, int cplx_y, const Op& op, const int* tsym) {
    auto id = [](C X, int, int) __attribute__((always_inline)) { return X; };
#pragma unroll 1
    for (int h = 0; h < 2; ++h) {
      __syncthreads();
      to_lds(L, F + h * NSH);
      __syncthreads();
      Lh::template mat_any<0>(L, Tm[0], Tp[0], cplx_x, id, tsym[0]);
      __syncthreads();
      from_lds(L, F + h * NSH);
    }
#pragma unroll 1
    for (int c = 0; c < 2; ++c) {
      __syncthreads();
      cols_to_lds(L, F, c);
      __syncthreads();
      Lv::template mat_any<1>(L, Tm[1], Tp[1], cplx_y, [&](C X, int l, int k) __attribute__((always_inline)) {
        return op(X, c * (NX / 2) + l, k);
      }, tsym[1]);
      __syncthreads();
      Lv::template mat_any<1>(L, Tm[3], Tp[3], cplx_y, id, tsym[3]);
      __syncthreads();
      cols_from_lds(L, F, c);
    }
#pragma unroll 1
    for (int h = 0; h < 2; ++h) {
      __syncthreads();
      to_lds(L, F + h * NSH);
      __syncthreads();
      Lh::template mat_any<0>(L, Tm[2], Tp[2], cplx_x, id, tsym[2]);
      __syncthreads();
      from_lds(L, F + h * NSH);
    }
    __syncthreads();
  }

  // F -> T^-1 op(T F) in place (F one plane in device memory).  twx: W_NX^m,
  // twy: W_NY^m (m < NY), twh: W_H^m (m < H).  op(X, kx, ky) as in spectral().
  // Five plane transfers: the second half-plane's forward x result stays in
  // LDS for the y split, and the first half-plane's merge result goes straight
  // into LDS for its inverse x pass (every LDS position is read and rewritten
  // by the one thread that owns its site, between barriers).
  template <class Op>
  __device__ static __forceinline__ void propagate(C* L, C* F, const C* twx, const C* twy, const C* twh,
                                                    const Op& op) {
    // forward x on each half-plane; the second stays in LDS
#pragma unroll 1
    for (int h = 0; h < 2; ++h) {
      __syncthreads();
      to_lds(L, F + h * NSH);
      __syncthreads();
      Lh::template dif1<0>(L, twx);
      __syncthreads();
      Lh::template dif2<0>(L);
      __syncthreads();
      if (h == 0) from_lds(L, F);
    }
    // y: radix-2 split across the halves (first half from F, second from LDS,
    // in place); odd part parked in registers
    C w[2 * HPAIRS];
#pragma unroll
    for (int i = 0; i < HPAIRS; ++i) {
      const int s = site(i), y = s / NX, x = s - y * NX;
      const C t = twy[y];
#pragma unroll
      for (int v = 0; v < 2; ++v) {
        const C a0 = F[s + v], a1 = L[y * RS + x + v];
        L[y * RS + x + v] = a0 + a1;
        w[2 * i + v] = (a0 - a1) * t;
      }
    }
    half_y(L, twh, op, 0);
    // even result out to the first half of F, odd input in
#pragma unroll
    for (int i = 0; i < HPAIRS; ++i) {
      const int s = site(i), y = s / NX, x = s - y * NX;
#pragma unroll
      for (int v = 0; v < 2; ++v) {
        F[s + v] = L[y * RS + x + v];
        L[y * RS + x + v] = w[2 * i + v];
      }
    }
    half_y(L, twh, op, 1);
    // radix-2 merge: a[y] = e + W^-y o into LDS (in place of o), a[y+H] = e - W^-y o
    // to the second half of F
#pragma unroll
    for (int i = 0; i < HPAIRS; ++i) {
      const int s = site(i), y = s / NX, x = s - y * NX;
      const C t = conjc(twy[y]);
#pragma unroll
      for (int v = 0; v < 2; ++v) {
        const C e = F[s + v], o = L[y * RS + x + v] * t;
        L[y * RS + x + v] = e + o;
        F[NSH + s + v] = e - o;
      }
    }
    // inverse x on each half-plane (the first is in LDS already)
#pragma unroll 1
    for (int h = 0; h < 2; ++h) {
      __syncthreads();
      if (h == 1) {
        to_lds(L, F + NSH);
        __syncthreads();
      }
      Lh::template dita<0>(L, twx);
      __syncthreads();
      Lh::template ditb<0>(L);
      __syncthreads();
      from_lds(L, F + h * NSH);
    }
    __syncthreads();
  }

  // half-length column transforms of the half-plane in LDS (parity 0: ky = 2k,
  // parity 1: ky = 2k + 1), fused with op and the inverse
  template <class Op>
  __device__ static __forceinline__ void half_y(C* L, const C* twh, const Op& op, int parity) {
    __syncthreads();
    Lh::template dif1<1>(L, twh);
    __syncthreads();
    Lh::template mid<1>(L, twh, [&](C X, int l, int k) __attribute__((always_inline)) {
      return op(X, Lh::kx_of_pos(l % NX), 2 * k + parity);
    });
    __syncthreads();
    Lh::template ditb<1>(L);
    __syncthreads();
  }
};

// Work items and instances as in spde_path; the workspace of instance i has
// a region for every workgroup of the grid (A.work, A.work_planes planes each).
// MAT: dense-transform propagator (homogeneous Neumann and periodic dimensions).
template <class R, int NX, int NY, int NT, bool MAT>
__global__ __launch_bounds__(NT) void spde_stream(const PathArgs* __restrict__ args, const int* __restrict__ wstart,
                                                  int ninst, int nwork) {
  using S = Stream<R, NX, NY, NT>;
  using C = cpx<R>;
  constexpr int NS = S::NS, PAIRS = S::PAIRS;
  __shared__ C L[S::LDS];
  __shared__ double red[2 * (NT / 64)];

#pragma unroll 1
  for (int wk = blockIdx.x; wk < nwork; wk += gridDim.x) {
    const int inst = ninst > 1 ? find_instance(wstart, ninst, wk) : 0;
    const PathArgs& A = args[inst];
    C* const F0 = static_cast<C*>(A.work) + (long long)blockIdx.x * A.work_planes * NS;
    C* const F1 = F0 + NS;
    C* const F2 = F1 + NS;  // corr observable scratch (present with corr or abs2k)
    // x-space k-space noise of the step (or the initial k-space randoms): the
    // last plane when the problem has knoises or krandoms
    C* const FK = A.knoises || A.krandoms ? F0 + (long long)(A.work_planes - 1) * NS : nullptr;
    const C* twx = static_cast<const C*>(A.twx);
    const C* twy = static_cast<const C*>(A.twy);
    const C* twh = static_cast<const C*>(A.twh);
    Drift<R> D;
    D.init(A);
    int pidx, nci, j, traj;
    path_identity<R>(A, wk - (ninst > 1 ? wstart[inst] : 0), pidx, nci, j, traj);
    const uint32_t nsp = (uint32_t)(A.seed + A.npar + (j - 1) * A.e3);
    const R hdt = (R)(nci ? A.hdt[1] : A.hdt[0]);
    const PropOp2<R> prop{static_cast<const C*>(nci ? A.px[1] : A.px[0]),
                         static_cast<const C*>(nci ? A.py[1] : A.py[0])};
    double* av = A.av + 2 * (long long)pidx * A.avlen;

    // k-space filtered noise (xnoise.m:39-50) or randoms (xrandom.m:24-29) into
    // FK: Z(k) = F0(k) z0(k) + i F1(k) z1(k) from the counter normals of every k
    // site (coarse: the mean of draws q and q + 1), inverse transformed.  The
    // propagation machinery does it: its middle operation returns Z(kx, ky) in
    // place of the plane's own spectrum (the filter tables hold the 1/N of the
    // inverse), so the plane's former content never reaches the result.
    auto kfield = [&](const C* __restrict__ Fk, int ncomp, uint32_t tag, uint32_t q, bool coarse)
        __attribute__((always_inline)) {
      auto zop = [&](C X, int kx, int ky) __attribute__((always_inline)) -> C {
        const int ks = kx + NX * ky;
        R z0, z1;
        knormals<R>(A, nsp, tag, q, (uint32_t)traj, ks, ncomp, z0, z1);
        if (coarse) {
          R u0, u1;
          knormals<R>(A, nsp, tag, q + 1, (uint32_t)traj, ks, ncomp, u0, u1);
          z0 = R(0.5) * (z0 + u0);
          z1 = R(0.5) * (z1 + u1);
        }
        const C f0 = Fk[ks];
        C Z = C{f0.x * z0, f0.y * z0};
        if (ncomp >= 2) {
          const C f1 = Fk[NS + ks];
          Z = Z + C{-f1.y * z1, f1.x * z1};
        }
        return Z;
      };
      if constexpr (!MAT) S::propagate(L, FK, twx, twy, twh, zop);
    };

    // ---- initial field (xensemble.m:53-65)
    __syncthreads();
    if (FK) {
#pragma unroll 1
      for (int i = 0; i < PAIRS; ++i) {  // finite content for the discarded forward pass
        const int s = S::site(i);
        FK[s] = FK[s + 1] = C{R(0), R(0)};
      }
      if (A.krandoms && !A.a0) kfield(static_cast<const C*>(A.kr_filter), A.krandoms, TAG_KRAND_INIT, 0u, false);
    }
#pragma unroll 1
    for (int i = 0; i < PAIRS; ++i) {
      const int s = S::site(i);
#pragma unroll
      for (int v = 0; v < 2; ++v) {
        const int st = s + v, iy = st / NX, ix = st - iy * NX;
        if (A.a0) {
          const long long g = (long long)traj * NS + st;
          F0[st] = C{(R)A.a0[2 * g], (R)A.a0[2 * g + 1]};
        } else {
          F0[st] = initial_value<R>(A, nsp, (uint32_t)traj, st, ix, iy,
                                    A.krandoms ? FK[st] : C{R(0), R(0)});
        }
      }
    }

    auto noise = [&](int s, uint32_t q, bool coarse, C& w0, C& w1) __attribute__((always_inline)) {
      if (A.knoises) {  // the step's x-space k-noise, transformed once per step
        w0 = FK[s];
        w1 = FK[s + 1];
        return;
      }
      draw_pair<R>(A, nsp, q, (uint32_t)traj, s, w0, w1);
      if (coarse) {
        C u0, u1;
        draw_pair<R>(A, nsp, q + 1, (uint32_t)traj, s, u0, u1);
        w0 = R(0.5) * (w0 + u0);
        w1 = R(0.5) * (w1 + u1);
      }
    };
    // Dirichlet ends and boundary values (xsetbound.m:30-62, xpatch.m:24-62)
    // on the planes in device memory, with the register-resident kernel's
    // rules: time level lev in half fine steps, y values win at corners, the
    // patch pa = sum_d b_lower s0(x_d) + b_upper s1(x_d) around every
    // propagation and the Neumann correction lp after it.  Each thread touches
    // its own sites only.
    auto dir_end = [&](int st, int& dd, int& sd) __attribute__((always_inline)) -> bool {
      const int iy = st / NX, ix = st - iy * NX;
      dd = -1;
      sd = 0;
      if (ix == 0 && A.dir_lo[0]) dd = 0, sd = 0;
      if (ix == NX - 1 && A.dir_hi[0]) dd = 0, sd = 1;
      if (iy == 0 && A.dir_lo[1]) dd = 1, sd = 0;
      if (iy == NY - 1 && A.dir_hi[1]) dd = 1, sd = 1;
      return dd >= 0;
    };
    auto bvalue = [&](int lev, int dd, int sd, int o) __attribute__((always_inline)) -> C {
      const C* bv = static_cast<const C*>(A.bval);
      return bv[(long long)lev * A.bval_level + (dd * 2 + sd) * A.bval_nb + o];
    };
    // the value a Dirichlet site takes at level lev (0 without boundary values)
    auto dir_value = [&](int st, int dd, int sd, int lev) __attribute__((always_inline)) -> C {
      const int iy = st / NX, ix = st - iy * NX;
      return A.has_patch ? bvalue(lev, dd, sd, dd == 0 ? iy : ix) : C{R(0), R(0)};
    };
    auto setbound = [&](C* F, int lev) __attribute__((always_inline)) {
      if constexpr (!MAT) return;
      if (!A.dirichlet) return;
#pragma unroll 1
      for (int i = 0; i < PAIRS; ++i) {
        const int s = S::site(i);
#pragma unroll
        for (int v = 0; v < 2; ++v) {
          int dd, sd;
          if (dir_end(s + v, dd, sd)) F[s + v] = dir_value(s + v, dd, sd, lev);
        }
      }
    };
    auto patch = [&](C* F, R sign, bool with_lp, int lev) __attribute__((always_inline)) {
      if constexpr (!MAT) return;
      if (!A.has_patch) return;
      const R* sh = static_cast<const R*>(A.bshape);
#pragma unroll 1
      for (int i = 0; i < PAIRS; ++i) {
        const int s = S::site(i);
#pragma unroll
        for (int v = 0; v < 2; ++v) {
          const int st = s + v, iy = st / NX, ix = st - iy * NX;
          C acc = C{R(0), R(0)};
#pragma unroll
          for (int dd = 0; dd < 2; ++dd) {
            const int m = dd == 0 ? ix : iy, o = dd == 0 ? iy : ix;
            const C b0 = bvalue(lev, dd, 0, o), b1 = bvalue(lev, dd, 1, o);
            acc = acc + sign * (sh[(dd * 2) * A.bshape_n + m] * b0 + sh[(dd * 2 + 1) * A.bshape_n + m] * b1);
            if (with_lp) acc = acc + C{(R)A.lpf[nci][dd][0], (R)A.lpf[nci][dd][1]} * (b1 - b0);
          }
          F[st] = F[st] + acc;
        }
      }
    };
    // xprop.m:17-46: patch, transform, propagator, inverse transform, patch, setbound
    auto propagate = [&](C* F, int lev) __attribute__((always_inline)) {
      if (A.has_linear) {
        if constexpr (MAT) {
          patch(F, R(-1), false, lev);
          S::propagate_mat(L, F, A.tm, A.tmp, A.cplx[0], A.cplx[1], prop, A.tsym);
          patch(F, R(1), true, lev);
        } else {
          S::propagate(L, F, twx, twy, twh, prop);
        }
      }
      setbound(F, lev);
    };

    // lev: time level of the step's start in half fine steps (a coarse step
    // spans twice the levels of a fine one).  The propagations of both methods
    // run from one code site (two phases, each a pointwise stage and then one
    // or two propagations): the 2D transform is inlined once, not once per call.
    auto method_step = [&](uint32_t q, bool coarse, int lev) __attribute__((always_inline)) {
      const int hl = coarse ? 2 : 1;
      const bool rk = A.method == XSH_RK4;
      __syncthreads();
      if (A.knoises) kfield(static_cast<const C*>(A.kn_filter), A.knoises, TAG_KNOISE, q, coarse);
#pragma unroll 1
      for (int ph = 0; ph < 2; ++ph) {
        if (rk && ph == 0) {
          // RK4.m:9-14: a1 = a + D(a) dt into F1
          setbound(F0, lev);
#pragma unroll 1
          for (int i = 0; i < PAIRS; ++i) {
            const int s = S::site(i);
            C w[2];
            noise(s, q, coarse, w[0], w[1]);
#pragma unroll
            for (int v = 0; v < 2; ++v) F1[s + v] = F0[s + v] + hdt * D(F0[s + v], w[v]);
          }
        } else if (rk) {
          // RK4.m:18-32: a3 -> F0, R -> F1 (Dirichlet sites at the midpoint's values)
#pragma unroll 1
          for (int i = 0; i < PAIRS; ++i) {
            const int s = S::site(i);
            C w[2];
            noise(s, q, coarse, w[0], w[1]);
#pragma unroll
            for (int v = 0; v < 2; ++v) {
              int dd = -1, sd = 0;
              const bool dz = MAT && A.dirichlet && dir_end(s + v, dd, sd);
              const C A0 = F0[s + v], A1 = F1[s + v];
              C a2 = A0 + hdt * D(A1, w[v]);
              if (dz) a2 = dir_value(s + v, dd, sd, lev + hl);
              C a3 = A0 + (R(2) * hdt) * D(a2, w[v]);
              if (dz) a3 = dir_value(s + v, dd, sd, lev + hl);
              F1[s + v] = R(1.0 / 3.0) * (A1 + R(2) * a2 + a3 - A0);
              F0[s + v] = a3;
            }
          }
        } else if (ph == 1) {
          // MP.m:13-24 (after the first propagation)
#pragma unroll 1
          for (int i = 0; i < PAIRS; ++i) {
            const int s = S::site(i);
            C w[2];
            noise(s, q, coarse, w[0], w[1]);
#pragma unroll
            for (int v = 0; v < 2; ++v) {
              int dd = -1, sd = 0;
              const bool dz = MAT && A.dirichlet && dir_end(s + v, dd, sd);
              const C a0 = F0[s + v];
              C a = a0;
              for (int it = 0; it < A.iterations; ++it) {
                a = a0 + hdt * D(a, w[v]);
                if (dz) a = dir_value(s + v, dd, sd, lev + hl);
              }
              F0[s + v] = R(2) * a - a0;
            }
          }
        }
        // RK4: P a, P a1 (phase 0), P a3, P R (phase 1); MP: P a (phase 0), P (2a - a0) (phase 1)
#pragma unroll 1
        for (int k = 0; k < (rk ? 2 : 1); ++k) propagate(k ? F1 : F0, lev + ph * 2 * hl);
      }
      if (rk) {
        // RK4.m:33-41: a = P R + D(P a3) dt/3
        __syncthreads();
#pragma unroll 1
        for (int i = 0; i < PAIRS; ++i) {
          const int s = S::site(i);
          C w[2];
          noise(s, q, coarse, w[0], w[1]);
#pragma unroll
          for (int v = 0; v < 2; ++v) F0[s + v] = F1[s + v] + (hdt / R(3)) * D(F0[s + v], w[v]);
        }
        setbound(F0, lev + 2 * hl);
      }
    };

    auto observe = [&](int n) __attribute__((always_inline)) {
      __syncthreads();
      if (A.nonfinite) {  // divergence check over this thread's site pairs
        bool bad = false;
#pragma unroll 1
        for (int i = 0; i < PAIRS; ++i) {
          const int s = S::site(i);
          bad = bad || not_finite(F0[s]) || not_finite(F0[s + 1]);
        }
        if (bad) atomicOr(A.nonfinite, 1);
      }
      if (A.raw) {  // xpath.m:73-78
#pragma unroll 1
        for (int i = 0; i < PAIRS; ++i) {
          const int s = S::site(i);
#pragma unroll
          for (int v = 0; v < 2; ++v) {
            const long long g = (((long long)pidx * A.npts + n) * A.E + traj) * NS + s + v;
            A.raw[2 * g] = (double)F0[s + v].x;
            A.raw[2 * g + 1] = (double)F0[s + v].y;
          }
        }
      }
      for (int o = 0; o < A.nobs; ++o) {
        const int kind = A.obs[o];
        double* dst = av + 2 * A.obs_off[o];
        if (kind == XSH_OBS_ABS2K) {
#pragma unroll 1
          for (int i = 0; i < PAIRS; ++i) {
            const int s = S::site(i);
            F2[s] = F0[s];
            F2[s + 1] = F0[s + 1];
          }
          const ObsOp<R> op{3, R(0), NX, nullptr, nullptr, dst + 2 * (long long)n * NS, A.kscale * A.inv_E, traj, A.E};
          S::propagate(L, F2, twx, twy, twh, op);
        } else if (kind == XSH_OBS_A || kind == XSH_OBS_ABS2 || kind == XSH_OBS_CORR) {
          const C* src = F0;
          if (kind == XSH_OBS_CORR) {
#pragma unroll 1
            for (int i = 0; i < PAIRS; ++i) {
              const int s = S::site(i);
              F2[s] = F0[s];
              F2[s + 1] = F0[s + 1];
            }
            const ObsOp<R> op{0, R(1.0 / ((double)NS * (double)NS)), NX, nullptr, nullptr, nullptr, 0.0, 0, 0};
            S::propagate(L, F2, twx, twy, twh, op);
            src = F2;
          }
#pragma unroll 1
          for (int i = 0; i < PAIRS; ++i) {
            const int s = S::site(i);
#pragma unroll
            for (int v = 0; v < 2; ++v) {
              const C a = src[s + v];
              double re = (double)a.x, im = (double)a.y;
              if (kind == XSH_OBS_ABS2) {
                re = (double)a.x * a.x + (double)a.y * a.y;
                im = 0.0;
              }
              atomic_add_c(dst + 2 * ((long long)n * NS + s + v), re * A.inv_E, im * A.inv_E);
            }
          }
        } else {
          double re = 0.0, im = 0.0;
#pragma unroll 1
          for (int i = 0; i < PAIRS; ++i) {
            const int s = S::site(i);
#pragma unroll
            for (int v = 0; v < 2; ++v) {
              const C a = F0[s + v];
              const double m = (double)a.x * a.x + (double)a.y * a.y;
              if (kind == XSH_OBS_A_AVE) {
                re += (double)a.x;
                im += (double)a.y;
              } else if (kind == XSH_OBS_ABS4_AVE) {
                re += m * m;
              } else {
                re += m;
              }
            }
          }
          block_sum2<NT>(re, im, red);
          const double sc0 = kind == XSH_OBS_ABS2_INT ? A.dv : 1.0 / NS;
          const double sc = sc0 * A.inv_E;
          if (threadIdx.x == 0) {
            atomic_add_c(dst + 2 * (long long)n, re * sc, im * sc);
            if (A.traj) A.traj[(((long long)pidx * A.E + traj) * A.npts + n) * A.nobs + o] = re * sc0;
          }
        }
      }
    };

    // ---- the time loop of xpath.m:29-86
    uint32_t q = 0;
    int lev = 0;
    for (int n = 0; n < A.npts; ++n) {
      for (int st = 0; st < A.steps; ++st) {
        if (n > 0 || st > 0 || A.npts == 1) {
          // a coarse step (draws q, q + 1), or 1 + checks fine steps: one call site
          const bool coarse = A.checks && nci == 1;
#pragma unroll 1
          for (int c = 0; c < (coarse ? 1 : 1 + A.checks); ++c) {
            method_step(q, coarse, lev);
            q += coarse ? 2 : 1;
            lev += coarse ? 4 : 2;
          }
        }
        if (st == 0) observe(n);
      }
    }
    __syncthreads();
    if (A.a_out) {
#pragma unroll 1
      for (int i = 0; i < PAIRS; ++i) {
        const int s = S::site(i);
#pragma unroll
        for (int v = 0; v < 2; ++v) {
          const long long g = (long long)traj * NS + s + v;
          A.a_out[2 * g] = (double)F0[s + v].x;
          A.a_out[2 * g + 1] = (double)F0[s + v].y;
        }
      }
    }
  }
}

// propagation probe of the streaming transform: planes [nplanes][NS] in place
template <class R, int NX, int NY, int NT, bool MAT>
__global__ __launch_bounds__(NT) void probe_stream(const PathArgs* __restrict__ Ap, int nci, int nplanes,
                                                   cpx<R>* planes) {
  const PathArgs& A = *Ap;
  using S = Stream<R, NX, NY, NT>;
  using C = cpx<R>;
  __shared__ C L[S::LDS];
  if ((int)blockIdx.x >= nplanes) return;
  const PropOp2<R> prop{static_cast<const C*>(nci ? A.px[1] : A.px[0]),
                       static_cast<const C*>(nci ? A.py[1] : A.py[0])};
  if constexpr (MAT)
    S::propagate_mat(L, planes + (long long)blockIdx.x * S::NS, A.tm, A.tmp, A.cplx[0], A.cplx[1], prop, A.tsym);
  else
    S::propagate(L, planes + (long long)blockIdx.x * S::NS, static_cast<const C*>(A.twx),
                 static_cast<const C*>(A.twy), static_cast<const C*>(A.twh), prop);
}

template <class R, int NX, int NY, int NT, bool MAT>
static hipError_t launch_stream(const PathArgs* args, const int* wstart, int ninst, int nwork, int grid,
                                hipStream_t s) {
  hipLaunchKernelGGL((spde_stream<R, NX, NY, NT, MAT>), dim3(grid), dim3(NT), 0, s, args, wstart, ninst, nwork);
  return hipGetLastError();
}

template <class R, int NX, int NY, int NT, bool MAT>
static hipError_t launch_probe_stream(const PathArgs* A, int nci, int nplanes, void* planes, hipStream_t s) {
  hipLaunchKernelGGL((probe_stream<R, NX, NY, NT, MAT>), dim3(nplanes), dim3(NT), 0, s, A, nci, nplanes,
                     static_cast<cpx<R>*>(planes));
  return hipGetLastError();
}

#define XSH_STREAM_ENTRY(R, NX, NY, NT)                                                                        \
  {sizeof(R) == 8 ? XSH_FP64 : XSH_FP32, NX, NY, NT, 1, 0, &launch_stream<R, NX, NY, NT, false>,              \
   &launch_probe_stream<R, NX, NY, NT, false>, 1, Split<NX>::N1, 0}
#define XSH_STREAM_MAT_ENTRY(R, NX, NY, NT)                                                                   \
  {sizeof(R) == 8 ? XSH_FP64 : XSH_FP32, NX, NY, NT, 1, 1, &launch_stream<R, NX, NY, NT, true>,              \
   &launch_probe_stream<R, NX, NY, NT, true>, 1, 0, 0}

}  // namespace xsh
