// Path kernel templates and their launch helpers.  Each xsh_k_*.hip
// translation unit instantiates one family and exports its launch table, so
// the families compile in parallel.
//
//  sde_path    dimensions = 1 (no space): one trajectory per lane, the whole
//              path in registers (xpath.m:29-86 with RK4.m / MP.m on a scalar
//              field; xprop.m:33-35 reduces to a scalar factor).
//  spde_path   dimensions = 2, 3 on periodic lattices: one workgroup per TPB
//              trajectories, fields resident in VGPRs for the whole path.
//  probes      normals and propagation of given planes (parity hooks).
#pragma once
#include "xsh_path.hpp"
#include "xsh_plane.hpp"
#include "xsh_launch.hpp"

namespace xsh {

// Block -> (path, trajectory group).  Blocks of the fine runs (nc = 1, twice
// the steps of a coarse run) come first so the longest work starts first; the
// averages of path (j, nc) go to slot (j-1)*nchk + nc-1, the layout of
// xsh_read_averages.
template <class R>
__device__ __forceinline__ void path_identity(const PathArgs& A, int lb, int& pidx, int& nci, int& j,
                                              int& group) {
  group = lb % A.groups;
  const int b = lb / A.groups;
  if (A.j_only > 0) {
    j = A.j_only;
    nci = A.nc_only - 1;
    pidx = 0;
  } else {
    nci = b / A.e2;
    j = b % A.e2 + 1;
    pidx = (j - 1) * (1 + A.checks) + nci;
  }
}

// noise of one draw for VEC=2 consecutive sites s, s+1 (s even):
// L = site*noises + comp, group L/4, lane L%4
template <class R>
__device__ __forceinline__ void draw_pair(const PathArgs& A, uint32_t nsp, uint32_t q, uint32_t traj,
                                          int s, cpx<R>& w0, cpx<R>& w1) {
  const R sc = (R)A.noise_scale;
  const uint32_t k1 = key1(TAG_NOISE, (uint32_t)A.seq);
  if (A.noises >= 2) {
    R n[4];
    normals4<R>(u4{(uint32_t)(s * A.noises) >> 2, q, traj, (uint32_t)A.inst}, nsp, k1, n);
    w0 = cpx<R>{sc * n[0], sc * n[1]};
    w1 = cpx<R>{sc * n[2], sc * n[3]};
  } else if (A.noises == 1) {
    const u4 x = philox10(u4{(uint32_t)s >> 2, q, traj, (uint32_t)A.inst}, nsp, k1);
    R a, b;
    if ((s & 3) == 0)
      box_muller(x.x, x.y, a, b);
    else
      box_muller(x.z, x.w, a, b);
    w0 = cpx<R>{sc * a, R(0)};
    w1 = cpx<R>{sc * b, R(0)};
  } else {
    w0 = w1 = cpx<R>{R(0), R(0)};
  }
}

// the Gaussian noise (w1 + i w2, or w1) of one site: the same counter stream as
// draw_pair, for sites whose neighbour in the chunk is not the next site
template <class R>
__device__ __forceinline__ cpx<R> draw_site(const PathArgs& A, uint32_t nsp, uint32_t q, uint32_t traj, int site) {
  if (A.noises <= 0) return cpx<R>{R(0), R(0)};
  const R sc = (R)A.noise_scale;
  const uint32_t L = (uint32_t)(site * A.noises);
  R n[4];
  normals4<R>(u4{L >> 2, q, traj, (uint32_t)A.inst}, nsp, key1(TAG_NOISE, (uint32_t)A.seq), n);
  const uint32_t lane = L & 3u;
  const R a = lane == 0 ? n[0] : lane == 1 ? n[1] : lane == 2 ? n[2] : n[3];
  const R b = A.noises >= 2 ? (lane == 0 ? n[1] : n[3]) : R(0);
  return cpx<R>{sc * a, sc * b};
}

// the ncomp (1 or 2) normals of k site ks in a k-space draw (counter layout
// of every other stream: L = ks * ncomp + comp, lane L & 3)
template <class R>
__device__ __forceinline__ void knormals(const PathArgs& A, uint32_t nsp, uint32_t tag, uint32_t q,
                                         uint32_t traj, int ks, int ncomp, R& z0, R& z1) {
  R n[4];
  const uint32_t L = (uint32_t)(ks * ncomp);
  normals4<R>(u4{L >> 2, q, traj, (uint32_t)A.inst}, nsp, key1(tag, (uint32_t)A.seq), n);
  const uint32_t lane = L & 3u;
  // selects, not a runtime index: n stays in registers
  z0 = lane == 0 ? n[0] : lane == 1 ? n[1] : lane == 2 ? n[2] : n[3];
  z1 = ncomp < 2 ? R(0) : (lane == 0 ? n[1] : n[3]);
}

// one normal / uniform of the initial randoms (xrandom.m:18-33)
template <class R>
__device__ __forceinline__ R init_normal(const PathArgs& A, uint32_t nsp, uint32_t traj, int L) {
  R n[4];
  normals4<R>(u4{(uint32_t)L >> 2, 0u, traj, (uint32_t)A.inst}, nsp,
              key1(TAG_GAUSS_INIT, (uint32_t)A.seq), n);
  const int l = L & 3;  // select, not a dynamically indexed private array
  return l == 0 ? n[0] : l == 1 ? n[1] : l == 2 ? n[2] : n[3];
}
template <class R>
__device__ __forceinline__ R init_uniform(const PathArgs& A, uint32_t nsp, uint32_t traj, int L) {
  const u4 x = philox10(u4{(uint32_t)L >> 2, 0u, traj, (uint32_t)A.inst}, nsp,
                        key1(TAG_UNIFORM_INIT, (uint32_t)A.seq));
  const int l = L & 3;
  return uniform01<R>(l == 0 ? x.x : l == 1 ? x.y : l == 2 ? x.z : x.w);
}

// built-in initial forms (user-level p.initial handles, xensemble.m:60)
template <class R>
// kv: the k-space filtered randoms v1 + i v2 at this site (used when krandoms > 0)
__device__ __forceinline__ cpx<R> initial_value(const PathArgs& A, uint32_t nsp, uint32_t traj, int site,
                                                int ix, int iy, cpx<R> kv = cpx<R>{R(0), R(0)}) {
  const double x = A.x0 + ix * A.dx, y = A.y0 + iy * A.dy;
  const R isc = (R)A.init_scale;
  switch (A.initial) {
    case XSH_INIT_SECH: {
      cpx<R> a = {(R)(A.ip[0] / cosh((x - A.ip[3]) / A.ip[1])), R(0)};
      if (A.ip[2] != 0.0) {
        const R e = (R)(A.ip[2] * 0.70710678118654752440);
        if (A.krandoms) return cpx<R>{a.x + e * kv.x, a.y + e * kv.y};
        const R v0 = isc * init_normal<R>(A, nsp, traj, site * A.inrandoms);
        const R v1 = isc * init_normal<R>(A, nsp, traj, site * A.inrandoms + 1);
        a.x += e * v0;
        a.y += e * v1;
      }
      return a;
    }
    case XSH_INIT_GAUSS: {
      const double wy = A.ip[2] > 0.0 ? A.ip[2] : A.ip[1];
      const double r2 = x * x / (A.ip[1] * A.ip[1]) + (A.ny > 1 ? y * y / (wy * wy) : 0.0);
      return cpx<R>{(R)(A.ip[0] * exp(-0.5 * r2)), R(0)};
    }
    case XSH_INIT_PHASE: {
      const R u = init_uniform<R>(A, nsp, traj, site * A.urandoms);
      R s, c;
      if constexpr (sizeof(R) == 4)
        __sincosf(6.2831853071795865f * u, &s, &c);
      else
        sincos(6.283185307179586476925 * u, &s, &c);
      return cpx<R>{(R)A.ip[0] * c, (R)A.ip[0] * s};
    }
    case XSH_INIT_COHERENT: {
      if (A.krandoms) return cpx<R>{(R)A.ip[0] + R(0.5) * kv.x, (R)A.ip[1] + R(0.5) * kv.y};
      const R v0 = isc * init_normal<R>(A, nsp, traj, site * A.inrandoms);
      const R v1 = isc * init_normal<R>(A, nsp, traj, site * A.inrandoms + 1);
      return cpx<R>{(R)A.ip[0] + R(0.5) * v0, (R)A.ip[1] + R(0.5) * v1};
    }
    default:
      return cpx<R>{(R)A.ip[0], (R)A.ip[1]};
  }
}

// ===================================================================== SDE
template <class R>
__global__ __launch_bounds__(256) void sde_path(const PathArgs* __restrict__ Ap) {
  const PathArgs& A = *Ap;
  using C = cpx<R>;
  __shared__ double red[8];
  int pidx, nci, j, group;
  path_identity<R>(A, blockIdx.x, pidx, nci, j, group);
  const int traj = group * 256 + threadIdx.x;
  const bool active = traj < A.E;
  const uint32_t nsp = (uint32_t)(A.seed + A.npar + (j - 1) * A.e3);
  const R hdt = (R)(nci ? A.hdt[1] : A.hdt[0]);
  Drift<R> D;
  D.init(A);
  const C P = A.has_linear ? static_cast<const C*>(nci ? A.px[1] : A.px[0])[0] : C{R(1), R(0)};
  double* av = A.av + 2 * (long long)pidx * A.avlen;

  C a;
  if (A.a0) {
    const long long g = active ? traj : 0;
    a = C{(R)A.a0[2 * g], (R)A.a0[2 * g + 1]};
  } else
    a = initial_value<R>(A, nsp, (uint32_t)traj, 0, 0, 0);

  auto draw = [&](uint32_t q) __attribute__((always_inline)) -> C {
    const R sc = (R)A.noise_scale;
    if (A.noises <= 0) return C{R(0), R(0)};
    R n[4];
    normals4<R>(u4{0u, q, (uint32_t)traj, (uint32_t)A.inst}, nsp, key1(TAG_NOISE, (uint32_t)A.seq), n);
    return C{sc * n[0], A.noises >= 2 ? sc * n[1] : R(0)};
  };
  auto step = [&](C xi) __attribute__((always_inline)) {
    if (A.method == XSH_RK4) {
      C a1 = a + hdt * D(a, xi);
      a1 = P * a1;
      a = P * a;
      C a2 = a + hdt * D(a1, xi);
      C a3 = a + (R(2) * hdt) * D(a2, xi);
      a = R(1.0 / 3.0) * (a1 + R(2) * a2 + a3 - a);
      a = P * a;
      a3 = P * a3;
      a = a + (hdt / R(3)) * D(a3, xi);
    } else {
      a = P * a;
      const C a0 = a;
      for (int it = 0; it < A.iterations; ++it) a = a0 + hdt * D(a, xi);
      a = R(2) * a - a0;
      a = P * a;
    }
  };
  auto observe = [&](int n) __attribute__((always_inline)) {
    if (A.nonfinite && active && not_finite(a)) atomicOr(A.nonfinite, 1);
#pragma unroll
    for (int o = 0; o < XSH_MAX_OBS; ++o) {
      if (o >= A.nobs) break;
      if ((A.tspec_mask >> o) & 1) continue;  // taken of the time transform (sde_spectrum)
      double re = 0.0, im = 0.0;
      if (active) {
        switch (A.obs[o]) {
          case XSH_OBS_A:
          case XSH_OBS_A_AVE:
            re = (double)a.x;
            im = (double)a.y;
            break;
          case XSH_OBS_ABS4_AVE: {
            const double m = (double)norm2(a);
            re = m * m;
            break;
          }
          default:
            re = (double)norm2(a);
        }
      }
      block_sum2<256>(re, im, red);
      if (threadIdx.x == 0) atomic_add_c(av + 2 * (A.obs_off[o] + n), re * A.inv_E, im * A.inv_E);
    }
  };

  // the field at every step for the time-domain spectra (xpath.m:45-68):
  // sample n1 after step n1 (n1 = 0: the initial field); a fine run with checks
  // stores the mean of the step's two half-step results
  double* ft = A.ft && active ? A.ft + 2 * (long long)pidx * A.ft_len * A.E : nullptr;
  auto sample = [&](int n1, double re, double im) __attribute__((always_inline)) {
    if (ft) {
      ft[2 * ((long long)n1 * A.E + traj)] = re;
      ft[2 * ((long long)n1 * A.E + traj) + 1] = im;
    }
  };
  uint32_t q = 0;
  int n1 = 0;
  for (int n = 0; n < A.npts; ++n) {
    for (int st = 0; st < A.steps; ++st, ++n1) {
      if (n > 0 || st > 0 || A.npts == 1) {
        if (A.checks && nci == 1) {
          const C w0 = draw(q), w1 = draw(q + 1);
          q += 2;
          step(R(0.5) * (w0 + w1));
          sample(n1, (double)a.x, (double)a.y);
        } else if (A.checks) {
          step(draw(q++));
          const double hr = (double)a.x, hi = (double)a.y;
          step(draw(q++));
          sample(n1, 0.5 * ((double)a.x + hr), 0.5 * ((double)a.y + hi));
        } else {
          step(draw(q++));
          sample(n1, (double)a.x, (double)a.y);
        }
      } else {
        sample(n1, (double)a.x, (double)a.y);
      }
      if (st == 0) observe(n);
    }
  }
  if (A.a_out && active) {
    A.a_out[2 * (long long)traj] = (double)a.x;
    A.a_out[2 * (long long)traj + 1] = (double)a.y;
  }
}

// Coupled components of an SDE (p.fields = F > 1): one trajectory per lane with
// its F components in registers; the matrix drift
//   D_i(a, w) = sum_j M_ij a_j + g_i |a_i|^2 a_i + sum_k S_ik w_k
// (w: the noises <= 4 real components of one draw, xnoise.m:30-32, from one
// Philox block), the diagonal interaction picture propagator P_i, and RK4 /
// MP exactly as sde_path does them with vector-valued a (RK4.m:7-41, MP.m:10-26).
template <class R, int F>
__global__ __launch_bounds__(256) void sde_fields(const PathArgs* __restrict__ Ap) {
  const PathArgs& A = *Ap;
  using C = cpx<R>;
  __shared__ double red[8];
  int pidx, nci, j, group;
  path_identity<R>(A, blockIdx.x, pidx, nci, j, group);
  const int traj = group * 256 + threadIdx.x;
  const bool active = traj < A.E;
  const uint32_t nsp = (uint32_t)(A.seed + A.npar + (j - 1) * A.e3);
  const R hdt = (R)(nci ? A.hdt[1] : A.hdt[0]);
  double* av = A.av + 2 * (long long)pidx * A.avlen;
  C P[F];
#pragma unroll
  for (int i = 0; i < F; ++i) P[i] = C{(R)A.fp[nci][i][0], (R)A.fp[nci][i][1]};

  C a[F];
#pragma unroll
  for (int i = 0; i < F; ++i) {
    if (A.a0) {
      const long long g = (long long)(active ? traj : 0) * F + i;
      a[i] = C{(R)A.a0[2 * g], (R)A.a0[2 * g + 1]};
    } else {
      a[i] = C{(R)A.fc0[i][0], (R)A.fc0[i][1]};
    }
  }
  // one draw: noise components k < noises, scaled (xnoise.m:32)
  auto draw = [&](uint32_t q, R (&w)[4]) __attribute__((always_inline)) {
    const R sc = (R)A.noise_scale;
    R n[4];
    normals4<R>(u4{0u, q, (uint32_t)traj, (uint32_t)A.inst}, nsp, key1(TAG_NOISE, (uint32_t)A.seq), n);
#pragma unroll
    for (int k = 0; k < 4; ++k) w[k] = k < A.noises ? sc * n[k] : R(0);
  };
  auto D = [&](const C (&x)[F], const R (&w)[4], C (&out)[F]) __attribute__((always_inline)) {
#pragma unroll
    for (int i = 0; i < F; ++i) {
      C acc = C{R(0), R(0)};
#pragma unroll
      for (int jj = 0; jj < F; ++jj) acc = acc + C{(R)A.fm[i][jj][0], (R)A.fm[i][jj][1]} * x[jj];
      acc = acc + (norm2(x[i]) * C{(R)A.fg[i][0], (R)A.fg[i][1]}) * x[i];
      R sw = R(0);
#pragma unroll
      for (int k = 0; k < 4; ++k) sw += (R)A.fs[i][k] * w[k];
      out[i] = acc + C{sw, R(0)};
    }
  };
  auto step = [&](const R (&w)[4]) __attribute__((always_inline)) {
    C d[F];
    if (A.method == XSH_RK4) {
      C a1[F], a2[F], a3[F];
      D(a, w, d);
#pragma unroll
      for (int i = 0; i < F; ++i) {
        a1[i] = P[i] * (a[i] + hdt * d[i]);
        a[i] = P[i] * a[i];
      }
      D(a1, w, d);
#pragma unroll
      for (int i = 0; i < F; ++i) a2[i] = a[i] + hdt * d[i];
      D(a2, w, d);
#pragma unroll
      for (int i = 0; i < F; ++i) {
        a3[i] = a[i] + (R(2) * hdt) * d[i];
        a[i] = P[i] * (R(1.0 / 3.0) * (a1[i] + R(2) * a2[i] + a3[i] - a[i]));
        a3[i] = P[i] * a3[i];
      }
      D(a3, w, d);
#pragma unroll
      for (int i = 0; i < F; ++i) a[i] = a[i] + (hdt / R(3)) * d[i];
    } else {
      C a0[F];
#pragma unroll
      for (int i = 0; i < F; ++i) a0[i] = a[i] = P[i] * a[i];
      for (int it = 0; it < A.iterations; ++it) {
        D(a, w, d);
#pragma unroll
        for (int i = 0; i < F; ++i) a[i] = a0[i] + hdt * d[i];
      }
#pragma unroll
      for (int i = 0; i < F; ++i) a[i] = P[i] * (R(2) * a[i] - a0[i]);
    }
  };
  auto observe = [&](int n) __attribute__((always_inline)) {
    if (A.nonfinite && active) {
      bool bad = false;
#pragma unroll
      for (int i = 0; i < F; ++i) bad = bad || not_finite(a[i]);
      if (bad) atomicOr(A.nonfinite, 1);
    }
#pragma unroll
    for (int o = 0; o < XSH_MAX_OBS; ++o) {
      if (o >= A.nobs) break;
      double re = 0.0, im = 0.0;
      if (active) {
        C x = a[0];
#pragma unroll
        for (int i = 1; i < F; ++i)
          if (A.obs_field[o] == i) x = a[i];
        switch (A.obs[o]) {
          case XSH_OBS_A:
          case XSH_OBS_A_AVE:
            re = (double)x.x;
            im = (double)x.y;
            break;
          case XSH_OBS_ABS4_AVE: {
            const double m = (double)norm2(x);
            re = m * m;
            break;
          }
          default:
            re = (double)norm2(x);
        }
      }
      block_sum2<256>(re, im, red);
      if (threadIdx.x == 0) atomic_add_c(av + 2 * (A.obs_off[o] + n), re * A.inv_E, im * A.inv_E);
    }
  };

  uint32_t q = 0;
  for (int n = 0; n < A.npts; ++n) {
    for (int st = 0; st < A.steps; ++st) {
      if (n > 0 || st > 0 || A.npts == 1) {
        if (A.checks && nci == 1) {
          R w0[4], w1[4], w[4];
          draw(q, w0);
          draw(q + 1, w1);
          q += 2;
#pragma unroll
          for (int k = 0; k < 4; ++k) w[k] = R(0.5) * (w0[k] + w1[k]);
          step(w);
        } else {
          for (int c = 0; c <= A.checks; ++c) {
            R w[4];
            draw(q++, w);
            step(w);
          }
        }
      }
      if (st == 0) observe(n);
    }
  }
  if (A.a_out && active) {
#pragma unroll
    for (int i = 0; i < F; ++i) {
      const long long g = (long long)traj * F + i;
      A.a_out[2 * g] = (double)a[i].x;
      A.a_out[2 * g + 1] = (double)a[i].y;
    }
  }
}

// ==================================================================== SPDE
// Instance of work item wk in a batch: the last i with wstart[i] <= wk
// (wstart[0] = 0; uniform over the workgroup).
__device__ __forceinline__ int find_instance(const int* __restrict__ wstart, int ninst, int wk) {
  int lo = 0, hi = ninst - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (wstart[mid] <= wk) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// Lattices whose periodic transform runs on Plane (wave-owned rows, two
// barriers per transform): one 128 x 128 trajectory per 1024-thread workgroup.
template <int NX, int NY, int NT, int TPB, bool MAT, int RX, bool KN>
constexpr bool kPlaneFFT = !MAT && !KN && RX == 0 && NX == 128 && NY == 128 && NT == 1024 && TPB == 1;

// LDS elements of the transform workspace
template <class R, int NX, int NY, int NT, int TPB, bool MAT, int RX, bool KN>
constexpr int lds_elems() {
  if constexpr (kPlaneFFT<NX, NY, NT, TPB, MAT, RX, KN>)
    return Plane<R, NX, NT>::LDS > Lattice<R, NX, NY, NT, TPB, RX>::LDS ? Plane<R, NX, NT>::LDS
                                                                          : Lattice<R, NX, NY, NT, TPB, RX>::LDS;
  else
    return Lattice<R, NX, NY, NT, TPB, RX>::LDS;
}

// The propagator multiply of a lattice kernel (per-dimension factors).  A
// product table P[kx + nx ky] (one complex multiply instead of two) measured
// slower on the 128 x 128 kernel: its 16 per-lane loads per column item sit
// between the transform's barriers (1784 -> 2231 ms per bench launch), where
// Py[ky] is a wave-uniform scalar load.
template <class R, int NX, int NY, int NT, int TPB, bool MAT, int RX, bool KN>
__device__ __forceinline__ auto make_prop(const PathArgs& A, int nci) {
  using C = cpx<R>;
  if constexpr (NY > 1)
    return PropOp2<R>{static_cast<const C*>(A.px[nci]), static_cast<const C*>(A.py[nci])};
  else
    return PropOp<R>{static_cast<const C*>(A.px[nci]), nullptr};
}

// a -> T^-1 op(T a) of the planes held in f, periodic lattices
template <class R, int NX, int NY, int NT, int TPB, bool MAT, int RX, bool KN, class Op>
__device__ __forceinline__ void periodic_spectral(cpx<R>* L, cpx<R> (&f)[Lattice<R, NX, NY, NT, TPB, RX>::PPT],
                                                  const cpx<R>* twx, const cpx<R>* twy, const Op& op) {
  if constexpr (kPlaneFFT<NX, NY, NT, TPB, MAT, RX, KN>)
    Plane<R, NX, NT>::spectral(L, f, twx, twy, op);
  else
    Lattice<R, NX, NY, NT, TPB, RX>::spectral(L, f, twx, twy, op);
}

// One work item (a group of TPB trajectories of one path) of one instance.
template <class R, int NX, int NY, int NT, int TPB, bool MAT, int RX, bool KN>
__device__ __forceinline__ void spde_work(const PathArgs& A, int lb, cpx<R>* __restrict__ L, double* red) {
  using Lat = Lattice<R, NX, NY, NT, TPB, RX>;
  using C = cpx<R>;
  constexpr int CPT = Lat::CPT, PPT = Lat::PPT, NS = Lat::NS;
  // Dense-transform lattices may hold fewer physical sites than NX x NY (1D of
  // any size; 2D kernels of fewer than kPadMaxSites sites): the other sites are
  // padding.  Zero rows and columns of every transform matrix keep them out of
  // the physical sites, no output reads them, and the physical sites keep their
  // own indices x + nx * y for the counter streams and the outputs.  FFT
  // kernels and the 128 x 128 dense kernel: exactly NS.
  constexpr bool PAD2 = MAT && NY > 1 && NX * NY < kPadMaxSites;
  const int NXP = MAT ? A.nx : NX, NYP = MAT && NY > 1 ? A.ny : NY;
  const int NSP = MAT ? A.ns : NS;
  // physical index of lattice site s, -1 for padding
  auto phys = [&](int s) __attribute__((always_inline)) -> int {
    if constexpr (MAT) {
      const int y = s / NX, x = s - y * NX;
      return x < NXP && y < NYP ? x + NXP * y : -1;
    } else {
      return s;
    }
  };

  int pidx, nci, j, group;
  path_identity<R>(A, lb, pidx, nci, j, group);
  const uint32_t nsp = (uint32_t)(A.seed + A.npar + (j - 1) * A.e3);
  const R hdt = (R)(nci ? A.hdt[1] : A.hdt[0]);
  Drift<R> D;
  D.init(A);
  const auto prop = make_prop<R, NX, NY, NT, TPB, MAT, RX, KN>(A, nci);
  const C* twx = static_cast<const C*>(A.twx);
  const C* twy = static_cast<const C*>(A.twy);
  double* av = A.av + 2 * (long long)pidx * A.avlen;

  C f0[PPT], f1[PPT];
  // Every loop that indexes f0/f1/xi is a compile-time static_for: a runtime
  // index would demote the arrays from VGPRs to scratch memory.
#define XSH_FOR(VAR, B, E) static_for<B, E>([&](auto VAR##_) __attribute__((always_inline)) { \
    constexpr int VAR = decltype(VAR##_)::value;
#define XSH_END });

  // k-space filtered noise (xnoise.m:39-50) and randoms (xrandom.m:24-29) of
  // the ncomp components: Z(k) = F0(k) z0(k) + i F1(k) z1(k) from the counter
  // normals of every k site, written to LDS in the forward stages' frequency
  // order and inverse transformed there, so that LDS then holds w1 + i w2 (one
  // component: w1), the combination the drift and initial forms use.  Coarse
  // steps use the mean of draws q and q + 1 (linear: the mean of the k-space
  // values).  Runtime loops and no register arrays: the rest of the kernel's
  // register budget is not touched.
  auto kfield = [&](const C* __restrict__ F, int ncomp, uint32_t tag, uint32_t q, bool coarse)
      __attribute__((always_inline)) {
    __syncthreads();
#pragma unroll 1
    for (int i = 0; i < CPT; ++i) {
      const int pt = Lat::point(i, 0);
      const int p = pt / NS, s0 = pt - p * NS;
      const uint32_t traj = (uint32_t)(group * TPB + p);
      const int a = Lat::lds_of(pt);
#pragma unroll
      for (int v = 0; v < 2; ++v) {
        const int ks = Lat::k_site(s0 + v);
        R z0, z1;
        knormals<R>(A, nsp, tag, q, traj, ks, ncomp, z0, z1);
        if (coarse) {
          R u0, u1;
          knormals<R>(A, nsp, tag, q + 1, traj, ks, ncomp, u0, u1);
          z0 = R(0.5) * (z0 + u0);
          z1 = R(0.5) * (z1 + u1);
        }
        const C F0 = F[ks];
        C Z = C{F0.x * z0, F0.y * z0};
        if (ncomp >= 2) {
          const C F1 = F[NS + ks];
          Z = Z + C{-F1.y * z1, F1.x * z1};
        }
        L[a + v] = Z;
      }
    }
    __syncthreads();
    Lat::inverse_lds(L, twx, twy);
  };
  // this workgroup's planes of the step's x-space k-noise
  // (indexed by the instance's work item, not the block: in a batch's
  // persistent grid a block serves the items of several instances, and each
  // instance's planes cover its own items)
  C* const kws = KN && A.knoises ? static_cast<C*>(A.kn_work) + (long long)lb * TPB * NS : nullptr;

  XSH_FOR(k, 0, PPT) f1[k] = C{R(0), R(0)}; XSH_END
  if constexpr (!MAT && KN) {
    if (A.krandoms && !A.a0) {
      kfield(static_cast<const C*>(A.kr_filter), A.krandoms, TAG_KRAND_INIT, 0u, false);
      Lat::load(L, f1);
    }
  }

  // ---- initial fields (xensemble.m:53-65)
  XSH_FOR(i, 0, CPT)
    const int pt = Lat::point(i, 0);
    const int p = pt / NS, s = pt % NS;
    const int traj = group * TPB + p;
    XSH_FOR(v, 0, 2)
      const int site = s + v, iy = site / NX, ix = site % NX;
      const int ps = phys(site);
      if (A.a0) {
        const long long g = (long long)(traj < A.E ? traj : 0) * NSP + (!MAT || ps >= 0 ? ps : 0);
        f0[2 * i + v] = !MAT || ps >= 0 ? C{(R)A.a0[2 * g], (R)A.a0[2 * g + 1]} : C{R(0), R(0)};
      } else {
        f0[2 * i + v] = initial_value<R>(A, nsp, (uint32_t)traj, !MAT || ps >= 0 ? ps : 0, ix, iy, f1[2 * i + v]);
      }
      __builtin_amdgcn_sched_barrier(0);
    XSH_END
  XSH_END

  // Noise of draw q (coarse: mean of draws q, q+1, xnoise.m:33-34) for
  // register chunk i.  It is regenerated from the counter at each of the
  // method's drift stages instead of being held for the whole step: that keeps
  // a single field resident during every transform, the peak of the register
  // budget (holding the step's noise of a 128x128 plane adds 32 VGPRs per
  // thread to the two fields' 64, and the compiler then spills the fields
  // around every transform).
  auto noise_chunk = [&](int i, uint32_t q, bool coarse, C& w0, C& w1) __attribute__((always_inline)) {
    if (!MAT && KN && A.knoises) {  // the step's k-noise, transformed once per step
      const int pt = Lat::point(i, 0);
      w0 = kws[pt];
      w1 = kws[pt + 1];
      return;
    }
    const int pt = Lat::point(i, 0);
    const int p = pt / NS, s = pt % NS;
    const uint32_t traj = (uint32_t)(group * TPB + p);
    if constexpr (PAD2) {
      // padded 2D lattice: each site's own physical index (odd rows of an odd
      // width do not start on a counter group)
      const int p0 = phys(s), p1 = phys(s + 1);
      w0 = p0 >= 0 ? draw_site<R>(A, nsp, q, traj, p0) : C{R(0), R(0)};
      w1 = p1 >= 0 ? draw_site<R>(A, nsp, q, traj, p1) : C{R(0), R(0)};
      if (coarse) {
        if (p0 >= 0) w0 = R(0.5) * (w0 + draw_site<R>(A, nsp, q + 1, traj, p0));
        if (p1 >= 0) w1 = R(0.5) * (w1 + draw_site<R>(A, nsp, q + 1, traj, p1));
      }
      return;
    }
    draw_pair<R>(A, nsp, q, traj, s, w0, w1);
    if (coarse) {
      C u0, u1;
      draw_pair<R>(A, nsp, q + 1, traj, s, u0, u1);
      w0 = R(0.5) * (w0 + u0);
      w1 = R(0.5) * (w1 + u1);
    }
  };

  // Dirichlet ends of the lattice (xsetbound.m:40-53, default boundary value 0)
  auto dir_site = [&](int i, int v) __attribute__((always_inline)) -> bool {
    const int site = Lat::point(i, v) % NS, iy = site / NX, ix = site % NX;
    return (ix == 0 && A.dir_lo[0]) || (ix == NXP - 1 && A.dir_hi[0]) ||
           (NY > 1 && ((iy == 0 && A.dir_lo[1]) || (iy == NYP - 1 && A.dir_hi[1])));
  };
  // Boundary value of dimension dd, side sd at time level lev, at coordinate
  // index o along the other dimension (p.boundfun, xpatch.m:37, xsetbound.m:39)
  auto bvalue = [&](int lev, int dd, int sd, int o) __attribute__((always_inline)) -> C {
    const C* bv = static_cast<const C*>(A.bval);
    return bv[(long long)lev * A.bval_level + (dd * 2 + sd) * A.bval_nb + o];
  };
  // value of a Dirichlet site: x ends, then y ends (xsetbound.m sets dimension
  // by dimension, so y values win at corners)
  auto dir_value = [&](int i, int v, int lev) __attribute__((always_inline)) -> C {
    const int site = Lat::point(i, v) % NS, iy = site / NX, ix = site % NX;
    int dd = -1, sd = 0;
    if (ix == 0 && A.dir_lo[0]) dd = 0, sd = 0;
    if (ix == NXP - 1 && A.dir_hi[0]) dd = 0, sd = 1;
    if (NY > 1 && iy == 0 && A.dir_lo[1]) dd = 1, sd = 0;
    if (NY > 1 && iy == NYP - 1 && A.dir_hi[1]) dd = 1, sd = 1;
    return dd < 0 || !A.has_patch ? C{R(0), R(0)} : bvalue(lev, dd, sd, dd == 0 ? iy : ix);
  };
  // (only dense-transform lattices can have Dirichlet ends)
  auto setbound = [&](C (&f)[PPT], int lev) __attribute__((always_inline)) {
    if constexpr (!MAT) return;
    if (!A.dirichlet) return;
    XSH_FOR(i, 0, CPT)
      XSH_FOR(v, 0, 2)
        if (dir_site(i, v)) f[2 * i + v] = dir_value(i, v, lev);
      XSH_END
    XSH_END
  };
  // non-zero boundary values: a -> a - pa before the transforms, + pa + lp
  // after (xprop.m:17-47 with xpatch.m:24-62), pa and lp at time level lev
  auto patch = [&](C (&f)[PPT], R sign, bool with_lp, int lev) __attribute__((always_inline)) {
    if constexpr (!MAT) return;
    if (!A.has_patch) return;
    const R* sh = static_cast<const R*>(A.bshape);
    XSH_FOR(i, 0, CPT)
      XSH_FOR(v, 0, 2)
        const int site = Lat::point(i, v) % NS, iy = site / NX, ix = site % NX;
        C acc = C{R(0), R(0)};
        XSH_FOR(dd, 0, (NY > 1 ? 2 : 1))
          const int m = dd == 0 ? ix : iy, o = dd == 0 ? iy : ix;
          const C b0 = bvalue(lev, dd, 0, o), b1 = bvalue(lev, dd, 1, o);
          acc = acc + sign * (sh[(dd * 2) * A.bshape_n + m] * b0 + sh[(dd * 2 + 1) * A.bshape_n + m] * b1);
          if (with_lp) acc = acc + C{(R)A.lpf[nci][dd][0], (R)A.lpf[nci][dd][1]} * (b1 - b0);
        XSH_END
        f[2 * i + v] = f[2 * i + v] + acc;
      XSH_END
    XSH_END
  };
  // xprop.m:17-46: transform, propagator, inverse transform, then setbound
  auto propagate = [&](C (&f)[PPT], int lev) __attribute__((always_inline)) {
    if (A.has_linear) {
      if constexpr (MAT) {
        patch(f, R(-1), false, lev);
        Lat::spectral_mat(L, f, A.tm, A.tmp, A.cplx[0], A.cplx[1], prop, A.tsym);
        patch(f, R(1), true, lev);
      } else {
        periodic_spectral<R, NX, NY, NT, TPB, MAT, RX, KN>(L, f, twx, twy, prop);
      }
    }
    setbound(f, lev);
  };

  // All propagations of a step go through one code site that transforms f1
  // (the transform is by far the largest code; one site keeps the kernel
  // small).  f0 holds the state between steps; where the reference propagates
  // the other field, the two arrays swap roles (register moves).
  auto swap01 = [&]() __attribute__((always_inline)) {
    XSH_FOR(k, 0, PPT)
      const C t = f0[k];
      f0[k] = f1[k];
      f1[k] = t;
    XSH_END
  };

  // Pointwise stages visit the register chunks (2 sites each) of one or two
  // fields.  Other 2D lattices run a runtime loop that works on the first UR
  // chunks and then rotates the arrays by UR chunks (register moves, static
  // indices): UR copies of the noise generator and drift per stage instead of
  // CPT keep their larger time steps inside the instruction cache (the
  // generic-transform 128x128 kernel per launch: unrolled 3.23 s, UR = 1
  // 3.03 s, UR = 2 2.94 s, UR = 4 3.04 s).  The 128 x 128 FFT kernel's
  // transform (Plane) is small enough that its fully unrolled stages fit: no
  // rotation moves (1871 -> 1784 ms per bench launch; UR = 4 rotating
  // 1866 ms).  1D lattices, with few chunks, unroll.
  constexpr int UR = CPT % 2 == 0 ? 2 : 1;
  constexpr bool ROTATE = NY > 1 && !kPlaneFFT<NX, NY, NT, TPB, MAT, RX, KN>;
  auto rotate = [&](C (&f)[PPT]) __attribute__((always_inline)) {
    C t[2 * UR];
    XSH_FOR(k, 0, 2 * UR) t[k] = f[k]; XSH_END
    XSH_FOR(k, 0, PPT - 2 * UR) f[k] = f[k + 2 * UR]; XSH_END
    XSH_FOR(k, 0, 2 * UR) f[PPT - 2 * UR + k] = t[k]; XSH_END
  };
  // body(i, a0, a1, b0, b1): chunk i of fields a and b (distinct arrays), in place
  auto chunks2 = [&](C (&a)[PPT], C (&b)[PPT], auto&& body) __attribute__((always_inline)) {
    if constexpr (ROTATE) {
#pragma unroll 1
      for (int i = 0; i < CPT; i += UR) {
        XSH_FOR(u, 0, UR)
          body(i + u, a[2 * u], a[2 * u + 1], b[2 * u], b[2 * u + 1]);
        XSH_END
        rotate(a);
        rotate(b);
      }
    } else {
      XSH_FOR(i, 0, CPT)
        body(i, a[2 * i], a[2 * i + 1], b[2 * i], b[2 * i + 1]);
        __builtin_amdgcn_sched_barrier(0);
      XSH_END
    }
  };
  auto chunks1 = [&](C (&a)[PPT], auto&& body) __attribute__((always_inline)) {
    if constexpr (ROTATE) {
#pragma unroll 1
      for (int i = 0; i < CPT; i += UR) {
        XSH_FOR(u, 0, UR)
          body(i + u, a[2 * u], a[2 * u + 1]);
        XSH_END
        rotate(a);
      }
    } else {
      XSH_FOR(i, 0, CPT)
        body(i, a[2 * i], a[2 * i + 1]);
        __builtin_amdgcn_sched_barrier(0);
      XSH_END
    }
  };

  // lev: time level of the step's start in half fine steps; a coarse step
  // spans twice the levels of a fine one
  auto method_step = [&](uint32_t q, bool coarse, int lev) __attribute__((always_inline)) {
    const bool rk4 = A.method == XSH_RK4;
    const int hl = coarse ? 2 : 1;
    if constexpr (!MAT && KN) {
      if (A.knoises) {
        // the step's noise is fixed over its stages: transform it once and keep
        // it in this workgroup's planes (the next transform synchronises
        // before it writes LDS again)
        kfield(static_cast<const C*>(A.kn_filter), A.knoises, TAG_KNOISE, q, coarse);
#pragma unroll 1
        for (int i = 0; i < CPT; ++i) {
          const int pt = Lat::point(i, 0);
          const int a = Lat::lds_of(pt);
          kws[pt] = L[a];
          kws[pt + 1] = L[a + 1];
        }
      }
    }
    if (rk4) {
      // RK4.m:9-14: a1 = a + D(a) dt, into f1 (a stays in f0)
      setbound(f0, lev);
      chunks2(f0, f1, [&](int i, C& a0, C& a1, C& b0, C& b1) __attribute__((always_inline)) {
        C w0, w1;
        noise_chunk(i, q, coarse, w0, w1);
        b0 = a0 + hdt * D(a0, w0);
        b1 = a1 + hdt * D(a1, w1);
      });
    }
    // RK4: P a1 | swap | P a, stage 2 | P a3 | swap | P R
    // MP : swap | P a, midpoint passes | P a | swap
    // Dense-transform and 2D FFT kernels run every propagation through one
    // code site (a runtime loop, swapping roles): the step then fits the
    // instruction cache (measured at 128x128: 3.9 s -> 3.25 s per launch with
    // 1024 threads); 1D FFT kernels, whose transforms are small, unroll it.
    const int nprop = rk4 ? 4 : 2;
    // RK4.m:18-32 per site with A in `a` and A1 in `b`: a3 -> a, R -> b
    auto rk4_stage2 = [&](C (&a)[PPT], C (&b)[PPT]) __attribute__((always_inline)) {
      chunks2(a, b, [&](int i, C& a0, C& a1, C& b0, C& b1) __attribute__((always_inline)) {
        C w0, w1;
        noise_chunk(i, q, coarse, w0, w1);
        auto site = [&](int v, C& av, C& bv, C w) __attribute__((always_inline)) {
          const C A0 = av, A1 = bv;
          const bool dz = MAT && A.dirichlet && dir_site(i, v);
          C a2 = A0 + hdt * D(A1, w);
          if (dz) a2 = dir_value(i, v, lev + hl);
          C a3 = A0 + (R(2) * hdt) * D(a2, w);
          if (dz) a3 = dir_value(i, v, lev + hl);
          bv = R(1.0 / 3.0) * (A1 + R(2) * a2 + a3 - A0);
          av = a3;
        };
        site(0, a0, b0, w0);
        site(1, a1, b1, w1);
      });
    };
    // MP.m:13-24: fixed-point passes at the midpoint, per site, on f
    auto mp_stage = [&](C (&f)[PPT]) __attribute__((always_inline)) {
      chunks1(f, [&](int i, C& x0, C& x1) __attribute__((always_inline)) {
        C w0, w1;
        noise_chunk(i, q, coarse, w0, w1);
        auto site = [&](int v, C& xv, C w) __attribute__((always_inline)) {
          const C a0 = xv;
          const bool dz = MAT && A.dirichlet && dir_site(i, v);
          C a = a0;
          for (int it = 0; it < A.iterations; ++it) {
            a = a0 + hdt * D(a, w);
            if (dz) a = dir_value(i, v, lev + hl);
          }
          xv = R(2) * a - a0;
        };
        site(0, x0, w0);
        site(1, x1, w1);
      });
    };
    // FFT kernels: A = P a in f0, A1 = P a1 in f1 -> a3 in f0, R in f1
    auto swap_free_stage2 = [&]() __attribute__((always_inline)) { rk4_stage2(f0, f1); };
    auto one = [&](int sp) __attribute__((always_inline)) {
      if (!rk4 && sp == 0) swap01();
      // RK4: P a1, P a at the start, P a3, P R at the end; MP: start, end
      propagate(f1, (rk4 ? sp >= 2 : sp == 1) ? lev + 2 * hl : lev);
      if (rk4 ? (sp % 2 == 0) : (sp == 1)) swap01();
      if (rk4 && sp == 1) rk4_stage2(f1, f0);
      else if (!rk4 && sp == 0) mp_stage(f1);
    };
    if constexpr (MAT || NY > 1) {
      for (int sp = 0; sp < nprop; ++sp) one(sp);
    } else if (rk4) {
      // the same sequence written out, with no role swaps
      propagate(f0, lev);
      propagate(f1, lev);
      swap_free_stage2();
      propagate(f0, lev + 2 * hl);  // P a3
      propagate(f1, lev + 2 * hl);  // P R
    } else {
      propagate(f0, lev);
      mp_stage(f0);
      propagate(f0, lev + 2 * hl);
    }
    if (rk4) {
      // RK4.m:33-41: a = P R + D(P a3) dt/3 with P a3 in f0, P R in f1
      chunks2(f0, f1, [&](int i, C& a0, C& a1, C& b0, C& b1) __attribute__((always_inline)) {
        C w0, w1;
        noise_chunk(i, q, coarse, w0, w1);
        a0 = b0 + (hdt / R(3)) * D(a0, w0);
        a1 = b1 + (hdt / R(3)) * D(a1, w1);
      });
      setbound(f0, lev + 2 * hl);
    }
  };

  // one transform instance for the observable-side ops (corr, noise modes) of f1
  auto obs_transform = [&](int mode, double* acc) __attribute__((always_inline)) {
    const R fs = (R)A.feat_scale;
    ObsOp<R> op{mode, mode == 0 ? R(1.0 / ((double)NS * (double)NS)) : (mode == 2 ? fs * fs : fs), NX, A.fmap,
                A.feat + ((long long)pidx * A.E + group * TPB) * 2 * A.nmodes, acc, A.kscale * A.inv_E,
                group * TPB, A.E};
    periodic_spectral<R, NX, NY, NT, TPB, MAT, RX, KN>(L, f1, twx, twy, op);
  };

  // Fourier modes of the initial noise (coherent initial form: a = alpha + n/2)
  if (A.nmodes > 0) {
    const C alpha = C{(R)A.ip[0], (R)A.ip[1]};
    XSH_FOR(k, 0, PPT) f1[k] = R(2) * (f0[k] - alpha); XSH_END
    obs_transform(1, nullptr);
  }

  auto observe = [&](int n) __attribute__((always_inline)) {
    if (A.nonfinite) {  // divergence check: any non-finite site of this thread's chunks
      bool bad = false;
      XSH_FOR(k, 0, PPT) bad = bad || not_finite(f0[k]); XSH_END
      if (bad) atomicOr(A.nonfinite, 1);
    }
    if (A.raw) {  // xpath.m:73-78: the fields themselves at this point
      XSH_FOR(i, 0, CPT)
        const int pt = Lat::point(i, 0);
        const int p = pt / NS, s = pt % NS;
        const int traj = group * TPB + p;
        if (traj < A.E) {
          XSH_FOR(v, 0, 2)
            const int ps = phys(s + v);
            if (!MAT || ps >= 0) {
              const long long g = (((long long)pidx * A.npts + n) * A.E + traj) * NSP + ps;
              A.raw[2 * g] = (double)f0[2 * i + v].x;
              A.raw[2 * g + 1] = (double)f0[2 * i + v].y;
            }
          XSH_END
        }
      XSH_END
    }
    for (int o = 0; o < A.nobs; ++o) {
      const int kind = A.obs[o];
      double* dst = av + 2 * A.obs_off[o];
      if (kind == XSH_OBS_ABS2K) {
        // k-space |a_k|^2 (xnft.m): accumulated by the transform's middle operation
        XSH_FOR(k, 0, PPT) f1[k] = f0[k]; XSH_END
        obs_transform(3, dst + 2 * (long long)n * NS);
      } else if (kind == XSH_OBS_A || kind == XSH_OBS_ABS2 || kind == XSH_OBS_CORR) {
        if (kind == XSH_OBS_CORR) {
          XSH_FOR(k, 0, PPT) f1[k] = f0[k]; XSH_END
          obs_transform(0, nullptr);
        }
        XSH_FOR(i, 0, CPT)
          const int pt = Lat::point(i, 0);
          const int p = pt / NS, s = pt % NS;
          if (group * TPB + p < A.E) {
            XSH_FOR(v, 0, 2)
              // select values, not array addresses: a pointer select would keep
              // f0 and f1 in scratch memory instead of VGPRs
              const C a0v = f0[2 * i + v], a1v = f1[2 * i + v];
              const C a = kind == XSH_OBS_CORR ? a1v : a0v;
              double re = (double)a.x, im = (double)a.y;
              if (kind == XSH_OBS_ABS2) {
                re = (double)a.x * a.x + (double)a.y * a.y;
                im = 0.0;
              }
              const int ps = phys(s + v);
              if (!MAT || ps >= 0) atomic_add_c(dst + 2 * ((long long)n * NSP + ps), re * A.inv_E, im * A.inv_E);
            XSH_END
          }
        XSH_END
      } else {
        if (kind == XSH_OBS_LOWK) {
          XSH_FOR(k, 0, PPT) f1[k] = f0[k]; XSH_END
          obs_transform(2, nullptr);
        }
        double re = 0.0, im = 0.0;
        XSH_FOR(i, 0, CPT)
          const int pt = Lat::point(i, 0);
          const int p = pt / NS, s = pt % NS;
          if (group * TPB + p < A.E) {
            XSH_FOR(v, 0, 2)
              const C a = !MAT || phys(s + v) >= 0 ? f0[2 * i + v] : C{R(0), R(0)};
              const double m = (double)a.x * a.x + (double)a.y * a.y;
              if (kind == XSH_OBS_LOWK) {
                if (s + v == 0) re += (double)f1[2 * i + v].x;
              } else if (kind == XSH_OBS_A_AVE) {
                re += (double)a.x;
                im += (double)a.y;
              } else if (kind == XSH_OBS_ABS4_AVE) {
                re += m * m;
              } else {
                re += m;
              }
            XSH_END
          }
        XSH_END
        block_sum2<NT>(re, im, red);
        const double sc0 = kind == XSH_OBS_ABS2_INT ? A.dv : (kind == XSH_OBS_LOWK ? 1.0 : 1.0 / NSP);
        const double sc = sc0 * A.inv_E;
        if (threadIdx.x == 0) {
          atomic_add_c(dst + 2 * (long long)n, re * sc, im * sc);
          if (A.traj && TPB == 1 && group < A.E)  // per-trajectory lattice mean
            A.traj[(((long long)pidx * A.E + group) * A.npts + n) * A.nobs + o] = re * sc0;
        }
      }
    }
  };

  // ---- the time loop of xpath.m:29-86.  One method_step code site for the
  // fine and the coarse runs (coarse: one step per two fine draws, xnoise.m:33-34):
  // the step is the kernel's largest code, and one copy keeps it in the
  // instruction cache.
  const bool coarse = A.checks && nci == 1;
  const int nsub = coarse ? 1 : 1 + A.checks;
  uint32_t q = 0;
  int lev = 0;
  for (int n = 0; n < A.npts; ++n) {
    for (int st = 0; st < A.steps; ++st) {
      if (n > 0 || st > 0 || A.npts == 1) {
#pragma unroll 1
        for (int c = 0; c < nsub; ++c) {
          method_step(q, coarse, lev);
          q += coarse ? 2 : 1;
          lev += coarse ? 4 : 2;
        }
      }
      if (st == 0) observe(n);
    }
  }
  if (A.a_out) {
    XSH_FOR(i, 0, CPT)
      const int pt = Lat::point(i, 0);
      const int p = pt / NS, s = pt % NS;
      const int traj = group * TPB + p;
      if (traj < A.E) {
        XSH_FOR(v, 0, 2)
          const int ps = phys(s + v);
          if (!MAT || ps >= 0) {
            const long long g = (long long)traj * NSP + ps;
            A.a_out[2 * g] = (double)f0[2 * i + v].x;
            A.a_out[2 * g + 1] = (double)f0[2 * i + v].y;
          }
        XSH_END
      }
    XSH_END
  }
#undef XSH_FOR
#undef XSH_END
}

// Whole-path kernel over a batch of instances sharing one lattice kernel:
// work items [wstart[i], wstart[i+1]) belong to instance i (args[i]); a
// persistent grid strides over all nwork items.  One instance: ninst = 1,
// wstart unused.
template <class R, int NX, int NY, int NT, int TPB, bool MAT, int RX, bool KN>
__global__ __launch_bounds__(NT) void spde_path(const PathArgs* __restrict__ args, const int* __restrict__ wstart,
                                                int ninst, int nwork) {
  __shared__ cpx<R> L[lds_elems<R, NX, NY, NT, TPB, MAT, RX, KN>()];
  __shared__ double red[2 * (NT / 64)];
#pragma unroll 1
  for (int wk = blockIdx.x; wk < nwork; wk += gridDim.x) {
    const int inst = ninst > 1 ? find_instance(wstart, ninst, wk) : 0;
    spde_work<R, NX, NY, NT, TPB, MAT, RX, KN>(args[inst], wk - (ninst > 1 ? wstart[inst] : 0), L, red);
    __syncthreads();
  }
}

// propagation probe: planes [nplanes][NS] complex R in place
template <class R, int NX, int NY, int NT, int TPB, bool MAT, int RX>
__global__ __launch_bounds__(NT) void probe_propagate(const PathArgs* __restrict__ Ap, int nci, int nplanes, cpx<R>* planes) {
  const PathArgs& A = *Ap;
  using Lat = Lattice<R, NX, NY, NT, TPB, RX>;
  using C = cpx<R>;
  __shared__ C L[lds_elems<R, NX, NY, NT, TPB, MAT, RX, false>()];
  C f[Lat::PPT];
  const long long base = (long long)blockIdx.x * Lat::NPTS;
#pragma unroll
  for (int i = 0; i < Lat::CPT; ++i)
#pragma unroll
    for (int v = 0; v < 2; ++v) {
      const long long g = base + Lat::point(i, v);
      f[2 * i + v] = g < (long long)nplanes * Lat::NS ? planes[g] : C{R(0), R(0)};
    }
  const auto prop = make_prop<R, NX, NY, NT, TPB, MAT, RX, false>(A, nci);
  if constexpr (MAT)
    Lat::spectral_mat(L, f, A.tm, A.tmp, A.cplx[0], A.cplx[1], prop, A.tsym);
  else
    periodic_spectral<R, NX, NY, NT, TPB, MAT, RX, false>(L, f, static_cast<const C*>(A.twx),
                                                          static_cast<const C*>(A.twy), prop);
#pragma unroll
  for (int i = 0; i < Lat::CPT; ++i)
#pragma unroll
    for (int v = 0; v < 2; ++v) {
      const long long g = base + Lat::point(i, v);
      if (g < (long long)nplanes * Lat::NS) planes[g] = f[2 * i + v];
    }
}

template <class R>
__global__ void probe_normals(uint32_t nsp, uint32_t tag, uint32_t seq, uint32_t inst, uint32_t draw,
                              int traj0, int ntraj, int count, double* out) {
  const int id = blockIdx.x * blockDim.x + threadIdx.x;
  const int groups = (count + 3) / 4;
  if (id >= ntraj * groups) return;
  const int t = id / groups, g = id % groups;
  R n[4];
  normals4<R>(u4{(uint32_t)g, draw, (uint32_t)(traj0 + t), inst}, nsp, key1(tag, seq), n);
  for (int l = 0; l < 4; ++l)
    if (4 * g + l < count) out[(long long)t * count + 4 * g + l] = (double)n[l];
}

// ================================================================ launchers
template <class R>
static hipError_t launch_sde(const PathArgs* A, int nblocks, hipStream_t s) {
  hipLaunchKernelGGL(sde_path<R>, dim3(nblocks), dim3(256), 0, s, A);
  return hipGetLastError();
}

template <class R, int F>
static hipError_t launch_sde_f(const PathArgs* A, int nblocks, hipStream_t s) {
  hipLaunchKernelGGL((sde_fields<R, F>), dim3(nblocks), dim3(256), 0, s, A);
  return hipGetLastError();
}

template <class R, int NX, int NY, int NT, int TPB, bool MAT, int RX, bool KN>
static hipError_t launch_spde(const PathArgs* args, const int* wstart, int ninst, int nwork, int grid,
                              hipStream_t s) {
  hipLaunchKernelGGL((spde_path<R, NX, NY, NT, TPB, MAT, RX, KN>), dim3(grid), dim3(NT), 0, s, args, wstart, ninst,
                     nwork);
  return hipGetLastError();
}

template <class R, int NX, int NY, int NT, int TPB, bool MAT, int RX>
static hipError_t launch_probe(const PathArgs* A, int nci, int nplanes, void* planes, hipStream_t s) {
  const int nb = (nplanes + TPB - 1) / TPB;
  hipLaunchKernelGGL((probe_propagate<R, NX, NY, NT, TPB, MAT, RX>), dim3(nb), dim3(NT), 0, s, A, nci, nplanes,
                     static_cast<cpx<R>*>(planes));
  return hipGetLastError();
}

#define XSH_ENTRY(R, NX, NY, NT, TPB) XSH_VAR_ENTRY(R, NX, NY, NT, TPB, 0)
// a variant with first x radix RX (0: default split)
#define XSH_VAR_ENTRY(R, NX, NY, NT, TPB, RX)                                                                    \
  {sizeof(R) == 8 ? XSH_FP64 : XSH_FP32, NX, NY, NT, TPB, 0, &launch_spde<R, NX, NY, NT, TPB, false, RX, true>, \
   &launch_probe<R, NX, NY, NT, TPB, false, RX>, 0, Split<NX, RX>::N1, 1}
// without the k-space noise code (its register and instruction footprint
// costs the periodic 128x128 FP32 kernel 1 %): looked up first, skipped for
// problems with knoises or krandoms
#define XSH_FAST_ENTRY(R, NX, NY, NT, TPB)                                                                    \
  {sizeof(R) == 8 ? XSH_FP64 : XSH_FP32, NX, NY, NT, TPB, 0, &launch_spde<R, NX, NY, NT, TPB, false, 0, false>, \
   &launch_probe<R, NX, NY, NT, TPB, false, 0>, 0, Split<NX>::N1, 0}
#define XSH_MAT_ENTRY(R, NX, NY, NT, TPB)                                                                   \
  {sizeof(R) == 8 ? XSH_FP64 : XSH_FP32, NX, NY, NT, TPB, 1, &launch_spde<R, NX, NY, NT, TPB, true, 0, false>, \
   &launch_probe<R, NX, NY, NT, TPB, true, 0>, 0, 0, 0}

}  // namespace xsh
