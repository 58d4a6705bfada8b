// C ABI of the ensemble integration path (include/xspde_hip.h).
//
// Host side of the drop-in: descriptor validation, the lattice and scale
// derivations of xpreferences.m:133-175 / spdpreferences.m:16-72, the
// separable propagator tables of xpropfactor.m:26-39, and the launch of
// whole-path kernels for one path (p.path, xensemble.m:79) or all
// ensembles(2) x checks paths of one parallel ensemble (xensemble.m:25-80).
#include <hip/hip_runtime.h>

#include <cmath>
#include <complex>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/xspde_hip.h"
#include "xsh_launch.hpp"
#include "xsh_path.hpp"

using cd = std::complex<double>;

namespace {
thread_local std::string g_err;

int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

#define HIPCHK(expr)                                                                  \
  do {                                                                                \
    hipError_t e_ = (expr);                                                           \
    if (e_ != hipSuccess) return fail(XSH_EDEVICE, "%s: %s", #expr, hipGetErrorString(e_)); \
  } while (0)

bool is_pow2(int n) { return n > 0 && (n & (n - 1)) == 0; }

// clears device doubles (a kernel, not a memset node, so captured graphs replay it)
__global__ void zero_doubles(double* p, long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.0;
}
}  // namespace

struct xsh_handle {
  xsh_desc d;
  int nd, nx, ny, ns, E, e2, e3, nchk;
  double dt, dtr[2], dx[2], dv, x0[2], nscale, noisefactor;
  const xsh::KernelEntry* kern;
  int groups, threads, tpb;
  long long avlen;
  long long obs_off[XSH_MAX_OBS];
  int obs_ns[XSH_MAX_OBS];
  size_t esize;
  void* d_tables;   // px[2], py[2], twx, twy, transform matrices, half-length twiddles,
                    // matrix-core operand copies of the real matrices
  size_t tab_off[19];
  bool has_tmp[4];
  int tsym[4];         // per packed matrix: even/odd split (1 DCT-I, 2 DST-I) or 0
  bool has_patch;      // non-zero boundary values on a non-periodic dimension
  int bval_level;      // boundary-value table: elements per time level (0: one level)
  int bval_nb;         // values per boundary line
  int bshape_n;        // entries per patch shape function
  double lpf[2][2][2]; // [nc][dim] Neumann correction per unit of b_upper - b_lower
  void* d_work;      // streaming kernels: field planes per workgroup
  void* d_kwork;     // k-space noise: x-space noise plane per trajectory of a launch
  int work_planes;
  int btype[2];      // boundary type per space dimension (0 periodic)
  int npad[2];       // sites per dimension of the kernel's lattice (> nx, ny: padded dense-transform lattice)
  int mat;           // dense per-dimension transforms in use
  int cplx[2];       // per dimension: complex (DFT) matrix
  double* d_av;      // [e2*nchk][avlen] complex
  double* d_path_av; // [avlen] complex (single path)
  double* d_a;       // [E][ns] complex (single path in/out)
  double* d_traj;    // [e2*nchk][E][points][nobs] per-trajectory lattice means
  double* d_raw;     // [e2*nchk][points][E][ns] complex raw fields
  float* d_feat;     // [e2*nchk][E][2*nmodes] noise modes
  int* d_fmap;       // [nx*ny] mode index per FFT position
  double* d_ft;      // SDE spectra: [e2*nchk][points*steps][E] complex step fields
  double* d_twt;     // [points*steps] complex exp(2 pi i k / M)
  int* d_nonfinite;  // [2] ints: [0] set by the kernels when a stored field value is not finite
  int tspec_mask;    // bit o: observable o of the time transform
  double fsc;        // sqrt(2 pi) / dw (xpreferences.m:199-200)
  hipStream_t stream;
  hipEvent_t ev0, ev1;
  xsh::PathArgs* d_args;   // kernel arguments in device memory
  xsh::PathArgs* h_args;   // pinned staging copy
  hipEvent_t ev_args;      // recorded after the last argument upload
  bool args_pending;
  bool ran;
  int device;              // device current at xsh_create
};

extern "C" int xsh_version(void) { return XSH_ABI_VERSION; }
extern "C" const char* xsh_last_error(void) { return g_err.c_str(); }

static int validate(const xsh_desc* d) {
  if (!d) return fail(XSH_EINVAL, "descriptor is NULL");
  if (d->struct_size != (int)sizeof(xsh_desc))
    return fail(XSH_EINVAL, "descriptor size %d != %d (ABI mismatch)", d->struct_size, (int)sizeof(xsh_desc));
  if (d->precision != XSH_FP32 && d->precision != XSH_FP64)
    return fail(XSH_EINVAL, "precision must be XSH_FP32 or XSH_FP64");
  if (d->dimensions < 1 || d->dimensions > 3) return fail(XSH_EUNSUPPORTED, "dimensions %d not on the GPU path (1..3)", d->dimensions);
  if (d->points < 1) return fail(XSH_EINVAL, "points must be >= 1");
  if (d->steps < 1) return fail(XSH_EINVAL, "steps must be >= 1");
  if (d->checks != 0 && d->checks != 1) return fail(XSH_EINVAL, "checks(1) must be 0 or 1");
  if (d->method != XSH_RK4 && d->method != XSH_MP) return fail(XSH_EUNSUPPORTED, "method %d is not RK4 or MP", d->method);
  if (d->method == XSH_MP && d->iterations < 1) return fail(XSH_EINVAL, "iterations must be >= 1");
  for (int i = 0; i < 3; ++i)
    if (d->ensembles[i] < 1) return fail(XSH_EINVAL, "ensembles(%d) must be >= 1", i + 1);
  if (d->fields < 0 || d->fields > XSH_MAX_FIELDS)
    return fail(XSH_EUNSUPPORTED, "fields %d: the GPU path integrates 1..%d components", d->fields, XSH_MAX_FIELDS);
  const int nf = d->fields > 1 ? d->fields : 1;
  if (d->noises < 0 || d->noises > (nf > 1 ? 4 : 2))
    return fail(XSH_EUNSUPPORTED, "noises %d: built-in forms use 0..%d", d->noises, nf > 1 ? 4 : 2);
  if (nf > 1) {
    // coupled components: the matrix drift on SDEs (lattice projects with several
    // components split them into one-component descriptors)
    if (d->dimensions != 1)
      return fail(XSH_EUNSUPPORTED, "coupled field components run on SDEs (dimensions = 1); a lattice project's "
                                    "uncoupled components run as one descriptor each");
    if (d->deriv != XSH_DERIV_MATRIX) return fail(XSH_EINVAL, "fields > 1 needs the matrix drift form");
    if (d->initial != XSH_INIT_CONST) return fail(XSH_EUNSUPPORTED, "coupled components start from constants (or a0)");
    for (int o = 0; o < d->nobserve && o < XSH_MAX_OBS; ++o) {
      const int k = d->observe[o];
      if (k != XSH_OBS_A && k != XSH_OBS_A_AVE && k != XSH_OBS_ABS2 && k != XSH_OBS_ABS2_AVE && k != XSH_OBS_ABS4_AVE)
        return fail(XSH_EUNSUPPORTED, "observe form %d: coupled components observe a_i, |a_i|^2 or |a_i|^4", k);
      if (d->obs_field[o] < 0 || d->obs_field[o] >= nf)
        return fail(XSH_EINVAL, "obs_field[%d] = %d outside 0..%d", o, d->obs_field[o], nf - 1);
      if (d->tspec[o]) return fail(XSH_EUNSUPPORTED, "time-domain spectra of coupled components are not on the GPU path");
    }
  } else if (d->deriv == XSH_DERIV_MATRIX) {
    return fail(XSH_EINVAL, "the matrix drift form couples fields >= 2 components");
  }
  if (d->inrandoms < 0 || d->urandoms < 0) return fail(XSH_EINVAL, "negative random count");
  if (d->deriv < XSH_DERIV_ZERO || d->deriv > XSH_DERIV_MATRIX) return fail(XSH_EUNSUPPORTED, "unknown drift form %d", d->deriv);
  if (d->deriv == XSH_DERIV_KUBO && d->noises < 1) return fail(XSH_EINVAL, "Kubo drift needs noises >= 1");
  if (d->initial < XSH_INIT_CONST || d->initial > XSH_INIT_COHERENT) return fail(XSH_EUNSUPPORTED, "unknown initial form %d", d->initial);
  if (d->knoises < 0 || d->knoises > 2) return fail(XSH_EUNSUPPORTED, "knoises %d: built-in forms use 0..2", d->knoises);
  if (d->krandoms != 0 && d->krandoms != 2)
    return fail(XSH_EUNSUPPORTED, "krandoms %d: the initial forms read v1 + i v2 (0 or 2)", d->krandoms);
  if (d->knoises && d->noises)
    return fail(XSH_EUNSUPPORTED, "Gaussian and k-space noise cells together are not a built-in drift form");
  if (d->krandoms && d->inrandoms)
    return fail(XSH_EUNSUPPORTED, "Gaussian and k-space random cells together are not a built-in initial form");
  if ((d->knoises || d->krandoms) && d->dimensions < 2) return fail(XSH_EINVAL, "k-space noises need a lattice");
  if ((d->knoises && !d->nfilter) || (d->krandoms && !d->rfilter))
    return fail(XSH_EINVAL, "k-space noises need their filter table");
  if ((d->initial == XSH_INIT_COHERENT || (d->initial == XSH_INIT_SECH && d->init_par[2] != 0.0)) &&
      d->inrandoms < 2 && d->krandoms < 2)
    return fail(XSH_EINVAL, "initial form needs inrandoms >= 2 (or krandoms = 2)");
  if (d->initial == XSH_INIT_PHASE && d->urandoms < 1) return fail(XSH_EINVAL, "phase initial form needs urandoms >= 1");
  if (d->initial == XSH_INIT_SECH && !(d->init_par[1] > 0)) return fail(XSH_EINVAL, "sech width must be > 0");
  if (d->initial == XSH_INIT_GAUSS && (!(d->init_par[1] > 0) || d->init_par[2] < 0))
    return fail(XSH_EINVAL, "gauss widths must be > 0");
  if (d->nobserve < 1 || d->nobserve > XSH_MAX_OBS) return fail(XSH_EINVAL, "nobserve must be 1..%d", XSH_MAX_OBS);
  for (int o = 0; o < d->nobserve; ++o) {
    if (d->observe[o] == XSH_OBS_LOWK && d->nmodes < 1) return fail(XSH_EINVAL, "observe form lowk needs noise modes");
    if (d->observe[o] < XSH_OBS_A || d->observe[o] > XSH_OBS_ABS2K) return fail(XSH_EUNSUPPORTED, "unknown observe form %d", d->observe[o]);
  }
  const int sdim = d->dimensions - 1;
  const int n[2] = {d->nx, d->ny};
  for (int k = 0; k < 2; ++k) {
    if (k < sdim) {
      if (n[k] < 2) return fail(XSH_EINVAL, "lattice size %d must be >= 2", n[k]);
      const int b0 = d->boundaries[k][0], b1 = d->boundaries[k][1];
      if (b0 < -1 || b0 > 1 || b1 < -1 || b1 > 1) return fail(XSH_EINVAL, "Invalid boundary: [%d,%d]", b0, b1);
      const int t = b0 + 2 * b1;  // type index of spdpreferences.m:53-54
      if (t != -3 && t != -1 && t != 0 && t != 1 && t != 3)
        return fail(XSH_EINVAL, "Invalid boundary type: [%d,%d]", b0, b1);
      // sizes without a compiled kernel run padded on a dense-transform kernel
      // (1D up to 256 sites, 2D up to 64 x 64, 128 x 64 and 64 x 128; kernel
      // selection reports the rest)
      if (n[k] % 2 && d->dimensions == 3 && n[k] > 128)
        return fail(XSH_EUNSUPPORTED, "odd lattice size %d above 128 is not on the GPU path in 2D", n[k]);
      if (n[k] < 2) return fail(XSH_EINVAL, "lattice size %d: at least 2 points", n[k]);
      if (!(d->ranges[k + 1] > 0)) return fail(XSH_EINVAL, "ranges(%d) must be > 0", k + 2);
    } else if (n[k] != 1) {
      return fail(XSH_EINVAL, "lattice size beyond dimensions must be 1");
    }
  }
  if (!(d->ranges[0] > 0) && d->points > 1) return fail(XSH_EINVAL, "ranges(1) must be > 0");
  if (d->traj_out != 0 && d->traj_out != 1) return fail(XSH_EINVAL, "traj_out must be 0 or 1");
  if (d->raw != 0 && d->raw != 1) return fail(XSH_EINVAL, "raw must be 0 or 1");
  if (d->raw && d->dimensions < 2) return fail(XSH_EUNSUPPORTED, "raw fields are kept for lattice problems");
  if (d->boundtab) {
    const long long calls = d->points > 1 ? (long long)d->points * d->steps - 1 : d->steps;
    const long long full = 2 * calls * (1 + d->checks) + 1;
    if (d->boundlevels != 1 && d->boundlevels != full)
      return fail(XSH_EINVAL, "boundlevels %d: must be 1 or %lld (one per half fine step)", d->boundlevels, full);
    const long long nb = d->dimensions == 3 ? std::max(n[0], n[1]) : 1;
    if ((long long)d->boundlevels * 4 * nb * 16 > (1LL << 31))
      return fail(XSH_EUNSUPPORTED, "boundary-value table beyond 2 GiB");
  }
  int tspec = 0;
  for (int o = 0; o < d->nobserve; ++o) {
    if (d->tspec[o] != 0 && d->tspec[o] != 1) return fail(XSH_EINVAL, "tspec[%d] must be 0 or 1", o);
    tspec |= d->tspec[o];
  }
  if (tspec && d->dimensions != 1)
    return fail(XSH_EUNSUPPORTED, "time-domain spectra (p.transforms{o}(1)) run on SDEs (dimensions = 1)");
  if (d->nmodes < 0 || d->nmodes > XSH_MAX_MODES) return fail(XSH_EINVAL, "nmodes must be 0..%d", XSH_MAX_MODES);
  if (d->nmodes > 0) {
    if (d->initial != XSH_INIT_COHERENT) return fail(XSH_EUNSUPPORTED, "noise modes need the coherent initial form");
    if (d->dimensions < 2) return fail(XSH_EINVAL, "noise modes need a lattice");
    for (int m = 0; m < d->nmodes; ++m) {
      const int kx = d->modes[m][0], ky = d->modes[m][1];
      if (kx < 0 || kx >= d->nx || ky < 0 || ky >= (d->dimensions > 2 ? d->ny : 1))
        return fail(XSH_EINVAL, "mode %d (%d,%d) outside the lattice", m, kx, ky);
      for (int m2 = 0; m2 < m; ++m2)
        if (d->modes[m2][0] == kx && d->modes[m2][1] == ky) return fail(XSH_EINVAL, "mode %d repeated", m);
    }
  }
  return XSH_OK;
}

template <class R>
static void pack(std::vector<char>& buf, size_t off, const std::vector<cd>& v) {
  R* p = reinterpret_cast<R*>(buf.data() + off);
  for (size_t i = 0; i < v.size(); ++i) {
    p[2 * i] = (R)v[i].real();
    p[2 * i + 1] = (R)v[i].imag();
  }
}

template <class R>
static void pack_real(std::vector<char>& buf, size_t off, const std::vector<double>& v) {
  R* p = reinterpret_cast<R*>(buf.data() + off);
  for (size_t i = 0; i < v.size(); ++i) p[i] = (R)v[i];
}

// Matrices of the reference's spectral transforms on N points, M = N - 1
// intervals, row-major T[k*N + n] so that out[k] = sum_n T[k][n] in[n]:
//   DCT1 xdct1.m:14-20   (1/sqrt(2M)) c_n cos(pi n k / M), c_0 = c_M = 1, else 2
//   DST1 xdst1.m:8-18    sqrt(2/M) sin(pi n k / M), interior n and k only
//   DCT2 xdct2.m:15-29   sqrt(2/M) cos(pi k (n+1/2) / M), n, k < M
//   DCT3 xdct3.m:17-33   sqrt(2/M) [n = 0: 1/2, else cos(pi n (k+1/2) / M)], n, k < M
//   DST2 xdst2.m:17-33   sqrt(2/M) sin(pi (n-1/2) k / M), n, k >= 1
//   DST3 xdst3.m:16-35   sqrt(2/M) [n = M: (-1)^(k-1)/2, else sin(pi n (k-1/2) / M)], n, k >= 1
enum { XT_DCT1 = 1, XT_DCT2, XT_DCT3, XT_DST1, XT_DST2, XT_DST3 };
static std::vector<double> trig_matrix(int kind, int N) {
  const int M = N - 1;
  std::vector<double> T((size_t)N * N, 0.0);
  const double s = std::sqrt(2.0 / M);
  for (int k = 0; k < N; ++k)
    for (int n = 0; n < N; ++n) {
      double v = 0.0;
      switch (kind) {
        case XT_DCT1:
          v = (n == 0 || n == M ? 1.0 : 2.0) * std::cos(M_PI * n * k / M) / std::sqrt(2.0 * M);
          break;
        case XT_DST1:
          if (n >= 1 && n <= M - 1 && k >= 1 && k <= M - 1) v = s * std::sin(M_PI * n * k / M);
          break;
        case XT_DCT2:
          if (n < M && k < M) v = s * std::cos(M_PI * k * (n + 0.5) / M);
          break;
        case XT_DCT3:
          if (n < M && k < M) v = s * (n == 0 ? 0.5 : std::cos(M_PI * n * (k + 0.5) / M));
          break;
        case XT_DST2:
          if (n >= 1 && k >= 1) v = s * std::sin(M_PI * (n - 0.5) * k / M);
          break;
        case XT_DST3:
          if (n >= 1 && k >= 1) v = s * (n == M ? 0.5 * ((k - 1) % 2 ? -1.0 : 1.0) : std::sin(M_PI * n * (k - 0.5) / M));
          break;
      }
      T[(size_t)k * N + n] = v;
    }
  return T;
}
// DFT on N points as an interleaved complex matrix: forward exp(-2 pi i nk/N),
// inverse exp(+2 pi i nk/N)/N (fft / ifft of xtransform.m:45,60)
static std::vector<double> dft_matrix(int N, bool inverse) {
  std::vector<double> T(2 * (size_t)N * N);
  for (int k = 0; k < N; ++k)
    for (int n = 0; n < N; ++n) {
      const long long m = ((long long)n * k) % N;
      cd w = std::polar(1.0, (inverse ? 2.0 : -2.0) * M_PI * m / N);
      if (inverse) w /= (double)N;
      T[2 * ((size_t)k * N + n)] = w.real();
      T[2 * ((size_t)k * N + n) + 1] = w.imag();
    }
  return T;
}

// Everything xsh_create derives on the host before it touches a device:
// the lattice (xpreferences.m:133-175, spdpreferences.m:16-72), the kernel
// for it, the averages layout and every device table (propagators
// xpropfactor.m:23-58, twiddles, transform matrices and their matrix-core
// operand copies, boundary patches and values, noise filters) packed into
// host.  No HIP call: xsh_plan runs it alone.
static int plan_handle(xsh_handle* h, const xsh_desc* desc, std::vector<char>& host, size_t& total) {
  h->d = *desc;
  const xsh_desc& d = h->d;
  h->nd = d.dimensions;
  h->nx = d.dimensions > 1 ? d.nx : 1;
  h->ny = d.dimensions > 2 ? d.ny : 1;
  h->ns = h->nx * h->ny;
  if (d.dimensions == 1 && d.fields > 1) h->ns = d.fields;  // a0, a_out: [E][fields]
  h->E = d.ensembles[0];
  h->e2 = d.ensembles[1];
  h->e3 = d.ensembles[2];
  h->nchk = 1 + d.checks;
  // xpreferences.m:138-141,174-175 ; xensemble.m:41-45
  h->dt = d.points > 1 ? d.ranges[0] / (d.points - 1) : d.ranges[0];
  h->dtr[0] = d.checks ? h->dt / (2.0 * d.steps) : h->dt / d.steps;
  h->dtr[1] = h->dt / d.steps;
  const int npt[2] = {h->nx, h->ny};
  h->dv = 1.0;
  for (int k = 0; k < 2; ++k) {
    h->dx[k] = (k < d.dimensions - 1) ? d.ranges[k + 1] / (npt[k] - 1) : 1.0;
    h->x0[k] = (k < d.dimensions - 1) ? d.origins[k + 1] : 0.0;
    if (k < d.dimensions - 1) h->dv *= h->dx[k];
  }
  h->noisefactor = 1.0 / h->dv;
  h->nscale = d.steps * (1 + d.checks) / h->dt;
  h->esize = d.precision == XSH_FP64 ? 16 : 8;

  if (d.dimensions > 1) {
    for (int k = 0; k < d.dimensions - 1; ++k) {
      h->btype[k] = d.boundaries[k][0] + 2 * d.boundaries[k][1];
      if (h->btype[k] != 0) h->mat = 1;
    }
    h->kern = xsh::find_kernel(d.precision, h->nx, h->ny, h->mat, d.knoises || d.krandoms);
    if (!h->kern && !h->mat) h->kern = xsh::find_stream_kernel(d.precision, h->nx, h->ny, 0);
    if (!h->kern && !h->mat) {  // periodic lattice without an FFT kernel: dense DFT
      h->mat = 1;
      h->kern = xsh::find_kernel(d.precision, h->nx, h->ny, 1);
    }
    if (!h->kern && h->mat) {
      // fields too large for the chip: dense transforms on the streaming kernel
      h->kern = xsh::find_stream_kernel(d.precision, h->nx, h->ny, 1);
    }
    if (!h->kern && h->mat && d.dimensions == 2) h->kern = xsh::find_padded_kernel(d.precision, h->nx);
    if (!h->kern && h->mat && d.dimensions == 3) h->kern = xsh::find_padded_kernel2d(d.precision, h->nx, h->ny);
    if (h->kern && (h->kern->nx > h->nx || h->kern->ny > h->ny)) {
      // padded 1D lattice: the outputs that transform the field periodically
      // (corr, abs2k, lowk, noise modes) would see the padding
      bool fft_obs = d.nmodes > 0;
      for (int o = 0; o < d.nobserve; ++o)
        fft_obs = fft_obs || d.observe[o] == XSH_OBS_CORR || d.observe[o] == XSH_OBS_ABS2K ||
                  d.observe[o] == XSH_OBS_LOWK;
      if (fft_obs) {
        const int np = h->kern->nx;
        return fail(XSH_EUNSUPPORTED, "corr, abs2k, lowk and noise modes need a lattice of a compiled size "
                                      "(%d x %d sites run padded to %d)", npt[0], npt[1], np);
      }
    }
    if (!h->kern) {
      return fail(XSH_EUNSUPPORTED, "no %s %s kernel for a %dx%d lattice", d.precision == XSH_FP64 ? "fp64" : "fp32",
                  h->mat ? "non-periodic" : "periodic", npt[0], npt[1]);
    }
    h->threads = h->kern->threads;
    h->tpb = h->kern->tpb;
    h->npad[0] = h->kern->nx;
    h->npad[1] = h->kern->ny;
    if ((d.knoises || d.krandoms) && (h->mat || !(h->kern->kn || h->kern->stream))) {
      return fail(XSH_EUNSUPPORTED, "k-space noises run on periodic lattices (not the dense-transform kernel)");
    }
    // corr, lowk and the noise modes are periodic FFTs of the field whatever the
    // boundaries (the dense-transform kernels have power-of-two lattices and the
    // FFT tables too); the streaming kernel computes corr only
    if (h->kern->stream) {
      bool lowk = false;
      for (int o = 0; o < d.nobserve; ++o) lowk = lowk || d.observe[o] == XSH_OBS_LOWK;
      if (d.nmodes > 0 || lowk) {
        return fail(XSH_EUNSUPPORTED, "noise modes and the lowk observable are not on the streaming %dx%d kernel",
                    npt[0], npt[1]);
      }
    }
  } else {
    h->threads = 256;
    h->tpb = 256;
    h->npad[0] = h->nx;
    h->npad[1] = h->ny;
  }
  h->groups = (h->E + h->tpb - 1) / h->tpb;

  // averages layout: observable o -> points x sites_o complex values
  long long off = 0;
  for (int o = 0; o < d.nobserve; ++o) {
    const int kind = d.observe[o];
    const bool site = kind == XSH_OBS_A || kind == XSH_OBS_ABS2 || kind == XSH_OBS_CORR || kind == XSH_OBS_ABS2K;
    h->obs_ns[o] = site && !(d.dimensions == 1 && d.fields > 1) ? h->ns : 1;
    h->obs_off[o] = off;
    off += (long long)d.points * h->obs_ns[o];
  }
  h->avlen = off;

  // tables: per nc the separable propagator factors (xpropfactor.m:37-39):
  //   P(kx,ky) = exp(dt l0) exp(dt (l1x i kx + l2x (i kx)^2)) exp(dt (l1y i ky + l2y (i ky)^2))
  // with dt = dtr/ipsteps, ipsteps = 2 and the k-grid of each boundary type
  // (spdpreferences.m:56-67).  FFT lattices: k in FFT order, inverse-FFT 1/N
  // folded in, then the twiddles W_N^m.  Dense-transform lattices: k in the
  // natural order of the transform, normalisation inside the matrices.
  const cd l0(d.lin_l0[0], d.lin_l0[1]);
  std::vector<std::vector<cd>> tabs(6);
  for (int nci = 0; nci < 2; ++nci) {
    const double dtp = h->dtr[nci] / 2.0;
    for (int k = 0; k < 2; ++k) {
      const int n = npt[k];
      const bool sp = k < d.dimensions - 1;
      std::vector<cd> t(n);
      const cd l1(d.lin_l1[k][0], d.lin_l1[k][1]), l2(d.lin_l2[k][0], d.lin_l2[k][1]);
      const double dk = 2.0 * M_PI / (n * h->dx[k]);
      const double dkp = n > 1 ? M_PI / ((n - 1) * h->dx[k]) : 0.0;
      for (int m = 0; m < n; ++m) {
        double kk = 0.0;
        if (sp) {
          switch (h->btype[k]) {
            case 0: kk = (m < (n + 1) / 2 ? m : m - n) * dk; break;  // ifftshift((-N/2..N/2-1))
            case -3:
            case 3: kk = m * dkp; break;
            case -1: kk = m == 0 ? 0.0 : (m - 0.5) * dkp; break;
            case 1: kk = m == n - 1 ? 0.0 : (m + 0.5) * dkp; break;
          }
        }
        cd L = l1 * cd(0, kk) + l2 * cd(-kk * kk, 0);
        if (k == 0) L += l0;
        t[m] = std::exp(L * dtp);
        if (sp && !h->mat) t[m] /= (double)n;
      }
      t.resize(h->npad[k], cd(0.0, 0.0));  // padding sites: propagator 0
      tabs[2 * nci + k] = t;
    }
  }
  for (int k = 0; k < 2; ++k) {
    const int n = npt[k];
    std::vector<cd> t(n);
    for (int m = 0; m < n; ++m) t[m] = std::polar(1.0, -2.0 * M_PI * m / n);
    tabs[4 + k] = t;
  }
  // streaming kernels: twiddles of the half-length column transforms
  std::vector<cd> twh(std::max(1, npt[1] / 2));
  for (int m = 0; m < (int)twh.size(); ++m) twh[m] = std::polar(1.0, -2.0 * M_PI * m / (npt[1] / 2 > 0 ? npt[1] / 2 : 1));
  // dense transforms (xtransform.m:33-70): forward x, forward y, inverse x, inverse y;
  // real matrices are stored with a zero imaginary part dropped (one value per entry)
  std::vector<std::vector<double>> mats(4);
  int cplx[2] = {0, 0};
  if (h->mat) {
    for (int k = 0; k < d.dimensions - 1; ++k) {
      const int n = npt[k];
      const int t = h->btype[k];
      if (t == 0) {
        cplx[k] = 1;
        mats[k] = dft_matrix(n, false);
        mats[2 + k] = dft_matrix(n, true);
      } else {
        const int fwd = t == -3 ? XT_DCT1 : t == -1 ? XT_DST3 : t == 1 ? XT_DCT3 : XT_DST1;
        const int inv = t == -3 ? XT_DCT1 : t == -1 ? XT_DST2 : t == 1 ? XT_DCT2 : XT_DST1;
        mats[k] = trig_matrix(fwd, n);
        mats[2 + k] = trig_matrix(inv, n);
      }
    }
    // padded lattice: the n x n matrices in the top-left corner of npad x npad
    // zeros, so the padding sites neither feed nor receive the physical ones
    for (int k = 0; k < d.dimensions - 1; ++k) {
      const int n = npt[k], np = h->npad[k], w = cplx[k] ? 2 : 1;
      if (np == n) continue;
      for (int i : {k, 2 + k}) {
        std::vector<double> m((size_t)np * np * w, 0.0);
        for (int r = 0; r < n; ++r)
          for (int c = 0; c < n * w; ++c) m[((size_t)r * np) * w + c] = mats[i][((size_t)r * n) * w + c];
        mats[i].swap(m);
      }
    }
  }
  h->cplx[0] = cplx[0];
  h->cplx[1] = cplx[1];
  total = 0;
  for (int i = 0; i < 6; ++i) {
    h->tab_off[i] = total;
    total += ((tabs[i].size() * h->esize + 255) / 256) * 256;
  }
  for (int i = 0; i < 4; ++i) {
    h->tab_off[6 + i] = total;
    total += ((mats[i].size() * (h->esize / 2) + 255) / 256) * 256;
  }
  h->tab_off[10] = total;
  total += ((twh.size() * h->esize + 255) / 256) * 256;
  // Real matrices in the operand order of the matrix-core instruction of the
  // precision (used by kernels that run those passes on MFMA):
  //   FP32, N % 32 == 0, 32x32x2: Tp[kb][kk/2][lane] = T[32 kb + lane%32][kk + lane/32]
  //   FP64, N % 16 == 0, 16x16x4: Tp[kb][kk/4][lane] = T[16 kb + lane%16][kk + lane/16]
  // A DCT-I (NN) or DST-I (DD) matrix of an unpadded dimension with N % (2 tile)
  // == 0 is packed as its even/odd split instead, half the GEMM depth:
  // T[k][N-1-n] = (-1)^k T[k][n] (DCT-I), -(-1)^k T[k][n] (DST-I), so the even
  // rows act on x_n +- x_{N-1-n} and the odd rows on x_n -+ x_{N-1-n}, n < N/2:
  //   Ts[r][n] = T[2r][n] (r < N/2), T[2(r - N/2) + 1][n] (r >= N/2), n < N/2
  std::vector<std::vector<double>> mpk(4);
  const int tile = d.precision == XSH_FP64 ? 16 : 32, kstep = 64 / tile;
  for (int i = 0; i < 4; ++i) {
    const int k = i % 2, n = h->npad[k];
    h->tsym[i] = 0;
    h->has_tmp[i] = h->mat && k < d.dimensions - 1 && !cplx[k] && n % tile == 0 && mats[i].size() == (size_t)n * n;
    if (!h->has_tmp[i]) continue;
    const int bt = h->btype[k];
    if ((bt == -3 || bt == 3) && npt[k] == n && n % (2 * tile) == 0) h->tsym[i] = bt == -3 ? 1 : 2;
    const int K = h->tsym[i] ? n / 2 : n;
    auto Tsrc = [&](int r, int c) -> double {
      if (!h->tsym[i]) return mats[i][(size_t)r * n + c];
      const int row = r < n / 2 ? 2 * r : 2 * (r - n / 2) + 1;
      return mats[i][(size_t)row * n + c];
    };
    mpk[i].resize((size_t)n * K);
    for (int kb = 0; kb < n / tile; ++kb)
      for (int kk = 0; kk < K; kk += kstep)
        for (int lane = 0; lane < 64; ++lane)
          mpk[i][((size_t)kb * (K / kstep) + kk / kstep) * 64 + lane] = Tsrc(tile * kb + lane % tile, kk + lane / tile);
  }
  for (int i = 0; i < 4; ++i) {
    h->tab_off[11 + i] = total;
    total += ((mpk[i].size() * (h->esize / 2) + 255) / 256) * 256;
  }
  // Boundary values (the boundval constants, or the caller's boundfun table per
  // time level) and the patch of xpatch.m:24-62: pa(x, y) = sum over the
  // non-periodic dimensions of b_lower s0(x_d) + b_upper s1(x_d) with the shape
  // functions of each boundary type, and the Neumann correction
  // lp = (b_upper - b_lower) / range * da, da = -(L(k1) - L(k0)) dt range^2 / pi^2
  // from the first two k values of the dimension (xpropfactor.m:40-51), per
  // check (dt = dtr/2).
  const int nsh = std::max(h->npad[0], h->npad[1]);
  const int nbv = d.dimensions == 3 ? nsh : 1;                      // device row width (kernel lattice)
  const int nb_in = d.dimensions == 3 ? std::max(h->nx, h->ny) : 1;  // caller's row width (physical lattice)
  const int nlev = d.boundtab ? d.boundlevels : 1;
  std::vector<double> bshape((size_t)4 * nsh, 0.0);
  std::vector<cd> bvals((size_t)nlev * 4 * nbv, cd(0.0, 0.0));
  h->has_patch = false;
  for (int nci = 0; nci < 2; ++nci)
    for (int k = 0; k < 2; ++k) h->lpf[nci][k][0] = h->lpf[nci][k][1] = 0.0;
  for (int k = 0; k < d.dimensions - 1; ++k) {
    const int t = h->btype[k], n = npt[k];
    if (t == 0) continue;
    const double rng = d.ranges[k + 1];
    for (int m = 0; m < n; ++m) {
      const double x = m * h->dx[k];
      double s0 = 0.0, s1 = 0.0;
      switch (t) {
        case -3: s0 = x - 0.5 * x * x / rng; s1 = 0.5 * x * x / rng; break;
        case -1: s0 = 1.0; s1 = x; break;
        case 1: s0 = x - rng; s1 = 1.0; break;
        case 3: s0 = 1.0 - x / rng; s1 = x / rng; break;
      }
      bshape[(size_t)(2 * k) * nsh + m] = s0;
      bshape[(size_t)(2 * k + 1) * nsh + m] = s1;
    }
    const int nother = d.dimensions == 3 ? npt[1 - k] : 1;
    for (int lev = 0; lev < nlev; ++lev)
      for (int side = 0; side < 2; ++side)
        for (int o = 0; o < nother; ++o) {
          const size_t line = ((size_t)lev * 2 + k) * 2 + side;
          const size_t at = line * nbv + o, at_in = line * nb_in + o;
          const cd v = d.boundtab ? cd(d.boundtab[2 * at_in], d.boundtab[2 * at_in + 1])
                                  : cd(d.boundval[k][side][0], d.boundval[k][side][1]);
          bvals[at] = v;
          if (v != cd(0.0, 0.0)) h->has_patch = true;
        }
    if (t == -3 && d.has_linear) {
      const double dkp = M_PI / ((n - 1) * h->dx[k]);
      const double k0 = 0.0, k1 = dkp;  // NN k-grid: m * dkp (spdpreferences.m:56-57)
      const cd l1(d.lin_l1[k][0], d.lin_l1[k][1]), l2(d.lin_l2[k][0], d.lin_l2[k][1]);
      auto f = [&](double kk) { return l1 * cd(0, kk) + l2 * cd(-kk * kk, 0); };
      for (int nci = 0; nci < 2; ++nci) {
        const double dtp = h->dtr[nci] / 2.0;
        const cd da = -(f(k1) - f(k0)) * dtp * rng * rng / (M_PI * M_PI);
        h->lpf[nci][k][0] = (da / rng).real();
        h->lpf[nci][k][1] = (da / rng).imag();
      }
    }
  }
  h->d.boundtab = nullptr;  // copied: the caller's table is not referenced again
  h->bval_level = nlev > 1 ? 4 * nbv : 0;
  h->bval_nb = nbv;
  h->bshape_n = nsh;
  h->tab_off[15] = total;
  total += ((bshape.size() * h->esize / 2 + 255) / 256) * 256;
  h->tab_off[16] = total;
  total += ((bvals.size() * h->esize + 255) / 256) * 256;
  // k-space filters with the scales of xnoise.m:42 and xrandom.m:27 and the
  // inverse transform's 1/N folded in (knoisefactor = N / dv, xpreferences.m:142)
  std::vector<cd> kfil((size_t)d.knoises * h->ns), rfil((size_t)d.krandoms * h->ns);
  {
    const double knf = (double)h->ns / h->dv;
    const double kns = std::sqrt(h->nscale * knf) / h->ns, krs = std::sqrt(knf) / h->ns;
    for (size_t i = 0; i < kfil.size(); ++i) kfil[i] = cd(d.nfilter[2 * i], d.nfilter[2 * i + 1]) * kns;
    for (size_t i = 0; i < rfil.size(); ++i) rfil[i] = cd(d.rfilter[2 * i], d.rfilter[2 * i + 1]) * krs;
  }
  h->d.nfilter = h->d.rfilter = nullptr;  // copied
  h->tab_off[17] = total;
  total += ((kfil.size() * h->esize + 255) / 256) * 256;
  h->tab_off[18] = total;
  total += ((rfil.size() * h->esize + 255) / 256) * 256;
  host.assign(total + 256, 0);
  for (int i = 0; i < 6; ++i) {
    if (d.precision == XSH_FP64)
      pack<double>(host, h->tab_off[i], tabs[i]);
    else
      pack<float>(host, h->tab_off[i], tabs[i]);
  }
  for (int i = 0; i < 4; ++i) {
    if (d.precision == XSH_FP64)
      pack_real<double>(host, h->tab_off[6 + i], mats[i]);
    else
      pack_real<float>(host, h->tab_off[6 + i], mats[i]);
  }
  if (d.precision == XSH_FP64)
    pack<double>(host, h->tab_off[10], twh);
  else
    pack<float>(host, h->tab_off[10], twh);
  for (int i = 0; i < 4; ++i) {
    if (!h->has_tmp[i]) continue;
    if (d.precision == XSH_FP64)
      pack_real<double>(host, h->tab_off[11 + i], mpk[i]);
    else
      pack_real<float>(host, h->tab_off[11 + i], mpk[i]);
  }
  if (d.precision == XSH_FP64) {
    pack_real<double>(host, h->tab_off[15], bshape);
    pack<double>(host, h->tab_off[16], bvals);
    pack<double>(host, h->tab_off[17], kfil);
    pack<double>(host, h->tab_off[18], rfil);
  } else {
    pack_real<float>(host, h->tab_off[15], bshape);
    pack<float>(host, h->tab_off[16], bvals);
    pack<float>(host, h->tab_off[17], kfil);
    pack<float>(host, h->tab_off[18], rfil);
  }
  // per-trajectory outputs index trajectories by workgroup
  if ((d.traj_out || d.nmodes) && h->tpb != 1)
    return fail(XSH_EUNSUPPORTED, "per-trajectory outputs need one trajectory per workgroup (2D lattices)");
  return XSH_OK;
}

extern "C" int xsh_create(const xsh_desc* desc, xsh_handle** out) {
  if (!out) return fail(XSH_EINVAL, "out is NULL");
  *out = nullptr;
  int rc = validate(desc);
  if (rc) return rc;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(XSH_EDEVICE, "no HIP device");
  auto* h = new xsh_handle();
  std::memset(h, 0, sizeof(*h));
  if (hipGetDevice(&h->device) != hipSuccess) {
    delete h;
    return fail(XSH_EDEVICE, "hipGetDevice");
  }
  std::vector<char> host;
  size_t total = 0;
  rc = plan_handle(h, desc, host, total);
  if (rc) {
    delete h;
    return rc;
  }
  const xsh_desc& d = h->d;
  auto cleanup = [&](int code) {
    xsh_destroy(h);
    return code;
  };
  if (hipMalloc(&h->d_tables, total) != hipSuccess) return cleanup(fail(XSH_ENOMEM, "hipMalloc tables"));
  if (hipMemcpy(h->d_tables, host.data(), total, hipMemcpyHostToDevice) != hipSuccess)
    return cleanup(fail(XSH_EDEVICE, "copy tables"));
  const size_t avbytes = (size_t)h->e2 * h->nchk * h->avlen * 16;
  if (hipMalloc(&h->d_av, avbytes) != hipSuccess)
    return cleanup(fail(XSH_ENOMEM, "hipMalloc averages (%zu bytes)", avbytes));
  if (hipMalloc(&h->d_path_av, (size_t)h->avlen * 16) != hipSuccess) return cleanup(fail(XSH_ENOMEM, "hipMalloc path averages"));
  if (hipMalloc(&h->d_a, (size_t)h->E * h->ns * 16) != hipSuccess) return cleanup(fail(XSH_ENOMEM, "hipMalloc fields"));
  const size_t paths = (size_t)h->e2 * h->nchk;
  if (d.traj_out &&
      hipMalloc(&h->d_traj, paths * h->E * d.points * d.nobserve * sizeof(double)) != hipSuccess)
    return cleanup(fail(XSH_ENOMEM, "hipMalloc per-trajectory observables"));
  if (d.raw) {
    const size_t rbytes = paths * d.points * (size_t)h->E * h->ns * 16;
    if (hipMalloc(&h->d_raw, rbytes) != hipSuccess)
      return cleanup(fail(XSH_ENOMEM, "hipMalloc raw fields (%zu bytes)", rbytes));
  }
  if (d.nmodes) {
    if (hipMalloc(&h->d_feat, paths * h->E * 2 * d.nmodes * sizeof(float)) != hipSuccess ||
        hipMalloc(&h->d_fmap, (size_t)h->ns * sizeof(int)) != hipSuccess)
      return cleanup(fail(XSH_ENOMEM, "hipMalloc noise modes"));
    std::vector<int> fmap(h->ns, -1);
    for (int m = 0; m < d.nmodes; ++m) fmap[d.modes[m][0] + h->nx * d.modes[m][1]] = m;
    if (hipMemcpy(h->d_fmap, fmap.data(), fmap.size() * sizeof(int), hipMemcpyHostToDevice) != hipSuccess)
      return cleanup(fail(XSH_EDEVICE, "copy mode map"));
  }
  for (int o = 0; o < d.nobserve; ++o)
    if (d.tspec[o]) h->tspec_mask |= 1 << o;
  if (h->tspec_mask) {
    // xpath.m:18 aft = [fields, points * steps, ens] per path; the frequency
    // grid of xpreferences.m:199-200
    const long long M = (long long)d.points * d.steps;
    const size_t fbytes = paths * (size_t)M * h->E * 16;
    if (hipMalloc(&h->d_ft, fbytes) != hipSuccess)
      return cleanup(fail(XSH_ENOMEM, "hipMalloc spectrum step fields (%zu bytes)", fbytes));
    std::vector<double> tw(2 * (size_t)M);
    for (long long k = 0; k < M; ++k) {
      const cd w = std::polar(1.0, 2.0 * M_PI * (double)k / (double)M);
      tw[2 * k] = w.real();
      tw[2 * k + 1] = w.imag();
    }
    if (hipMalloc(&h->d_twt, tw.size() * sizeof(double)) != hipSuccess)
      return cleanup(fail(XSH_ENOMEM, "hipMalloc spectrum twiddles"));
    if (hipMemcpy(h->d_twt, tw.data(), tw.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
      return cleanup(fail(XSH_EDEVICE, "copy spectrum twiddles"));
    const double dw = 2.0 * M_PI / (d.ranges[0] + h->dt);
    h->fsc = std::sqrt(2.0 * M_PI) / dw;
  }
  if (hipMalloc(&h->d_nonfinite, 2 * sizeof(int)) != hipSuccess) return cleanup(fail(XSH_ENOMEM, "hipMalloc flags"));
  if (hipMemset(h->d_nonfinite, 0, 2 * sizeof(int)) != hipSuccess) return cleanup(fail(XSH_EDEVICE, "clear flags"));
  if (h->kern && h->kern->stream) {
    // one workspace per resident workgroup of the persistent grid
    bool corr = false;
    for (int o = 0; o < d.nobserve; ++o)
      corr = corr || d.observe[o] == XSH_OBS_CORR || d.observe[o] == XSH_OBS_ABS2K;
    h->work_planes = (corr ? 3 : 2) + (d.knoises || d.krandoms ? 1 : 0);
    const long long blocks = std::min<long long>((long long)paths * h->groups, xsh::kStreamMaxBlocks);
    const size_t wbytes = (size_t)blocks * h->work_planes * h->ns * h->esize;
    if (hipMalloc(&h->d_work, wbytes) != hipSuccess)
      return cleanup(fail(XSH_ENOMEM, "hipMalloc streaming workspace (%zu bytes)", wbytes));
  }
  if (d.knoises && !h->kern->stream) {
    // x-space noise planes per work item of the largest launch (every (j, nc)
    // path of an ensemble), for whichever compiled variant of this lattice runs
    // it (a batch's autotuner times them all): ceil(E / TPB) items of TPB planes
    // per path, at most E + TPB - 1 planes
    int tpb_max = h->tpb;
    const xsh::KernelEntry* v[16];
    const int nv = xsh::kernel_variants(d.precision, h->nx, h->ny, h->mat, v, 16);
    for (int k = 0; k < nv; ++k) tpb_max = std::max(tpb_max, v[k]->tpb);
    const size_t kbytes = paths * ((size_t)h->E + tpb_max - 1) * h->ns * h->esize;
    if (hipMalloc(&h->d_kwork, kbytes) != hipSuccess)
      return cleanup(fail(XSH_ENOMEM, "hipMalloc k-space noise planes (%zu bytes)", kbytes));
  }
  if (hipStreamCreate(&h->stream) != hipSuccess) return cleanup(fail(XSH_EDEVICE, "hipStreamCreate"));
  if (hipEventCreate(&h->ev0) != hipSuccess || hipEventCreate(&h->ev1) != hipSuccess ||
      hipEventCreateWithFlags(&h->ev_args, hipEventDisableTiming) != hipSuccess)
    return cleanup(fail(XSH_EDEVICE, "hipEventCreate"));
  if (hipMalloc(&h->d_args, sizeof(xsh::PathArgs)) != hipSuccess) return cleanup(fail(XSH_ENOMEM, "hipMalloc arguments"));
  if (hipHostMalloc(&h->h_args, sizeof(xsh::PathArgs), hipHostMallocDefault) != hipSuccess)
    return cleanup(fail(XSH_ENOMEM, "hipHostMalloc arguments"));
  *out = h;
  return XSH_OK;
}

extern "C" int xsh_plan(const xsh_desc* desc, xsh_info* info) {
  if (!info) return fail(XSH_EINVAL, "info is NULL");
  int rc = validate(desc);
  if (rc) return rc;
  auto* h = new xsh_handle();
  std::memset(h, 0, sizeof(*h));
  std::vector<char> host;
  size_t total = 0;
  rc = plan_handle(h, desc, host, total);
  if (!rc) rc = xsh_query(h, info);
  delete h;
  return rc;
}

extern "C" void xsh_destroy(xsh_handle* h) {
  if (!h) return;
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  if (h->d_tables) (void)hipFree(h->d_tables);
  if (h->d_av) (void)hipFree(h->d_av);
  if (h->d_path_av) (void)hipFree(h->d_path_av);
  if (h->d_a) (void)hipFree(h->d_a);
  if (h->d_traj) (void)hipFree(h->d_traj);
  if (h->d_raw) (void)hipFree(h->d_raw);
  if (h->d_kwork) (void)hipFree(h->d_kwork);
  if (h->d_feat) (void)hipFree(h->d_feat);
  if (h->d_fmap) (void)hipFree(h->d_fmap);
  if (h->d_ft) (void)hipFree(h->d_ft);
  if (h->d_nonfinite) (void)hipFree(h->d_nonfinite);
  if (h->d_twt) (void)hipFree(h->d_twt);
  if (h->d_work) (void)hipFree(h->d_work);
  if (h->ev0) (void)hipEventDestroy(h->ev0);
  if (h->ev1) (void)hipEventDestroy(h->ev1);
  if (h->ev_args) (void)hipEventDestroy(h->ev_args);
  if (h->d_args) (void)hipFree(h->d_args);
  if (h->h_args) (void)hipHostFree(h->h_args);
  if (h->stream) (void)hipStreamDestroy(h->stream);
  delete h;
}

extern "C" int xsh_query(const xsh_handle* h, xsh_info* info) {
  if (!h || !info) return fail(XSH_EINVAL, "NULL argument");
  std::memset(info, 0, sizeof(*info));
  info->nsites = h->ns;
  info->nchk = h->nchk;
  info->avlen = h->avlen;
  for (int o = 0; o < h->d.nobserve; ++o) {
    info->obs_offset[o] = h->obs_off[o];
    info->obs_sites[o] = h->obs_ns[o];
  }
  // xpath.m:29-34: points*steps - 1 steps per path; fine runs take (1+checks) per step
  const long long per = (long long)h->d.points * h->d.steps - (h->d.points > 1 ? 1 : 0);
  info->traj_steps = (long long)h->E * h->e2 * per * (h->d.checks ? 3 : 1);
  info->threads = h->threads;
  info->traj_per_block = h->tpb;
  info->kernel_kind = h->nd == 1 ? 3 : (h->kern && h->kern->stream) ? 1 : h->mat ? 2 : 0;
  for (int i = 0; i < 3; ++i) info->ensembles[i] = h->d.ensembles[i];
  info->points = h->d.points;
  return XSH_OK;
}

static xsh::PathArgs make_args(const xsh_handle* h, int npar) {
  xsh::PathArgs A;
  std::memset(&A, 0, sizeof(A));
  const xsh_desc& d = h->d;
  A.nx = h->nx;
  A.ny = h->ny;
  A.ns = h->ns;
  A.x0 = h->x0[0];
  A.dx = h->dx[0];
  A.y0 = h->x0[1];
  A.dy = h->dx[1];
  A.dv = h->dv;
  A.npts = d.points;
  A.steps = d.steps;
  A.checks = d.checks;
  A.method = d.method;
  A.iterations = d.iterations;
  A.hdt[0] = h->dtr[0] / 2.0;
  A.hdt[1] = h->dtr[1] / 2.0;
  A.E = h->E;
  A.e2 = h->e2;
  A.e3 = h->e3;
  A.npar = npar;
  A.seed = d.seed;
  A.seq = d.seq;
  A.inst = d.inst;
  A.groups = h->groups;
  A.noises = d.noises;
  A.inrandoms = d.inrandoms;
  A.urandoms = d.urandoms;
  A.knoises = d.knoises;
  A.krandoms = d.krandoms;
  A.kn_filter = static_cast<char*>(h->d_tables) + h->tab_off[17];
  A.kr_filter = static_cast<char*>(h->d_tables) + h->tab_off[18];
  A.kn_work = h->d_kwork;
  A.noise_scale = std::sqrt(h->nscale * h->noisefactor);
  A.init_scale = std::sqrt(h->noisefactor);
  A.deriv = d.deriv;
  for (int i = 0; i < 6; ++i) A.dp[i] = d.deriv_par[i];
  A.initial = d.initial;
  for (int i = 0; i < 6; ++i) A.ip[i] = d.init_par[i];
  A.nobs = d.nobserve;
  for (int o = 0; o < d.nobserve; ++o) {
    A.obs[o] = d.observe[o];
    A.obs_off[o] = h->obs_off[o];
  }
  A.avlen = h->avlen;
  A.inv_E = 1.0 / h->E;
  A.ft = h->d_ft;
  A.nonfinite = h->d_nonfinite;
  A.ft_len = d.points * d.steps;
  A.tspec_mask = h->tspec_mask;
  A.has_linear = d.has_linear;
  char* t = static_cast<char*>(h->d_tables);
  A.px[0] = t + h->tab_off[0];
  A.py[0] = t + h->tab_off[1];
  A.px[1] = t + h->tab_off[2];
  A.py[1] = t + h->tab_off[3];
  A.twx = t + h->tab_off[4];
  A.twy = t + h->tab_off[5];
  A.mat = h->mat;
  for (int k = 0; k < 2; ++k) {
    A.cplx[k] = h->cplx[k];
    A.dir_lo[k] = k < d.dimensions - 1 && d.boundaries[k][0] == 1;
    A.dir_hi[k] = k < d.dimensions - 1 && d.boundaries[k][1] == 1;
    if (A.dir_lo[k] || A.dir_hi[k]) A.dirichlet = 1;
  }
  for (int i = 0; i < 4; ++i) A.tm[i] = t + h->tab_off[6 + i];
  A.twh = t + h->tab_off[10];
  A.has_patch = h->has_patch;
  A.bshape = t + h->tab_off[15];
  A.bshape_n = h->bshape_n;
  A.fields = d.fields > 1 ? d.fields : 1;
  if (A.fields > 1) {
    for (int i = 0; i < A.fields; ++i) {
      for (int jj = 0; jj < A.fields; ++jj)
        for (int c = 0; c < 2; ++c) A.fm[i][jj][c] = d.fmat[i][jj][c];
      for (int c = 0; c < 2; ++c) {
        A.fg[i][c] = d.fnl[i][c];
        A.fc0[i][c] = d.fc0[i][c];
      }
      for (int k = 0; k < 4; ++k) A.fs[i][k] = k < d.noises ? d.fnoise[i][k] : 0.0;
      // per-component interaction picture propagator exp(l0_i dtr / ipsteps)
      // (xpropfactor.m:26-39 with ipsteps = 2)
      for (int nci = 0; nci < 2; ++nci) {
        const cd P = d.has_linear ? std::exp(cd(d.fl0[i][0], d.fl0[i][1]) * (h->dtr[nci] / 2.0)) : cd(1.0, 0.0);
        A.fp[nci][i][0] = P.real();
        A.fp[nci][i][1] = P.imag();
      }
    }
    for (int o = 0; o < d.nobserve; ++o) A.obs_field[o] = d.obs_field[o];
  }
  A.bval = t + h->tab_off[16];
  A.bval_level = h->bval_level;
  A.bval_nb = h->bval_nb;
  for (int nci = 0; nci < 2; ++nci)
    for (int k = 0; k < 2; ++k)
      for (int c = 0; c < 2; ++c) A.lpf[nci][k][c] = h->lpf[nci][k][c];
  for (int i = 0; i < 4; ++i) {
    A.tmp[i] = h->has_tmp[i] ? t + h->tab_off[11 + i] : nullptr;
    A.tsym[i] = h->tsym[i];
  }
  A.work = h->d_work;
  A.work_planes = h->work_planes;
  A.traj = h->d_traj;
  A.raw = h->d_raw;
  A.nmodes = d.nmodes;
  A.fmap = h->d_fmap;
  A.feat = h->d_feat;
  A.feat_scale = std::sqrt(h->dv / h->ns);
  A.kscale = 1.0;
  for (int k = 0; k < d.dimensions - 1; ++k) A.kscale *= h->dx[k] * h->dx[k] / (2.0 * M_PI);
  return A;
}

// upload the arguments on the handle's stream (the staging buffer is reused
// only after the previous upload has completed)
static int upload_args(xsh_handle* h, const xsh::PathArgs& A) {
  if (h->args_pending) HIPCHK(hipEventSynchronize(h->ev_args));
  *h->h_args = A;
  HIPCHK(hipMemcpyAsync(h->d_args, h->h_args, sizeof(xsh::PathArgs), hipMemcpyHostToDevice, h->stream));
  HIPCHK(hipEventRecord(h->ev_args, h->stream));
  h->args_pending = true;
  return XSH_OK;
}

static int launch(xsh_handle* h, const xsh::PathArgs& A, int nblocks) {
  int rc = upload_args(h, A);
  if (rc) return rc;
  hipError_t e;
  if (h->nd == 1) {
    e = h->d.fields > 1 ? xsh::launch_sde_fields(h->d.precision, h->d.fields, h->d_args, nblocks, h->stream)
                        : xsh::launch_sde_path(h->d.precision, h->d_args, nblocks, h->stream);
    if (e == hipSuccess && h->tspec_mask)
      e = xsh::launch_sde_spectrum(h->d_args, h->d_twt, h->fsc, h->E, h->d.points, A.j_only > 0 ? 1 : h->e2 * h->nchk,
                                   h->stream);
  } else {
    const int grid = h->kern->stream ? std::min(nblocks, xsh::kStreamMaxBlocks) : nblocks;
    e = h->kern->launch(h->d_args, nullptr, 1, nblocks, grid, h->stream);
  }
  if (e != hipSuccess) return fail(XSH_EDEVICE, "kernel launch: %s", hipGetErrorString(e));
  return XSH_OK;
}

// The kernels accumulate each observable as [points][sites] (sites fastest:
// neighbouring lanes update neighbouring addresses).  The host buffers are the
// reference's p.d.av{o} = [lines, points, space] in MATLAB column-major order
// (xpreferences.m:340): time fastest.  One path's averages, device -> host order.
static void av_to_host_order(const xsh_handle* h, const double* src, double* dst) {
  const int npts = h->d.points;
  for (int o = 0; o < h->d.nobserve; ++o) {
    const long long off = h->obs_off[o];
    const int ns = h->obs_ns[o];
    for (int n = 0; n < npts; ++n)
      for (int s = 0; s < ns; ++s) {
        const long long a = off + (long long)n * ns + s, b = off + n + (long long)npts * s;
        dst[2 * b] = src[2 * a];
        dst[2 * b + 1] = src[2 * a + 1];
      }
  }
}

// copy n_paths consecutive path averages from the device to host order
static int copy_averages(xsh_handle* h, const double* dev, int n_paths, double* out, hipStream_t s) {
  const size_t per = (size_t)h->avlen * 2;
  std::vector<double> tmp((size_t)n_paths * per);
  HIPCHK(hipMemcpyAsync(tmp.data(), dev, tmp.size() * sizeof(double), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  for (int p = 0; p < n_paths; ++p) av_to_host_order(h, tmp.data() + p * per, out + p * per);
  return XSH_OK;
}

extern "C" int xsh_run_path(xsh_handle* h, int npar, int j, int nc, const double* a0, double* a_out,
                            double* av_out) {
  if (!h) return fail(XSH_EINVAL, "handle is NULL");
  if (npar < 1 || npar > h->e3) return fail(XSH_EINVAL, "npar %d outside 1..%d", npar, h->e3);
  if (j < 1 || j > h->e2) return fail(XSH_EINVAL, "j %d outside 1..%d", j, h->e2);
  if (nc < 1 || nc > h->nchk) return fail(XSH_EINVAL, "nc %d outside 1..%d", nc, h->nchk);
  if (!av_out) return fail(XSH_EINVAL, "av_out is NULL");
  xsh::PathArgs A = make_args(h, npar);
  A.j_only = j;
  A.nc_only = nc;
  A.av = h->d_path_av;
  A.traj = nullptr;  // per-trajectory outputs belong to xsh_run_ensemble
  A.raw = nullptr;
  A.nmodes = 0;
  const size_t fbytes = (size_t)h->E * h->ns * 16;
  HIPCHK(hipMemsetAsync(h->d_path_av, 0, (size_t)h->avlen * 16, h->stream));
  HIPCHK(hipMemsetAsync(h->d_nonfinite, 0, 2 * sizeof(int), h->stream));
  if (a0) {
    HIPCHK(hipMemcpyAsync(h->d_a, a0, fbytes, hipMemcpyHostToDevice, h->stream));
    A.a0 = h->d_a;
  }
  A.a_out = a_out ? h->d_a : nullptr;
  int rc = launch(h, A, h->groups);
  if (rc) return rc;
  if (a_out) HIPCHK(hipMemcpyAsync(a_out, h->d_a, fbytes, hipMemcpyDeviceToHost, h->stream));
  return copy_averages(h, h->d_path_av, 1, av_out, h->stream);
}

// ---------------------------------------------------------------------------
// xsh_run_path_noalloc: xsh_run_path for callers that cannot allocate device
// memory.  Its scratch - the kernel arguments, the averages as the kernels
// accumulate them, the streaming kernel's field planes and the k-space noise
// planes - lives in one file-scope device array of XSH_SCRATCH_BYTES, which
// every device the library runs on holds until the process ends.  Host code
// reaches it through the address hipGetSymbolAddress returns for the current
// device (kept per device), never through the array's name.
__device__ __align__(256) unsigned char scratch[XSH_SCRATCH_BYTES];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned char* g_scratch[kMaxDevices];

// address of this device's scratch array (looked up once per device)
int scratch_address(int dev, unsigned char** out) {
  if (dev < 0 || dev >= kMaxDevices) return fail(XSH_EDEVICE, "device %d beyond the scratch table", dev);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) return fail(XSH_EDEVICE, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
    g_scratch[dev] = static_cast<unsigned char*>(p);
  }
  *out = g_scratch[dev];
  return XSH_OK;
}

size_t align256(size_t n) { return (n + 255) / 256 * 256; }

// observable blocks of one path: offsets and site counts
struct AvLayout {
  long long off[XSH_MAX_OBS];
  int ns[XSH_MAX_OBS];
  int nobs, npts;
};

// one path's averages, kernel order [points][sites] -> p.d.av order (time fastest)
__global__ void av_reorder(const double* __restrict__ src, double* __restrict__ dst, AvLayout lay, long long avlen) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= avlen) return;
  int o = 0;
  while (o + 1 < lay.nobs && i >= lay.off[o + 1]) ++o;
  const long long r = i - lay.off[o];
  const long long n = r / lay.ns[o], st = r - n * lay.ns[o];
  const long long j = lay.off[o] + n + (long long)lay.npts * st;
  dst[2 * j] = src[2 * i];
  dst[2 * j + 1] = src[2 * i + 1];
}
}  // namespace

extern "C" int xsh_run_path_noalloc(xsh_handle* h, int npar, int j, int nc, const double* a0_dev, double* a_out_dev,
                                    double* av_dev, void* stream) {
  if (!h) return fail(XSH_EINVAL, "handle is NULL");
  if (npar < 1 || npar > h->e3) return fail(XSH_EINVAL, "npar %d outside 1..%d", npar, h->e3);
  if (j < 1 || j > h->e2) return fail(XSH_EINVAL, "j %d outside 1..%d", j, h->e2);
  if (nc < 1 || nc > h->nchk) return fail(XSH_EINVAL, "nc %d outside 1..%d", nc, h->nchk);
  if (!av_dev) return fail(XSH_EINVAL, "av_dev is NULL");
  if (h->tspec_mask)
    return fail(XSH_EUNSUPPORTED, "time-domain spectra keep every step's fields: not in the scratch of this entry");
  if (h->d.fields > 1) return fail(XSH_EUNSUPPORTED, "coupled SDE components are not on this entry");
  int dev = -1;
  HIPCHK(hipGetDevice(&dev));
  if (dev != h->device) return fail(XSH_ESTATE, "handle belongs to device %d, current device is %d", h->device, dev);
  // scratch layout: kernel arguments | averages (kernel order) | streaming planes | k-noise planes
  const size_t args_bytes = align256(sizeof(xsh::PathArgs));
  const size_t av_bytes = align256((size_t)h->avlen * 16);
  const int items = h->groups;
  int grid = items;
  size_t work_bytes = 0;
  if (h->kern && h->kern->stream) {
    const size_t per_block = (size_t)h->work_planes * h->ns * h->esize;
    const size_t room = XSH_SCRATCH_BYTES > args_bytes + av_bytes ? XSH_SCRATCH_BYTES - args_bytes - av_bytes : 0;
    grid = (int)std::min<size_t>({(size_t)items, (size_t)xsh::kStreamMaxBlocks, room / per_block});
    if (grid < 1)
      return fail(XSH_EINVAL, "lattice too large for the %d-byte scratch (one workgroup needs %zu bytes)",
                  (int)XSH_SCRATCH_BYTES, args_bytes + av_bytes + per_block);
    work_bytes = align256((size_t)grid * per_block);
  }
  const size_t kn_bytes =
      h->d.knoises && !h->kern->stream ? align256((size_t)items * h->tpb * h->ns * h->esize) : 0;
  const size_t need = args_bytes + av_bytes + work_bytes + kn_bytes;
  if (need > (size_t)XSH_SCRATCH_BYTES)
    return fail(XSH_EINVAL, "path needs %zu bytes of scratch, more than the %d available (averages %zu, "
                            "k-space noise planes %zu)", need, (int)XSH_SCRATCH_BYTES, av_bytes, kn_bytes);
  unsigned char* base = nullptr;
  int rc = scratch_address(dev, &base);
  if (rc) return rc;
  hipStream_t s = stream ? static_cast<hipStream_t>(stream) : h->stream;
  xsh::PathArgs* d_args = reinterpret_cast<xsh::PathArgs*>(base);
  double* d_av = reinterpret_cast<double*>(base + args_bytes);
  xsh::PathArgs A = make_args(h, npar);
  A.j_only = j;
  A.nc_only = nc;
  A.av = d_av;
  A.traj = nullptr;  // per-trajectory outputs belong to xsh_run_ensemble
  A.raw = nullptr;
  A.nmodes = 0;
  A.a0 = a0_dev;
  A.a_out = a_out_dev;
  A.nonfinite = nullptr;  // the divergence flag is device memory of the handle's own runs
  A.work = work_bytes ? base + args_bytes + av_bytes : nullptr;
  A.kn_work = kn_bytes ? base + args_bytes + av_bytes + work_bytes : nullptr;
  // arguments through the handle's pinned staging copy (reused once its last upload is done)
  if (h->args_pending) HIPCHK(hipEventSynchronize(h->ev_args));
  *h->h_args = A;
  HIPCHK(hipMemcpyAsync(d_args, h->h_args, sizeof(xsh::PathArgs), hipMemcpyHostToDevice, s));
  HIPCHK(hipEventRecord(h->ev_args, s));
  h->args_pending = true;
  // the averages accumulate with atomics: clear them (a fill kernel, no memset node)
  const long long nd = h->avlen * 2;
  hipLaunchKernelGGL(zero_doubles, dim3((unsigned)((nd + 255) / 256)), dim3(256), 0, s, d_av, nd);
  HIPCHK(hipGetLastError());
  hipError_t e;
  if (h->nd == 1)
    e = xsh::launch_sde_path(h->d.precision, d_args, items, s);
  else
    e = h->kern->launch(d_args, nullptr, 1, items, grid, s);
  if (e != hipSuccess) return fail(XSH_EDEVICE, "kernel launch: %s", hipGetErrorString(e));
  AvLayout lay{};
  lay.nobs = h->d.nobserve;
  lay.npts = h->d.points;
  for (int o = 0; o < h->d.nobserve; ++o) {
    lay.off[o] = h->obs_off[o];
    lay.ns[o] = h->obs_ns[o];
  }
  hipLaunchKernelGGL(av_reorder, dim3((unsigned)((h->avlen + 255) / 256)), dim3(256), 0, s, d_av, av_dev, lay, h->avlen);
  HIPCHK(hipGetLastError());
  return XSH_OK;
}

extern "C" int xsh_run_ensemble(xsh_handle* h, int npar) {
  if (!h) return fail(XSH_EINVAL, "handle is NULL");
  if (npar < 1 || npar > h->e3) return fail(XSH_EINVAL, "npar %d outside 1..%d", npar, h->e3);
  xsh::PathArgs A = make_args(h, npar);
  A.av = h->d_av;
  HIPCHK(hipMemsetAsync(h->d_av, 0, (size_t)h->e2 * h->nchk * h->avlen * 16, h->stream));
  HIPCHK(hipMemsetAsync(h->d_nonfinite, 0, 2 * sizeof(int), h->stream));
  HIPCHK(hipEventRecord(h->ev0, h->stream));
  int rc = launch(h, A, h->groups * h->e2 * h->nchk);
  if (rc) return rc;
  HIPCHK(hipEventRecord(h->ev1, h->stream));
  h->ran = true;
  return XSH_OK;
}

extern "C" int xsh_synchronize(xsh_handle* h, double* ms) {
  if (!h) return fail(XSH_EINVAL, "handle is NULL");
  HIPCHK(hipStreamSynchronize(h->stream));
  if (ms) {
    float f = 0.f;
    if (!h->ran) return fail(XSH_ESTATE, "nothing has run on this handle");
    HIPCHK(hipEventElapsedTime(&f, h->ev0, h->ev1));
    *ms = f;
  }
  return XSH_OK;
}

extern "C" int xsh_read_averages(xsh_handle* h, double* av_out) {
  if (!h || !av_out) return fail(XSH_EINVAL, "NULL argument");
  if (!h->ran) return fail(XSH_ESTATE, "xsh_run_ensemble has not run");
  return copy_averages(h, h->d_av, h->e2 * h->nchk, av_out, h->stream);
}

extern "C" int xsh_nonfinite(xsh_handle* h, int* flag) {
  if (!h || !flag) return fail(XSH_EINVAL, "NULL argument");
  HIPCHK(hipStreamSynchronize(h->stream));
  HIPCHK(hipMemcpy(flag, h->d_nonfinite, sizeof(int), hipMemcpyDeviceToHost));
  return XSH_OK;
}

extern "C" int xsh_read_raw(xsh_handle* h, double* out) {
  if (!h || !out) return fail(XSH_EINVAL, "NULL argument");
  if (!h->ran) return fail(XSH_ESTATE, "xsh_run_ensemble has not run");
  if (!h->d_raw) return fail(XSH_ESTATE, "descriptor did not ask for raw fields");
  // device [path][points][E][sites] -> per path p.d.raw = [1, points, nx, ny, E]
  // column-major (xpreferences.m:148): time fastest, then sites, then trajectories
  const size_t npts = h->d.points, E = h->E, ns = h->ns, per = npts * E * ns;
  const size_t paths = (size_t)h->e2 * h->nchk;
  std::vector<double> tmp(2 * paths * per);
  HIPCHK(hipMemcpyAsync(tmp.data(), h->d_raw, tmp.size() * sizeof(double), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  for (size_t p = 0; p < paths; ++p)
    for (size_t n = 0; n < npts; ++n)
      for (size_t e = 0; e < E; ++e)
        for (size_t s = 0; s < ns; ++s) {
          const size_t a = p * per + (n * E + e) * ns + s, b = p * per + n + npts * (s + ns * e);
          out[2 * b] = tmp[2 * a];
          out[2 * b + 1] = tmp[2 * a + 1];
        }
  return XSH_OK;
}

extern "C" int xsh_read_trajectories(xsh_handle* h, double* obs, float* feat) {
  if (!h) return fail(XSH_EINVAL, "handle is NULL");
  if (!h->ran) return fail(XSH_ESTATE, "xsh_run_ensemble has not run");
  const size_t paths = (size_t)h->e2 * h->nchk;
  if (obs) {
    if (!h->d_traj) return fail(XSH_ESTATE, "descriptor did not ask for traj_out");
    HIPCHK(hipMemcpyAsync(obs, h->d_traj, paths * h->E * h->d.points * h->d.nobserve * sizeof(double),
                          hipMemcpyDeviceToHost, h->stream));
  }
  if (feat) {
    if (!h->d_feat) return fail(XSH_ESTATE, "descriptor did not ask for noise modes");
    HIPCHK(hipMemcpyAsync(feat, h->d_feat, paths * h->E * 2 * h->d.nmodes * sizeof(float), hipMemcpyDeviceToHost,
                          h->stream));
  }
  HIPCHK(hipStreamSynchronize(h->stream));
  return XSH_OK;
}

extern "C" int xsh_probe_normals(xsh_handle* h, int nsp, int tag, int draw, int traj0, int ntraj, int count,
                                 double* out) {
  if (!h || !out) return fail(XSH_EINVAL, "NULL argument");
  if (ntraj < 1 || count < 1 || traj0 < 0) return fail(XSH_EINVAL, "bad probe sizes");
  double* d = nullptr;
  const size_t bytes = (size_t)ntraj * count * sizeof(double);
  HIPCHK(hipMalloc(&d, bytes));
  hipError_t e = xsh::launch_probe_normals(h->d.precision, (uint32_t)nsp, (uint32_t)tag, (uint32_t)h->d.seq,
                                           (uint32_t)h->d.inst, (uint32_t)draw, traj0, ntraj, count, d, h->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(out, d, bytes, hipMemcpyDeviceToHost, h->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  (void)hipFree(d);
  if (e != hipSuccess) return fail(XSH_EDEVICE, "probe normals: %s", hipGetErrorString(e));
  return XSH_OK;
}

extern "C" int xsh_probe_propagate(xsh_handle* h, int nc, int nplanes, double* planes) {
  if (!h || !planes) return fail(XSH_EINVAL, "NULL argument");
  if (h->nd < 2) return fail(XSH_EINVAL, "no lattice to propagate (dimensions = 1)");
  if (nc < 1 || nc > 2 || nplanes < 1) return fail(XSH_EINVAL, "bad probe arguments");
  xsh::PathArgs A = make_args(h, 1);
  // kernel planes hold npad[0] x npad[1] sites (padded lattices: zeros around the physical ones)
  const size_t kns = (size_t)h->npad[0] * h->npad[1];
  const size_t n = (size_t)nplanes * kns;
  std::vector<char> buf(n * h->esize, 0);
  for (int pl = 0; pl < nplanes; ++pl)
    for (int sx = 0; sx < h->ns; ++sx) {
      const size_t src = (size_t)pl * h->ns + sx, dst = (size_t)pl * kns + sx % h->nx + (size_t)h->npad[0] * (sx / h->nx);
      if (h->d.precision == XSH_FP64) {
        reinterpret_cast<double*>(buf.data())[2 * dst] = planes[2 * src];
        reinterpret_cast<double*>(buf.data())[2 * dst + 1] = planes[2 * src + 1];
      } else {
        reinterpret_cast<float*>(buf.data())[2 * dst] = (float)planes[2 * src];
        reinterpret_cast<float*>(buf.data())[2 * dst + 1] = (float)planes[2 * src + 1];
      }
    }
  void* d = nullptr;
  HIPCHK(hipMalloc(&d, buf.size()));
  int rc = upload_args(h, A);
  if (rc) {
    (void)hipFree(d);
    return rc;
  }
  hipError_t e = hipMemcpyAsync(d, buf.data(), buf.size(), hipMemcpyHostToDevice, h->stream);
  if (e == hipSuccess) e = h->kern->probe(h->d_args, nc - 1, nplanes, d, h->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(buf.data(), d, buf.size(), hipMemcpyDeviceToHost, h->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  (void)hipFree(d);
  if (e != hipSuccess) return fail(XSH_EDEVICE, "probe propagate: %s", hipGetErrorString(e));
  for (int pl = 0; pl < nplanes; ++pl)
    for (int sx = 0; sx < h->ns; ++sx) {
      const size_t dst = (size_t)pl * h->ns + sx, src = (size_t)pl * kns + sx % h->nx + (size_t)h->npad[0] * (sx / h->nx);
      if (h->d.precision == XSH_FP64) {
        planes[2 * dst] = reinterpret_cast<double*>(buf.data())[2 * src];
        planes[2 * dst + 1] = reinterpret_cast<double*>(buf.data())[2 * src + 1];
      } else {
        planes[2 * dst] = reinterpret_cast<float*>(buf.data())[2 * src];
        planes[2 * dst + 1] = reinterpret_cast<float*>(buf.data())[2 * src + 1];
      }
    }
  return XSH_OK;
}

// ======================================================================
// Batches: many independent instances (uneven lattices, ensembles and drift
// parameters) in one graph launch.  Instances that share a lattice kernel
// form a group; each group runs as one persistent kernel over the work items
// of all its instances (spde_path / spde_stream with an instance table), and
// the groups run concurrently on forked streams inside one captured HIP
// graph, after fill kernels that clear the averages.
namespace {

struct BatchGroup {
  int precision, nx, ny, mat;
  bool need_tpb1;
  std::vector<int> inst;
  std::vector<const xsh::KernelEntry*> cand;
  const xsh::KernelEntry* kern;
  int grid, nwork;
  std::vector<int> wstart;
  xsh::PathArgs* d_args;
  int* d_wstart;
  void* d_work;
  size_t work_bytes;
  hipStream_t s;
  hipEvent_t join;
};

}  // namespace

struct xsh_batch {
  std::vector<xsh_handle*> h;
  std::vector<BatchGroup> g;
  hipStream_t s0;
  hipEvent_t ev0, ev1, fork;
  hipGraph_t graph;
  hipGraphExec_t exec;
  xsh::PathArgs* h_args;  // pinned staging, one per instance
  bool ran;
  bool serial;            // families one after another instead of concurrently
};

static void batch_free(xsh_batch* b) {
  if (!b) return;
  if (b->s0) (void)hipStreamSynchronize(b->s0);
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  for (auto& g : b->g) {
    if (g.d_args) (void)hipFree(g.d_args);
    if (g.d_wstart) (void)hipFree(g.d_wstart);
    if (g.d_work) (void)hipFree(g.d_work);
    if (g.join) (void)hipEventDestroy(g.join);
    if (g.s) (void)hipStreamDestroy(g.s);
  }
  for (auto* h : b->h) xsh_destroy(h);
  if (b->h_args) (void)hipHostFree(b->h_args);
  if (b->ev0) (void)hipEventDestroy(b->ev0);
  if (b->ev1) (void)hipEventDestroy(b->ev1);
  if (b->fork) (void)hipEventDestroy(b->fork);
  if (b->s0) (void)hipStreamDestroy(b->s0);
  delete b;
}

// work items of instance h under kernel k
static long long batch_items(const xsh_handle* h, const xsh::KernelEntry* k) {
  return (long long)((h->E + k->tpb - 1) / k->tpb) * h->e2 * h->nchk;
}

// (re)size a group for its chosen kernel: work table, grid, streaming workspace
static int batch_layout(BatchGroup& g, const std::vector<xsh_handle*>& hs, int grid) {
  g.wstart.assign(1, 0);
  long long tot = 0;
  for (int i : g.inst) {
    tot += batch_items(hs[i], g.kern);
    if (tot > (1LL << 30)) return fail(XSH_EINVAL, "batch group too large");
    g.wstart.push_back((int)tot);
  }
  g.nwork = (int)tot;
  g.grid = std::max(1, std::min(grid > 0 ? grid : g.nwork, g.nwork));
  if (g.kern->stream) {
    g.grid = std::min(g.grid, xsh::kStreamMaxBlocks);
    const size_t need = (size_t)g.grid * 3 * g.nx * g.ny * (g.precision == XSH_FP64 ? 16 : 8);
    if (need > g.work_bytes) {
      if (g.d_work) (void)hipFree(g.d_work);
      g.d_work = nullptr;
      g.work_bytes = 0;
      if (hipMalloc(&g.d_work, need) != hipSuccess) return fail(XSH_ENOMEM, "batch streaming workspace");
      g.work_bytes = need;
    }
  }
  HIPCHK(hipMemcpy(g.d_wstart, g.wstart.data(), g.wstart.size() * sizeof(int), hipMemcpyHostToDevice));
  return XSH_OK;
}

// kernel arguments of every instance of group g for parallel ensemble npar
static void batch_args(xsh_batch* b, BatchGroup& g, int npar, xsh::PathArgs* out, int npts_cap, int steps_cap) {
  for (size_t k = 0; k < g.inst.size(); ++k) {
    const xsh_handle* h = b->h[g.inst[k]];
    xsh::PathArgs A = make_args(h, npar);
    A.av = h->d_av;
    A.groups = (h->E + g.kern->tpb - 1) / g.kern->tpb;
    if (g.kern->stream) {
      A.work = g.d_work;
      A.work_planes = 3;
    }
    if (npts_cap > 0) A.npts = std::min(A.npts, npts_cap);
    if (steps_cap > 0) A.steps = std::min(A.steps, steps_cap);
    out[k] = A;
  }
}

static int batch_upload(xsh_batch* b, int npar) {
  HIPCHK(hipStreamSynchronize(b->s0));  // the staging buffer is free again
  int off = 0;
  for (auto& g : b->g) {
    batch_args(b, g, npar, b->h_args + off, 0, 0);
    HIPCHK(hipMemcpyAsync(g.d_args, b->h_args + off, g.inst.size() * sizeof(xsh::PathArgs), hipMemcpyHostToDevice,
                          b->s0));
    off += (int)g.inst.size();
  }
  return XSH_OK;
}

// time one shortened launch of group g under its current kernel and grid (ms per trajectory-step)
static int batch_time(xsh_batch* b, BatchGroup& g, double* ms_per_step) {
  std::vector<xsh::PathArgs> A(g.inst.size());
  batch_args(b, g, 1, A.data(), 2, 4);
  long long steps = 0;
  for (size_t k = 0; k < g.inst.size(); ++k) {
    const xsh_handle* h = b->h[g.inst[k]];
    steps += (long long)h->E * h->e2 * (2 * A[k].steps - 1) * (h->d.checks ? 3 : 1);
  }
  HIPCHK(hipMemcpy(g.d_args, A.data(), A.size() * sizeof(xsh::PathArgs), hipMemcpyHostToDevice));
  float best = 1e30f;
  for (int rep = 0; rep < 3; ++rep) {  // first launch loads the code object
    HIPCHK(hipEventRecord(b->ev0, b->s0));
    hipError_t e = g.kern->launch(g.d_args, g.d_wstart, (int)g.inst.size(), g.nwork, g.grid, b->s0);
    if (e != hipSuccess) return fail(XSH_EDEVICE, "batch tuning launch: %s", hipGetErrorString(e));
    HIPCHK(hipEventRecord(b->ev1, b->s0));
    HIPCHK(hipEventSynchronize(b->ev1));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, b->ev0, b->ev1));
    if (rep > 0) best = std::min(best, ms);
  }
  *ms_per_step = best / (double)std::max(1LL, steps);
  return XSH_OK;
}

static int batch_capture(xsh_batch* b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
  HIPCHK(hipStreamBeginCapture(b->s0, hipStreamCaptureModeThreadLocal));
  hipError_t e = hipSuccess;
  for (auto* h : b->h) {
    const long long n = (long long)h->e2 * h->nchk * h->avlen * 2;
    hipLaunchKernelGGL(zero_doubles, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, b->s0, h->d_av, n);
    hipLaunchKernelGGL(zero_doubles, dim3(1), dim3(256), 0, b->s0, reinterpret_cast<double*>(h->d_nonfinite), 1LL);
  }
  if (b->serial) {
    for (auto& g : b->g)
      if (e == hipSuccess) e = g.kern->launch(g.d_args, g.d_wstart, (int)g.inst.size(), g.nwork, g.grid, b->s0);
  } else {
    e = hipEventRecord(b->fork, b->s0);
    for (auto& g : b->g) {
      if (e == hipSuccess) e = hipStreamWaitEvent(g.s, b->fork, 0);
      if (e == hipSuccess) e = g.kern->launch(g.d_args, g.d_wstart, (int)g.inst.size(), g.nwork, g.grid, g.s);
      if (e == hipSuccess) e = hipEventRecord(g.join, g.s);
      if (e == hipSuccess) e = hipStreamWaitEvent(b->s0, g.join, 0);
    }
  }
  hipGraph_t graph = nullptr;
  hipError_t e2 = hipStreamEndCapture(b->s0, &graph);
  if (e != hipSuccess || e2 != hipSuccess) {
    if (graph) (void)hipGraphDestroy(graph);
    return fail(XSH_EDEVICE, "batch graph capture: %s", hipGetErrorString(e != hipSuccess ? e : e2));
  }
  b->graph = graph;
  HIPCHK(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
  return XSH_OK;
}

extern "C" int xsh_batch_create(int n, const xsh_desc* descs, int flags, xsh_batch** out) {
  if (!out || !descs) return fail(XSH_EINVAL, "NULL argument");
  *out = nullptr;
  if (n < 1) return fail(XSH_EINVAL, "a batch needs at least one instance");
  auto* b = new xsh_batch();
  b->serial = (flags & XSH_BATCH_SERIAL) != 0;
  b->h.assign(n, nullptr);
  for (int i = 0; i < n; ++i) {
    if (descs[i].dimensions < 2) {
      batch_free(b);
      return fail(XSH_EUNSUPPORTED, "instance %d: batches hold lattice problems (dimensions >= 2)", i);
    }
    int rc = xsh_create(&descs[i], &b->h[i]);
    if (rc) {
      const std::string msg = g_err;
      batch_free(b);
      return fail(rc, "instance %d: %s", i, msg.c_str());
    }
  }
  // group the instances by lattice kernel family
  for (int i = 0; i < n; ++i) {
    const xsh_handle* h = b->h[i];
    const bool tpb1 = h->d.traj_out || h->d.nmodes;
    BatchGroup* gp = nullptr;
    for (auto& g : b->g)
      if (g.precision == h->d.precision && g.nx == h->nx && g.ny == h->ny && g.mat == h->mat && g.kern == h->kern)
        gp = &g;
    if (!gp) {
      BatchGroup g{};
      g.precision = h->d.precision;
      g.nx = h->nx;
      g.ny = h->ny;
      g.mat = h->mat;
      g.kern = h->kern;
      b->g.push_back(g);
      gp = &b->g.back();
    }
    gp->need_tpb1 = gp->need_tpb1 || tpb1;
    gp->inst.push_back(i);
  }
  auto bail = [&](int rc) {
    batch_free(b);
    return rc;
  };
  if (hipStreamCreate(&b->s0) != hipSuccess || hipEventCreate(&b->ev0) != hipSuccess ||
      hipEventCreate(&b->ev1) != hipSuccess || hipEventCreateWithFlags(&b->fork, hipEventDisableTiming) != hipSuccess)
    return bail(fail(XSH_EDEVICE, "batch streams/events"));
  if (hipHostMalloc(&b->h_args, n * sizeof(xsh::PathArgs), hipHostMallocDefault) != hipSuccess)
    return bail(fail(XSH_ENOMEM, "batch argument staging"));
  for (auto& g : b->g) {
    if (hipStreamCreate(&g.s) != hipSuccess || hipEventCreateWithFlags(&g.join, hipEventDisableTiming) != hipSuccess)
      return bail(fail(XSH_EDEVICE, "batch group stream"));
    if (hipMalloc(&g.d_args, g.inst.size() * sizeof(xsh::PathArgs)) != hipSuccess ||
        hipMalloc(&g.d_wstart, (g.inst.size() + 1) * sizeof(int)) != hipSuccess)
      return bail(fail(XSH_ENOMEM, "batch tables"));
    // candidate kernels: every compiled variant of this lattice (streaming kernels have one)
    const xsh::KernelEntry* v[16];
    const int nv = g.kern->stream ? 0 : xsh::kernel_variants(g.precision, g.nx, g.ny, g.mat, v, 16);
    g.cand.assign(1, g.kern);
    for (int k = 0; k < nv; ++k)
      if (v[k] != g.kern && (!g.need_tpb1 || v[k]->tpb == 1) && (!g.kern->kn || v[k]->kn)) g.cand.push_back(v[k]);
    int rc = batch_layout(g, b->h, 0);
    if (rc) return bail(rc);
  }
  if (flags & XSH_BATCH_AUTOTUNE) {
    // per group: every variant x persistent grid, timed on a shortened run
    int cus = 256;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, 0) == hipSuccess && prop.multiProcessorCount > 0) cus = prop.multiProcessorCount;
    for (auto& g : b->g) {
      const xsh::KernelEntry* best_k = g.kern;
      int best_grid = 0;
      double best = 1e300;
      for (const auto* k : g.cand) {
        g.kern = k;
        const int grids[] = {0, cus, 2 * cus, 4 * cus, 8 * cus};
        for (int gr : grids) {
          int rc = batch_layout(g, b->h, gr);
          if (rc) return bail(rc);
          if (gr > 0 && g.grid >= g.nwork) continue;  // same as one item per workgroup
          double t = 0.0;
          rc = batch_time(b, g, &t);
          if (rc) return bail(rc);
          if (t < best) {
            best = t;
            best_k = k;
            best_grid = gr;
          }
        }
      }
      g.kern = best_k;
      int rc = batch_layout(g, b->h, best_grid);
      if (rc) return bail(rc);
    }
  }
  int rc = batch_capture(b);
  if (rc) return bail(rc);
  *out = b;
  return XSH_OK;
}

extern "C" void xsh_batch_destroy(xsh_batch* b) { batch_free(b); }

extern "C" int xsh_batch_run(xsh_batch* b, int npar) {
  if (!b) return fail(XSH_EINVAL, "batch is NULL");
  for (auto* h : b->h)
    if (npar < 1 || npar > h->e3) return fail(XSH_EINVAL, "npar %d outside 1..ensembles(3)", npar);
  int rc = batch_upload(b, npar);
  if (rc) return rc;
  HIPCHK(hipEventRecord(b->ev0, b->s0));
  HIPCHK(hipGraphLaunch(b->exec, b->s0));
  HIPCHK(hipEventRecord(b->ev1, b->s0));
  for (auto* h : b->h) h->ran = false;  // the handles' own streams did not run
  b->ran = true;
  return XSH_OK;
}

extern "C" int xsh_batch_synchronize(xsh_batch* b, double* ms) {
  if (!b) return fail(XSH_EINVAL, "batch is NULL");
  HIPCHK(hipStreamSynchronize(b->s0));
  if (ms) {
    if (!b->ran) return fail(XSH_ESTATE, "nothing has run on this batch");
    float f = 0.f;
    HIPCHK(hipEventElapsedTime(&f, b->ev0, b->ev1));
    *ms = f;
  }
  return XSH_OK;
}

extern "C" int xsh_batch_read_averages(xsh_batch* b, int i, double* av_out) {
  if (!b || !av_out) return fail(XSH_EINVAL, "NULL argument");
  if (i < 0 || i >= (int)b->h.size()) return fail(XSH_EINVAL, "instance %d outside 0..%d", i, (int)b->h.size() - 1);
  if (!b->ran) return fail(XSH_ESTATE, "xsh_batch_run has not run");
  xsh_handle* h = b->h[i];
  HIPCHK(hipStreamSynchronize(b->s0));
  return copy_averages(h, h->d_av, h->e2 * h->nchk, av_out, b->s0);
}

extern "C" int xsh_batch_nonfinite(xsh_batch* b, int i, int* flag) {
  if (!b || !flag) return fail(XSH_EINVAL, "NULL argument");
  if (i < 0 || i >= (int)b->h.size()) return fail(XSH_EINVAL, "instance %d outside 0..%d", i, (int)b->h.size() - 1);
  HIPCHK(hipStreamSynchronize(b->s0));
  HIPCHK(hipMemcpy(flag, b->h[i]->d_nonfinite, sizeof(int), hipMemcpyDeviceToHost));
  return XSH_OK;
}

extern "C" int xsh_batch_query(const xsh_batch* b, int i, xsh_info* info) {
  if (!b || !info) return fail(XSH_EINVAL, "NULL argument");
  if (i < 0 || i >= (int)b->h.size()) return fail(XSH_EINVAL, "instance %d outside 0..%d", i, (int)b->h.size() - 1);
  return xsh_query(b->h[i], info);
}

extern "C" int xsh_batch_plan(const xsh_batch* b, int* ngroups, int* plan) {
  if (!b || !ngroups) return fail(XSH_EINVAL, "NULL argument");
  *ngroups = (int)b->g.size();
  if (plan)
    for (size_t k = 0; k < b->g.size(); ++k) {
      const BatchGroup& g = b->g[k];
      int* p = plan + 8 * k;
      p[0] = g.precision;
      p[1] = g.nx;
      p[2] = g.ny;
      p[3] = (int)g.inst.size();
      p[4] = g.kern->threads;
      p[5] = g.kern->tpb;
      p[6] = g.kern->radix;
      p[7] = g.grid;
    }
  return XSH_OK;
}
