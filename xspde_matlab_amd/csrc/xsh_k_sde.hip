// SDE kernels, noise probe, and the lookup over every family's launch table.
#include "xsh_kernels.hpp"

namespace xsh {

const KernelEntry* table_f32_1d(int* n);
const KernelEntry* table_f32_2d(int* n);
const KernelEntry* table_f64(int* n);
const KernelEntry* table_mat_f32(int* n);
const KernelEntry* table_mat_f32_1d(int* n);
const KernelEntry* table_mat_f32_1d256(int* n);
const KernelEntry* table_f32_var2d(int* n);
const KernelEntry* table_mat_f64_1d256(int* n);
const KernelEntry* table_kn_f32(int* n);
const KernelEntry* table_kn_f64(int* n);
const KernelEntry* table_kn_f32_2d(int* n);
const KernelEntry* table_mat_f64(int* n);
const KernelEntry* table_f32_var(int* n);
const KernelEntry* table_mat_pad(int* n);

int kernel_variants(int precision, int nx, int ny, int mat, const KernelEntry** out, int max) {
  const KernelEntry* (*tables[])(int*) = {table_f32_1d,     table_f32_2d,        table_f64,
                                          table_mat_f32,    table_mat_f32_1d,    table_mat_f32_1d256,
                                          table_mat_f64,    table_mat_f64_1d256, table_f32_var,
                                          table_f32_var2d,  table_kn_f32,        table_kn_f64,
                                          table_kn_f32_2d,  table_mat_pad};
  int k = 0;
  for (auto tab : tables) {
    int n = 0;
    const KernelEntry* e = tab(&n);
    for (int i = 0; i < n && k < max; ++i)
      if (e[i].precision == precision && e[i].nx == nx && e[i].ny == ny && e[i].mat == mat) out[k++] = &e[i];
  }
  return k;
}

const KernelEntry* find_kernel(int precision, int nx, int ny, int mat, int kn) {
  const KernelEntry* (*tables[])(int*) = {table_f32_1d,     table_f32_2d,        table_f64,
                                          table_mat_f32,    table_mat_f32_1d,    table_mat_f32_1d256,
                                          table_mat_f64,    table_mat_f64_1d256, table_kn_f32,
                                          table_kn_f64,     table_kn_f32_2d,     table_mat_pad};
  for (auto tab : tables) {
    int n = 0;
    const KernelEntry* e = tab(&n);
    for (int i = 0; i < n; ++i)
      if (e[i].precision == precision && e[i].nx == nx && e[i].ny == ny && e[i].mat == mat && (!kn || e[i].kn))
        return &e[i];
  }
  return nullptr;
}

const KernelEntry* find_padded_kernel(int precision, int nx) {
  const KernelEntry* (*tables[])(int*) = {table_mat_f32_1d, table_mat_f32_1d256, table_mat_f64,
                                          table_mat_f64_1d256};
  const KernelEntry* best = nullptr;
  for (auto tab : tables) {
    int n = 0;
    const KernelEntry* e = tab(&n);
    for (int i = 0; i < n; ++i)
      if (e[i].precision == precision && e[i].ny == 1 && e[i].mat == 1 && e[i].nx >= nx && (!best || e[i].nx < best->nx))
        best = &e[i];
  }
  return best;
}

const KernelEntry* find_padded_kernel2d(int precision, int nx, int ny) {
  const KernelEntry* (*tables[])(int*) = {table_mat_f32, table_mat_f64, table_mat_pad};
  const KernelEntry* best = nullptr;
  for (auto tab : tables) {
    int n = 0;
    const KernelEntry* e = tab(&n);
    for (int i = 0; i < n; ++i)
      if (e[i].precision == precision && e[i].mat == 1 && e[i].ny > 1 && e[i].nx * e[i].ny < kPadMaxSites &&
          e[i].nx >= nx && e[i].ny >= ny && (!best || e[i].nx * e[i].ny < best->nx * best->ny))
        best = &e[i];
  }
  return best;
}

hipError_t launch_sde_path(int precision, const PathArgs* A, int nblocks, hipStream_t s) {
  return precision == XSH_FP64 ? launch_sde<double>(A, nblocks, s) : launch_sde<float>(A, nblocks, s);
}

hipError_t launch_sde_fields(int precision, int fields, const PathArgs* A, int nblocks, hipStream_t s) {
  const bool f64 = precision == XSH_FP64;
  switch (fields) {
    case 2: return f64 ? launch_sde_f<double, 2>(A, nblocks, s) : launch_sde_f<float, 2>(A, nblocks, s);
    case 3: return f64 ? launch_sde_f<double, 3>(A, nblocks, s) : launch_sde_f<float, 3>(A, nblocks, s);
    case 4: return f64 ? launch_sde_f<double, 4>(A, nblocks, s) : launch_sde_f<float, 4>(A, nblocks, s);
    default: return hipErrorInvalidValue;
  }
}

// Time-domain spectra of an SDE ensemble (xpath.m:90-106): per trajectory and
// kept frequency m, a(w_m) = fsc/M sum_n a_n exp(+2 pi i j_m n / M) (MATLAB's
// ifft times fsc, fftshifted, at specpts: j_m = m - floor(points/2) mod M) over
// the M step fields sde_path stored, then the observables flagged in
// tspec_mask of it, reduced over the trajectories like xdata (FP64
// throughout).  Grid (trajectory blocks, points, paths); tw[k] = exp(2 pi i k / M).
__global__ __launch_bounds__(256) void sde_spectrum(const PathArgs* __restrict__ Ap, const double* __restrict__ tw,
                                                    double fsc) {
  const PathArgs& A = *Ap;
  __shared__ double red[8];
  const int traj = blockIdx.x * 256 + threadIdx.x;
  const int m = blockIdx.y, path = blockIdx.z;
  const bool active = traj < A.E;
  const int M = A.ft_len;
  const int jm = (((m - A.npts / 2) % M) + M) % M;
  double xr = 0.0, xi = 0.0;
  if (active) {
    const double* f = A.ft + 2 * (long long)path * M * A.E;
    int k = 0;
    for (int n = 0; n < M; ++n) {
      const long long at = 2 * ((long long)n * A.E + traj);
      const double ar = f[at], ai = f[at + 1], wr = tw[2 * k], wi = tw[2 * k + 1];
      xr += ar * wr - ai * wi;
      xi += ar * wi + ai * wr;
      k += jm;
      if (k >= M) k -= M;
    }
    xr *= fsc / M;
    xi *= fsc / M;
  }
  double* av = A.av + 2 * (long long)path * A.avlen;
  for (int o = 0; o < A.nobs; ++o) {
    if (!((A.tspec_mask >> o) & 1)) continue;
    double re = 0.0, im = 0.0;
    if (active) {
      const double n2 = xr * xr + xi * xi;
      switch (A.obs[o]) {
        case XSH_OBS_A:
        case XSH_OBS_A_AVE:
          re = xr;
          im = xi;
          break;
        case XSH_OBS_ABS4_AVE:
          re = n2 * n2;
          break;
        default:
          re = n2;
      }
    }
    block_sum2<256>(re, im, red);
    if (threadIdx.x == 0) atomic_add_c(av + 2 * (A.obs_off[o] + m), re * A.inv_E, im * A.inv_E);
  }
}

hipError_t launch_sde_spectrum(const PathArgs* A, const double* tw, double fsc, int E, int npts, int npaths,
                               hipStream_t s) {
  hipLaunchKernelGGL(sde_spectrum, dim3((unsigned)((E + 255) / 256), (unsigned)npts, (unsigned)npaths), dim3(256), 0,
                     s, A, tw, fsc);
  return hipGetLastError();
}

hipError_t launch_probe_normals(int precision, uint32_t nsp, uint32_t tag, uint32_t seq, uint32_t inst,
                                uint32_t draw, int traj0, int ntraj, int count, double* out, hipStream_t s) {
  const int groups = (count + 3) / 4;
  const int total = ntraj * groups;
  const int nb = (total + 255) / 256;
  if (precision == XSH_FP64)
    hipLaunchKernelGGL(probe_normals<double>, dim3(nb), dim3(256), 0, s, nsp, tag, seq, inst, draw, traj0,
                       ntraj, count, out);
  else
    hipLaunchKernelGGL(probe_normals<float>, dim3(nb), dim3(256), 0, s, nsp, tag, seq, inst, draw, traj0,
                       ntraj, count, out);
  return hipGetLastError();
}

}  // namespace xsh
