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
