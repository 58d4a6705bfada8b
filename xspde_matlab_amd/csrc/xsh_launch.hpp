// Launch table shared by the kernels and the C ABI implementation.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace xsh {
struct PathArgs;

// PathArgs pointers are device pointers (the arguments live in device memory)
typedef hipError_t (*SpdeLaunch)(const PathArgs*, int nblocks, hipStream_t);
typedef hipError_t (*ProbeLaunch)(const PathArgs*, int nci, int nplanes, void* planes, hipStream_t);

struct KernelEntry {
  int precision, nx, ny, threads, tpb;
  int mat;  // dense per-dimension transforms (non-periodic lattices)
  SpdeLaunch launch;
  ProbeLaunch probe;
};

const KernelEntry* find_kernel(int precision, int nx, int ny, int mat);
hipError_t launch_sde_path(int precision, const PathArgs* A, int nblocks, hipStream_t s);
hipError_t launch_probe_normals(int precision, uint32_t nsp, uint32_t tag, uint32_t seq, uint32_t inst,
                                uint32_t draw, int traj0, int ntraj, int count, double* out, hipStream_t s);
}  // namespace xsh
