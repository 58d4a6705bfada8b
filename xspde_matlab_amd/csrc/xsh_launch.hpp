// Launch table shared by the kernels and the C ABI implementation.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace xsh {
struct PathArgs;

// PathArgs pointers are device pointers (the arguments live in device memory)
// args[ninst] instances; work items [wstart[i], wstart[i+1]) belong to instance i
// (wstart may be NULL for one instance); grid workgroups stride over nwork items
typedef hipError_t (*SpdeLaunch)(const PathArgs* args, const int* wstart, int ninst, int nwork, int grid,
                                 hipStream_t);
typedef hipError_t (*ProbeLaunch)(const PathArgs*, int nci, int nplanes, void* planes, hipStream_t);

struct KernelEntry {
  int precision, nx, ny, threads, tpb;
  int mat;  // dense per-dimension transforms (non-periodic lattices)
  SpdeLaunch launch;
  ProbeLaunch probe;
  int stream;  // fields in a device-memory workspace (lattices too large for the chip)
  int radix;   // first radix of the x transforms (FFT kernels)
  int kn;      // has the k-space filtered noise code (knoises, krandoms)
};

// persistent grid of the streaming kernels: at most this many workgroups,
// each with its own workspace planes
constexpr int kStreamMaxBlocks = 1024;
// 2D dense-transform kernels of fewer sites than this handle padded lattices
// (the 128 x 128 kernels keep their compile-time size)
constexpr int kPadMaxSites = 128 * 128;

// kn: the problem has k-space filtered noises or randoms (entries without that
// code, compiled for speed, are skipped)
const KernelEntry* find_kernel(int precision, int nx, int ny, int mat, int kn = 0);
// all compiled variants (threads, trajectories per workgroup, radix) of a lattice kernel
int kernel_variants(int precision, int nx, int ny, int mat, const KernelEntry** out, int max);
const KernelEntry* find_stream_kernel(int precision, int nx, int ny, int mat);
// 1D dense-transform kernel of the smallest compiled size >= nx (the lattice is
// padded to that size: the padding sites take no part in the physical ones)
const KernelEntry* find_padded_kernel(int precision, int nx);
// 2D dense-transform kernel of the fewest sites with both sizes >= nx, ny among
// those that handle padding (fewer than kPadMaxSites: 64 x 64, 128 x 64, 64 x 128)
const KernelEntry* find_padded_kernel2d(int precision, int nx, int ny);
hipError_t launch_sde_path(int precision, const PathArgs* A, int nblocks, hipStream_t s);
// coupled SDE components (A->fields = fields, 2..XSH_MAX_FIELDS)
hipError_t launch_sde_fields(int precision, int fields, const PathArgs* A, int nblocks, hipStream_t s);
// time-domain spectra of the step fields sde_path stored (tw: exp(2 pi i k / M), k < M)
hipError_t launch_sde_spectrum(const PathArgs* A, const double* tw, double fsc, int E, int npts, int npaths,
                               hipStream_t s);
hipError_t launch_probe_normals(int precision, uint32_t nsp, uint32_t tag, uint32_t seq, uint32_t inst,
                                uint32_t draw, int traj0, int ntraj, int count, double* out, hipStream_t s);
}  // namespace xsh
