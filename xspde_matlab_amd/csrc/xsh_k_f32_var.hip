// Fp32 fft kernel variants for the batch autotuner: other tiles (threads,
// trajectories per workgroup) and first radices of the x transform than the
// default entries of the 1D and 2D tables.
#include "xsh_kernels.hpp"

namespace xsh {

static const KernelEntry kEntries[] = {
    // 1D: tile (TPB) and radix variants
    XSH_VAR_ENTRY(float, 32, 1, 256, 16, 0), XSH_VAR_ENTRY(float, 32, 1, 256, 32, 8),
    XSH_VAR_ENTRY(float, 64, 1, 256, 16, 0), XSH_VAR_ENTRY(float, 64, 1, 256, 32, 16),
    XSH_VAR_ENTRY(float, 128, 1, 256, 16, 0), XSH_VAR_ENTRY(float, 128, 1, 256, 32, 8),
    XSH_VAR_ENTRY(float, 256, 1, 256, 8, 0), XSH_VAR_ENTRY(float, 256, 1, 256, 16, 8),
    XSH_VAR_ENTRY(float, 256, 1, 256, 16, 32),
    XSH_VAR_ENTRY(float, 512, 1, 256, 4, 0), XSH_VAR_ENTRY(float, 512, 1, 256, 8, 32),
    // 2D: threads per trajectory
    XSH_VAR_ENTRY(float, 32, 32, 256, 2, 0), XSH_VAR_ENTRY(float, 64, 64, 256, 1, 0),
    XSH_VAR_ENTRY(float, 64, 64, 1024, 1, 0), XSH_VAR_ENTRY(float, 128, 128, 512, 1, 0),
};

const KernelEntry* table_f32_var(int* n) {
  *n = (int)(sizeof(kEntries) / sizeof(kEntries[0]));
  return kEntries;
}

}  // namespace xsh
