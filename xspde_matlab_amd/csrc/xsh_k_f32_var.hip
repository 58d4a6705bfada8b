// Fp32 fft kernel variants for the batch autotuner: other tiles (threads,
// trajectories per workgroup) and first radices of the x transform than the
// default entries of the 1D table (the 2D variants are in xsh_k_f32_var2d.hip).
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
};

const KernelEntry* table_f32_var(int* n) {
  *n = (int)(sizeof(kEntries) / sizeof(kEntries[0]));
  return kEntries;
}

}  // namespace xsh
