// Fp32 fft kernel variants of 2D lattices for the batch autotuner: other
// thread counts per trajectory than the default entries of the 2D table.
#include "xsh_kernels.hpp"

namespace xsh {

static const KernelEntry kEntries[] = {
    XSH_VAR_ENTRY(float, 32, 32, 256, 2, 0), XSH_VAR_ENTRY(float, 64, 64, 256, 1, 0),
    XSH_VAR_ENTRY(float, 64, 64, 1024, 1, 0), XSH_VAR_ENTRY(float, 128, 128, 512, 1, 0),
};

const KernelEntry* table_f32_var2d(int* n) {
  *n = (int)(sizeof(kEntries) / sizeof(kEntries[0]));
  return kEntries;
}

}  // namespace xsh
