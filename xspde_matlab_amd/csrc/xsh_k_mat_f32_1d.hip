// Fp32 dense-transform kernels of 1D lattices (non-periodic, mixed, padded): 64 and 128 sites.
#include "xsh_kernels.hpp"

namespace xsh {

static const KernelEntry kEntries[] = {
    XSH_MAT_ENTRY(float, 64, 1, 256, 64), XSH_MAT_ENTRY(float, 128, 1, 256, 64),
};

const KernelEntry* table_mat_f32_1d(int* n) {
  *n = (int)(sizeof(kEntries) / sizeof(kEntries[0]));
  return kEntries;
}

}  // namespace xsh
