// Fp32 dense-transform kernels of 2D lattices (non-periodic and mixed).
#include "xsh_kernels.hpp"

namespace xsh {

static const KernelEntry kEntries[] = {
    XSH_MAT_ENTRY(float, 64, 64, 512, 1), XSH_MAT_ENTRY(float, 128, 128, 512, 1),
};

const KernelEntry* table_mat_f32(int* n) {
  *n = (int)(sizeof(kEntries) / sizeof(kEntries[0]));
  return kEntries;
}

}  // namespace xsh
