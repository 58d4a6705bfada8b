// Fp32 dense-transform kernel of 1D lattices of 256 sites (its own translation
// unit: the largest kernel, compiled in parallel with the others).
#include "xsh_kernels.hpp"

namespace xsh {

static const KernelEntry kEntries[] = {
    XSH_MAT_ENTRY(float, 256, 1, 512, 64),
};

const KernelEntry* table_mat_f32_1d256(int* n) {
  *n = (int)(sizeof(kEntries) / sizeof(kEntries[0]));
  return kEntries;
}

}  // namespace xsh
