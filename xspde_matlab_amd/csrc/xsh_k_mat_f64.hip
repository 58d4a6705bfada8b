// Fp64 dense-transform kernels (non-periodic and mixed lattices).
#include "xsh_kernels.hpp"

namespace xsh {

static const KernelEntry kEntries[] = {
    XSH_MAT_ENTRY(double, 64, 1, 256, 64), XSH_MAT_ENTRY(double, 128, 1, 512, 64),
    XSH_MAT_ENTRY(double, 64, 64, 512, 1),
};

const KernelEntry* table_mat_f64(int* n) {
  *n = (int)(sizeof(kEntries) / sizeof(kEntries[0]));
  return kEntries;
}

}  // namespace xsh
