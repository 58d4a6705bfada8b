// Rectangular dense-transform kernels (128 x 64 and 64 x 128, FP32 and FP64):
// 2D lattices wider than 64 sites in one dimension, padded per dimension.
#include "xsh_kernels.hpp"

namespace xsh {

static const KernelEntry kEntries[] = {
    XSH_MAT_ENTRY(float, 128, 64, 512, 1),   XSH_MAT_ENTRY(float, 64, 128, 512, 1),
    XSH_MAT_ENTRY(double, 128, 64, 512, 1),  XSH_MAT_ENTRY(double, 64, 128, 512, 1),
};

const KernelEntry* table_mat_pad(int* n) {
  *n = (int)(sizeof(kEntries) / sizeof(kEntries[0]));
  return kEntries;
}

}  // namespace xsh
