// Fp32 fft kernels of 2D lattices with the k-space filtered noise code
// (knoises, krandoms): looked up only for problems that use it.
#include "xsh_kernels.hpp"

namespace xsh {

static const KernelEntry kEntries[] = {
    XSH_ENTRY(float, 32, 32, 256, 1), XSH_ENTRY(float, 64, 64, 512, 1),
    XSH_ENTRY(float, 128, 128, 1024, 1),
};

const KernelEntry* table_kn_f32_2d(int* n) {
  *n = (int)(sizeof(kEntries) / sizeof(kEntries[0]));
  return kEntries;
}

}  // namespace xsh
