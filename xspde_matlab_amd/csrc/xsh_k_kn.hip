// Fp32 fft kernels of 1D lattices with the k-space filtered noise code
// (knoises, krandoms): looked up only for problems that use it.
#include "xsh_kernels.hpp"

namespace xsh {

static const KernelEntry kEntries[] = {
    XSH_ENTRY(float, 32, 1, 256, 32), XSH_ENTRY(float, 64, 1, 256, 32),
    XSH_ENTRY(float, 128, 1, 256, 32), XSH_ENTRY(float, 256, 1, 256, 16),
    XSH_ENTRY(float, 512, 1, 256, 8), XSH_ENTRY(float, 1024, 1, 256, 4),
};

const KernelEntry* table_kn_f32(int* n) {
  *n = (int)(sizeof(kEntries) / sizeof(kEntries[0]));
  return kEntries;
}

}  // namespace xsh
