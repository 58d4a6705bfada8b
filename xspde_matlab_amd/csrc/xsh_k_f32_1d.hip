// Fp32 fft kernels, 1d lattices (compiled without the k-space noise code; the
// entries with it are in xsh_k_kn.hip).
#include "xsh_kernels.hpp"

namespace xsh {

static const KernelEntry kEntries[] = {
    XSH_FAST_ENTRY(float, 32, 1, 256, 32), XSH_FAST_ENTRY(float, 64, 1, 256, 32),
    XSH_FAST_ENTRY(float, 128, 1, 256, 32), XSH_FAST_ENTRY(float, 256, 1, 256, 16),
    XSH_FAST_ENTRY(float, 512, 1, 256, 8), XSH_FAST_ENTRY(float, 1024, 1, 256, 4),
};

const KernelEntry* table_f32_1d(int* n) {
  *n = (int)(sizeof(kEntries) / sizeof(kEntries[0]));
  return kEntries;
}

}  // namespace xsh
