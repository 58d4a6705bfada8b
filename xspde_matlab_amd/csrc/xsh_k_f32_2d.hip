// Fp32 fft kernels, 2d lattices (compiled without the k-space noise code; the
// entries with it are in xsh_k_kn2d.hip).
#include "xsh_kernels.hpp"

namespace xsh {

static const KernelEntry kEntries[] = {
    XSH_FAST_ENTRY(float, 32, 32, 256, 1), XSH_FAST_ENTRY(float, 64, 64, 512, 1),
    XSH_FAST_ENTRY(float, 128, 128, 1024, 1),
};

const KernelEntry* table_f32_2d(int* n) {
  *n = (int)(sizeof(kEntries) / sizeof(kEntries[0]));
  return kEntries;
}

}  // namespace xsh
