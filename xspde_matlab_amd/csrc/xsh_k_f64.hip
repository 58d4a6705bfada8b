// Fp64 fft kernels (compiled without the k-space noise code; the entries with
// it are in xsh_k_kn.hip and xsh_k_kn64.hip).
#include "xsh_kernels.hpp"

namespace xsh {

static const KernelEntry kEntries[] = {
    XSH_FAST_ENTRY(double, 32, 1, 256, 16), XSH_FAST_ENTRY(double, 64, 1, 256, 16),
    XSH_FAST_ENTRY(double, 128, 1, 256, 16), XSH_FAST_ENTRY(double, 256, 1, 256, 8),
    XSH_FAST_ENTRY(double, 512, 1, 256, 4), XSH_FAST_ENTRY(double, 1024, 1, 256, 2),
    XSH_FAST_ENTRY(double, 32, 32, 256, 1), XSH_FAST_ENTRY(double, 64, 64, 512, 1),
};

const KernelEntry* table_f64(int* n) {
  *n = (int)(sizeof(kEntries) / sizeof(kEntries[0]));
  return kEntries;
}

}  // namespace xsh
