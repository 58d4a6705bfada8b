// Fp64 dense-transform kernel of 1D lattices of 256 sites: 32 trajectories per
// workgroup (64 would need 256 KB of LDS).
#include "xsh_kernels.hpp"

namespace xsh {

static const KernelEntry kEntries[] = {
    XSH_MAT_ENTRY(double, 256, 1, 512, 32),
};

const KernelEntry* table_mat_f64_1d256(int* n) {
  *n = (int)(sizeof(kEntries) / sizeof(kEntries[0]));
  return kEntries;
}

}  // namespace xsh
