// Streaming kernel with dense transforms (non-periodic FP64 128x128 lattices):
// its own translation unit, compiled in parallel with the others.
#include "xsh_stream.hpp"

namespace xsh {

static const KernelEntry kEntries[] = {
    XSH_STREAM_MAT_ENTRY(double, 128, 128, 1024),
};

const KernelEntry* table_stream_mat(int* n) {
  *n = (int)(sizeof(kEntries) / sizeof(kEntries[0]));
  return kEntries;
}

}  // namespace xsh
