// Streaming kernels: lattices whose fields do not fit on chip (fields in a
// device-memory workspace, FFT through LDS half a plane at a time).
#include "xsh_stream.hpp"

namespace xsh {

static const KernelEntry kStreamEntries[] = {
    XSH_STREAM_ENTRY(double, 128, 128, 512),
};

const KernelEntry* find_stream_kernel(int precision, int nx, int ny) {
  for (const auto& e : kStreamEntries)
    if (e.precision == precision && e.nx == nx && e.ny == ny) return &e;
  return nullptr;
}

}  // namespace xsh
