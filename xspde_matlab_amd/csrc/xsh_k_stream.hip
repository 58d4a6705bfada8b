// Streaming kernels: lattices whose fields do not fit on chip (fields in a
// device-memory workspace, FFT or dense transforms through LDS half a plane at
// a time).
#include "xsh_stream.hpp"

namespace xsh {

static const KernelEntry kStreamEntries[] = {
    XSH_STREAM_ENTRY(double, 128, 128, 1024),
};

const KernelEntry* table_stream_mat(int* n);

const KernelEntry* find_stream_kernel(int precision, int nx, int ny, int mat) {
  for (const auto& e : kStreamEntries)
    if (e.precision == precision && e.nx == nx && e.ny == ny && e.mat == mat) return &e;
  int n = 0;
  const KernelEntry* e = table_stream_mat(&n);
  for (int i = 0; i < n; ++i)
    if (e[i].precision == precision && e[i].nx == nx && e[i].ny == ny && e[i].mat == mat) return &e[i];
  return nullptr;
}

}  // namespace xsh
