// Chunk-path kernel instantiations for element type HEX8.
#include "femhot_chunk_kernels.h"

namespace femhot {
template <>
int launch_chunk_et<HEX8>(int dtype, int ph, int ngp, int mode, int block, dim3 grid, size_t shm, hipStream_t s,
                         const ChunkArgs &A) {
  return launch_chunk_et_impl<HEX8>(dtype, ph, ngp, mode, block, grid, shm, s, A);
}
}  // namespace femhot
