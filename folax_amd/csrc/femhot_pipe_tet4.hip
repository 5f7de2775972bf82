// Pipelined (LDS-DMA) chunk-kernel instantiations for element type TET4.
#include "femhot_chunk_pipe.h"

namespace femhot {
template <>
int launch_chunk_pipe_et<TET4>(int ph, int ngp, int mode, int block, dim3 grid, size_t shm, hipStream_t s,
                             const ChunkArgs &A) {
  return launch_chunk_pipe_et_impl<TET4>(ph, ngp, mode, block, grid, shm, s, A);
}
}  // namespace femhot
