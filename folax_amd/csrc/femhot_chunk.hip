#include "femhot_chunk.h"

namespace femhot {
int run_chunk(Ctx *, int, int, int, const void *, long long, const void *, long long, const void *, long long, void *,
              long long, void *, long long, double *, unsigned, hipStream_t, int, int) {
  set_error("chunk scatter not built yet");
  return FEMHOT_ERR_UNSUPPORTED;
}
}  // namespace femhot
