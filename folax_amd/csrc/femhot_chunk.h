// LDS-staged chunk assembly (persistent launch over a precomputed chunk plan).
#pragma once
#include "femhot_internal.h"

namespace femhot {

// Runs MODE_RES / MODE_VJP / MODE_EGRAD over the chunk plan of `epc` elements per chunk
// (0 = autotuned default) with `block` threads per workgroup (0 = default).
int run_chunk(Ctx *c, int dtype, int mode, int batch, const void *field, long long fs, const void *u, long long us,
              const void *w, long long ws, void *ou, long long ous, void *od, long long ods, double *energy,
              unsigned flags, hipStream_t stream, int epc, int block);

}  // namespace femhot
