// LDS-staged chunk assembly (persistent launch over a precomputed chunk plan).
#pragma once
#include "femhot_internal.h"

namespace femhot {

// Runs MODE_RES / MODE_VJP / MODE_EGRAD over the chunk plan of `epc` elements per chunk
// (0 = autotuned default) with `block` threads per workgroup (0 = default).
int run_chunk(Ctx *c, int dtype, int mode, int batch, const void *field, long long fs, const void *u, long long us,
              const void *w, long long ws, void *ou, long long ous, void *od, long long ods, double *energy,
              unsigned flags, hipStream_t stream, int epc, int block);

// Times chunk configurations (elements per chunk x workgroup size x grid shape) on the given
// inputs with HIP events and records the fastest for (ctx, mode, dtype, batch, field location).
// Outputs are scratch while tuning.  Synchronises the stream; not for use under graph capture.
int autotune_chunk(Ctx *c, int dtype, int mode, int batch, const void *field, long long fs, const void *u,
                   long long us, const void *w, long long ws, void *ou, long long ous, void *od, long long ods,
                   double *energy, unsigned flags, hipStream_t s, int *best_epc, int *best_block, int *best_grid,
                   float *best_ms);

int set_chunk_config(Ctx *c, int mode, int dtype, int batch, unsigned flags, int epc, int block, int grid_per_cu);

void forget_tuning(const void *ctx);

// True when a call of this shape should take the plain path even though FEMHOT_SCATTER_CHUNK is set:
// the autotuner found the plain path faster, or, untuned, the call is FP64 Neo-Hookean (its element
// routine needs most of the register file, so only small chunks with small workgroups keep the chunk
// kernel ahead; femhot_autotune times both and records the winner).
bool chunk_prefers_plain(const Ctx *c, int dtype, int mode, int batch, unsigned flags);

// Records that the plain path won the autotuner for this call shape.
void record_plain(const Ctx *c, int dtype, int mode, int batch, unsigned flags);

}  // namespace femhot
