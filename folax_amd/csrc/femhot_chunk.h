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

// The chunk path on the library's fixed device scratch: allocates nothing, requires the call shape's
// plan to be built, rejects calls whose partials exceed the fixed size.
int run_chunk_fixed_scratch(Ctx *c, int dtype, int mode, int batch, const void *field, long long fs, const void *u,
                            long long us, void *ou, long long ous, void *od, long long ods, double *energy,
                            unsigned flags, hipStream_t s);

void forget_tuning(const void *ctx);

// The scatter a call takes when FEMHOT_SCATTER_CHUNK leaves the choice to the library: the LDS-staged
// chunk kernel, the element-parallel kernel with hardware atomics, or the element-parallel kernel one
// colour class at a time with plain read-modify-writes.
enum ScatterPath { PATH_CHUNK = 0, PATH_ATOMIC = 1, PATH_COLOUR = 2, PATH_ROWS = 3 /* tangent only */ };
// mode of the tangent assembly in the tuning table (next to MODE_RES / MODE_VJP / MODE_EGRAD)
constexpr int MODE_TANGENT = 3;

// The path femhot_autotune measured fastest for this call shape; untuned: the chunk kernel, except
// FP64 Neo-Hookean (its element routine needs most of the register file, so only small chunks with
// small workgroups keep the chunk kernel ahead) which takes the atomic path, and the tangent, which
// takes the row-owner assembly (colour-ordered adds for Neo-Hookean).
int tuned_path(const Ctx *c, int dtype, int mode, int batch, unsigned flags);

// Records the path that won the autotuner for this call shape.
void record_path(const Ctx *c, int dtype, int mode, int batch, unsigned flags, int path);

}  // namespace femhot
