// C-ABI compute entry points: residual/energy, residual VJP, potential gradient, tangent -> CSR.
// Each validates its arguments on the host, clears its outputs with a fill kernel (graph-safe,
// no memset nodes), picks the scatter strategy and launches on the caller's stream.
#include <algorithm>

#include "femhot_chunk.h"
#include "femhot_launch.h"

namespace femhot {

static int launch_elem_any(int et, const ElemLaunch &L) {
  switch (et) {
    case QUAD4: return launch_elem<QUAD4>(L);
    case HEX8: return launch_elem<HEX8>(L);
    default: return launch_elem<TET4>(L);
  }
}

static int launch_tangent_any(int et, const TangentLaunch &L) {
  switch (et) {
    case QUAD4: return launch_tangent<QUAD4>(L);
    case HEX8: return launch_tangent<HEX8>(L);
    default: return launch_tangent<TET4>(L);
  }
}

struct CallIO {
  int dtype = 0, batch = 1, mode = 0;
  const void *field = nullptr;
  long long fs = 0;
  const void *u = nullptr;
  long long us = 0;
  const void *w = nullptr;
  long long ws = 0;
  void *ou = nullptr;
  long long ous = 0;
  void *od = nullptr;
  long long ods = 0;
  double *energy = nullptr;
  unsigned flags = 0;
  hipStream_t stream = nullptr;
};

static int validate(Ctx *c, const CallIO &io, const char *fn) {
  const bool elemental = (io.flags & FEMHOT_ELEMENT_FIELD) != 0;
  const long long fsize = elemental ? c->Ne : c->Nn;
  std::string f(fn);
  FH_CHECK_ARG(io.dtype == FEMHOT_F32 || io.dtype == FEMHOT_F64, FEMHOT_ERR_ARG, f + ": dtype must be F32 or F64");
  FH_CHECK_ARG(io.batch >= 1 && io.batch <= 65535, FEMHOT_ERR_SHAPE, f + ": batch must be in [1, 65535]");
  FH_CHECK_ARG(io.field && io.u, FEMHOT_ERR_ARG, f + ": null field/u");
  if (io.batch > 1) {
    FH_CHECK_ARG(io.fs >= fsize && io.us >= c->Ndof, FEMHOT_ERR_SHAPE, f + ": input strides smaller than rows");
    FH_CHECK_ARG(!io.ou || io.ous >= c->Ndof, FEMHOT_ERR_SHAPE, f + ": output stride smaller than n_dofs");
    FH_CHECK_ARG(!io.od || io.ods >= fsize, FEMHOT_ERR_SHAPE, f + ": g_field stride smaller than field size");
    FH_CHECK_ARG(!io.w || io.ws >= c->Ndof, FEMHOT_ERR_SHAPE, f + ": cotangent stride smaller than n_dofs");
  }
  return FEMHOT_OK;
}

static int run_elem(Ctx *c, const CallIO &io_in) {
  CallIO io = io_in;
  const bool elemental = (io.flags & FEMHOT_ELEMENT_FIELD) != 0;
  const long long fsize = elemental ? c->Ne : c->Nn;
  const bool acc = (io.flags & FEMHOT_ACCUMULATE) != 0;
  int rc;
  // FEMHOT_SCATTER_CHUNK leaves the scatter to the library: the path tuned for this call shape
  bool chunk = false;
  if (io.flags & FEMHOT_SCATTER_CHUNK) {
    const int path = tuned_path(c, io.dtype, io.mode, io.batch, io.flags);
    chunk = path == PATH_CHUNK;
    if (path == PATH_COLOUR) io.flags |= FEMHOT_SCATTER_COLOUR;
  }
  // clear outputs (the chunk path writes every node itself and clears only shared nodes)
  if (!acc && !chunk) {
    if (io.dtype == FEMHOT_F32) {
      if ((rc = fill<float>((float *)io.ou, c->Ndof, io.batch > 1 ? io.ous : c->Ndof, io.batch, 0.f, io.stream)))
        return rc;
      if (!elemental || io.mode == MODE_RES)
        if ((rc = fill<float>((float *)io.od, fsize, io.batch > 1 ? io.ods : fsize, io.batch, 0.f, io.stream)))
          return rc;
    } else {
      if ((rc = fill<double>((double *)io.ou, c->Ndof, io.batch > 1 ? io.ous : c->Ndof, io.batch, 0.0, io.stream)))
        return rc;
      if (!elemental || io.mode == MODE_RES)
        if ((rc = fill<double>((double *)io.od, fsize, io.batch > 1 ? io.ods : fsize, io.batch, 0.0, io.stream)))
          return rc;
    }
  }
  if (!acc && !chunk && (rc = fill<double>(io.energy, (long long)io.batch * c->nseg, (long long)io.batch * c->nseg, 1, 0.0,
                                 io.stream)))
    return rc;
  if (chunk) {
    return run_chunk(c, io.dtype, io.mode, io.batch, io.field, io.fs, io.u, io.us, io.w, io.ws, io.ou, io.ous, io.od,
                     io.ods, io.energy, io.flags, io.stream, 0, 0);
  }
  ElemLaunch L{};
  L.dtype = io.dtype;
  L.ph = c->ph;
  L.ngp = c->ngp;
  L.mode = io.mode;
  L.stream = io.stream;
  KArgs &A = L.args;
  A.coords = c->d_coords64;
  A.conn = c->d_conn;
  A.Ne = c->Ne;
  A.Nn = c->Nn;
  A.field = io.field;
  A.fstride = io.batch > 1 ? io.fs : 0;
  A.u = io.u;
  A.ustride = io.batch > 1 ? io.us : 0;
  A.w = io.w;
  A.wstride = io.batch > 1 ? io.ws : 0;
  A.out_u = io.ou;
  A.oustride = io.batch > 1 ? io.ous : 0;
  A.out_d = io.od;
  A.odstride = io.batch > 1 ? io.ods : 0;
  A.energy = io.energy;
  A.dmask = c->d_dmask;
  A.seg_off = c->d_seg_off;
  A.nseg = c->nseg;
  A.flags = io.flags;
  A.pp = c->pp;
  if ((rc = timing_mark(c, io.stream))) return rc;
  if (io.flags & FEMHOT_SCATTER_COLOUR) {
    if ((rc = ensure_colouring(c))) return rc;
    L.plain = true;
    for (int k = 0; k < c->n_colours; ++k) {
      const long long cnt = c->h_colour_off[k + 1] - c->h_colour_off[k];
      if (!cnt) continue;
      A.list = c->d_colour_elems + c->h_colour_off[k];
      A.count = cnt;
      L.grid = dim3((unsigned)((cnt + 255) / 256), io.batch);
      if ((rc = launch_elem_any(c->et, L))) return rc;
    }
    return timing_mark(c, io.stream);
  }
  L.plain = false;
  A.list = nullptr;
  A.count = c->Ne;
  L.grid = dim3((unsigned)((c->Ne + 255) / 256), io.batch);
  if ((rc = launch_elem_any(c->et, L))) return rc;
  return timing_mark(c, io.stream);
}

}  // namespace femhot

using namespace femhot;

extern "C" {

int femhot_residual(femhot_ctx *h, int dtype, int batch, const void *field, long long fs, const void *u,
                    long long us, void *R, long long Rs, double *energy, unsigned flags, void *stream) {
  FH_CHECK_ARG(h, FEMHOT_ERR_ARG, "null context");
  CallIO io;
  io.dtype = dtype;
  io.batch = batch;
  io.mode = MODE_RES;
  io.field = field;
  io.fs = fs;
  io.u = u;
  io.us = us;
  io.ou = R;
  io.ous = Rs;
  io.energy = energy;
  io.flags = flags;
  io.stream = (hipStream_t)stream;
  int rc = validate(&h->c, io, "femhot_residual");
  if (rc) return rc;
  return run_elem(&h->c, io);
}

int femhot_residual_vjp(femhot_ctx *h, int dtype, int batch, const void *field, long long fs, const void *u,
                        long long us, const void *w, long long ws, void *g_u, long long gus, void *g_field,
                        long long gfs, unsigned flags, void *stream) {
  FH_CHECK_ARG(h, FEMHOT_ERR_ARG, "null context");
  FH_CHECK_ARG(w, FEMHOT_ERR_ARG, "femhot_residual_vjp: null cotangent");
  CallIO io;
  io.dtype = dtype;
  io.batch = batch;
  io.mode = MODE_VJP;
  io.field = field;
  io.fs = fs;
  io.u = u;
  io.us = us;
  io.w = w;
  io.ws = ws;
  io.ou = g_u;
  io.ous = gus;
  io.od = g_field;
  io.ods = gfs;
  io.flags = flags;
  io.stream = (hipStream_t)stream;
  int rc = validate(&h->c, io, "femhot_residual_vjp");
  if (rc) return rc;
  return run_elem(&h->c, io);
}

int femhot_energy_grad(femhot_ctx *h, int dtype, int batch, const void *field, long long fs, const void *u,
                       long long us, double *energy, void *g_u, long long gus, void *g_field, long long gfs,
                       unsigned flags, void *stream) {
  FH_CHECK_ARG(h, FEMHOT_ERR_ARG, "null context");
  Ctx *c = &h->c;
  FH_CHECK_ARG(!(c->ph == THERMAL && c->pp.beta != 0.0), FEMHOT_ERR_UNSUPPORTED,
               "femhot_energy_grad: thermal conduction with beta != 0 has no potential");
  CallIO io;
  io.dtype = dtype;
  io.batch = batch;
  io.mode = MODE_EGRAD;
  io.field = field;
  io.fs = fs;
  io.u = u;
  io.us = us;
  io.ou = g_u;
  io.ous = gus;
  io.od = g_field;
  io.ods = gfs;
  io.energy = energy;
  io.flags = flags;
  io.stream = (hipStream_t)stream;
  int rc = validate(c, io, "femhot_energy_grad");
  if (rc) return rc;
  return run_elem(c, io);
}

int femhot_energy_grad_noalloc(femhot_ctx *h, int dtype, int batch, const void *field, long long fs, const void *u,
                               long long us, double *energy, void *g_u, long long gus, void *g_field, long long gfs,
                               unsigned flags, void *stream) {
  FH_CHECK_ARG(h, FEMHOT_ERR_ARG, "null context");
  Ctx *c = &h->c;
  FH_CHECK_ARG(!(c->ph == THERMAL && c->pp.beta != 0.0), FEMHOT_ERR_UNSUPPORTED,
               "femhot_energy_grad_noalloc: thermal conduction with beta != 0 has no potential");
  CallIO io;
  io.dtype = dtype;
  io.batch = batch;
  io.mode = MODE_EGRAD;
  io.field = field;
  io.fs = fs;
  io.u = u;
  io.us = us;
  io.ou = g_u;
  io.ous = gus;
  io.od = g_field;
  io.ods = gfs;
  io.energy = energy;
  io.flags = flags;
  int rc = validate(c, io, "femhot_energy_grad_noalloc");
  if (rc) return rc;
  return run_chunk_fixed_scratch(c, dtype, MODE_EGRAD, batch, field, batch > 1 ? fs : 0, u, batch > 1 ? us : 0, g_u,
                                 gus, g_field, gfs, energy, flags & ~(FEMHOT_SCATTER_COLOUR | FEMHOT_SCATTER_CHUNK),
                                 (hipStream_t)stream);
}

int femhot_tangent_csr(femhot_ctx *h, int dtype, const void *field, const void *u, void *vals, void *R,
                       unsigned flags, void *stream) {
  FH_CHECK_ARG(h && field && u && vals, FEMHOT_ERR_ARG, "femhot_tangent_csr: null argument");
  FH_CHECK_ARG(dtype == FEMHOT_F32 || dtype == FEMHOT_F64, FEMHOT_ERR_ARG, "femhot_tangent_csr: bad dtype");
  Ctx *c = &h->c;
  int rc = ensure_pattern(c);
  if (rc) return rc;
  hipStream_t s = (hipStream_t)stream;
  const bool acc = (flags & FEMHOT_ACCUMULATE) != 0;
  // FEMHOT_SCATTER_CHUNK leaves the assembly order to the library: the path femhot_autotune
  // measured fastest for the tangent (untuned: row-owner, colour-ordered for Neo-Hookean)
  if ((flags & FEMHOT_SCATTER_CHUNK) && !(flags & (FEMHOT_SCATTER_COLOUR | FEMHOT_SCATTER_ROWS))) {
    const int path = tuned_path(c, dtype, MODE_TANGENT, 1, flags);
    if (path == PATH_COLOUR) flags |= FEMHOT_SCATTER_COLOUR;
    if (path == PATH_ROWS) flags |= FEMHOT_SCATTER_ROWS;
  }
  // row-owner assembly covers thermal and linear elasticity (and grids of up to 2^26 nodes of 64
  // lanes); otherwise the call takes colour-ordered adds
  const bool rows = (flags & FEMHOT_SCATTER_ROWS) && c->ph != NEOHOOKE && c->Nn <= (1LL << 26);
  if ((flags & FEMHOT_SCATTER_ROWS) && !rows) flags |= FEMHOT_SCATTER_COLOUR;
  if (!acc && !rows) {
    rc = dtype == FEMHOT_F32 ? fill<float>((float *)vals, c->nnz, c->nnz, 1, 0.f, s)
                             : fill<double>((double *)vals, c->nnz, c->nnz, 1, 0.0, s);
    if (rc) return rc;
  }
  TangentLaunch L{};
  L.dtype = dtype;
  L.ph = c->ph;
  L.ngp = c->ngp;
  L.stream = s;
  TArgs &A = L.args;
  A.coords = c->d_coords64;
  A.conn = c->d_conn;
  A.field = field;
  A.u = u;
  A.vals = vals;
  A.row_ptr = c->d_row_ptr;
  A.slot = c->d_slot;
  A.dmask = c->d_dmask;
  A.ne_off = c->d_ne_off;
  A.ne_list = c->d_ne_list;
  A.n_nodes = c->Nn;
  A.flags = flags;
  A.pp = c->pp;
  const int n = c->n;
  L.rows = false;
  if (rows) {  // one 64-lane workgroup per node
    L.rows = true;
    L.plain = true;
    A.list = nullptr;
    A.count = c->Ne;
    L.grid = dim3((unsigned)c->Nn);
    if (c->Nn && (rc = launch_tangent_any(c->et, L))) return rc;
  } else if ((flags & FEMHOT_SCATTER_COLOUR) && !c->repeated_nodes) {
    // colour-ordered plain adds need distinct nodes within an element (lanes of one element would
    // otherwise update the same entries); meshes with collapsed elements assemble with atomics
    if ((rc = ensure_colouring(c))) return rc;
    L.plain = true;
    for (int k = 0; k < c->n_colours; ++k) {
      const long long cnt = c->h_colour_off[k + 1] - c->h_colour_off[k];
      if (!cnt) continue;
      A.list = c->d_colour_elems + c->h_colour_off[k];
      A.count = cnt;
      L.grid = dim3((unsigned)((cnt * n * c->ndpn + 255) / 256));
      if ((rc = launch_tangent_any(c->et, L))) return rc;
    }
  } else {
    L.plain = false;
    A.list = nullptr;
    A.count = c->Ne;
    L.grid = dim3((unsigned)((c->Ne * n * c->ndpn + 255) / 256));
    if ((rc = launch_tangent_any(c->et, L))) return rc;
  }
  if (R) {
    CallIO io;
    io.dtype = dtype;
    io.batch = 1;
    io.mode = MODE_RES;
    io.field = field;
    io.u = u;
    io.ou = R;
    // the residual takes the scatter the call asks for (FEMHOT_SCATTER_CHUNK: its tuned path)
    io.flags = (flags & (FEMHOT_ELEMENT_FIELD | FEMHOT_ACCUMULATE | FEMHOT_SCATTER_COLOUR | FEMHOT_SCATTER_CHUNK)) |
               FEMHOT_MASK_DIRICHLET;
    io.stream = s;
    if ((rc = run_elem(c, io))) return rc;
  }
  return FEMHOT_OK;
}

int femhot_autotune(femhot_ctx *h, int op, int dtype, int batch, const void *field, long long fs, const void *u,
                    long long us, const void *w, long long ws, void *out_u, long long ous, void *out_field,
                    long long ofs, double *energy, unsigned flags, int *epc, int *block, int *grid_per_cu,
                    float *ms, void *stream) {
  FH_CHECK_ARG(h, FEMHOT_ERR_ARG, "null context");
  FH_CHECK_ARG(op >= FEMHOT_OP_RESIDUAL && op <= FEMHOT_OP_TANGENT, FEMHOT_ERR_ARG, "femhot_autotune: bad op");
  FH_CHECK_ARG(op != FEMHOT_OP_VJP || w, FEMHOT_ERR_ARG, "femhot_autotune: VJP needs a cotangent");
  Ctx *c = &h->c;
  hipEvent_t e0, e1;
  // mean time of `reps` calls of f after one warm-up call, with HIP events on the stream
  auto time_it = [&](auto &&f, float *ms_out) -> int {
    int r = f();
    if (r) return r;
    FH_HIP(hipEventRecord(e0, (hipStream_t)stream));
    constexpr int reps = 3;
    for (int k = 0; k < reps && !r; ++k) r = f();
    FH_HIP(hipEventRecord(e1, (hipStream_t)stream));
    FH_HIP(hipEventSynchronize(e1));
    if (r) return r;
    float t = 0.f;
    FH_HIP(hipEventElapsedTime(&t, e0, e1));
    *ms_out = t / reps;
    return FEMHOT_OK;
  };
  if (op == FEMHOT_OP_TANGENT) {  // atomic, colour-ordered and row-owner assembly; out_u holds the values
    FH_CHECK_ARG(field && u && out_u, FEMHOT_ERR_ARG, "femhot_autotune: tangent needs field, u and values");
    const unsigned tf = flags & (FEMHOT_ELEMENT_FIELD | FEMHOT_TRANSPOSE);
    FH_HIP(hipEventCreate(&e0));
    FH_HIP(hipEventCreate(&e1));
    const bool rows_ok = c->ph != NEOHOOKE && c->Nn <= (1LL << 26);
    float t_atomic = 0.f, t_colour = 0.f, t_rows = 0.f;
    int rc = time_it([&] { return femhot_tangent_csr(h, dtype, field, u, out_u, nullptr, tf, stream); }, &t_atomic);
    if (!rc)
      rc = time_it([&] { return femhot_tangent_csr(h, dtype, field, u, out_u, nullptr, tf | FEMHOT_SCATTER_COLOUR,
                                                   stream); }, &t_colour);
    if (!rc && rows_ok)
      rc = time_it([&] { return femhot_tangent_csr(h, dtype, field, u, out_u, nullptr, tf | FEMHOT_SCATTER_ROWS,
                                                   stream); }, &t_rows);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    int path = t_colour <= t_atomic ? PATH_COLOUR : PATH_ATOMIC;
    float best = std::min(t_colour, t_atomic);
    if (rows_ok && t_rows <= best) path = PATH_ROWS, best = t_rows;
    record_path(c, dtype, MODE_TANGENT, 1, flags, path);
    if (epc) *epc = 0;
    if (block)
      *block = path == PATH_ROWS ? FEMHOT_PATH_ROWS : path == PATH_COLOUR ? FEMHOT_PATH_COLOUR : FEMHOT_PATH_ATOMIC;
    if (grid_per_cu) *grid_per_cu = 0;
    if (ms) *ms = best;
    return FEMHOT_OK;
  }
  CallIO io;
  io.dtype = dtype;
  io.batch = batch;
  io.mode = op;  // FEMHOT_OP_* equal MODE_*
  io.field = field;
  io.fs = fs;
  io.u = u;
  io.us = us;
  io.w = w;
  io.ws = ws;
  io.ou = out_u;
  io.ous = ous;
  io.od = out_field;
  io.ods = ofs;
  io.energy = energy;
  io.flags = (flags & ~FEMHOT_ACCUMULATE) | FEMHOT_SCATTER_CHUNK;
  io.stream = (hipStream_t)stream;
  int rc = validate(c, io, "femhot_autotune");
  if (rc) return rc;
  float chunk_ms = 0.f;
  rc = autotune_chunk(c, dtype, op, batch, field, io.batch > 1 ? fs : 0, u, io.batch > 1 ? us : 0, w,
                      io.batch > 1 ? ws : 0, out_u, ous, out_field, ofs, energy, io.flags, io.stream, epc, block,
                      grid_per_cu, &chunk_ms);
  if (rc) return rc;
  // the element-parallel paths compete too (hardware atomics, and colour-ordered plain adds); when
  // one wins, FEMHOT_SCATTER_CHUNK calls of this shape take it
  CallIO pio = io;
  pio.flags &= ~FEMHOT_SCATTER_CHUNK;
  CallIO cio = pio;
  cio.flags |= FEMHOT_SCATTER_COLOUR;
  FH_HIP(hipEventCreate(&e0));
  FH_HIP(hipEventCreate(&e1));
  float atomic_ms = 0.f, colour_ms = 0.f;
  rc = time_it([&] { return run_elem(c, pio); }, &atomic_ms);
  if (!rc) rc = time_it([&] { return run_elem(c, cio); }, &colour_ms);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc) return rc;
  int path = PATH_CHUNK;
  float best = chunk_ms;
  if (atomic_ms < best) path = PATH_ATOMIC, best = atomic_ms;
  if (colour_ms < best) path = PATH_COLOUR, best = colour_ms;
  record_path(c, dtype, op, batch, io.flags, path);
  if (path != PATH_CHUNK) {
    if (epc) *epc = 0;
    if (block) *block = path == PATH_COLOUR ? FEMHOT_PATH_COLOUR : FEMHOT_PATH_ATOMIC;
    if (grid_per_cu) *grid_per_cu = 0;
  }
  if (ms) *ms = best;
  return run_elem(c, io);  // leave valid outputs
}

int femhot_set_chunk_config(femhot_ctx *h, int op, int dtype, int batch, unsigned flags, int epc, int block,
                            int grid_per_cu) {
  FH_CHECK_ARG(h, FEMHOT_ERR_ARG, "null context");
  FH_CHECK_ARG(op >= FEMHOT_OP_RESIDUAL && op <= FEMHOT_OP_TANGENT, FEMHOT_ERR_ARG, "bad op");
  FH_CHECK_ARG(dtype == FEMHOT_F32 || dtype == FEMHOT_F64, FEMHOT_ERR_ARG, "bad dtype");
  return set_chunk_config(&h->c, op, dtype, batch, flags, epc, block, grid_per_cu);
}

}  // extern "C"
