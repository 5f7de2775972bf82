// LDS-staged chunk assembly: the batched / persistent path of the residual, residual-VJP and
// potential-gradient kernels.
//
// One workgroup owns one chunk of spatially compact elements at a time (grid-stride loop over
// chunk x sample-block work items, so the launch can be persistent).  Per work item:
//   1. stage the chunk's node coordinates (precomputed chunk-relative, contiguous -> coalesced)
//      and the chunk's nodal solution / cotangent / field into LDS;
//   2. one lane per element: element routine reading LDS; each corner's results go to an LDS
//      buffer at the corner's position in the chunk's node-sorted incidence list (the plan's
//      per-corner rank; no atomics), per-element results straight to HBM;
//   3. one lane per chunk node sums its contiguous run of that buffer in a fixed order and writes
//      the sum once with a plain store: to the output when the node belongs to this chunk only, to
//      the node's partial slot when it is shared with other chunks.
// k_reduce then sums each shared node's partials in chunk order and the per-chunk energies per
// segment, so the path uses no atomics (device-scope atomics resolve beyond the per-XCD L2) and
// its results are bitwise reproducible run to run.
#include <algorithm>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "femhot_chunk.h"
#include "femhot_chunk_kernels.h"
#include "femhot_chunk_pipe.h"

namespace femhot {

// Both deterministic reductions in one launch.  Blocks [0, nb_nodes) sum, for each shared node,
// the partials of every chunk containing it, in chunk order; the last nseg * batch blocks each sum
// the per-chunk energies of one (sample, segment) in a fixed order.  Shared node k's partial from
// the o-th chunk containing it is slot k (o = 0), n_shared + k (o = 1) or, for nodes in 3+ chunks,
// 2 n_shared + red_off[k] - 2 k + o - 2: every shared node has the first two, and their addresses
// need no table, so they load together with the node's ids (consecutive threads, consecutive memory).
template <class R> struct ReduceArgs {
  long long n_shared, n_slots;
  int P, batch, nb_nodes, nseg, n_chunks;
  const int *sh_nodes, *red_off, *seg_chunk_off;
  const R *part_u, *part_d;
  R *ou, *od;
  long long ous, ods;
  const uint8_t *dmask;
  const double *chunk_energy;
  double *energy;
  unsigned flags;
};

constexpr int REDUCE_NPT = 2;  // shared nodes per thread of k_reduce

template <class R, int P>
__global__ __launch_bounds__(256) void k_reduce(ReduceArgs<R> A) {
  const bool accumulate = (A.flags & FEMHOT_ACCUMULATE) != 0;
  if ((int)blockIdx.x >= A.nb_nodes) {
    __shared__ double red[4];
    const int q = (int)blockIdx.x - A.nb_nodes;
    const int b = q / A.nseg, sg = q - b * A.nseg;
    double v = 0.0;
    for (int c = A.seg_chunk_off[sg] + (int)threadIdx.x; c < A.seg_chunk_off[sg + 1]; c += 256)
      v += A.chunk_energy[(long long)b * A.n_chunks + c];
    v = wave_sum(v);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
      const double t = (red[0] + red[1]) + (red[2] + red[3]);
      double *o = A.energy + (long long)b * A.nseg + sg;
      *o = accumulate ? *o + t : t;
    }
    return;
  }
  // REDUCE_NPT nodes per thread (node i and i + 256 of the block's range): the index and partial loads of
  // both are issued before either sum, and the launch needs half as many threads
  const long long tot = A.n_shared * A.batch;
  const bool mask = (A.flags & FEMHOT_MASK_DIRICHLET) != 0;
  bool ok[REDUCE_NPT];
  long long bu[REDUCE_NPT];
  int gu[REDUCE_NPT], nu[REDUCE_NPT];
  long long ku[REDUCE_NPT], xu[REDUCE_NPT];
#pragma unroll
  for (int u = 0; u < REDUCE_NPT; ++u) {
    const long long i = (long long)blockIdx.x * (256 * REDUCE_NPT) + u * 256 + threadIdx.x;
    ok[u] = i < tot;
    const long long ic = ok[u] ? i : tot - 1;  // clamped: the loads stay in range, the stores are skipped
    bu[u] = ic / A.n_shared;
    ku[u] = ic - bu[u] * A.n_shared;
    gu[u] = A.sh_nodes[ku[u]];
    const int r0 = A.red_off[ku[u]];
    nu[u] = A.red_off[ku[u] + 1] - r0;  // a shared node has a slot in >= 2 chunks
    xu[u] = 2 * A.n_shared + r0 - 2 * ku[u] - 2;  // + o: its o-th slot, o >= 2
  }
  // Every load a node needs is issued before the first add: its first two slots (at fixed places),
  // the Dirichlet mask and, when accumulating, the old outputs.  Further slots (nodes in 3+ chunks)
  // follow in a loop.  The sum runs over the slots in chunk order.
  if (A.ou) {
    R v0[REDUCE_NPT][P], v1[REDUCE_NPT][P], old[REDUCE_NPT][P];
    bool masked[REDUCE_NPT][P];
#pragma unroll
    for (int u = 0; u < REDUCE_NPT; ++u) {
      const R *pb = A.part_u + bu[u] * A.n_slots * P;
      const R *o = A.ou + bu[u] * A.ous + (long long)gu[u] * P;
#pragma unroll
      for (int c = 0; c < P; ++c) {
        v0[u][c] = pb[ku[u] * P + c];
        v1[u][c] = pb[(A.n_shared + ku[u]) * P + c];
        masked[u][c] = mask && A.dmask[(long long)gu[u] * P + c];
        old[u][c] = accumulate ? o[c] : R(0);
      }
    }
#pragma unroll
    for (int u = 0; u < REDUCE_NPT; ++u) {
      const R *pb = A.part_u + bu[u] * A.n_slots * P;
      R *o = A.ou + bu[u] * A.ous + (long long)gu[u] * P;
      R s[P];
#pragma unroll
      for (int c = 0; c < P; ++c) s[c] = (R(-0.0) + v0[u][c]) + v1[u][c];
      for (int j = 2; j < nu[u]; ++j)
#pragma unroll
        for (int c = 0; c < P; ++c) s[c] += pb[(xu[u] + j) * P + c];
      if (ok[u]) {
#pragma unroll
        for (int c = 0; c < P; ++c)
          if (!(masked[u][c] && accumulate)) o[c] = masked[u][c] ? R(0) : (accumulate ? old[u][c] + s[c] : s[c]);
      }
    }
  }
  if (A.od) {
    R v0[REDUCE_NPT], v1[REDUCE_NPT], old[REDUCE_NPT];
#pragma unroll
    for (int u = 0; u < REDUCE_NPT; ++u) {
      const R *pd = A.part_d + bu[u] * A.n_slots;
      v0[u] = pd[ku[u]];
      v1[u] = pd[A.n_shared + ku[u]];
      old[u] = accumulate ? A.od[bu[u] * A.ods + gu[u]] : R(0);
    }
#pragma unroll
    for (int u = 0; u < REDUCE_NPT; ++u) {
      const R *pd = A.part_d + bu[u] * A.n_slots;
      R s = (R(-0.0) + v0[u]) + v1[u];
      for (int j = 2; j < nu[u]; ++j) s += pd[xu[u] + j];
      if (ok[u]) A.od[bu[u] * A.ods + gu[u]] = accumulate ? old[u] + s : s;
    }
  }
}

// Clears the outputs of nodes in no element (the chunk path writes every other node).
template <class R>
__global__ void k_zero_nodes(const int *nodes, long long n, int P, int batch, R *ou, long long ous, R *od,
                             long long ods) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n * batch) return;
  const long long b = i / n;
  const int g = nodes[i - b * n];
  if (ou)
    for (int c = 0; c < P; ++c) ou[b * ous + (long long)g * P + c] = R(0);
  if (od) od[b * ods + g] = R(0);
}

// grid_per_cu 0: one workgroup per work item; k > 0: k workgroups per CU; k < 0: -k workgroups in
// all (persistent).  sb: samples per work item (a work item stages its
// chunk's plan once and loops over sb samples); 0 picks enough items for about 8 per CU.
struct ChunkConfig {
  int epc = 1024, block = 256, grid_per_cu = 0, sb = 0;
  int path = PATH_CHUNK;  // PATH_ATOMIC / PATH_COLOUR: an element-parallel path won the autotuner
  bool dma = false;       // FP32 only: the pipelined LDS-DMA kernel (k_chunk_pipe) instead of k_chunk
};

static std::mutex g_tune_mu;
static std::map<std::tuple<const void *, int, int, int, int>, ChunkConfig> g_tuned;  // ctx,mode,dtype,batch,flags

static int cu_count() {
  static int cus = 0;
  if (!cus) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
      cus = 256;
  }
  return cus;
}

// Largest dynamic LDS a workgroup may request on this device (less the kernel's static LDS).
static size_t lds_limit() {
  static size_t lim = 0;
  if (!lim) {
    int dev = 0, v = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess || v <= 0)
      v = 64 * 1024;
    lim = (size_t)v - 256;
  }
  return lim;
}

// fixed_ws: caller-provided scratch of fixed_bytes (the allocation-free entry point); then the plan
// must already be built, nothing is allocated and no timing events are recorded.
static int run_chunk_cfg(Ctx *c, int dtype, int mode, int batch, const void *field, long long fs, const void *u,
                         long long us, const void *w, long long ws, void *ou, long long ous, void *od, long long ods,
                         double *energy, unsigned flags, hipStream_t s, const ChunkConfig &cfg,
                         void *fixed_ws = nullptr, size_t fixed_bytes = 0) {
  const ChunkPlan *plan = nullptr;
  int rc;
  if (fixed_ws) {
    plan = find_chunks(c, cfg.epc);
    if (!plan) {
      set_error("femhot_energy_grad_noalloc: the chunk plan of " + std::to_string(cfg.epc) +
                " elements per chunk is not built; build it first with femhot_set_chunk_config or a call of "
                "femhot_energy_grad with FEMHOT_SCATTER_CHUNK");
      return FEMHOT_ERR_ARG;
    }
  } else if ((rc = ensure_chunks(c, cfg.epc, s, &plan))) {
    return rc;
  }
  const ChunkPlan &p = *plan;
  const bool elemental = (flags & FEMHOT_ELEMENT_FIELD) != 0;
  const int P = c->ndpn;
  const size_t rs = dtype == FEMHOT_F32 ? sizeof(float) : sizeof(double);
  const bool nodal_gd = od && !elemental && mode != MODE_RES;
  const size_t cap4 = ((size_t)p.max_nodes + 3) & ~(size_t)3;
  if (cfg.dma && dtype != FEMHOT_F32) {
    set_error("run_chunk: the LDS-DMA chunk kernel is FP32 only");
    return FEMHOT_ERR_UNSUPPORTED;
  }
  // The staged TET4 kernel's corner buffer takes one spare row per chunk node (a node gather without
  // LDS bank conflicts) unless those rows would cost a workgroup per CU: compared in 512-byte LDS
  // allocation units, against the workgroups per CU the configuration asks for (any number for one
  // workgroup per item).
  const size_t shm_base = cap4 * (3 + P + (mode == MODE_VJP ? P : 0) + (elemental ? 0 : 1)) * rs +
                          (size_t)p.max_elems * c->n * (P + (nodal_gd ? 1 : 0)) * rs + 16;
  const size_t shm_rows = cap4 * (P + (nodal_gd ? 1 : 0)) * rs;
  bool pad = false;
  if (!cfg.dma && c->et == TET4) {
    const size_t lds_cu = lds_limit() + 256;
    auto per_cu = [&](size_t b) {
      const size_t alloc = (b + (size_t)cfg.block / 64 * sizeof(double) + 511) / 512 * 512;
      return (long long)(lds_cu / alloc);
    };
    const long long want = cfg.grid_per_cu > 0 ? cfg.grid_per_cu : (1LL << 30);
    pad = shm_base + shm_rows <= lds_limit() &&
          per_cu(shm_base + shm_rows) >= std::min(per_cu(shm_base), want);
  }
  const size_t shm =
      cfg.dma ? (size_t)pipe_layout(p.max_nodes, p.max_elems, c->n, P, mode, elemental, ou != nullptr, nodal_gd,
                                    p.perm_identity)
                    .total
              : shm_base + (pad ? shm_rows : 0);
  if (shm > lds_limit()) {
    set_error("run_chunk: chunk too large for LDS; use fewer elements per chunk");
    return FEMHOT_ERR_UNSUPPORTED;
  }
  const bool acc = (flags & FEMHOT_ACCUMULATE) != 0;
  const bool want_gd_nodal = od && !elemental && mode != MODE_RES;
  // scratch: shared-node partials (batch x n_slots x P, + batch x n_slots for the nodal field
  // gradient) and per-chunk energies (batch x n_chunks); grown here, never inside a capture
  const size_t pu_bytes = ou ? ((size_t)batch * p.n_slots * P * rs + 255) / 256 * 256 : 0;
  const size_t pd_bytes = want_gd_nodal ? ((size_t)batch * p.n_slots * rs + 255) / 256 * 256 : 0;
  const size_t ce_bytes = energy ? ((size_t)batch * p.n_chunks * sizeof(double) + 255) / 256 * 256 : 0;
  const size_t need = pu_bytes + pd_bytes + ce_bytes + 256;
  if (fixed_ws && need > fixed_bytes) {
    set_error("femhot_energy_grad_noalloc: this call needs " + std::to_string(need) + " bytes of scratch (shared-node " +
              "partials and per-chunk energies), more than the library's fixed " + std::to_string(fixed_bytes) +
              " bytes; use femhot_energy_grad, or fewer samples per call");
    return FEMHOT_ERR_SHAPE;
  }
  if (!fixed_ws && c->chunk_ws.bytes < need) {
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    FH_HIP(hipStreamIsCapturing(s, &cs));
    if (cs != hipStreamCaptureStatusNone) {
      set_error("run_chunk: scratch must be sized by an uncaptured call with the same batch before capture");
      return FEMHOT_ERR_ARG;
    }
    FH_HIP(hipStreamSynchronize(s));
    if (c->chunk_ws.ptr) FH_HIP(hipFree(c->chunk_ws.ptr));
    c->chunk_ws.ptr = nullptr;
    c->chunk_ws.bytes = 0;
    FH_HIP(hipMalloc(&c->chunk_ws.ptr, need));
    c->chunk_ws.bytes = need;
  }
  char *wsp = fixed_ws ? (char *)fixed_ws : (char *)c->chunk_ws.ptr;
  void *part_u = pu_bytes ? (void *)wsp : nullptr;
  void *part_d = pd_bytes ? (void *)(wsp + pu_bytes) : nullptr;
  double *chunk_energy = ce_bytes ? (double *)(wsp + pu_bytes + pd_bytes) : nullptr;
  // clear nodes in no element
  if (!acc && p.n_zero && (ou || want_gd_nodal)) {
    const long long tot = p.n_zero * batch;
    const dim3 g((unsigned)((tot + 255) / 256));
    if (dtype == FEMHOT_F32)
      hipLaunchKernelGGL(k_zero_nodes<float>, g, dim3(256), 0, s, p.zero_nodes, p.n_zero, P, batch, (float *)ou,
                         batch > 1 ? ous : c->Ndof, want_gd_nodal ? (float *)od : nullptr, batch > 1 ? ods : c->Nn);
    else
      hipLaunchKernelGGL(k_zero_nodes<double>, g, dim3(256), 0, s, p.zero_nodes, p.n_zero, P, batch, (double *)ou,
                         batch > 1 ? ous : c->Ndof, want_gd_nodal ? (double *)od : nullptr, batch > 1 ? ods : c->Nn);
    FH_HIP(hipGetLastError());
  }
  ChunkArgs A{};
  A.elem_off = p.chunk_elem_off;
  A.node_off = p.chunk_node_off;
  A.perm = p.perm;
  A.nodes = p.chunk_nodes;
  A.node_slot = p.node_slot;
  A.lconn = p.lconn;
  A.inc = p.inc;
  A.inc_start = p.inc_start;
  A.rank = p.rank;
  A.xyz = cfg.dma ? (const void *)p.chunk_xyz32
                  : dtype == FEMHOT_F32 ? (const void *)p.chunk_xyz12 : (const void *)p.chunk_xyz64;
  A.n_chunks = p.n_chunks;
  if (cfg.sb > 0) {
    A.sb = std::min(cfg.sb, batch);
  } else {  // largest sb that still leaves ~8 work items per CU
    const long long target = 8LL * cu_count();
    A.sb = (int)std::max<long long>(1, std::min<long long>(batch, (long long)p.n_chunks * batch / target));
  }
  A.n_sblocks = (batch + A.sb - 1) / A.sb;
  A.batch = batch;
  A.cap = p.max_nodes;
  A.emax = p.max_elems;
  A.perm_identity = p.perm_identity ? 1 : 0;
  A.pad = pad ? 1 : 0;
  A.n_slots = p.n_slots;
  A.field = field;
  A.fs = batch > 1 ? fs : 0;
  A.u = u;
  A.us = batch > 1 ? us : 0;
  A.w = w;
  A.ws = batch > 1 ? ws : 0;
  A.ou = ou;
  A.ous = batch > 1 ? ous : c->Ndof;
  A.od = mode == MODE_RES ? nullptr : od;
  A.ods = batch > 1 ? ods : (elemental ? c->Ne : c->Nn);
  A.part_u = part_u;
  A.part_d = part_d;
  A.chunk_energy = chunk_energy;
  A.dmask = c->d_dmask;
  A.flags = flags;
  A.pp = c->pp;
  const long long items = (long long)A.n_chunks * A.n_sblocks;
  long long grid = items;
  if (cfg.grid_per_cu > 0) grid = std::min<long long>(items, (long long)cfg.grid_per_cu * cu_count());
  if (cfg.grid_per_cu < 0) grid = std::min<long long>(items, -(long long)cfg.grid_per_cu);
  if (grid < 1) return FEMHOT_OK;
  if (!fixed_ws && (rc = timing_mark(c, s))) return rc;
  const dim3 gd((unsigned)grid);
  if (cfg.dma) {
    switch (c->et) {
      case QUAD4: rc = launch_chunk_pipe_et<QUAD4>(c->ph, c->ngp, mode, cfg.block, gd, shm, s, A); break;
      case HEX8: rc = launch_chunk_pipe_et<HEX8>(c->ph, c->ngp, mode, cfg.block, gd, shm, s, A); break;
      default: rc = launch_chunk_pipe_et<TET4>(c->ph, c->ngp, mode, cfg.block, gd, shm, s, A); break;
    }
  } else {
    switch (c->et) {
      case QUAD4: rc = launch_chunk_et<QUAD4>(dtype, c->ph, c->ngp, mode, cfg.block, gd, shm, s, A); break;
      case HEX8: rc = launch_chunk_et<HEX8>(dtype, c->ph, c->ngp, mode, cfg.block, gd, shm, s, A); break;
      default: rc = launch_chunk_et<TET4>(dtype, c->ph, c->ngp, mode, cfg.block, gd, shm, s, A); break;
    }
  }
  if (rc) return rc;
  // deterministic reductions of shared nodes and per-chunk energies (one launch)
  {
    const bool nodes_part = p.n_shared && (ou || want_gd_nodal);
    const long long tot = nodes_part ? p.n_shared * batch : 0;
    const int nb_nodes = (int)((tot + 256 * REDUCE_NPT - 1) / (256 * REDUCE_NPT));
    const int nb_energy = energy ? c->nseg * batch : 0;
    if (nb_nodes + nb_energy > 0) {
      const long long os = batch > 1 ? ous : c->Ndof, dstr = batch > 1 ? ods : c->Nn;
      auto launch = [&](auto zero) -> int {
        using R = decltype(zero);
        ReduceArgs<R> ra{};
        ra.n_shared = nodes_part ? p.n_shared : 0;
        ra.n_slots = p.n_slots;
        ra.P = P;
        ra.batch = batch;
        ra.nb_nodes = nb_nodes;
        ra.nseg = c->nseg;
        ra.n_chunks = p.n_chunks;
        ra.sh_nodes = p.sh_nodes;
        ra.red_off = p.red_off;
        ra.seg_chunk_off = p.seg_chunk_off;
        ra.part_u = (const R *)part_u;
        ra.part_d = (const R *)part_d;
        ra.ou = nodes_part ? (R *)ou : nullptr;
        ra.od = nodes_part && want_gd_nodal ? (R *)od : nullptr;
        ra.ous = os;
        ra.ods = dstr;
        ra.dmask = c->d_dmask;
        ra.chunk_energy = chunk_energy;
        ra.energy = energy;
        ra.flags = flags;
        const dim3 rg((unsigned)(nb_nodes + nb_energy));
        if (P == 1)
          hipLaunchKernelGGL((k_reduce<R, 1>), rg, dim3(256), 0, s, ra);
        else if (P == 2)
          hipLaunchKernelGGL((k_reduce<R, 2>), rg, dim3(256), 0, s, ra);
        else
          hipLaunchKernelGGL((k_reduce<R, 3>), rg, dim3(256), 0, s, ra);
        FH_HIP(hipGetLastError());
        return FEMHOT_OK;
      };
      rc = dtype == FEMHOT_F32 ? launch(0.0f) : launch(0.0);
      if (rc) return rc;
    }
  }
  return fixed_ws ? FEMHOT_OK : timing_mark(c, s);  // the timed region: chunk kernel and its reductions
}

// Default chunk size: the largest power of two <= 1024 whose per-corner LDS buffer stays within
// 48 KB, so several workgroups share a CU.
static ChunkConfig default_config(Ctx *c, int dtype) {
  ChunkConfig cfg;
  const size_t per_elem = (size_t)c->n * ((c->ndpn + 1) * (dtype == FEMHOT_F64 ? 8 : 4) + 2);
  cfg.epc = 1024;
  while (cfg.epc > 64 && cfg.epc * per_elem > 48 * 1024) cfg.epc /= 2;
  cfg.block = 256;
  return cfg;
}
int run_chunk(Ctx *c, int dtype, int mode, int batch, const void *field, long long fs, const void *u, long long us,
              const void *w, long long ws, void *ou, long long ous, void *od, long long ods, double *energy,
              unsigned flags, hipStream_t stream, int epc, int block) {
  ChunkConfig cfg = default_config(c, dtype);
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_tuned.find({c, mode, dtype, batch, (int)(flags & FEMHOT_ELEMENT_FIELD)});
    if (it != g_tuned.end()) cfg = it->second;
  }
  if (block > 0) cfg.block = block;
  if (epc > 0) {
    cfg.epc = epc;
    return run_chunk_cfg(c, dtype, mode, batch, field, fs, u, us, w, ws, ou, ous, od, ods, energy, flags, stream, cfg);
  }
  // a chunk that does not fit LDS (or the 16-bit incidence range) is retried with half the size;
  // the size that worked is remembered for this call shape
  for (bool halved = false;;) {
    int rc = run_chunk_cfg(c, dtype, mode, batch, field, fs, u, us, w, ws, ou, ous, od, ods, energy, flags, stream, cfg);
    if (rc == FEMHOT_OK && halved) {
      std::lock_guard<std::mutex> lk(g_tune_mu);
      g_tuned[{c, mode, dtype, batch, (int)(flags & FEMHOT_ELEMENT_FIELD)}] = cfg;
    }
    if (rc != FEMHOT_ERR_UNSUPPORTED || cfg.epc <= 32) return rc;
    cfg.epc /= 2;
    halved = true;
  }
}

int autotune_chunk(Ctx *c, int dtype, int mode, int batch, const void *field, long long fs, const void *u,
                   long long us, const void *w, long long ws, void *ou, long long ous, void *od, long long ods,
                   double *energy, unsigned flags, hipStream_t s, int *best_epc, int *best_block, int *best_grid,
                   float *best_ms) {
  static const int epcs[] = {128, 256, 512, 1024, 2048};
  // staged kernel: 3 workgroup sizes x {one workgroup per item, 4 or 8 per CU}; pipelined kernel
  // (FP32): 2 workgroup sizes x persistent grids only (its prefetch needs a next step)
  struct Shape {
    bool dma;
    int block, grid;
  };
  static const Shape shapes[] = {{false, 128, 0}, {false, 128, 4}, {false, 128, 8}, {false, 256, 0},
                                 {false, 256, 4}, {false, 256, 8}, {false, 512, 0}, {false, 512, 4},
                                 {false, 512, 8}, {true, 256, 1},  {true, 256, 2},  {true, 256, 4},
                                 {true, 512, 1},  {true, 512, 2},  {true, 512, 4}};
  const int sbs[] = {0, 1};  // automatic, and one sample per work item (only differs when batch > 1)
  hipEvent_t e0, e1;
  FH_HIP(hipEventCreate(&e0));
  FH_HIP(hipEventCreate(&e1));
  ChunkConfig best;
  float best_t = 1e30f;
  for (int epc : epcs) {
    if ((long long)epc > c->Ne * 2 && epc != 128) continue;
    for (const Shape &sh : shapes)
      for (int sb : sbs) {
        if (sb == 1 && batch == 1) continue;
        if (sh.dma && dtype != FEMHOT_F32) continue;
        ChunkConfig cfg;
        cfg.epc = epc;
        cfg.block = sh.block;
        cfg.grid_per_cu = sh.grid;
        cfg.dma = sh.dma;
        cfg.sb = sb;
        int rc = run_chunk_cfg(c, dtype, mode, batch, field, fs, u, us, w, ws, ou, ous, od, ods, energy, flags, s, cfg);
        if (rc == FEMHOT_ERR_UNSUPPORTED) continue;  // chunk does not fit LDS
        if (rc) return rc;
        FH_HIP(hipEventRecord(e0, s));
        constexpr int reps = 3;
        for (int r = 0; r < reps; ++r) {
          rc = run_chunk_cfg(c, dtype, mode, batch, field, fs, u, us, w, ws, ou, ous, od, ods, energy, flags, s, cfg);
          if (rc) return rc;
        }
        FH_HIP(hipEventRecord(e1, s));
        FH_HIP(hipEventSynchronize(e1));
        float ms = 0.f;
        FH_HIP(hipEventElapsedTime(&ms, e0, e1));
        ms /= reps;
        if (ms < best_t) {
          best_t = ms;
          best = cfg;
        }
      }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (best_t >= 1e30f) {
    set_error("autotune: no chunk configuration fits");
    return FEMHOT_ERR_UNSUPPORTED;
  }
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tuned[{c, mode, dtype, batch, (int)(flags & FEMHOT_ELEMENT_FIELD)}] = best;
  }
  int rc = ensure_chunks(c, best.epc, nullptr, nullptr);  // leave the winning plan resident
  if (rc) return rc;
  if (best_epc) *best_epc = best.epc;
  if (best_block) *best_block = best.block | (best.dma ? (int)FEMHOT_CHUNK_DMA : 0);
  if (best_grid) *best_grid = best.grid_per_cu;
  if (best_ms) *best_ms = best_t;
  return FEMHOT_OK;
}

int set_chunk_config(Ctx *c, int mode, int dtype, int batch, unsigned flags, int epc, int block, int grid_per_cu) {
  if (epc == 0 || mode == MODE_TANGENT) {  // an element-parallel path, as femhot_autotune reports it
    if (block != FEMHOT_PATH_ATOMIC && block != FEMHOT_PATH_COLOUR &&
        !(mode == MODE_TANGENT && block == FEMHOT_PATH_ROWS)) {
      set_error("femhot_set_chunk_config: epc 0 (or the tangent) takes block FEMHOT_PATH_ATOMIC or FEMHOT_PATH_COLOUR "
                "(the tangent also FEMHOT_PATH_ROWS)");
      return FEMHOT_ERR_ARG;
    }
    record_path(c, dtype, mode, mode == MODE_TANGENT ? 1 : batch, flags,
                block == FEMHOT_PATH_ROWS ? PATH_ROWS : block == FEMHOT_PATH_COLOUR ? PATH_COLOUR : PATH_ATOMIC);
    return FEMHOT_OK;
  }
  const bool dma = (block & (int)FEMHOT_CHUNK_DMA) != 0;
  block &= ~(int)FEMHOT_CHUNK_DMA;
  if (epc < 1 || epc > 8192 || (block != 128 && block != 256 && block != 512) ||
      (dma && (dtype != FEMHOT_F32 || block == 128))) {
    set_error("femhot_set_chunk_config: epc in [1,8192], block in {128,256,512} (FEMHOT_CHUNK_DMA: FP32, "
              "{256,512})");
    return FEMHOT_ERR_ARG;
  }
  ChunkConfig cfg;
  cfg.epc = epc;
  cfg.block = block;
  cfg.grid_per_cu = grid_per_cu;
  cfg.dma = dma;
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tuned[{c, mode, dtype, batch, (int)(flags & FEMHOT_ELEMENT_FIELD)}] = cfg;
  }
  return ensure_chunks(c, epc, nullptr, nullptr);
}

int tuned_path(const Ctx *c, int dtype, int mode, int batch, unsigned flags) {
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_tuned.find({c, mode, dtype, batch, (int)(flags & FEMHOT_ELEMENT_FIELD)});
    if (it != g_tuned.end()) return it->second.path;
  }
  if (mode == MODE_TANGENT) return c->ph == NEOHOOKE ? PATH_COLOUR : PATH_ROWS;
  return (dtype == FEMHOT_F64 && c->ph == NEOHOOKE) ? PATH_ATOMIC : PATH_CHUNK;
}

void record_path(const Ctx *c, int dtype, int mode, int batch, unsigned flags, int path) {
  ChunkConfig cfg;
  std::lock_guard<std::mutex> lk(g_tune_mu);
  auto key = std::make_tuple((const void *)c, mode, dtype, batch, (int)(flags & FEMHOT_ELEMENT_FIELD));
  auto it = g_tuned.find(key);
  if (it != g_tuned.end()) cfg = it->second;  // keeps the best chunk configuration beside the path
  cfg.path = path;
  g_tuned[key] = cfg;
}

void forget_tuning(const void *ctx) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  for (auto it = g_tuned.begin(); it != g_tuned.end();)
    if (std::get<0>(it->first) == ctx)
      it = g_tuned.erase(it);
    else
      ++it;
}

// ---------------------------------------------------------------------------------------------
// Fixed scratch of the allocation-free entry point: the chunk path's shared-node partials and
// per-chunk energies in a file-scope device array, one copy per GPU, held until the process ends.
// Host code never passes `scratch` itself to a kernel (on the host it is only a stand-in symbol):
// its device address is resolved once per device with hipGetSymbolAddress and kept here.
// ---------------------------------------------------------------------------------------------
constexpr size_t kFixedScratchBytes = 64ull << 20;  // 64 MiB per GPU
static __device__ __align__(256) unsigned char scratch[kFixedScratchBytes];

constexpr int kMaxDevices = 64;
static void *g_scratch_ptr[kMaxDevices];  // device address of `scratch` on each device, once resolved
static std::mutex g_scratch_mu;

int init_fixed_scratch() {
  int dev = 0;
  FH_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) {
    set_error("init_fixed_scratch: device index beyond the library's table");
    return FEMHOT_ERR_UNSUPPORTED;
  }
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_ptr[dev]) return FEMHOT_OK;
  void *p = nullptr;
  FH_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  g_scratch_ptr[dev] = p;
  return FEMHOT_OK;
}

static int fixed_scratch(void **out) {
  int dev = 0;
  FH_HIP(hipGetDevice(&dev));
  {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (dev >= 0 && dev < kMaxDevices && g_scratch_ptr[dev]) {
      *out = g_scratch_ptr[dev];
      return FEMHOT_OK;
    }
  }
  int rc = init_fixed_scratch();  // a context created on another device: resolve this one now
  if (rc) return rc;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  *out = g_scratch_ptr[dev];
  return FEMHOT_OK;
}

int run_chunk_fixed_scratch(Ctx *c, int dtype, int mode, int batch, const void *field, long long fs, const void *u,
                            long long us, void *ou, long long ous, void *od, long long ods, double *energy,
                            unsigned flags, hipStream_t s) {
  void *ws = nullptr;
  int rc = fixed_scratch(&ws);
  if (rc) return rc;
  ChunkConfig cfg = default_config(c, dtype);
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_tuned.find({c, mode, dtype, batch, (int)(flags & FEMHOT_ELEMENT_FIELD)});
    if (it != g_tuned.end() && it->second.path == PATH_CHUNK) cfg = it->second;
  }
  return run_chunk_cfg(c, dtype, mode, batch, field, fs, u, us, nullptr, 0, ou, ous, od, ods, energy, flags, s, cfg,
                       ws, kFixedScratchBytes);
}

}  // namespace femhot
