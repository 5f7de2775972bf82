// LDS-staged chunk assembly: the batched / persistent path of the residual, residual-VJP and
// potential-gradient kernels.
//
// One workgroup owns one chunk of spatially compact elements at a time (grid-stride loop over
// chunk x sample-block work items, so the launch can be persistent).  Per work item:
//   1. stage the chunk's node coordinates (precomputed chunk-relative, contiguous -> coalesced)
//      and gather the chunk's nodal solution / cotangent / field (node ids sorted) into LDS;
//   2. one lane per element: element routine reading LDS, nodal results accumulated with LDS
//      float atomics (ds_add), per-element results written straight to HBM;
//   3. every chunk node is written once: plain store when the node belongs to this chunk only,
//      hardware float atomic when it is shared with another chunk (those, and nodes in no element,
//      are cleared first by k_zero_nodes).
// HBM traffic per element is therefore the chunk-local node data once per chunk (instead of once
// per element-node incidence) plus the element's own connectivity (uint16 local ids) and field.
#include <algorithm>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "femhot_chunk.h"
#include "femhot_launch.h"

namespace femhot {

struct ChunkArgs {
  const int *elem_off, *node_off, *seg, *perm, *nodes;
  const uint16_t *lconn;
  const void *xyz;  // chunk-relative coordinates in the arithmetic type
  int n_chunks, n_sblocks, sb, batch, cap;
  const void *field;
  long long fs;
  const void *u;
  long long us;
  const void *w;
  long long ws;
  void *ou;
  long long ous;
  void *od;
  long long ods;
  double *energy;
  int nseg;
  const uint8_t *dmask;
  unsigned flags;
  PhysParams pp;
};

template <int N> __device__ __forceinline__ void load_lconn(const uint16_t *p, int (&l)[N]) {
  if constexpr (N == 4) {
    const uint2 v = *reinterpret_cast<const uint2 *>(p);
    l[0] = v.x & 0xffff;
    l[1] = v.x >> 16;
    l[2] = v.y & 0xffff;
    l[3] = v.y >> 16;
  } else {
    const uint4 v = *reinterpret_cast<const uint4 *>(p);
    l[0] = v.x & 0xffff;
    l[1] = v.x >> 16;
    l[2] = v.y & 0xffff;
    l[3] = v.y >> 16;
    l[4] = v.z & 0xffff;
    l[5] = v.z >> 16;
    l[6] = v.w & 0xffff;
    l[7] = v.w >> 16;
  }
}

template <class R, int ET, int PH, int NGP, int MODE, int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_chunk(ChunkArgs A) {
  using E = Elem<R, ET, PH>;
  constexpr int N = E::N, D = E::D, P = E::P, NL = E::NL;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  __shared__ double red[BLOCK / 64];
  const bool elemental = (A.flags & FEMHOT_ELEMENT_FIELD) != 0;
  const bool mask = (A.flags & FEMHOT_MASK_DIRICHLET) != 0;
  const bool accumulate = (A.flags & FEMHOT_ACCUMULATE) != 0;
  const bool picard = (A.flags & FEMHOT_PICARD) != 0;
  const int cap = A.cap;
  // LDS layout (all arrays sized for `cap` nodes)
  R *X = reinterpret_cast<R *>(smem_raw);
  R *U = X + 3 * cap;
  R *W = U + P * cap;                              // cotangent (VJP only)
  R *FD = W + (MODE == MODE_VJP ? P * cap : 0);    // nodal field
  R *ACC = FD + (elemental ? 0 : cap);             // nodal output accumulators
  R *GD = ACC + P * cap;                           // nodal field-gradient accumulators
  const bool want_gd = (MODE != MODE_RES) && A.od != nullptr;
  const bool want_ou = A.ou != nullptr;
  const int tid = threadIdx.x;
  const int n_items = A.n_chunks * A.n_sblocks;
  for (int item = blockIdx.x; item < n_items; item += gridDim.x) {
    const int chunk = item / A.n_sblocks;
    const int sbi = item - chunk * A.n_sblocks;
    const int e0 = A.elem_off[chunk], e1 = A.elem_off[chunk + 1];
    const int n0 = A.node_off[chunk], nl = A.node_off[chunk + 1] - n0;
    const R *xyz = reinterpret_cast<const R *>(A.xyz) + 3LL * n0;
    for (int i = tid; i < 3 * nl; i += BLOCK) X[i] = xyz[i];
    const int b0 = sbi * A.sb, b1 = min(A.batch, b0 + A.sb);
    for (int b = b0; b < b1; ++b) {
      const R *Ub = reinterpret_cast<const R *>(A.u) + (long long)b * A.us;
      const R *Wb = reinterpret_cast<const R *>(A.w) + (long long)b * A.ws;
      const R *Fb = reinterpret_cast<const R *>(A.field) + (long long)b * A.fs;
      for (int i = tid; i < nl; i += BLOCK) {
        const int g = A.nodes[n0 + i] & 0x7fffffff;
#pragma unroll
        for (int c = 0; c < P; ++c) {
          U[i * P + c] = Ub[(long long)g * P + c];
          if constexpr (MODE == MODE_VJP) W[i * P + c] = Wb[(long long)g * P + c];
          ACC[i * P + c] = R(0);
        }
        if (!elemental) {
          FD[i] = Fb[g];
          GD[i] = R(0);
        }
      }
      __syncthreads();
      double en = 0.0;
      for (int p = e0 + tid; p < e1; p += BLOCK) {
        const int e = A.perm[p];
        int l[N];
        load_lconn<N>(A.lconn + (long long)p * N, l);
        E el;
#pragma unroll
        for (int i = 0; i < D; ++i) {
          const R x0 = X[l[0] * 3 + i];
#pragma unroll
          for (int a = 0; a < N; ++a) el.x[a][i] = X[l[a] * 3 + i] - x0;
        }
        el.elemental = elemental;
        if (elemental) {
          el.de = Fb[e];
#pragma unroll
          for (int a = 0; a < N; ++a) el.d[a] = el.de;
        } else {
          el.de = R(0);
#pragma unroll
          for (int a = 0; a < N; ++a) el.d[a] = FD[l[a]];
        }
#pragma unroll
        for (int a = 0; a < N; ++a)
#pragma unroll
          for (int c = 0; c < P; ++c) el.u[a * P + c] = U[l[a] * P + c];
        R ro[NL], gd[N];
        if constexpr (MODE == MODE_RES) {
          R energy;
          elem_residual<R, ET, PH, NGP>(el, A.pp, ro, energy);
          en += (double)energy;
        } else if constexpr (MODE == MODE_VJP) {
          R w[NL];
#pragma unroll
          for (int a = 0; a < N; ++a)
#pragma unroll
            for (int c = 0; c < P; ++c) w[a * P + c] = W[l[a] * P + c];
          elem_vjp<R, ET, PH, NGP>(el, w, A.pp, picard, ro, gd);
        } else {
          R pi;
          elem_egrad<R, ET, PH, NGP>(el, A.pp, ro, gd, pi);
          en += (double)pi;
        }
        if (want_ou) {
#pragma unroll
          for (int a = 0; a < N; ++a)
#pragma unroll
            for (int c = 0; c < P; ++c) atomicAdd(&ACC[l[a] * P + c], ro[a * P + c]);
        }
        if constexpr (MODE != MODE_RES) {
          if (want_gd) {
            if (elemental) {
              R *OD = reinterpret_cast<R *>(A.od) + (long long)b * A.ods;
              OD[e] = accumulate ? OD[e] + gd[0] : gd[0];
            } else {
#pragma unroll
              for (int a = 0; a < N; ++a) atomicAdd(&GD[l[a]], gd[a]);
            }
          }
        }
      }
      if (A.energy) {
        en = wave_sum(en);
        if ((tid & 63) == 0) red[tid >> 6] = en;
      }
      __syncthreads();
      if (A.energy && tid == 0) {
        double s = 0.0;
#pragma unroll
        for (int k = 0; k < BLOCK / 64; ++k) s += red[k];
        unsafeAtomicAdd(&A.energy[(long long)b * A.nseg + A.seg[chunk]], s);
      }
      R *OU = reinterpret_cast<R *>(A.ou) + (long long)b * A.ous;
      R *OD = reinterpret_cast<R *>(A.od) + (long long)b * A.ods;
      for (int i = tid; i < nl; i += BLOCK) {
        const int raw = A.nodes[n0 + i];
        const int g = raw & 0x7fffffff;
        const bool shared = raw < 0;
        if (want_ou) {
#pragma unroll
          for (int c = 0; c < P; ++c) {
            const long long dof = (long long)g * P + c;
            const bool masked = mask && A.dmask[dof];
            const R v = ACC[i * P + c];
            if (shared) {
              if (!masked) unsafeAtomicAdd(OU + dof, v);
            } else if (!(masked && accumulate)) {
              OU[dof] = masked ? R(0) : (accumulate ? OU[dof] + v : v);
            }
          }
        }
        if (want_gd && !elemental) {
          const R v = GD[i];
          if (shared)
            unsafeAtomicAdd(OD + g, v);
          else
            OD[g] = accumulate ? OD[g] + v : v;
        }
      }
      __syncthreads();
    }
  }
}

// Clears the outputs of shared nodes and of nodes in no element (the chunk kernel writes every
// other node exactly once).
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

struct ChunkConfig {
  int epc = 1024, block = 256, grid_per_cu = 0, sb = 0;  // grid_per_cu 0: one workgroup per item
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

// The dynamic-LDS limit is raised once per kernel (to the full 160 KiB) on its first launch, so
// later launches (e.g. under graph capture) make no attribute calls.
template <class K> static int raise_lds_limit(K k) {
  static bool done = false;
  if (!done) {
    FH_HIP(hipFuncSetAttribute((const void *)k, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
    done = true;
  }
  return FEMHOT_OK;
}

template <class R, int ET, int PH, int NGP, int MODE>
static int launch_chunk_block(int block, dim3 grid, size_t shm, hipStream_t s, const ChunkArgs &A) {
  int rc;
  if (block == 128) {
    auto k = k_chunk<R, ET, PH, NGP, MODE, 128>;
    if ((rc = raise_lds_limit(k))) return rc;
    hipLaunchKernelGGL(k, grid, dim3(128), shm, s, A);
  } else if (block == 512) {
    auto k = k_chunk<R, ET, PH, NGP, MODE, 512>;
    if ((rc = raise_lds_limit(k))) return rc;
    hipLaunchKernelGGL(k, grid, dim3(512), shm, s, A);
  } else {
    auto k = k_chunk<R, ET, PH, NGP, MODE, 256>;
    if ((rc = raise_lds_limit(k))) return rc;
    hipLaunchKernelGGL(k, grid, dim3(256), shm, s, A);
  }
  FH_HIP(hipGetLastError());
  return FEMHOT_OK;
}

template <int ET>
static int launch_chunk_et(int dtype, int ph, int ngp, int mode, int block, dim3 grid, size_t shm, hipStream_t s,
                           const ChunkArgs &A) {
  return dispatch_r_ph_ngp(dtype, ph, ngp, [&](auto rp, auto phc, auto ngpc) -> int {
    using R = std::remove_pointer_t<decltype(rp)>;
    constexpr int PH = decltype(phc)::value, NGP = decltype(ngpc)::value;
    if (mode == MODE_RES) return launch_chunk_block<R, ET, PH, NGP, MODE_RES>(block, grid, shm, s, A);
    if (mode == MODE_VJP) return launch_chunk_block<R, ET, PH, NGP, MODE_VJP>(block, grid, shm, s, A);
    return launch_chunk_block<R, ET, PH, NGP, MODE_EGRAD>(block, grid, shm, s, A);
  });
}

static int run_chunk_cfg(Ctx *c, int dtype, int mode, int batch, const void *field, long long fs, const void *u,
                         long long us, const void *w, long long ws, void *ou, long long ous, void *od, long long ods,
                         double *energy, unsigned flags, hipStream_t s, const ChunkConfig &cfg) {
  int rc = ensure_chunks(c, cfg.epc);
  if (rc) return rc;
  const ChunkPlan &p = c->chunk;
  const bool elemental = (flags & FEMHOT_ELEMENT_FIELD) != 0;
  const int P = c->ndpn;
  const size_t rs = dtype == FEMHOT_F32 ? sizeof(float) : sizeof(double);
  const int per_node = 3 + P + (mode == MODE_VJP ? P : 0) + (elemental ? 0 : 1) + P + (elemental ? 0 : 1);
  const size_t shm = (size_t)p.max_nodes * per_node * rs;
  if (shm > 160 * 1024) {
    set_error("run_chunk: chunk too large for LDS; use fewer elements per chunk");
    return FEMHOT_ERR_UNSUPPORTED;
  }
  const bool acc = (flags & FEMHOT_ACCUMULATE) != 0;
  const bool want_gd_nodal = od && !elemental && mode != MODE_RES;
  // clear shared / untouched nodes
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
  A.seg = p.chunk_seg;
  A.perm = p.perm;
  A.nodes = p.chunk_nodes;
  A.lconn = p.lconn;
  A.xyz = dtype == FEMHOT_F32 ? (const void *)p.chunk_xyz32 : (const void *)p.chunk_xyz64;
  A.n_chunks = p.n_chunks;
  A.sb = cfg.sb > 0 ? std::min(cfg.sb, batch) : batch;
  A.n_sblocks = (batch + A.sb - 1) / A.sb;
  A.batch = batch;
  A.cap = p.max_nodes;
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
  A.energy = energy;
  A.nseg = c->nseg;
  A.dmask = c->d_dmask;
  A.flags = flags;
  A.pp = c->pp;
  const long long items = (long long)A.n_chunks * A.n_sblocks;
  long long grid = items;
  if (cfg.grid_per_cu > 0) grid = std::min<long long>(items, (long long)cfg.grid_per_cu * cu_count());
  if (grid < 1) return FEMHOT_OK;
  if ((rc = timing_mark(c, s))) return rc;
  switch (c->et) {
    case QUAD4: rc = launch_chunk_et<QUAD4>(dtype, c->ph, c->ngp, mode, cfg.block, dim3((unsigned)grid), shm, s, A); break;
    case HEX8: rc = launch_chunk_et<HEX8>(dtype, c->ph, c->ngp, mode, cfg.block, dim3((unsigned)grid), shm, s, A); break;
    default: rc = launch_chunk_et<TET4>(dtype, c->ph, c->ngp, mode, cfg.block, dim3((unsigned)grid), shm, s, A); break;
  }
  if (rc) return rc;
  return timing_mark(c, s);
}

static ChunkConfig default_config(Ctx *c, int dtype) {
  ChunkConfig cfg;
  const bool big = c->et == HEX8 || (dtype == FEMHOT_F64 && c->ph == NEOHOOKE);
  cfg.epc = big ? 256 : 1024;
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
  if (epc > 0) cfg.epc = epc;
  if (block > 0) cfg.block = block;
  return run_chunk_cfg(c, dtype, mode, batch, field, fs, u, us, w, ws, ou, ous, od, ods, energy, flags, stream, cfg);
}

int autotune_chunk(Ctx *c, int dtype, int mode, int batch, const void *field, long long fs, const void *u,
                   long long us, const void *w, long long ws, void *ou, long long ous, void *od, long long ods,
                   double *energy, unsigned flags, hipStream_t s, int *best_epc, int *best_block, int *best_grid,
                   float *best_ms) {
  static const int epcs[] = {256, 512, 1024, 2048, 4096};
  static const int blocks[] = {128, 256, 512};
  static const int grids[] = {0, 4, 8};
  hipEvent_t e0, e1;
  FH_HIP(hipEventCreate(&e0));
  FH_HIP(hipEventCreate(&e1));
  ChunkConfig best;
  float best_t = 1e30f;
  for (int epc : epcs) {
    if ((long long)epc > c->Ne * 2 && epc != 256) continue;
    for (int blk : blocks)
      for (int gp : grids) {
        ChunkConfig cfg;
        cfg.epc = epc;
        cfg.block = blk;
        cfg.grid_per_cu = gp;
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
  int rc = ensure_chunks(c, best.epc);  // leave the winning plan resident
  if (rc) return rc;
  if (best_epc) *best_epc = best.epc;
  if (best_block) *best_block = best.block;
  if (best_grid) *best_grid = best.grid_per_cu;
  if (best_ms) *best_ms = best_t;
  return FEMHOT_OK;
}

void forget_tuning(const void *ctx) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  for (auto it = g_tuned.begin(); it != g_tuned.end();)
    if (std::get<0>(it->first) == ctx)
      it = g_tuned.erase(it);
    else
      ++it;
}

}  // namespace femhot
