// Software-pipelined FP32 chunk kernel: the same work as k_chunk (femhot_chunk_kernels.h), with
// the global->LDS staging done by LDS DMA (global_load_lds) one step ahead of the arithmetic.
//
// A persistent workgroup walks its steps (chunk, sample).  While it computes the elements of
// step s, the plan of step s+1 (coordinates, node ids, partial slots, incidence list, local
// connectivity) streams into the second of two plan buffers; while it gathers the per-corner
// results of step s into nodes and writes them out, the nodal solution / cotangent / field of
// step s+1 streams into the data buffer that step s has finished reading.  DMA loads take no
// registers and need no store instructions, so the staging of the next step costs the
// arithmetic of this one nothing but issue slots; each step waits only for what it reads.
//
// LDS DMA on gfx950: lane i of a wave writes its `size` loaded bytes at (wave-uniform LDS base,
// taken from M0) + i * size; the load is counted by vmcnt, so every consumer below is preceded
// by an explicit s_waitcnt and a barrier.
// Every DMA'd LDS region is padded to whole 64-entry wave rows, so no lane of a partially
// active wave can address past its region.
#pragma once
#include "femhot_chunk_kernels.h"

namespace femhot {

typedef __attribute__((address_space(3))) void lds_void_t;

// One LDS-DMA load per lane.  Written as inline assembly rather than __builtin_amdgcn_global_load_lds:
// the compiler treats every LDS access after a builtin LDS DMA as possibly aliasing it and waits
// for the DMA (vmcnt) before the access, which would stall the arithmetic of this step on the
// prefetch of the next.  Here the kernel orders every consumer itself (wait_all_memory + barrier).
// M0 (the wave's LDS base) is saved and restored around the load; s_nop 0 is the one wait state
// between an SALU write of M0 and the DMA that reads it.
#define FEMHOT_LDS_DMA(INSN)                                                                  \
  {                                                                                           \
    const unsigned lds = (unsigned)(uintptr_t)(lds_void_t *)wave_dst;                         \
    unsigned saved;                                                                           \
    asm volatile("s_mov_b32 %0, m0\n\ts_mov_b32 m0, %2\n\ts_nop 0\n\t" INSN " %1, off\n\t" \
                 "s_mov_b32 m0, %0"                                                           \
                 : "=&s"(saved)                                                               \
                 : "v"(g), "s"(lds)                                                           \
                 : "memory");                                                                 \
  }
__device__ __forceinline__ void dma4(const void *g, void *wave_dst) FEMHOT_LDS_DMA("global_load_lds_dword")
__device__ __forceinline__ void dma16(const void *g, void *wave_dst) FEMHOT_LDS_DMA("global_load_lds_dwordx4")
#undef FEMHOT_LDS_DMA
__device__ __forceinline__ void wait_all_memory() { __builtin_amdgcn_s_waitcnt(0); }

__host__ __device__ constexpr int round64(int x) { return (x + 63) & ~63; }

// Byte offsets of the LDS regions of k_chunk_pipe.  The plan regions are those of plan buffer 0;
// plan buffer 1 follows at + plan (a stride, so a runtime buffer index costs no indexed registers).
struct PipeLayout {
  int x, nid, slot, ist, lc, inc, perm, plan;
  int u, w, fd, ef, er, erd, total;
  int cap64;  // node rows per plan / data region
};

__host__ __device__ inline PipeLayout pipe_layout(int cap, int emax, int N, int P, int mode, bool elemental,
                                                   bool want_ou, bool nodal_gd, bool identity) {
  PipeLayout L{};
  const int cap64 = round64(cap), emax64 = round64(emax);
  const int ist_words = round64(cap / 2 + 2);  // inc_start read as whole 32-bit words
  const int lc_words = round64(N * emax / 2);  // uint16 per corner, N even
  int off = 0;
  L.x = off, off += cap64 * 16;
  L.nid = off, off += cap64 * 4;
  L.slot = off, off += cap64 * 4;
  L.ist = off, off += ist_words * 4;
  L.lc = off, off += lc_words * 4;
  L.inc = off, off += (want_ou || nodal_gd) ? lc_words * 4 : 0;
  L.perm = off, off += identity ? 0 : emax64 * 4;
  L.plan = off;
  off *= 2;
  L.u = off, off += P * cap64 * 4;
  L.w = off, off += mode == MODE_VJP ? P * cap64 * 4 : 0;
  L.fd = off, off += elemental ? 0 : cap64 * 4;
  L.ef = off, off += elemental ? emax64 * 4 : 0;
  L.er = off, off += want_ou ? ((P * N * emax * 4 + 15) & ~15) : 0;
  L.erd = off, off += nodal_gd ? ((N * emax * 4 + 15) & ~15) : 0;
  L.total = off;
  L.cap64 = cap64;
  return L;
}

struct StepGeom {
  int e0, ne, n0, nl;
};

__device__ __forceinline__ StepGeom step_geom(const ChunkArgs &A, int chunk) {
  StepGeom g;
  g.e0 = A.elem_off[chunk];
  g.ne = A.elem_off[chunk + 1] - g.e0;
  g.n0 = A.node_off[chunk];
  g.nl = A.node_off[chunk + 1] - g.n0;
  return g;
}

// A step of a workgroup's walk: work item (chunk, block of samples) and the sample within it.
struct PipeStep {
  int item, chunk, b, b1;
};

__device__ __forceinline__ PipeStep first_step(const ChunkArgs &A, int item) {
  PipeStep s;
  s.item = item;
  s.chunk = item / A.n_sblocks;
  s.b = (item - s.chunk * A.n_sblocks) * A.sb;
  s.b1 = min(A.batch, s.b + A.sb);
  return s;
}

// The step after s: the next sample of the work item, else the first sample of the workgroup's
// next work item (item >= the number of work items when s is the workgroup's last step).
// Returned by value and computed unconditionally, so it stays in scalar registers.
__device__ __forceinline__ PipeStep next_step(const ChunkArgs &A, const PipeStep &s) {
  PipeStep n = s;
  n.b = s.b + 1;
  if (n.b >= s.b1) n = first_step(A, s.item + (int)gridDim.x);
  return n;
}

// DMA of one chunk's plan into plan buffer k.  wv: wave index (uniform), lane: lane in wave.
template <int N, int BLOCK>
__device__ __forceinline__ void pipe_issue_plan(const ChunkArgs &A, const PipeLayout &L, unsigned char *sm, int k,
                                                const StepGeom &G, bool want_inc, int wv, int lane) {
  const float *xyz = reinterpret_cast<const float *>(A.xyz);
  for (int base = wv * 64; base < G.nl; base += BLOCK) {
    const int i = base + lane;
    if (i < G.nl) {
      dma16(xyz + 4LL * (G.n0 + i), sm + (L.x + k * L.plan) + base * 16);
      dma4(A.nodes + G.n0 + i, sm + (L.nid + k * L.plan) + base * 4);
      dma4(A.node_slot + G.n0 + i, sm + (L.slot + k * L.plan) + base * 4);
    }
  }
  {  // inc_start is uint16: whole words from the one holding entry n0 to the one holding n0 + nl - 1
    const uint32_t *src = reinterpret_cast<const uint32_t *>(A.inc_start) + (G.n0 >> 1);
    const int nw = ((G.n0 + G.nl - 1) >> 1) - (G.n0 >> 1) + 1;
    for (int base = wv * 64; base < nw; base += BLOCK) {
      const int i = base + lane;
      if (i < nw) dma4(src + i, sm + (L.ist + k * L.plan) + base * 4);
    }
  }
  {  // local connectivity and incidence list: N uint16 per element, N even, 8-byte aligned
    const int nw = N * G.ne / 2;
    const uint32_t *lc = reinterpret_cast<const uint32_t *>(A.lconn) + (long long)N * G.e0 / 2;
    const uint32_t *inc = reinterpret_cast<const uint32_t *>(A.inc) + (long long)N * G.e0 / 2;
    for (int base = wv * 64; base < nw; base += BLOCK) {
      const int i = base + lane;
      if (i < nw) {
        dma4(lc + i, sm + (L.lc + k * L.plan) + base * 4);
        if (want_inc) dma4(inc + i, sm + (L.inc + k * L.plan) + base * 4);
      }
    }
  }
  if (!A.perm_identity) {
    for (int base = wv * 64; base < G.ne; base += BLOCK) {
      const int i = base + lane;
      if (i < G.ne) dma4(A.perm + G.e0 + i, sm + (L.perm + k * L.plan) + base * 4);
    }
  }
}

// DMA of one step's per-sample data (solution, cotangent, field), addressed through the node ids
// of plan buffer k, into the data buffer.
template <int P, int MODE, int BLOCK>
__device__ __forceinline__ void pipe_issue_data(const ChunkArgs &A, const PipeLayout &L, unsigned char *sm, int k,
                                                const StepGeom &G, int b, bool elemental, int wv, int lane) {
  const float *Ub = reinterpret_cast<const float *>(A.u) + (long long)b * A.us;
  const float *Wb = reinterpret_cast<const float *>(A.w) + (long long)b * A.ws;
  const float *Fb = reinterpret_cast<const float *>(A.field) + (long long)b * A.fs;
  const int *nid = reinterpret_cast<const int *>(sm + (L.nid + k * L.plan));
  const int cap64 = L.cap64;
  for (int base = wv * 64; base < G.nl; base += BLOCK) {
    const int i = base + lane;
    if (i < G.nl) {
      const long long g = nid[i];
#pragma unroll
      for (int c = 0; c < P; ++c) dma4(Ub + g * P + c, sm + L.u + (c * cap64 + base) * 4);
      if constexpr (MODE == MODE_VJP) {
#pragma unroll
        for (int c = 0; c < P; ++c) dma4(Wb + g * P + c, sm + L.w + (c * cap64 + base) * 4);
      }
      if (!elemental) dma4(Fb + g, sm + L.fd + base * 4);
    }
  }
  if (elemental) {
    const int *perm = reinterpret_cast<const int *>(sm + (L.perm + k * L.plan));
    for (int base = wv * 64; base < G.ne; base += BLOCK) {
      const int i = base + lane;
      if (i < G.ne) dma4(Fb + (A.perm_identity ? G.e0 + i : perm[i]), sm + L.ef + base * 4);
    }
  }
}

template <int ET, int PH, int NGP, int MODE, int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_chunk_pipe(ChunkArgs A) {
  using R = float;
  using E = Elem<R, ET, PH>;
  constexpr int N = E::N, D = E::D, P = E::P, NL = E::NL;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  __shared__ double red[BLOCK / 64];
  const bool elemental = (A.flags & FEMHOT_ELEMENT_FIELD) != 0;
  const bool mask = (A.flags & FEMHOT_MASK_DIRICHLET) != 0;
  const bool accumulate = (A.flags & FEMHOT_ACCUMULATE) != 0;
  const bool picard = (A.flags & FEMHOT_PICARD) != 0;
  const bool want_gd = (MODE != MODE_RES) && A.od != nullptr;
  const bool want_ou = A.ou != nullptr;
  const bool nodal_gd = want_gd && !elemental;
  const bool want_inc = want_ou || nodal_gd;
  const PipeLayout L =
      pipe_layout(A.cap, A.emax, N, P, MODE, elemental, want_ou, nodal_gd, A.perm_identity != 0);
  const int cap64 = L.cap64;
  const int tid = threadIdx.x, lane = tid & 63;
  const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int n_items = A.n_chunks * A.n_sblocks;
  if ((int)blockIdx.x >= n_items) return;  // uniform over the workgroup
  __shared__ __align__(16) int geo[4];     // chunk offsets of the step after next (DMA)
  PipeStep S = first_step(A, blockIdx.x);
  StepGeom G = step_geom(A, S.chunk);
  PipeStep NS = next_step(A, S);
  bool has_next = NS.item < n_items;
  StepGeom NG = G;
  if (has_next && NS.chunk != S.chunk) NG = step_geom(A, NS.chunk);
  int cur = 0;
  pipe_issue_plan<N, BLOCK>(A, L, smem_raw, cur, G, want_inc, wv, lane);
  wait_all_memory();
  __syncthreads();
  pipe_issue_data<P, MODE, BLOCK>(A, L, smem_raw, cur, G, S.b, elemental, wv, lane);
  wait_all_memory();
  __syncthreads();
  for (;;) {
    const int chunk = S.chunk, b = S.b;
    const bool new_plan = has_next && NS.chunk != chunk;
    const int nxt = new_plan ? cur ^ 1 : cur;
    if (new_plan) pipe_issue_plan<N, BLOCK>(A, L, smem_raw, nxt, NG, want_inc, wv, lane);
    // ---- elements of this step (LDS only) ----
    const float4 *X = reinterpret_cast<const float4 *>(smem_raw + (L.x + cur * L.plan));
    const uint16_t *LC = reinterpret_cast<const uint16_t *>(smem_raw + (L.lc + cur * L.plan));
    const int *PERM = reinterpret_cast<const int *>(smem_raw + (L.perm + cur * L.plan));
    const R *U = reinterpret_cast<const R *>(smem_raw + L.u);
    const R *W = reinterpret_cast<const R *>(smem_raw + L.w);
    const R *FD = reinterpret_cast<const R *>(smem_raw + L.fd);
    const R *EF = reinterpret_cast<const R *>(smem_raw + L.ef);
    R *ER = reinterpret_cast<R *>(smem_raw + L.er);
    R *ERD = reinterpret_cast<R *>(smem_raw + L.erd);
    R *ODb = reinterpret_cast<R *>(A.od) + (long long)b * A.ods;
    double en = 0.0;
    for (int pos = tid; pos < G.ne; pos += BLOCK) {
      int l[N];
      load_lconn<N>(LC + pos * N, l);
      E el;
      float4 xr[N];
#pragma unroll
      for (int a = 0; a < N; ++a) xr[a] = X[l[a]];
#pragma unroll
      for (int a = 0; a < N; ++a) {
        el.x[a][0] = xr[a].x - xr[0].x;
        if constexpr (D > 1) el.x[a][1] = xr[a].y - xr[0].y;
        if constexpr (D > 2) el.x[a][2] = xr[a].z - xr[0].z;
      }
      el.elemental = elemental;
      if (elemental) {
        const R dv = EF[pos];
        el.de = dv;
#pragma unroll
        for (int a = 0; a < N; ++a) el.d[a] = dv;
      } else {
        el.de = R(0);
#pragma unroll
        for (int a = 0; a < N; ++a) el.d[a] = FD[l[a]];
      }
#pragma unroll
      for (int a = 0; a < N; ++a)
#pragma unroll
        for (int c = 0; c < P; ++c) el.u[a * P + c] = U[c * cap64 + l[a]];
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
          for (int c = 0; c < P; ++c) w[a * P + c] = W[c * cap64 + l[a]];
        elem_vjp<R, ET, PH, NGP>(el, w, A.pp, picard, ro, gd);
      } else {
        R pi;
        elem_egrad<R, ET, PH, NGP>(el, A.pp, ro, gd, pi);
        en += (double)pi;
      }
      if (want_ou) store_corner_results<R, NL>(ER + pos * NL, ro);
      if constexpr (MODE != MODE_RES) {
        if (want_gd) {
          if (elemental) {
            const int ee = A.perm_identity ? G.e0 + pos : PERM[pos];
            if (accumulate)
              unsafeAtomicAdd(&ODb[ee], gd[0]);
            else
              ODb[ee] = gd[0];
          } else {
            store_corner_results<R, N>(ERD + pos * N, gd);
          }
        }
      }
    }
    if (A.chunk_energy) {
      en = wave_sum_to_last(en);
      if (lane == 63) red[wv] = en;
    }
    wait_all_memory();  // the next plan has landed (this wave's share)
    __syncthreads();    // ... every wave's share; the data buffer is free
    if (A.chunk_energy && tid == 0) {
      double s = 0.0;
#pragma unroll
      for (int k = 0; k < BLOCK / 64; ++k) s += red[k];
      A.chunk_energy[(long long)b * A.n_chunks + chunk] = s;
    }
    if (has_next) pipe_issue_data<P, MODE, BLOCK>(A, L, smem_raw, nxt, NG, NS.b, elemental, wv, lane);
    // one step further ahead: its chunk offsets stream into `geo` during the gather below
    const PipeStep NNS = next_step(A, NS);
    const bool has_nn = has_next && NNS.item < n_items;
    const bool nn_plan = has_nn && NNS.chunk != NS.chunk;
    if (nn_plan && wv == 0 && lane < 2) {
      dma4(A.elem_off + NNS.chunk + lane, geo);
      dma4(A.node_off + NNS.chunk + lane, geo + 2);
    }
    // ---- node-centric gather of the corner results (fixed order) and write-out ----
    if (want_inc) {
      const int *NID = reinterpret_cast<const int *>(smem_raw + (L.nid + cur * L.plan));
      const int *SLOT = reinterpret_cast<const int *>(smem_raw + (L.slot + cur * L.plan));
      const uint16_t *IST = reinterpret_cast<const uint16_t *>(smem_raw + (L.ist + cur * L.plan)) + (G.n0 & 1);
      const uint16_t *INC = reinterpret_cast<const uint16_t *>(smem_raw + (L.inc + cur * L.plan));
      R *OU = reinterpret_cast<R *>(A.ou) + (long long)b * A.ous;
      R *PU = reinterpret_cast<R *>(A.part_u) + (long long)b * A.n_slots * P;
      R *PD = reinterpret_cast<R *>(A.part_d) + (long long)b * A.n_slots;
      for (int i = tid; i < G.nl; i += BLOCK) {
        const int g = NID[i];
        const int sl = SLOT[i];
        const int js = IST[i];
        const int je = (i + 1 < G.nl) ? (int)IST[i + 1] : N * G.ne;
        if (want_ou) {
          R s[P];
#pragma unroll
          for (int c = 0; c < P; ++c) s[c] = R(-0.0);
          for (int j = js; j < je; ++j) {
            const R *e = ER + INC[j] * P;  // entry q = position * N + corner
#pragma unroll
            for (int c = 0; c < P; ++c) s[c] += e[c];
          }
#pragma unroll
          for (int c = 0; c < P; ++c) {
            if (sl >= 0) {
              PU[(long long)sl * P + c] = s[c];
            } else {
              const long long dof = (long long)g * P + c;
              bool masked = false;
              if (mask) masked = A.dmask[dof];
              if (accumulate) {
                if (!masked) unsafeAtomicAdd(OU + dof, s[c]);
              } else {
                OU[dof] = masked ? R(0) : s[c];
              }
            }
          }
        }
        if (nodal_gd) {
          R s = R(0);
          for (int j = js; j < je; ++j) s += ERD[INC[j]];
          if (sl >= 0)
            PD[sl] = s;
          else if (accumulate)
            unsafeAtomicAdd(&ODb[g], s);
          else
            ODb[g] = s;
        }
      }
    }
    wait_all_memory();  // the next step's data has landed
    __syncthreads();    // ... for every wave; this step's plan buffer and results are free
    if (!has_next) break;
    S = NS;
    G = NG;
    if (nn_plan) {
      NG.e0 = geo[0];
      NG.ne = geo[1] - NG.e0;
      NG.n0 = geo[2];
      NG.nl = geo[3] - NG.n0;
    }
    NS = NNS;
    has_next = has_nn;
    cur = nxt;  // `geo` is rewritten only after the next barrier, when every wave has read it
  }
}

template <int ET, int PH, int NGP, int MODE>
static int launch_pipe_block(int block, dim3 grid, size_t shm, hipStream_t s, const ChunkArgs &A) {
  if (block == 512) {
    hipLaunchKernelGGL((k_chunk_pipe<ET, PH, NGP, MODE, 512>), grid, dim3(512), shm, s, A);
  } else {
    hipLaunchKernelGGL((k_chunk_pipe<ET, PH, NGP, MODE, 256>), grid, dim3(256), shm, s, A);
  }
  FH_HIP(hipGetLastError());
  return FEMHOT_OK;
}

template <int ET>
int launch_chunk_pipe_et_impl(int ph, int ngp, int mode, int block, dim3 grid, size_t shm, hipStream_t s,
                              const ChunkArgs &A) {
  return dispatch_ph_ngp<float>(ph, ngp, [&](auto, auto phc, auto ngpc) -> int {
    constexpr int PH = decltype(phc)::value, NGP = decltype(ngpc)::value;
    if (mode == MODE_RES) return launch_pipe_block<ET, PH, NGP, MODE_RES>(block, grid, shm, s, A);
    if (mode == MODE_VJP) return launch_pipe_block<ET, PH, NGP, MODE_VJP>(block, grid, shm, s, A);
    return launch_pipe_block<ET, PH, NGP, MODE_EGRAD>(block, grid, shm, s, A);
  });
}

template <int ET>
int launch_chunk_pipe_et(int ph, int ngp, int mode, int block, dim3 grid, size_t shm, hipStream_t s,
                         const ChunkArgs &A);

}  // namespace femhot
