// Chunk-path kernel template and its launcher (instantiated per element type in
// femhot_chunk_<type>.hip so the build compiles them in parallel).
#pragma once
#include "femhot_launch.h"

namespace femhot {

struct ChunkArgs {
  const int *elem_off, *node_off, *perm, *nodes, *node_slot;
  const uint16_t *lconn, *inc, *inc_start;
  const uint16_t *rank;  // per corner: its position in the chunk's incidence list
  const void *xyz;  // chunk-relative coordinates in the arithmetic type
  int n_chunks, n_sblocks, sb, batch, cap, emax;
  int perm_identity;  // perm[p] == p: the original element id is the position
  // 1: row r of node q's run sits at r + q in ER / ERD (one spare row per node), so that the lanes of
  // a node-gather wave, whose runs all have about the same length, fall on different LDS banks (TET4)
  int pad;
  long long n_slots;
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
  void *part_u;          // (batch, n_slots, P) partials of shared nodes
  void *part_d;          // (batch, n_slots) nodal field-gradient partials
  double *chunk_energy;  // (batch, n_chunks)
  const uint8_t *dmask;
  unsigned flags;
  PhysParams pp;
};

template <int N, class UN> __device__ __forceinline__ void unpack_lconn(const UN &v, int (&l)[N]) {
  if constexpr (N == 4) {
    l[0] = v.x & 0xffff;
    l[1] = v.x >> 16;
    l[2] = v.y & 0xffff;
    l[3] = v.y >> 16;
  } else {
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

// Per-node LDS records are unpadded (3 coordinates, P solution values): the lanes of a wave read
// records of randomly indexed nodes, and a 12-byte stride spreads them over every bank where a
// 16-byte one would map node l and node l + 16 to the same banks.

// An element's K per-corner results (K a multiple of 16 bytes) written with 16-byte LDS stores.
// Consecutive lanes hold consecutive elements, so with K * sizeof(R) = 48 / 96 bytes each group
// of 8 lanes covers 8 distinct 4-bank groups.
template <class R, int K> __device__ __forceinline__ void store_corner_results(R *dst, const R (&v)[K]) {
  static_assert((K * sizeof(R)) % 16 == 0, "per-element record must be whole 16-byte units");
  if constexpr (sizeof(R) == 4) {
#pragma unroll
    for (int i = 0; i < K; i += 4)
      reinterpret_cast<float4 *>(dst)[i / 4] = make_float4(v[i], v[i + 1], v[i + 2], v[i + 3]);
  } else {
#pragma unroll
    for (int i = 0; i < K; i += 2) reinterpret_cast<double2 *>(dst)[i / 2] = make_double2(v[i], v[i + 1]);
  }
}

// LDS layout of k_chunk (in units of R unless noted; cap4 = max chunk nodes rounded up to 4,
// emax = max chunk elements):
//   X   [cap4 * 3]      chunk-relative coordinates (x, y, z)
//   U   [cap4 * P]      nodal solution
//   W   [cap4 * P]      cotangent (VJP only)
//   FD  [cap4]          nodal field (nodal fields only)
//   ER  [(emax * N + cap4 * pad) * P]  per element corner: residual / g_u, at [incidence-list position
//                       (the plan's rank of the corner) + pad * chunk node][component], so each node's
//                       corners are one contiguous run
//   ERD [emax * N + cap4 * pad]  per element corner, at the same row: nodal field-gradient contribution
// Waves per SIMD the compiler must fit: 4 for FP32 TET4 thermal and linear elasticity at 256 threads
// (their LDS holds 4 workgroups of 512-tet chunks per CU, so more registers would only cost waves),
// else its own choice.
template <class R, int ET, int PH, int BLOCK> constexpr int chunk_min_waves() {
  return (sizeof(R) == 4 && ET == TET4 && PH != NEOHOOKE && BLOCK == 256) ? 4 : 1;
}

template <class R, int ET, int PH, int NGP, int MODE, int BLOCK>
__global__ __launch_bounds__(BLOCK, (chunk_min_waves<R, ET, PH, BLOCK>())) void k_chunk(ChunkArgs A) {
  using E = Elem<R, ET, PH>;
  constexpr int N = E::N, D = E::D, P = E::P, NL = E::NL;
  // nodes / elements per lane per round of loads: enough that one round covers 512 nodes and
  // 1024 elements (a 1024-element chunk's nodes and elements) without idle registers at 512 threads
  // (TET4 at 256 threads: 512 elements per round, as below)
  constexpr int KN = BLOCK >= 512 ? 1 : (BLOCK == 256 ? 2 : 4);
  // Light element routines (several elements per lane per round) keep the first round of node
  // metadata, element connectivity and per-element field in registers from the plan load on; heavy
  // ones (HEX8, high-order rules) need those registers for the element arithmetic and reload them
  // where used.  At 256 threads TET4 takes 2 elements per lane and the prefetch (119 VGPRs, still 4
  // waves per SIMD: 2.5 us faster on the uneven TET4 batch), QUAD4 4 and none (the prefetch measured
  // 1 us slower on the QUAD4 thermal batch).
  constexpr bool PF256 = ET == TET4 && BLOCK == 256;
  constexpr int KE = (ET != HEX8 && NGP <= 2) ? ((BLOCK >= 512 || PF256) ? 2 : 4) : 1;
  constexpr bool PF = KE > 1 && (BLOCK >= 512 || PF256);
  extern __shared__ __align__(16) unsigned char smem_raw[];
  __shared__ double red[BLOCK / 64];
  const bool elemental = (A.flags & FEMHOT_ELEMENT_FIELD) != 0;
  const bool mask = (A.flags & FEMHOT_MASK_DIRICHLET) != 0;
  const bool accumulate = (A.flags & FEMHOT_ACCUMULATE) != 0;
  const bool picard = (A.flags & FEMHOT_PICARD) != 0;
  const bool want_gd = (MODE != MODE_RES) && A.od != nullptr;
  const bool want_ou = A.ou != nullptr;
  const bool nodal_gd = want_gd && !elemental;
  // Corner results go to ER / ERD at their position in the chunk's incidence list (the plan's
  // per-corner rank), so a node sums a contiguous run and the gather reads no list.
  const bool by_rank = want_ou || nodal_gd;
  const int cap4 = (A.cap + 3) & ~3, emax = A.emax;
  R *X = reinterpret_cast<R *>(smem_raw);
  R *U = X + 3 * cap4;
  R *W = U + P * cap4;
  R *FD = W + (MODE == MODE_VJP ? P * cap4 : 0);
  R *ER = FD + (elemental ? 0 : cap4);
  // (TET4 only: its interior nodes' runs of 24 corners x 3 components put a wave's lanes 72 words
  // apart, on 4 of 32 banks; QUAD4 and HEX8 measured no gain from the spare rows)
  const int pad = ET == TET4 ? A.pad : 0;
  R *ERD = ER + (long)(emax * N + cap4 * pad) * P;
  const int tid = threadIdx.x;
  const int n_items = A.n_chunks * A.n_sblocks;
  using UN = typename std::conditional<N == 4, uint2, uint4>::type;  // one element's N uint16
  // Workgroups go to the 8 XCDs in turn; neighbouring chunks share nodes (their solution gathers, and
  // the partial slots of their shared nodes, which the plan numbers in node order, fall on the same
  // lines), so XCD x walks the x-th eighth of the items in order and finds those lines in its own L2.  Items [i0, i1) of XCD x are split over
  // its nx workgroups with stride nx (every item is taken exactly once).
  // (Grids of fewer than 8 workgroups walk all items with the plain stride.)
  const bool xmap = gridDim.x >= 8;
  const int xcd = blockIdx.x % 8, nx = gridDim.x / 8 + (xcd < (int)(gridDim.x % 8) ? 1 : 0);
  const int ipx = n_items / 8, irem = n_items % 8;
  const int i0 = xmap ? xcd * ipx + min(xcd, irem) : 0;
  const int i1 = xmap ? i0 + ipx + (xcd < irem ? 1 : 0) : n_items;
  const int istride = xmap ? nx : (int)gridDim.x;
  for (int item = i0 + (int)(xmap ? blockIdx.x / 8 : blockIdx.x); item < i1; item += istride) {
    const int chunk = item / A.n_sblocks;
    const int sbi = item - chunk * A.n_sblocks;
    const int e0 = A.elem_off[chunk], e1 = A.elem_off[chunk + 1];
    const int ne = e1 - e0;
    const int n0 = A.node_off[chunk], nl = A.node_off[chunk + 1] - n0;
    // The first round of this lane's nodes: id, partial slot and incidence range, loaded with the
    // plan and kept in registers for the solution gather and the write-out (indices clamped, so
    // the loads of a round issue back to back).
    int g0[KN], sl0[KN], js0[KN], je0[KN];
#pragma unroll
    for (int k = 0; k < KN; ++k) {
      const int i = min(k * BLOCK + tid, nl - 1);
      g0[k] = A.nodes[n0 + i];
      sl0[k] = A.node_slot[n0 + i];
      js0[k] = A.inc_start[n0 + i];
      je0[k] = (i + 1 < nl) ? (int)A.inc_start[n0 + i + 1] : N * ne;
    }
    {  // coordinates (one (x, y, z) record per node)
      const R *xyz = reinterpret_cast<const R *>(A.xyz) + 3LL * n0;
      for (int i = tid; i < nl; i += BLOCK) {
        const R x = xyz[3 * i], y = xyz[3 * i + 1], z = xyz[3 * i + 2];
        X[3 * i] = x;
        X[3 * i + 1] = y;
        X[3 * i + 2] = z;
      }
    }
    const int b0 = sbi * A.sb, b1 = min(A.batch, b0 + A.sb);
    for (int b = b0; b < b1; ++b) {
      const R *Ub = reinterpret_cast<const R *>(A.u) + (long long)b * A.us;
      const R *Wb = reinterpret_cast<const R *>(A.w) + (long long)b * A.ws;
      const R *Fb = reinterpret_cast<const R *>(A.field) + (long long)b * A.fs;
      R uv[KN][P], wv[KN][P], fv[KN];
      // stage nodal solution / cotangent / field of one round of nodes into LDS
      auto store_round = [&](int base) {
#pragma unroll
        for (int k = 0; k < KN; ++k) {
          const int i = base + k * BLOCK + tid;
          if (i < nl) {
#pragma unroll
            for (int c = 0; c < P; ++c) {
              U[i * P + c] = uv[k][c];
              if constexpr (MODE == MODE_VJP) W[i * P + c] = wv[k][c];
            }
            if (!elemental) FD[i] = fv[k];
          }
        }
      };
      auto load_round = [&](const int (&g)[KN]) {
#pragma unroll
        for (int k = 0; k < KN; ++k) {
#pragma unroll
          for (int c = 0; c < P; ++c) {
            uv[k][c] = Ub[(long long)g[k] * P + c];
            if constexpr (MODE == MODE_VJP) wv[k][c] = Wb[(long long)g[k] * P + c];
          }
          fv[k] = elemental ? R(0) : Fb[g[k]];
        }
      };
      load_round(g0);
      // the first round of elements' local connectivity, original ids and per-element field, in
      // flight together with the solution gather above
      UN lw0[KE], rk0[KE];
      int ee0[KE];
      R dev0[KE];
#pragma unroll
      for (int k = 0; k < KE; ++k) {
        const int p = min(e0 + k * BLOCK + tid, e1 - 1);
        lw0[k] = *reinterpret_cast<const UN *>(A.lconn + (long long)p * N);
        if (by_rank) rk0[k] = *reinterpret_cast<const UN *>(A.rank + (long long)p * N);
        ee0[k] = A.perm_identity ? p : A.perm[p];
      }
#pragma unroll
      for (int k = 0; k < KE; ++k) dev0[k] = R(0);
      if (elemental) {
#pragma unroll
        for (int k = 0; k < KE; ++k) dev0[k] = Fb[ee0[k]];
      }
      store_round(0);
      for (int base = KN * BLOCK; base < nl; base += KN * BLOCK) {  // chunks of more nodes
        int g[KN];
#pragma unroll
        for (int k = 0; k < KN; ++k) g[k] = A.nodes[n0 + min(base + k * BLOCK + tid, nl - 1)];
        load_round(g);
        store_round(base);
      }
      __syncthreads();
      double en = 0.0;
      R *ODb = reinterpret_cast<R *>(A.od) + (long long)b * A.ods;
      for (int base = e0; base < e1; base += KE * BLOCK) {
        int l[KE][N], rk[KE][N], ee[KE];
        R dev[KE];
        if (PF && base == e0) {
#pragma unroll
          for (int k = 0; k < KE; ++k) {
            unpack_lconn<N>(lw0[k], l[k]);
            if (by_rank) unpack_lconn<N>(rk0[k], rk[k]);
            ee[k] = ee0[k];
            dev[k] = dev0[k];
          }
        } else {
#pragma unroll
          for (int k = 0; k < KE; ++k) {
            const int p = min(base + k * BLOCK + tid, e1 - 1);
            load_lconn<N>(A.lconn + (long long)p * N, l[k]);
            if (by_rank) load_lconn<N>(A.rank + (long long)p * N, rk[k]);
            ee[k] = A.perm_identity ? p : A.perm[p];
          }
          if (elemental) {
#pragma unroll
            for (int k = 0; k < KE; ++k) dev[k] = Fb[ee[k]];
          }
        }
#pragma unroll
        for (int k = 0; k < KE; ++k) {
          const int pos = base + k * BLOCK + tid - e0;
          if (pos >= ne) continue;
          E el;
          R xr[N][3];
#pragma unroll
          for (int a = 0; a < N; ++a)
#pragma unroll
            for (int i = 0; i < 3; ++i) xr[a][i] = X[3 * l[k][a] + i];
#pragma unroll
          for (int a = 0; a < N; ++a)
#pragma unroll
            for (int i = 0; i < D; ++i) el.x[a][i] = xr[a][i] - xr[0][i];
          el.elemental = elemental;
          if (elemental) {
            el.de = dev[k];
#pragma unroll
            for (int a = 0; a < N; ++a) el.d[a] = dev[k];
          } else {
            el.de = R(0);
#pragma unroll
            for (int a = 0; a < N; ++a) el.d[a] = FD[l[k][a]];
          }
#pragma unroll
          for (int a = 0; a < N; ++a)
#pragma unroll
            for (int c = 0; c < P; ++c) el.u[a * P + c] = U[l[k][a] * P + c];
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
              for (int c = 0; c < P; ++c) w[a * P + c] = W[l[k][a] * P + c];
            elem_vjp<R, ET, PH, NGP>(el, w, A.pp, picard, ro, gd);
          } else {
            R pi;
            elem_egrad<R, ET, PH, NGP>(el, A.pp, ro, gd, pi);
            en += (double)pi;
          }
          if (want_ou) {  // each corner at its place in the incidence list
#pragma unroll
            for (int a = 0; a < N; ++a)
#pragma unroll
              for (int c = 0; c < P; ++c) ER[(rk[k][a] + pad * l[k][a]) * P + c] = ro[a * P + c];
          }
          if constexpr (MODE != MODE_RES) {
            if (want_gd) {
              if (elemental) {
                if (accumulate)
                  unsafeAtomicAdd(&ODb[ee[k]], gd[0]);  // RMW the compiler cannot hoist as a load
                else
                  ODb[ee[k]] = gd[0];
              } else {
#pragma unroll
                for (int a = 0; a < N; ++a) ERD[rk[k][a] + pad * l[k][a]] = gd[a];
              }
            }
          }
        }
      }
      if (A.chunk_energy) {
        en = wave_sum_to_last(en);
        if ((tid & 63) == 63) red[tid >> 6] = en;
      }
      __syncthreads();
      if (A.chunk_energy && tid == 0) {
        double s = 0.0;
#pragma unroll
        for (int k = 0; k < BLOCK / 64; ++k) s += red[k];
        A.chunk_energy[(long long)b * A.n_chunks + chunk] = s;
      }
      // node-centric gather of the corner results (fixed order) and write-out
      R *OU = reinterpret_cast<R *>(A.ou) + (long long)b * A.ous;
      R *PU = reinterpret_cast<R *>(A.part_u) + (long long)b * A.n_slots * P;
      R *PD = reinterpret_cast<R *>(A.part_d) + (long long)b * A.n_slots;
      if (want_ou || nodal_gd) {
        for (int base = 0; base < nl; base += KN * BLOCK) {
          int g[KN], sl[KN], js[KN], je[KN];
          if (PF && base == 0) {
#pragma unroll
            for (int k = 0; k < KN; ++k) g[k] = g0[k], sl[k] = sl0[k], js[k] = js0[k], je[k] = je0[k];
          } else {
#pragma unroll
            for (int k = 0; k < KN; ++k) {
              const int i = min(base + k * BLOCK + tid, nl - 1);
              g[k] = A.nodes[n0 + i];
              sl[k] = A.node_slot[n0 + i];
              js[k] = A.inc_start[n0 + i];
              je[k] = (i + 1 < nl) ? (int)A.inc_start[n0 + i + 1] : N * ne;
            }
          }
#pragma unroll
          for (int k = 0; k < KN; ++k) {
            const int i = base + k * BLOCK + tid;
            if (i >= nl) continue;
            if (want_ou) {
              R s[P];
#pragma unroll
              for (int c = 0; c < P; ++c) s[c] = R(-0.0);
              // the node's corner results lie at its list positions js .. je - 1 (+ i with the spare rows)
              for (int j = js[k] + pad * i; j < je[k] + pad * i; ++j) {
                const R *e = ER + j * P;
#pragma unroll
                for (int c = 0; c < P; ++c) s[c] += e[c];
              }
              if (sl[k] >= 0) {  // a shared node: its partial slot, the P values stored back to back
                R *pu = PU + (long long)sl[k] * P;
#pragma unroll
                for (int c = 0; c < P; ++c) pu[c] = s[c];
              } else {
#pragma unroll
                for (int c = 0; c < P; ++c) {
                  const long long dof = (long long)g[k] * P + c;
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
              for (int j = js[k] + pad * i; j < je[k] + pad * i; ++j) s += ERD[j];
              if (sl[k] >= 0)
                PD[sl[k]] = s;
              else if (accumulate)
                unsafeAtomicAdd(&ODb[g[k]], s);
              else
                ODb[g[k]] = s;
            }
          }
        }
      }
      // No barrier here, so a wave done with its gather goes on to the next sample's or item's loads.
      // That staging writes X, U, W and FD only, whose last reads (the element phase) precede the
      // barrier above; ER and ERD are rewritten only after the next staging barrier, which every
      // wave reaches after its gather; `red` is rewritten after that barrier too.
    }
  }
}

template <class R, int ET, int PH, int NGP, int MODE>
static int launch_chunk_block(int block, dim3 grid, size_t shm, hipStream_t s, const ChunkArgs &A) {
  if (block == 128) {
    auto k = k_chunk<R, ET, PH, NGP, MODE, 128>;
    hipLaunchKernelGGL(k, grid, dim3(128), shm, s, A);
  } else if (block == 512) {
    auto k = k_chunk<R, ET, PH, NGP, MODE, 512>;
    hipLaunchKernelGGL(k, grid, dim3(512), shm, s, A);
  } else {
    auto k = k_chunk<R, ET, PH, NGP, MODE, 256>;
    hipLaunchKernelGGL(k, grid, dim3(256), shm, s, A);
  }
  FH_HIP(hipGetLastError());
  return FEMHOT_OK;
}

template <int ET>
int launch_chunk_et_impl(int dtype, int ph, int ngp, int mode, int block, dim3 grid, size_t shm, hipStream_t s,
                           const ChunkArgs &A) {
  return dispatch_r_ph_ngp(dtype, ph, ngp, [&](auto rp, auto phc, auto ngpc) -> int {
    using R = std::remove_pointer_t<decltype(rp)>;
    constexpr int PH = decltype(phc)::value, NGP = decltype(ngpc)::value;
    if (mode == MODE_RES) return launch_chunk_block<R, ET, PH, NGP, MODE_RES>(block, grid, shm, s, A);
    if (mode == MODE_VJP) return launch_chunk_block<R, ET, PH, NGP, MODE_VJP>(block, grid, shm, s, A);
    return launch_chunk_block<R, ET, PH, NGP, MODE_EGRAD>(block, grid, shm, s, A);
  });
}

template <int ET>
int launch_chunk_et(int dtype, int ph, int ngp, int mode, int block, dim3 grid, size_t shm, hipStream_t s,
                    const ChunkArgs &A);

}  // namespace femhot
