// Element-parallel kernels ("plain" path): one lane per (element, sample), gather straight from
// HBM through the int32 connectivity, element routine in registers, scatter with hardware float
// atomics (global_atomic_add_f32/f64) or, for one colour class at a time, with plain
// read-modify-write (no two elements of a colour share a node).
#pragma once
#include "femhot_internal.h"
#include "femhot_physics.h"

namespace femhot {

enum { MODE_RES = 0, MODE_VJP = 1, MODE_EGRAD = 2 };

struct KArgs {
  const double *coords;  // (Nn*3) FP64 (edge vectors are formed in FP64, then rounded once)
  const int *conn;       // (Ne*n)
  long long Ne, Nn;
  const int *list;       // optional element list (colour class), else elements [0, count)
  long long count;
  const void *field;
  long long fstride;
  const void *u;
  long long ustride;
  const void *w;
  long long wstride;
  void *out_u;  // residual or g_u
  long long oustride;
  void *out_d;  // g_field
  long long odstride;
  double *energy;  // [b * nseg + seg]
  const uint8_t *dmask;
  const long long *seg_off;
  int nseg;
  unsigned flags;
  PhysParams pp;
};

template <class T> __device__ __forceinline__ void add_out(T *p, T v, bool plain) {
  if (plain)
    *p += v;
  else
    unsafeAtomicAdd(p, v);
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off);
  return v;
}

// v of the lane DPP control CTRL names, rows outside ROWS (a 4-bit row mask) reading 0.
template <int CTRL, int ROWS> __device__ __forceinline__ double dpp_f64(double v) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, ROWS, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, ROWS, 0xf, false);
  return __hiloint2double(hi, lo);
}

// Sum over the 64 lanes of a wave, valid in lane 63 only: pairs and quads (quad_perm), half rows and
// rows (row_half_mirror, row_mirror), then row 0 into row 1 and row 2 into row 3 (row_bcast:15), then
// rows 0-1 into rows 2-3 (row_bcast:31).  Register moves instead of the LDS round trips of a shuffle
// butterfly; a fixed order, so the sum is the same on every run.
__device__ __forceinline__ double wave_sum_to_last(double v) {
  v += dpp_f64<0xB1, 0xf>(v);   // quad_perm [1,0,3,2]
  v += dpp_f64<0x4E, 0xf>(v);   // quad_perm [2,3,0,1]
  v += dpp_f64<0x141, 0xf>(v);  // row_half_mirror
  v += dpp_f64<0x140, 0xf>(v);  // row_mirror
  v += dpp_f64<0x142, 0xa>(v);  // row_bcast:15 into rows 1 and 3
  v += dpp_f64<0x143, 0xc>(v);  // row_bcast:31 into rows 2 and 3
  return v;
}

// Sum `v` into energy[idx] with one atomic per wave when all active lanes share idx.
__device__ __forceinline__ void add_energy(double *energy, int idx, double v, bool active) {
  const unsigned long long act = __ballot(active);
  if (act == 0ULL) return;
  const int leader = __ffsll((long long)act) - 1;
  const int idx0 = __shfl(idx, leader);
  const bool same = __all(!active || idx == idx0);
  if (same) {
    const double s = wave_sum(active ? v : 0.0);
    if ((int)(threadIdx.x & 63) == leader) unsafeAtomicAdd(&energy[idx0], s);
  } else if (active) {
    unsafeAtomicAdd(&energy[idx], v);
  }
}

__device__ __forceinline__ int segment_of(const long long *off, int nseg, long long e) {
  int lo = 0, hi = nseg - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (off[mid] <= e)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

template <int N> __device__ __forceinline__ void load_conn(const int *conn, long long e, int (&nodes)[N]) {
  const int4 *p = reinterpret_cast<const int4 *>(conn + e * N);
  const int4 v0 = p[0];
  nodes[0] = v0.x;
  nodes[1] = v0.y;
  nodes[2] = v0.z;
  nodes[3] = v0.w;
  if constexpr (N == 8) {
    const int4 v1 = p[1];
    nodes[4] = v1.x;
    nodes[5] = v1.y;
    nodes[6] = v1.z;
    nodes[7] = v1.w;
  }
}

template <class R, int ET, int PH>
__device__ __forceinline__ void load_geometry(const double *coords, const int (&nodes)[ElemT<ET>::N],
                                              Elem<R, ET, PH> &el) {
  constexpr int N = ElemT<ET>::N, D = ElemT<ET>::D;
  double x0[D];
#pragma unroll
  for (int i = 0; i < D; ++i) x0[i] = coords[(size_t)nodes[0] * 3 + i];
#pragma unroll
  for (int a = 0; a < N; ++a)
#pragma unroll
    for (int i = 0; i < D; ++i) el.x[a][i] = R(coords[(size_t)nodes[a] * 3 + i] - x0[i]);
}

template <class R, int ET, int PH>
__device__ __forceinline__ void load_fields(const KArgs &A, long long e, int b, const int (&nodes)[ElemT<ET>::N],
                                            Elem<R, ET, PH> &el) {
  using E = Elem<R, ET, PH>;
  const R *F = reinterpret_cast<const R *>(A.field) + (long long)b * A.fstride;
  el.elemental = (A.flags & FEMHOT_ELEMENT_FIELD) != 0;
  if (el.elemental) {
    el.de = F[e];
#pragma unroll
    for (int a = 0; a < E::N; ++a) el.d[a] = el.de;
  } else {
    el.de = R(0);
#pragma unroll
    for (int a = 0; a < E::N; ++a) el.d[a] = F[nodes[a]];
  }
  const R *U = reinterpret_cast<const R *>(A.u) + (long long)b * A.ustride;
#pragma unroll
  for (int a = 0; a < E::N; ++a)
#pragma unroll
    for (int c = 0; c < E::P; ++c) el.u[a * E::P + c] = U[(size_t)nodes[a] * E::P + c];
}

template <class R, int ET, int PH, int NGP, int MODE, bool PLAIN>
__global__ __launch_bounds__(256) void k_elem(KArgs A) {
  using E = Elem<R, ET, PH>;
  constexpr int N = E::N, P = E::P, NL = E::NL;
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int b = blockIdx.y;
  const bool active = i < A.count;
  double en = 0.0;
  int seg = 0;
  if (active) {
    const long long e = A.list ? (long long)A.list[i] : i;
    int nodes[N];
    load_conn<N>(A.conn, e, nodes);
    E el;
    load_geometry<R, ET, PH>(A.coords, nodes, el);
    load_fields<R, ET, PH>(A, e, b, nodes, el);
    const bool mask = (A.flags & FEMHOT_MASK_DIRICHLET) != 0;
    R *OU = reinterpret_cast<R *>(A.out_u);
    R *OD = reinterpret_cast<R *>(A.out_d);
    if constexpr (MODE == MODE_RES) {
      R r[NL], energy;
      elem_residual<R, ET, PH, NGP>(el, A.pp, r, energy);
      en = (double)energy;
      if (OU) {
        OU += (long long)b * A.oustride;
#pragma unroll
        for (int a = 0; a < N; ++a)
#pragma unroll
          for (int c = 0; c < P; ++c) {
            const long long dof = (long long)nodes[a] * P + c;
            if (!(mask && A.dmask[dof])) add_out<R>(OU + dof, r[a * P + c], PLAIN);
          }
      }
    } else if constexpr (MODE == MODE_VJP) {
      const R *W = reinterpret_cast<const R *>(A.w) + (long long)b * A.wstride;
      R w[NL];
#pragma unroll
      for (int a = 0; a < N; ++a)
#pragma unroll
        for (int c = 0; c < P; ++c) w[a * P + c] = W[(size_t)nodes[a] * P + c];
      R gu[NL], gd[N];
      elem_vjp<R, ET, PH, NGP>(el, w, A.pp, (A.flags & FEMHOT_PICARD) != 0, gu, gd);
      if (OU) {
        OU += (long long)b * A.oustride;
#pragma unroll
        for (int a = 0; a < N; ++a)
#pragma unroll
          for (int c = 0; c < P; ++c) {
            const long long dof = (long long)nodes[a] * P + c;
            if (!(mask && A.dmask[dof])) add_out<R>(OU + dof, gu[a * P + c], PLAIN);
          }
      }
      if (OD) {
        OD += (long long)b * A.odstride;
        if (el.elemental)
          OD[e] = (A.flags & FEMHOT_ACCUMULATE) ? OD[e] + gd[0] : gd[0];
        else {
#pragma unroll
          for (int a = 0; a < N; ++a) add_out<R>(OD + nodes[a], gd[a], PLAIN);
        }
      }
    } else {
      R r[NL], gd[N], pi;
      elem_egrad<R, ET, PH, NGP>(el, A.pp, r, gd, pi);
      en = (double)pi;
      if (OU) {
        OU += (long long)b * A.oustride;
#pragma unroll
        for (int a = 0; a < N; ++a)
#pragma unroll
          for (int c = 0; c < P; ++c) {
            const long long dof = (long long)nodes[a] * P + c;
            if (!(mask && A.dmask[dof])) add_out<R>(OU + dof, r[a * P + c], PLAIN);
          }
      }
      if (OD) {
        OD += (long long)b * A.odstride;
        if (el.elemental)
          OD[e] = (A.flags & FEMHOT_ACCUMULATE) ? OD[e] + gd[0] : gd[0];
        else {
#pragma unroll
          for (int a = 0; a < N; ++a) add_out<R>(OD + nodes[a], gd[a], PLAIN);
        }
      }
    }
    if (A.nseg > 1) seg = segment_of(A.seg_off, A.nseg, e);
  }
  if (A.energy) add_energy(A.energy, b * A.nseg + seg, en, active);
}

struct TArgs {
  const double *coords;
  const int *conn;
  const int *list;
  long long count;
  const void *field;
  const void *u;
  void *vals;
  const long long *row_ptr;
  const uint8_t *slot;
  const uint8_t *dmask;
  const int *ne_off, *ne_list;  // row-owner assembly: node -> (element * n + local index)
  long long n_nodes;
  unsigned flags;
  PhysParams pp;
};

// One lane per (element, local node a, component c): row (a, c) of the element tangent, BC-masked as
// fe_loss.py:191-207 (Dirichlet rows keep only the element diagonal), added into the CSR values at
// row_ptr[row] + P*slot(e,a,b) + d.
template <class R, int ET, int PH, int NGP, bool PLAIN>
__global__ __launch_bounds__(256) void k_tangent(TArgs A) {
  using E = Elem<R, ET, PH>;
  constexpr int N = E::N, P = E::P, NL = E::NL;
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= A.count * N * P) return;
  const long long ei = i / (N * P);
  const int ac = (int)(i - ei * (N * P));
  const int a = ac / P, c = ac - a * P;
  const long long e = A.list ? (long long)A.list[ei] : ei;
  int nodes[N];
  load_conn<N>(A.conn, e, nodes);
  E el;
  load_geometry<R, ET, PH>(A.coords, nodes, el);
  el.elemental = (A.flags & FEMHOT_ELEMENT_FIELD) != 0;
  const R *F = reinterpret_cast<const R *>(A.field);
  const R *U = reinterpret_cast<const R *>(A.u);
  if (el.elemental) {
    el.de = F[e];
#pragma unroll
    for (int b = 0; b < N; ++b) el.d[b] = el.de;
  } else {
    el.de = R(0);
#pragma unroll
    for (int b = 0; b < N; ++b) el.d[b] = F[nodes[b]];
  }
  if constexpr (PH != MECH) {  // the linear-elastic tangent does not depend on u
#pragma unroll
    for (int b = 0; b < N; ++b)
#pragma unroll
      for (int d = 0; d < P; ++d) el.u[b * P + d] = U[(size_t)nodes[b] * P + d];
  }
  R K[NL];
  elem_tangent_row<R, ET, PH, NGP>(el, A.pp, a, c, K);
  int na = nodes[0];
#pragma unroll
  for (int b = 1; b < N; ++b) na = (b == a) ? nodes[b] : na;
  const uint8_t *sl = A.slot + ((size_t)e * N + a) * N;
  uint8_t ks[N];
#pragma unroll
  for (int b = 0; b < N; ++b) ks[b] = sl[b];
  R *V = reinterpret_cast<R *>(A.vals);
  const long long row = (long long)na * P + c;
  const long long rp = A.row_ptr[row];
  if (A.dmask[row]) {
    R kd = R(0);
    int kaa = 0;
#pragma unroll
    for (int b = 0; b < N; ++b) {
#pragma unroll
      for (int d = 0; d < P; ++d) kd = (b == a && d == c) ? K[b * P + d] : kd;  // selects: no dynamic index
      kaa = (b == a) ? ks[b] : kaa;
    }
    add_out<R>(V + rp + (long long)P * kaa + c, kd, PLAIN);
  } else {
#pragma unroll
    for (int b = 0; b < N; ++b)
#pragma unroll
      for (int d = 0; d < P; ++d) add_out<R>(V + rp + (long long)P * ks[b] + d, K[b * P + d], PLAIN);
  }
}

template <class T> __global__ void k_fill(T *p, long long n, T v) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long k = i; k < n; k += stride) p[k] = v;
}

// strided 2-D fill (batch rows of `n` entries with stride `ld`)
template <class T> __global__ void k_fill2(T *p, long long n, long long ld, int rows, T v) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long k = i; k < n * rows; k += stride) {
    const long long r = k / n, j = k - r * n;
    p[r * ld + j] = v;
  }
}

template <class T> inline int fill(T *p, long long n, long long ld, int rows, T v, hipStream_t s) {
  if (!p || n <= 0 || rows <= 0) return FEMHOT_OK;
  const long long tot = n * rows;
  const int blocks = (int)std::min<long long>((tot + 255) / 256, 8192);
  if (ld == n || rows == 1)
    hipLaunchKernelGGL(k_fill<T>, dim3(blocks), dim3(256), 0, s, p, tot, v);
  else
    hipLaunchKernelGGL(k_fill2<T>, dim3(blocks), dim3(256), 0, s, p, n, ld, rows, v);
  FH_HIP(hipGetLastError());
  return FEMHOT_OK;
}

}  // namespace femhot
