// Jacobi-preconditioned conjugate gradients on the assembled CSR tangent.
//
// Reference: fol/solvers/fe_solver.py:60-78 (LinearSolve: BiCGSTAB / direct solve of K du = -R)
// and fe_linear_residual_based_solver.py:15-22.  The reference has no CG; with the reference's
// BC treatment (Dirichlet rows zeroed except the diagonal, R_D = 0) the Dirichlet unknowns stay
// zero and the iteration runs on the SPD free block, so CG applies to the matrix as assembled.
//
// Design: an iteration is three stream-ordered kernels with no host round trip; the host reads the
// residual norm only every `check` iterations.  Dot products are reduced without atomics: every
// reducing kernel runs a fixed grid of NBR workgroups (grid-stride loops) and writes one FP64
// partial per workgroup; the kernel that needs the scalar sums the NBR partials itself, in a fixed
// order, in its prologue.  (One device-scope atomic per wave onto a single address serialises: on
// an 823,875-row system that alone cost ~10 ms per iteration.)  Results are bitwise reproducible.
// FP32 mode keeps the matrix and vectors in FP32, accumulates every dot product in FP64, and wraps
// the FP32 solves in FP64 iterative refinement r = b - A x with the FP64 matrix.
#include <cmath>

#include "femhot_kernels.h"

namespace femhot {

constexpr int NBR = 1024;  // workgroups of every reducing kernel (one FP64 partial each)

// Sum over the workgroup; the result is returned to every thread (blockDim.x == 256).
__device__ __forceinline__ double block_sum256(double v) {
  __shared__ double red[4];
  v = wave_sum(v);
  __syncthreads();  // red may still be read by a previous call
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  return (red[0] + red[1]) + (red[2] + red[3]);
}

// Sum of the NBR partials, in a fixed order, returned to every thread.
__device__ __forceinline__ double sum_partials(const double *part) {
  double v = 0.0;
#pragma unroll
  for (int k = 0; k < NBR / 256; ++k) v += part[k * 256 + threadIdx.x];
  return block_sum256(v);
}

// y = A x, LPR lanes per row, one pass (grid = rows * LPR / 256).  Each lane first loads up to U = 8
// of its row's (column, value) pairs, then gathers x for all of them, so a row costs three dependent
// memory round trips whatever its length up to 8 * LPR (longer rows continue in a loop).  When part
// is set, part[block] = this workgroup's share of x.y (FP64).
template <class T, int LPR>
__global__ __launch_bounds__(256) void k_spmv(long long nrows, const long long *__restrict__ rp,
                                              const int *__restrict__ ci, const T *__restrict__ v,
                                              const T *__restrict__ x, T *__restrict__ y, double *part,
                                              const int *stop) {
  if (stop && *stop) return;  // a device-side solve that has stopped (uniform over the grid)
  constexpr int U = 8;
  const long long row = ((long long)blockIdx.x * 256 + threadIdx.x) / LPR;
  const int lane = (int)(threadIdx.x % LPR);
  double d = 0.0;
  if (row < nrows) {
    const long long b = rp[row], e = rp[row + 1];
    int cols[U];
    T vals[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const long long k = b + lane + (long long)u * LPR;
      const bool in = k < e;
      cols[u] = in ? ci[k] : (int)row;
      vals[u] = in ? v[k] : T(0);
    }
    // accumulated in the arithmetic type (FP32 only in the mixed solve's inner iterations)
    T s = T(0);
#pragma unroll
    for (int u = 0; u < U; ++u) s += vals[u] * x[cols[u]];
    for (long long k = b + lane + (long long)U * LPR; k < e; k += LPR) s += v[k] * x[ci[k]];
#pragma unroll
    for (int off = LPR / 2; off >= 1; off >>= 1) s += __shfl_xor(s, off, LPR);
    if (lane == 0) {
      y[row] = s;
      d = (double)x[row] * (double)s;
    }
  }
  if (part) {
    const double t = block_sum256(d);
    if (threadIdx.x == 0) part[blockIdx.x] = t;
  }
}

// y = A x on a node-blocked pattern: P dofs per node, and every dof row of node i lists the same
// neighbours m, each as the P consecutive columns P*m .. P*m+P-1 (the pattern ensure_pattern builds).
// LPN lanes per node; a lane takes every LPN-th neighbour and reads its column index once for the
// whole P x P block (the other P*P-1 indices of the block are implied), then the P x values of the
// neighbour as one contiguous block.  Index traffic drops from 4 to 4/(P*P) bytes per nonzero.
// When part is set, part[block] = this workgroup's share of x.y (FP64).
template <class T, int P, int LPN>
__global__ __launch_bounds__(256) void k_spmv_node(long long nnodes, const long long *__restrict__ rp,
                                                   const int *__restrict__ ci, const T *__restrict__ v,
                                                   const T *__restrict__ x, T *__restrict__ y, double *part,
                                                   const int *stop) {
  if (stop && *stop) return;
  constexpr int U = 4;
  const long long node = ((long long)blockIdx.x * 256 + threadIdx.x) / LPN;
  const int lane = (int)(threadIdx.x % LPN);
  double d = 0.0;
  if (node < nnodes) {
    long long rb[P];
#pragma unroll
    for (int c = 0; c < P; ++c) rb[c] = rp[node * P + c];
    const int deg = (int)((rp[node * P + 1] - rb[0]) / P);
    // accumulated in the arithmetic type: FP32 for the inner iterations of the mixed solve (whose
    // FP64 outer refinement recomputes the residual), FP64 otherwise
    T s[P];
#pragma unroll
    for (int c = 0; c < P; ++c) s[c] = T(0);
    // slots past the node's last neighbour re-read that neighbour (deg >= 1 inside the loop) and skip
    // the add: unconditional loads let each row's P consecutive values merge into one wide load
    for (int k0 = lane; k0 < deg; k0 += U * LPN) {
      int col[U];
      T vv[U][P][P];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const long long k = min(k0 + u * LPN, deg - 1);
        col[u] = ci[rb[0] + P * k];
#pragma unroll
        for (int c = 0; c < P; ++c)
#pragma unroll
          for (int e = 0; e < P; ++e) vv[u][c][e] = v[rb[c] + P * k + e];
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        T xv[P];
#pragma unroll
        for (int e = 0; e < P; ++e) xv[e] = x[col[u] + e];
        if (k0 + u * LPN < deg) {
#pragma unroll
          for (int c = 0; c < P; ++c)
#pragma unroll
            for (int e = 0; e < P; ++e) s[c] += vv[u][c][e] * xv[e];
        }
      }
    }
#pragma unroll
    for (int c = 0; c < P; ++c)
#pragma unroll
      for (int off = LPN / 2; off >= 1; off >>= 1) s[c] += __shfl_xor(s[c], off, LPN);
    if (lane == 0) {
#pragma unroll
      for (int c = 0; c < P; ++c) {
        y[node * P + c] = s[c];
        d += (double)x[node * P + c] * (double)s[c];
      }
    }
  }
  if (part) {
    const double t = block_sum256(d);
    if (threadIdx.x == 0) part[blockIdx.x] = t;
  }
}

// out[0] = sum of part[0 .. n) in a fixed order (one workgroup of 1024; four independent
// accumulators per thread keep four loads in flight).
__global__ __launch_bounds__(1024) void k_sum_parts(const double *part, long long n, double *out,
                                                    const int *stop) {
  if (stop && *stop) return;
  __shared__ double red[16];
  double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
  long long i = threadIdx.x;
  for (; i + 3 * 1024 < n; i += 4 * 1024) {
    a0 += part[i];
    a1 += part[i + 1024];
    a2 += part[i + 2 * 1024];
    a3 += part[i + 3 * 1024];
  }
  for (; i < n; i += 1024) a0 += part[i];
  double v = wave_sum((a0 + a1) + (a2 + a3));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int k = 0; k < 16; ++k) t += red[k];
    out[0] = t;
  }
}

template <class T>
__global__ void k_diag_inv(long long nrows, const long long *rp, const int *ci, const double *v, T *dinv) {
  const long long r = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nrows) return;
  long long lo = rp[r], hi = rp[r + 1];
  double dv = 0.0;
  while (lo < hi) {
    const long long mid = (lo + hi) >> 1;
    if (ci[mid] < r)
      lo = mid + 1;
    else
      hi = mid;
  }
  if (lo < rp[r + 1] && ci[lo] == r) dv = v[lo];
  dinv[r] = (T)(dv != 0.0 ? 1.0 / dv : 1.0);
}

// The same on the context's own pattern, whose diagonal offsets are known: one value per row is read
// instead of a binary search over the row's column indices.
template <class T>
__global__ void k_diag_inv_pos(long long nrows, const long long *rp, const int *doff, const double *v, T *dinv) {
  const long long r = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nrows) return;
  const double dv = doff[r] >= 0 ? v[rp[r] + doff[r]] : 0.0;
  dinv[r] = (T)(dv != 0.0 ? 1.0 / dv : 1.0);
}

template <class T> __global__ void k_cast(long long n, const double *a, T *b) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long st = (long long)gridDim.x * blockDim.x;
  for (long long k = i; k < n; k += st) b[k] = (T)a[k];
}

// r = b - r (r holds A x on entry); p = Dinv r; partials of r.Dinv r and r.r
template <class T>
__global__ __launch_bounds__(256) void k_init(long long n, const double *b, T *r, T *p, const T *dinv, double *rz_part,
                                              double *rr_part) {
  double rz = 0.0, rr = 0.0;
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256) {
    const double ri = b[i] - (double)r[i];
    const T rt = (T)ri;
    r[i] = rt;
    const T z = rt * dinv[i];
    p[i] = z;
    rz += (double)rt * (double)z;
    rr += (double)rt * (double)rt;
  }
  rz = block_sum256(rz);
  rr = block_sum256(rr);
  if (threadIdx.x == 0) rz_part[blockIdx.x] = rz, rr_part[blockIdx.x] = rr;
}

// alpha = (r.z) / (p.Ap); x += alpha p; r -= alpha q; partials of the new r.z and r.r.  The
// thread's first element is loaded before the partial sums, so its loads overlap their latency.
template <class T>
__global__ __launch_bounds__(256) void k_update_xr(long long n, T *x, T *r, const T *p, const T *q, const T *dinv,
                                                   const double *rz_part, const double *pap, double *rz_next,
                                                   double *rr_next, const int *stop) {
  if (stop && *stop) return;
  const long long i0 = (long long)blockIdx.x * 256 + threadIdx.x, st = (long long)gridDim.x * 256;
  const long long j0 = i0 < n ? i0 : 0;  // n >= 1: index 0 is in bounds
  const T x0 = x[j0], p0 = p[j0], r0 = r[j0], q0 = q[j0], d0 = dinv[j0];
  const double alpha = sum_partials(rz_part) / pap[0];
  double rz = 0.0, rr = 0.0;
  auto step = [&](long long i, T xi, T pi, T rv, T qi, T di) {
    x[i] = (T)((double)xi + alpha * (double)pi);
    const T ri = (T)((double)rv - alpha * (double)qi);
    r[i] = ri;
    const T z = ri * di;
    rz += (double)ri * (double)z;
    rr += (double)ri * (double)ri;
  };
  if (i0 < n) step(i0, x0, p0, r0, q0, d0);
  for (long long i = i0 + st; i < n; i += st) step(i, x[i], p[i], r[i], q[i], dinv[i]);
  rz = block_sum256(rz);
  rr = block_sum256(rr);
  if (threadIdx.x == 0) rz_next[blockIdx.x] = rz, rr_next[blockIdx.x] = rr;
}

// beta = (r.z)_new / (r.z)_old; p = Dinv r + beta p (first element loaded before the partial sums)
template <class T>
__global__ __launch_bounds__(256) void k_update_p(long long n, T *p, const T *r, const T *dinv, const double *rz_new,
                                                  const double *rz_old, const int *stop) {
  if (stop && *stop) return;
  const long long i0 = (long long)blockIdx.x * 256 + threadIdx.x, st = (long long)gridDim.x * 256;
  const long long j0 = i0 < n ? i0 : 0;  // n >= 1: index 0 is in bounds
  const T r0 = r[j0], d0 = dinv[j0], p0 = p[j0];
  const double beta = sum_partials(rz_new) / sum_partials(rz_old);
  if (i0 < n) p[i0] = (T)((double)(r0 * d0) + beta * (double)p0);
  for (long long i = i0 + st; i < n; i += st) p[i] = (T)((double)(r[i] * dinv[i]) + beta * (double)p[i]);
}

template <class T> __global__ void k_axpy_to64(long long n, const T *d, double *x) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] += (double)d[i];
}

// Lanes per row: the smallest power of two (4..64) whose eight loads per lane cover an average row.
static int lanes_per_row(long long nnz, long long n) {
  const double avg = n ? (double)nnz / (double)n : 1.0;
  int l = 4;
  while (l < 64 && 8.0 * l < avg) l *= 2;
  return l;
}

// How an SpMV is launched: row-wise CSR with `lanes` lanes per row, or node-blocked (P = 2 or 3 dofs
// per node) with `lanes` lanes per node.  blocks = workgroups of 256 = number of x.y partials.
struct SpmvPlan {
  int P = 1;  // 1: row-wise CSR kernel
  int lanes = 4;
  unsigned blocks = 1;
};

static SpmvPlan spmv_plan(long long n, long long nnz, int P, bool node_blocked) {
  SpmvPlan sp;
  if (node_blocked && (P == 2 || P == 3) && n % P == 0) {
    const long long nodes = n / P;
    const double deg = nodes ? (double)nnz / (double)(P * P * nodes) : 1.0;  // neighbours per node
    int l = 2;
    while (l < 16 && 4.0 * l < deg) l *= 2;
    sp.P = P;
    sp.lanes = l;
    sp.blocks = (unsigned)((nodes * l + 255) / 256);
  } else {
    sp.lanes = lanes_per_row(nnz, n);
    sp.blocks = (unsigned)((n * sp.lanes + 255) / 256);
  }
  return sp;
}

template <class T, int P>
static void spmv_node_launch(long long nodes, const long long *rp, const int *ci, const T *v, const T *x, T *y,
                             double *part, const SpmvPlan &sp, hipStream_t s, const int *stop) {
  const dim3 grid(sp.blocks);
  switch (sp.lanes) {
    case 2: hipLaunchKernelGGL((k_spmv_node<T, P, 2>), grid, dim3(256), 0, s, nodes, rp, ci, v, x, y, part, stop); break;
    case 4: hipLaunchKernelGGL((k_spmv_node<T, P, 4>), grid, dim3(256), 0, s, nodes, rp, ci, v, x, y, part, stop); break;
    case 8: hipLaunchKernelGGL((k_spmv_node<T, P, 8>), grid, dim3(256), 0, s, nodes, rp, ci, v, x, y, part, stop); break;
    default: hipLaunchKernelGGL((k_spmv_node<T, P, 16>), grid, dim3(256), 0, s, nodes, rp, ci, v, x, y, part, stop); break;
  }
}

template <class T>
static void spmv_launch(long long n, const long long *rp, const int *ci, const T *v, const T *x, T *y, double *part,
                        const SpmvPlan &sp, hipStream_t s, const int *stop = nullptr) {
  if (sp.P == 3) return spmv_node_launch<T, 3>(n / 3, rp, ci, v, x, y, part, sp, s, stop);
  if (sp.P == 2) return spmv_node_launch<T, 2>(n / 2, rp, ci, v, x, y, part, sp, s, stop);
  const dim3 grid(sp.blocks);
  switch (sp.lanes) {
    case 4: hipLaunchKernelGGL((k_spmv<T, 4>), grid, dim3(256), 0, s, n, rp, ci, v, x, y, part, stop); break;
    case 8: hipLaunchKernelGGL((k_spmv<T, 8>), grid, dim3(256), 0, s, n, rp, ci, v, x, y, part, stop); break;
    case 16: hipLaunchKernelGGL((k_spmv<T, 16>), grid, dim3(256), 0, s, n, rp, ci, v, x, y, part, stop); break;
    case 32: hipLaunchKernelGGL((k_spmv<T, 32>), grid, dim3(256), 0, s, n, rp, ci, v, x, y, part, stop); break;
    default: hipLaunchKernelGGL((k_spmv<T, 64>), grid, dim3(256), 0, s, n, rp, ci, v, x, y, part, stop); break;
  }
}

// Host sum (fixed order) of NBR partials copied to pinned memory.
static int host_sum_partials(Ctx *c, const double *part, hipStream_t s, double *out) {
  FH_HIP(hipMemcpyAsync(c->h_pinned, part, NBR * sizeof(double), hipMemcpyDeviceToHost, s));
  FH_HIP(hipStreamSynchronize(s));
  double t = 0.0;
  for (int k = 0; k < NBR; ++k) t += c->h_pinned[k];
  *out = t;
  return FEMHOT_OK;
}

// Jacobi-PCG in type T on (rp, ci, v) for A x = b (b in FP64), x in T.  Partials in sc (3 x NBR
// doubles: r.z for two parities, r.r), pap_part (one per SpMV workgroup) and pap (the summed p.Ap).
// Returns iterations; *relres = ||r||/||b|| at exit (as tracked by the recurrence).
template <class T>
static int pcg_core(Ctx *c, long long n, const long long *rp, const int *ci, const T *v, const double *b, T *x,
                    T *r, T *p, T *q, T *dinv, double *sc, double *pap_part, double *pap, double bnorm, double tol,
                    int maxiter, int check, const SpmvPlan &sp, hipStream_t s, int *iters_out, double *relres,
                    bool *converged) {
  const dim3 grid(NBR);
  double *rz[2] = {sc, sc + NBR}, *rrp = sc + 2 * NBR;
  const unsigned nsp = sp.blocks;
  spmv_launch<T>(n, rp, ci, v, x, r, nullptr, sp, s);  // r = A x
  hipLaunchKernelGGL((k_init<T>), grid, dim3(256), 0, s, n, b, r, p, dinv, rz[0], rrp);
  double rr = 0.0;
  int rc = host_sum_partials(c, rrp, s, &rr);
  if (rc) return rc;
  *converged = false;
  *relres = bnorm > 0 ? std::sqrt(rr) / bnorm : 0.0;
  if (bnorm == 0.0 || *relres <= tol) {
    *converged = true;
    *iters_out = 0;
    return FEMHOT_OK;
  }
  int k = 0;
  while (k < maxiter) {
    const int par = k & 1;
    spmv_launch<T>(n, rp, ci, v, p, q, pap_part, sp, s);
    hipLaunchKernelGGL(k_sum_parts, dim3(1), dim3(1024), 0, s, pap_part, (long long)nsp, pap, (const int *)nullptr);
    hipLaunchKernelGGL((k_update_xr<T>), grid, dim3(256), 0, s, n, x, r, p, q, dinv, rz[par], pap, rz[1 - par], rrp,
                       (const int *)nullptr);
    hipLaunchKernelGGL((k_update_p<T>), grid, dim3(256), 0, s, n, p, r, dinv, rz[1 - par], rz[par],
                       (const int *)nullptr);
    ++k;
    if (k % check == 0 || k == maxiter) {
      rc = host_sum_partials(c, rrp, s, &rr);
      if (rc) return rc;
      *relres = std::sqrt(std::max(rr, 0.0)) / bnorm;
      if (!(std::isfinite(rr))) {
        set_error("femhot_pcg: non-finite residual");
        *iters_out = k;
        return FEMHOT_ERR_NONFINITE;
      }
      if (*relres <= tol) {
        *converged = true;
        break;
      }
    }
  }
  FH_HIP(hipGetLastError());
  *iters_out = k;
  return FEMHOT_OK;
}

// State of a device-side-stopped PCG (femhot_pcg_async): status 0 running, 1 converged, 2 non-finite
// residual, 3 iteration budget spent.  Every guarded kernel returns at once when status != 0.
struct PcgState {
  int status;
  int iters;
  double bnorm2;
  double rr;
};

// Sum (fixed order) of NBR partials by one workgroup of 1024 threads.
__device__ __forceinline__ double sum_nbr_1024(const double *part) {
  __shared__ double red[16];
  double v = part[threadIdx.x];
  v = wave_sum(v);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double t = 0.0;
  for (int k = 0; k < 16; ++k) t += red[k];
  return t;
}

// ||b||^2 and the initial ||r||^2 from their partials; converged at once when b = 0 or r is small enough.
// A nonzero *outer (the enclosing Newton solve has finished this load step, or failed) stops the solve
// before its first iteration with status 4.
__global__ __launch_bounds__(1024) void k_pcg_start(const double *bb_part, const double *rr_part, double tol2,
                                                    int maxiter, const int *outer, PcgState *st) {
  const double bb = sum_nbr_1024(bb_part);
  const double rr = sum_nbr_1024(rr_part);
  if (threadIdx.x == 0) {
    st->bnorm2 = bb;
    st->rr = rr;
    st->iters = 0;
    st->status = (outer && *outer)              ? 4
                 : !(rr == rr && rr < INFINITY) ? 2
                 : (bb == 0.0 || rr <= tol2 * bb) ? 1
                                                  : (maxiter <= 0 ? 3 : 0);
  }
}

// After one iteration's x, r update: count it, sum ||r||^2, stop on convergence, a non-finite residual or
// the iteration budget
__global__ __launch_bounds__(1024) void k_pcg_check(const double *rr_part, double tol2, int maxiter, PcgState *st) {
  if (st->status) return;
  const double rr = sum_nbr_1024(rr_part);
  if (threadIdx.x == 0) {
    const int it = st->iters + 1;
    st->iters = it;
    st->rr = rr;
    st->status = !(rr == rr && rr < INFINITY) ? 2 : rr <= tol2 * st->bnorm2 ? 1 : (it >= maxiter ? 3 : 0);
  }
}

__global__ void k_pcg_finish(const PcgState *st, int *iters, double *rel_res, int *status) {
  if (threadIdx.x == 0) {
    if (iters) *iters = st->iters;
    if (rel_res) *rel_res = st->bnorm2 > 0.0 ? sqrt(st->rr / st->bnorm2) : 0.0;
    if (status) *status = st->status;
  }
}

// part[block] = this workgroup's share of a.b
__global__ __launch_bounds__(256) void k_dot64(long long n, const double *a, const double *b, double *part) {
  double v = 0.0;
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256)
    v += a[i] * b[i];
  v = block_sum256(v);
  if (threadIdx.x == 0) part[blockIdx.x] = v;
}

// r = b - A x (FP64); part[block] = this workgroup's share of ||r||^2
__global__ __launch_bounds__(256) void k_resid64(long long n, const double *b, const double *ax, double *r,
                                                 double *part) {
  double v = 0.0;
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256) {
    const double ri = b[i] - ax[i];
    r[i] = ri;
    v += ri * ri;
  }
  v = block_sum256(v);
  if (threadIdx.x == 0) part[blockIdx.x] = v;
}

}  // namespace femhot

using namespace femhot;

extern "C" int femhot_spmv(femhot_ctx *h, const long long *row_ptr, const int *col_idx, const double *vals,
                           const double *x, double *y, void *stream) {
  FH_CHECK_ARG(h && row_ptr && col_idx && vals && x && y, FEMHOT_ERR_ARG, "null argument");
  Ctx *c = &h->c;
  int rc = ensure_pattern(c);
  if (rc) return rc;
  spmv_launch<double>(c->Ndof, row_ptr, col_idx, vals, x, y, nullptr, spmv_plan(c->Ndof, c->nnz, 1, false),
                      (hipStream_t)stream);
  FH_HIP(hipGetLastError());
  return FEMHOT_OK;
}

extern "C" int femhot_pcg(femhot_ctx *h, int mode, const long long *row_ptr, const int *col_idx, const double *vals,
                          const double *b, double *x, double tol, int maxiter, int *iters, double *rel_res,
                          void *stream) {
  FH_CHECK_ARG(h && row_ptr && col_idx && vals && b && x, FEMHOT_ERR_ARG, "null argument");
  const bool node_blocked = (mode & FEMHOT_PCG_CONTEXT_PATTERN) != 0;
  mode &= ~FEMHOT_PCG_CONTEXT_PATTERN;
  FH_CHECK_ARG(mode == FEMHOT_F64 || mode == FEMHOT_F32, FEMHOT_ERR_ARG,
               "mode must be FEMHOT_F64 or FEMHOT_F32 (| FEMHOT_PCG_CONTEXT_PATTERN)");
  FH_CHECK_ARG(tol > 0 && maxiter > 0, FEMHOT_ERR_ARG, "tol and maxiter must be positive");
  Ctx *c = &h->c;
  int rc = ensure_pattern(c);
  if (rc) return rc;
  hipStream_t s = (hipStream_t)stream;
  const long long n = c->Ndof, nnz = c->nnz;
  const dim3 grid((unsigned)((n + 255) / 256));
  const int check = 10;
  // workspace: reduction partials (5 x NBR doubles) + 5 FP64 vectors + (FP32 mode) 5 FP32 vectors + FP32 matrix
  const size_t vec64 = sizeof(double) * n, vec32 = sizeof(float) * n;
  const SpmvPlan sp = spmv_plan(n, nnz, c->ndpn, node_blocked);
  const size_t nsp = sp.blocks;
  size_t need = (4 * NBR + nsp + 8) * sizeof(double) + 1024 + 5 * vec64 + 1024;
  if (mode == FEMHOT_F32) need += 5 * vec32 + sizeof(float) * nnz + 1024;
  rc = ensure_workspace(c, need);
  if (rc) return rc;
  char *ws = (char *)c->d_ws;
  auto take = [&](size_t bytes) {
    char *p = ws;
    ws += (bytes + 255) & ~size_t(255);
    return (void *)p;
  };
  double *sc = (double *)take(3 * NBR * sizeof(double));
  double *aux = (double *)take(NBR * sizeof(double));
  double *pap_part = (double *)take(nsp * sizeof(double));
  double *pap = (double *)take(8 * sizeof(double));
  double *r64 = (double *)take(vec64), *p64 = (double *)take(vec64), *q64 = (double *)take(vec64);
  double *dinv64 = (double *)take(vec64), *ax64 = (double *)take(vec64);
  // ||b||
  hipLaunchKernelGGL(k_dot64, dim3(NBR), dim3(256), 0, s, n, b, b, aux);
  double bb = 0.0;
  rc = host_sum_partials(c, aux, s, &bb);
  if (rc) return rc;
  const double bnorm = std::sqrt(bb);
  if (node_blocked)
    hipLaunchKernelGGL((k_diag_inv_pos<double>), grid, dim3(256), 0, s, n, row_ptr, c->d_diag_off, vals, dinv64);
  else
    hipLaunchKernelGGL((k_diag_inv<double>), grid, dim3(256), 0, s, n, row_ptr, col_idx, vals, dinv64);
  int it = 0;
  double rel = 0.0;
  bool conv = false;
  if (mode == FEMHOT_F64) {
    rc = pcg_core<double>(c, n, row_ptr, col_idx, vals, b, x, r64, p64, q64, dinv64, sc, pap_part, pap, bnorm, tol,
                          maxiter, check, sp, s, &it, &rel, &conv);
    if (rc) return rc;
  } else {
    float *v32 = (float *)take(sizeof(float) * nnz);
    float *x32 = (float *)take(vec32), *r32 = (float *)take(vec32), *p32 = (float *)take(vec32);
    float *q32 = (float *)take(vec32), *dinv32 = (float *)take(vec32);
    const unsigned cg = (unsigned)std::min<long long>((nnz + 255) / 256, 65536);
    hipLaunchKernelGGL((k_cast<float>), dim3(cg), dim3(256), 0, s, nnz, vals, v32);
    if (node_blocked)
      hipLaunchKernelGGL((k_diag_inv_pos<float>), grid, dim3(256), 0, s, n, row_ptr, c->d_diag_off, vals, dinv32);
    else
      hipLaunchKernelGGL((k_diag_inv<float>), grid, dim3(256), 0, s, n, row_ptr, col_idx, vals, dinv32);
    // outer FP64 refinement: r = b - A x; solve A d = r in FP32 to inner_tol; x += d
    const double inner_tol = 1e-4;
    int outer = 0;
    while (true) {
      spmv_launch<double>(n, row_ptr, col_idx, vals, x, ax64, nullptr, sp, s);
      hipLaunchKernelGGL(k_resid64, dim3(NBR), dim3(256), 0, s, n, b, ax64, r64, aux);
      double r2 = 0.0;
      rc = host_sum_partials(c, aux, s, &r2);
      if (rc) return rc;
      rel = bnorm > 0 ? std::sqrt(r2) / bnorm : 0.0;
      if (!std::isfinite(rel)) {
        set_error("femhot_pcg: non-finite residual");
        return FEMHOT_ERR_NONFINITE;
      }
      if (rel <= tol) {
        conv = true;
        break;
      }
      if (it >= maxiter || outer >= 50) break;
      FH_CHECK_ARG(fill<float>(x32, n, n, 1, 0.0f, s) == FEMHOT_OK, FEMHOT_ERR_HIP, "fill failed");
      int k_in = 0;
      double rel_in = 0.0;
      bool cin = false;
      // inner solve on the FP64 residual r64 (FP32 iterates, FP64 dot products)
      rc = pcg_core<float>(c, n, row_ptr, col_idx, v32, r64, x32, r32, p32, q32, dinv32, sc, pap_part, pap, bnorm * rel,
                           std::max(inner_tol, 0.5 * tol / std::max(rel, 1e-300)), maxiter - it, check, sp, s,
                           &k_in, &rel_in, &cin);
      if (rc) return rc;
      it += std::max(k_in, 1);
      hipLaunchKernelGGL((k_axpy_to64<float>), grid, dim3(256), 0, s, n, x32, x);
      ++outer;
    }
  }
  FH_HIP(hipGetLastError());
  if (iters) *iters = it;
  if (rel_res) *rel_res = rel;
  if (!conv) {
    set_error("femhot_pcg: maximum iterations reached");
    return FEMHOT_ERR_NOT_CONVERGED;
  }
  return FEMHOT_OK;
}

// Workspace of femhot_pcg_async: the layout of femhot_pcg in FP64 mode plus the solver state, each piece
// 256-byte aligned.
static size_t pcg_async_ws_bytes(const Ctx *c, bool node_blocked) {
  auto r256 = [](size_t b) { return (b + 255) & ~size_t(255); };
  const size_t vec64 = sizeof(double) * c->Ndof;
  const size_t nsp = spmv_plan(c->Ndof, c->nnz, c->ndpn, node_blocked).blocks;
  return r256(3 * NBR * sizeof(double)) + r256(NBR * sizeof(double)) + r256(nsp * sizeof(double)) +
         r256(8 * sizeof(double)) + 5 * r256(vec64) + r256(sizeof(PcgState));
}

static int pcg_async_core(Ctx *c, bool node_blocked, const long long *row_ptr, const int *col_idx,
                          const double *vals, const double *b, double *x, double tol, int iterations,
                          const int *outer, const PcgState **state_out, int *iters, double *rel_res, int *status,
                          hipStream_t s);

extern "C" int femhot_pcg_async(femhot_ctx *h, int mode, const long long *row_ptr, const int *col_idx,
                                const double *vals, const double *b, double *x, double tol, int iterations,
                                int *iters, double *rel_res, int *status, void *stream) {
  FH_CHECK_ARG(h && row_ptr && col_idx && vals && b && x, FEMHOT_ERR_ARG, "femhot_pcg_async: null argument");
  const bool node_blocked = (mode & FEMHOT_PCG_CONTEXT_PATTERN) != 0;
  mode &= ~FEMHOT_PCG_CONTEXT_PATTERN;
  FH_CHECK_ARG(mode == FEMHOT_F64, FEMHOT_ERR_ARG, "femhot_pcg_async: mode must be FEMHOT_F64 (| FEMHOT_PCG_CONTEXT_PATTERN)");
  FH_CHECK_ARG(tol > 0 && iterations > 0, FEMHOT_ERR_ARG, "femhot_pcg_async: tol and iterations must be positive");
  Ctx *c = &h->c;
  FH_CHECK_ARG(c->have_pattern, FEMHOT_ERR_ARG,
               "femhot_pcg_async: build the context's pattern first (femhot_pattern or one tangent call)");
  return pcg_async_core(c, node_blocked, row_ptr, col_idx, vals, b, x, tol, iterations, nullptr, nullptr, iters,
                        rel_res, status, (hipStream_t)stream);
}

// The enqueued solve of femhot_pcg_async.  outer: a device flag that, when nonzero as the solve starts,
// skips it (status 4; the Newton driver's end-of-step flag); *state_out: where the solver state lives
// until the next solve on this context.
static int pcg_async_core(Ctx *c, bool node_blocked, const long long *row_ptr, const int *col_idx,
                          const double *vals, const double *b, double *x, double tol, int iterations,
                          const int *outer, const PcgState **state_out, int *iters, double *rel_res, int *status,
                          hipStream_t s) {
  const long long n = c->Ndof, nnz = c->nnz;
  const dim3 grid((unsigned)((n + 255) / 256)), nbr(NBR);
  const size_t vec64 = sizeof(double) * n;
  const SpmvPlan sp = spmv_plan(n, nnz, c->ndpn, node_blocked);
  const size_t nsp = sp.blocks;
  const size_t need = pcg_async_ws_bytes(c, node_blocked);
  if (c->ws_bytes < need) {  // sized once; never inside a graph capture
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    FH_HIP(hipStreamIsCapturing(s, &cs));
    FH_CHECK_ARG(cs == hipStreamCaptureStatusNone, FEMHOT_ERR_ARG,
                 "femhot_pcg_async: the solver workspace is not sized; make one uncaptured call (or femhot_pcg) on "
                 "this context first, so captured calls allocate nothing");
    int rc = ensure_workspace(c, need);
    if (rc) return rc;
  }
  char *ws = (char *)c->d_ws;
  auto take = [&](size_t bytes) {
    char *p = ws;
    ws += (bytes + 255) & ~size_t(255);
    return (void *)p;
  };
  double *sc = (double *)take(3 * NBR * sizeof(double));
  double *aux = (double *)take(NBR * sizeof(double));
  double *pap_part = (double *)take(nsp * sizeof(double));
  double *pap = (double *)take(8 * sizeof(double));
  double *r = (double *)take(vec64), *p = (double *)take(vec64), *q = (double *)take(vec64);
  double *dinv = (double *)take(vec64);
  (void)take(vec64);
  PcgState *st = (PcgState *)take(sizeof(PcgState));
  const int *stop = &st->status;
  double *rz[2] = {sc, sc + NBR}, *rrp = sc + 2 * NBR;
  const double tol2 = tol * tol;
  hipLaunchKernelGGL(k_dot64, nbr, dim3(256), 0, s, n, b, b, aux);
  if (node_blocked)
    hipLaunchKernelGGL((k_diag_inv_pos<double>), grid, dim3(256), 0, s, n, row_ptr, c->d_diag_off, vals, dinv);
  else
    hipLaunchKernelGGL((k_diag_inv<double>), grid, dim3(256), 0, s, n, row_ptr, col_idx, vals, dinv);
  spmv_launch<double>(n, row_ptr, col_idx, vals, x, r, nullptr, sp, s);  // r = A x
  hipLaunchKernelGGL((k_init<double>), nbr, dim3(256), 0, s, n, b, r, p, dinv, rz[0], rrp);
  hipLaunchKernelGGL(k_pcg_start, dim3(1), dim3(1024), 0, s, aux, rrp, tol2, iterations, outer, st);
  for (int k = 0; k < iterations; ++k) {  // the same kernels as femhot_pcg; a stopped solve skips them
    const int par = k & 1;
    spmv_launch<double>(n, row_ptr, col_idx, vals, p, q, pap_part, sp, s, stop);
    hipLaunchKernelGGL(k_sum_parts, dim3(1), dim3(1024), 0, s, pap_part, (long long)nsp, pap, stop);
    hipLaunchKernelGGL((k_update_xr<double>), nbr, dim3(256), 0, s, n, x, r, p, q, dinv, rz[par], pap, rz[1 - par],
                       rrp, stop);
    hipLaunchKernelGGL(k_pcg_check, dim3(1), dim3(1024), 0, s, rrp, tol2, iterations, st);
    hipLaunchKernelGGL((k_update_p<double>), nbr, dim3(256), 0, s, n, p, r, dinv, rz[1 - par], rz[par], stop);
  }
  hipLaunchKernelGGL(k_pcg_finish, dim3(1), dim3(64), 0, s, st, iters, rel_res, status);
  FH_HIP(hipGetLastError());
  if (state_out) *state_out = st;
  return FEMHOT_OK;
}

// ------------------------------------------------------------------------------------------
// Load-stepped Newton-Raphson driver (fol/solvers/fe_nonlinear_residual_based_solver.py:76-170)
// ------------------------------------------------------------------------------------------

// Dirichlet dofs of u set to scale * value (ApplyDirichletBCOnDofVector, fe_loss.py:91-92 with the
// load increment of the nonlinear solver).
__global__ void k_apply_dirichlet(long long n, double *u, const uint8_t *dmask, const double *dval, double scale) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && dmask[i]) u[i] = scale * dval[i];
}
__global__ void k_newton_rhs(long long n, const double *R, double *b, double *du) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    b[i] = -R[i];
    du[i] = 0.0;
  }
}
__global__ void k_add64(long long n, double *u, const double *du) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) u[i] += du[i];
}

extern "C" int femhot_newton(femhot_ctx *h, int mode, const double *field, double *u, unsigned flags,
                             int load_steps, int max_iter, double abs_tol, double rel_tol, double lin_tol,
                             int lin_maxiter, double *history, int *steps, void *stream) {
  FH_CHECK_ARG(h && field && u, FEMHOT_ERR_ARG, "femhot_newton: null argument");
  FH_CHECK_ARG(mode == FEMHOT_F64 || mode == FEMHOT_F32, FEMHOT_ERR_ARG,
               "femhot_newton: mode must be FEMHOT_F64 or FEMHOT_F32");
  FH_CHECK_ARG(load_steps > 0 && max_iter > 0 && lin_tol > 0 && lin_maxiter > 0, FEMHOT_ERR_ARG,
               "femhot_newton: load_steps, max_iter, lin_tol and lin_maxiter must be positive");
  FH_CHECK_ARG((flags & ~(FEMHOT_ELEMENT_FIELD | FEMHOT_SCATTER_COLOUR | FEMHOT_SCATTER_CHUNK | FEMHOT_SCATTER_ROWS)) ==
                   0,
               FEMHOT_ERR_ARG,
               "femhot_newton: flags may hold FEMHOT_ELEMENT_FIELD and FEMHOT_SCATTER_{COLOUR,CHUNK,ROWS} only");
  Ctx *c = &h->c;
  int rc = ensure_pattern(c);
  if (rc) return rc;
  hipStream_t s = (hipStream_t)stream;
  const long long n = c->Ndof, nnz = c->nnz;
  auto up = [](size_t b) { return (b + 255) & ~size_t(255); };
  const size_t need = up(sizeof(double) * std::max<long long>(nnz, 1)) + 4 * up(sizeof(double) * n) +
                      up(sizeof(double) * NBR);
  if (c->newton_ws.bytes < need) {
    FH_HIP(hipStreamSynchronize(s));
    if (c->newton_ws.ptr) FH_HIP(hipFree(c->newton_ws.ptr));
    c->newton_ws.ptr = nullptr;
    c->newton_ws.bytes = 0;
    c->newton_dval_ready = false;
    FH_HIP(hipMalloc(&c->newton_ws.ptr, need));
    c->newton_ws.bytes = need;
  }
  char *ws = (char *)c->newton_ws.ptr;
  auto take = [&](size_t bytes) {
    char *p = ws;
    ws += up(bytes);
    return (void *)p;
  };
  double *vals = (double *)take(sizeof(double) * std::max<long long>(nnz, 1));
  double *R = (double *)take(sizeof(double) * n), *b = (double *)take(sizeof(double) * n);
  double *du = (double *)take(sizeof(double) * n), *dval = (double *)take(sizeof(double) * n);
  double *part = (double *)take(sizeof(double) * NBR);
  if (!c->newton_dval_ready) {  // Dirichlet values as a dense dof vector (the last entry of a repeated dof wins)
    std::vector<double> hv((size_t)n, 0.0);
    for (size_t k = 0; k < c->h_dir_idx.size(); ++k) hv[(size_t)c->h_dir_idx[k]] = c->h_dir_val[k];
    FH_HIP(hipStreamSynchronize(s));
    FH_HIP(hipMemcpy(dval, hv.data(), sizeof(double) * n, hipMemcpyHostToDevice));
    c->newton_dval_ready = true;
  }
  if (history)
    for (long long k = 0; k < 2LL * load_steps * max_iter; ++k) history[k] = std::nan("");
  const dim3 grid((unsigned)((n + 255) / 256));
  for (int step = 1; step <= load_steps; ++step) {
    hipLaunchKernelGGL(k_apply_dirichlet, grid, dim3(256), 0, s, n, u, c->d_dmask, dval,
                       (double)step / load_steps);
    int it = 1;
    for (;; ++it) {
      rc = femhot_tangent_csr(h, FEMHOT_F64, field, u, vals, R, flags, stream);
      if (rc) return rc;
      hipLaunchKernelGGL(k_dot64, dim3(NBR), dim3(256), 0, s, n, R, R, part);
      double rr = 0.0;
      if ((rc = host_sum_partials(c, part, s, &rr))) return rc;
      const double res_norm = std::sqrt(rr);
      if (std::isnan(res_norm)) {
        if (steps) steps[step - 1] = it;
        set_error("femhot_newton: residual norm contains NaN values");
        return FEMHOT_ERR_NONFINITE;
      }
      hipLaunchKernelGGL(k_newton_rhs, grid, dim3(256), 0, s, n, R, b, du);
      int lit = 0;
      double lrel = 0.0;
      rc = femhot_pcg(h, mode | FEMHOT_PCG_CONTEXT_PATTERN, c->d_row_ptr, c->d_col_idx, vals, b, du, lin_tol,
                      lin_maxiter, &lit, &lrel, stream);
      if (rc && rc != FEMHOT_ERR_NOT_CONVERGED) return rc;  // an unconverged solve continues with its iterate
      hipLaunchKernelGGL(k_dot64, dim3(NBR), dim3(256), 0, s, n, du, du, part);
      double dd = 0.0;
      if ((rc = host_sum_partials(c, part, s, &dd))) return rc;
      const double du_norm = std::sqrt(dd);
      if (history) {
        history[2LL * ((long long)(step - 1) * max_iter + (it - 1))] = res_norm;
        history[2LL * ((long long)(step - 1) * max_iter + (it - 1)) + 1] = du_norm;
      }
      if (res_norm < abs_tol || du_norm < rel_tol || it == max_iter) break;
      hipLaunchKernelGGL(k_add64, grid, dim3(256), 0, s, n, u, du);
    }
    if (steps) steps[step - 1] = it;
  }
  FH_HIP(hipGetLastError());
  FH_HIP(hipStreamSynchronize(s));
  return FEMHOT_OK;
}

// ------------------------------------------------------------------------------------------
// The same load-stepped Newton solve with every decision on the device: nothing waits for the host,
// so a whole nonlinear solve can be captured into a HIP graph.
// ------------------------------------------------------------------------------------------

struct NewtonState {
  int halt;    // this load step has ended (or the solve failed): the guarded kernels return
  int failed;  // a NaN residual norm or a non-finite linear solve: every later kernel returns
  double res_norm;
};

__global__ void k_newton_init(NewtonState *st, double *history, long long n_hist, int *steps, int load_steps) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0) {
    st->halt = 0;
    st->failed = 0;
    st->res_norm = 0.0;
  }
  if (history && i < n_hist) history[i] = __builtin_nan("");
  if (steps && i < load_steps) steps[i] = 0;
}
__global__ void k_newton_begin_step(NewtonState *st) { st->halt = st->failed; }
__global__ void k_apply_dirichlet_if(long long n, double *u, const uint8_t *dmask, const double *dval, double scale,
                                     const int *failed) {
  if (*failed) return;
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && dmask[i]) u[i] = scale * dval[i];
}
__global__ void k_newton_rhs_if(long long n, const double *R, double *b, double *du, const int *halt) {
  if (*halt) return;
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    b[i] = -R[i];
    du[i] = 0.0;
  }
}
__global__ void k_add64_if(long long n, double *u, const double *du, const int *halt) {
  if (*halt) return;
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) u[i] += du[i];
}
// ||R|| of this iteration (from k_dot64's partials) into the state and the history; a NaN ends the solve
__global__ __launch_bounds__(1024) void k_newton_res(const double *part, NewtonState *st, double *hist,
                                                     int *step_iters, int it) {
  if (st->halt) return;
  const double rn = sqrt(sum_nbr_1024(part));
  if (threadIdx.x == 0) {
    st->res_norm = rn;
    if (hist) hist[0] = rn;
    if (rn != rn) {
      st->failed = 1;
      st->halt = 1;
      if (step_iters) *step_iters = it;
    }
  }
}
// ||du|| after the linear solve; the load step ends when ||R|| < abs_tol, ||du|| < rel_tol or at max_iter
// (the update of that last iteration is not applied), the solve when the linear solve was non-finite
__global__ __launch_bounds__(1024) void k_newton_du(const double *part, NewtonState *st, const PcgState *ps,
                                                    double *hist, int *step_iters, int it, int max_iter,
                                                    double abs_tol, double rel_tol) {
  if (st->halt) return;
  const double dn = sqrt(sum_nbr_1024(part));
  if (threadIdx.x == 0) {
    if (ps->status == 2) {
      st->failed = 1;
      st->halt = 1;
      if (step_iters) *step_iters = it;
      return;
    }
    if (hist) hist[1] = dn;
    if (st->res_norm < abs_tol || dn < rel_tol || it == max_iter) {
      st->halt = 1;
      if (step_iters) *step_iters = it;
    }
  }
}
__global__ void k_newton_finish(const NewtonState *st, int *status) {
  if (threadIdx.x == 0 && status) *status = st->failed ? FEMHOT_ERR_NONFINITE : FEMHOT_OK;
}

extern "C" int femhot_newton_async(femhot_ctx *h, const double *field, double *u, unsigned flags, int load_steps,
                                   int max_iter, double abs_tol, double rel_tol, double lin_tol, int lin_iterations,
                                   double *history, int *steps, int *status, void *stream) {
  FH_CHECK_ARG(h && field && u, FEMHOT_ERR_ARG, "femhot_newton_async: null argument");
  FH_CHECK_ARG(load_steps > 0 && max_iter > 0 && lin_tol > 0 && lin_iterations > 0, FEMHOT_ERR_ARG,
               "femhot_newton_async: load_steps, max_iter, lin_tol and lin_iterations must be positive");
  FH_CHECK_ARG((flags & ~(FEMHOT_ELEMENT_FIELD | FEMHOT_SCATTER_COLOUR | FEMHOT_SCATTER_CHUNK | FEMHOT_SCATTER_ROWS)) ==
                   0,
               FEMHOT_ERR_ARG,
               "femhot_newton_async: flags may hold FEMHOT_ELEMENT_FIELD and FEMHOT_SCATTER_{COLOUR,CHUNK,ROWS} only");
  Ctx *c = &h->c;
  hipStream_t s = (hipStream_t)stream;
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  FH_HIP(hipStreamIsCapturing(s, &cs));
  const bool capturing = cs != hipStreamCaptureStatusNone;
  auto up = [](size_t b) { return (b + 255) & ~size_t(255); };
  const long long n = c->Ndof;
  // the workspace of femhot_newton followed by the solve's state
  const size_t need = up(sizeof(double) * std::max<long long>(c->nnz, 1)) + 4 * up(sizeof(double) * n) +
                      up(sizeof(double) * NBR) + up(sizeof(NewtonState));
  if (!c->have_pattern || c->newton_ws.bytes < need || !c->newton_dval_ready ||
      c->ws_bytes < pcg_async_ws_bytes(c, true)) {
    FH_CHECK_ARG(!capturing, FEMHOT_ERR_ARG,
                 "femhot_newton_async: make one uncaptured call on this context first (it builds the pattern and sizes "
                 "the workspaces), so captured calls allocate nothing");
    int rc = ensure_pattern(c);
    if (rc) return rc;
    const size_t need_now = up(sizeof(double) * std::max<long long>(c->nnz, 1)) + 4 * up(sizeof(double) * n) +
                            up(sizeof(double) * NBR) + up(sizeof(NewtonState));
    if (c->newton_ws.bytes < need_now) {
      FH_HIP(hipStreamSynchronize(s));
      if (c->newton_ws.ptr) FH_HIP(hipFree(c->newton_ws.ptr));
      c->newton_ws.ptr = nullptr;
      c->newton_ws.bytes = 0;
      c->newton_dval_ready = false;
      FH_HIP(hipMalloc(&c->newton_ws.ptr, need_now));
      c->newton_ws.bytes = need_now;
    }
    if ((rc = ensure_workspace(c, pcg_async_ws_bytes(c, true)))) return rc;
  }
  char *ws = (char *)c->newton_ws.ptr;
  auto take = [&](size_t bytes) {
    char *p = ws;
    ws += up(bytes);
    return (void *)p;
  };
  double *vals = (double *)take(sizeof(double) * std::max<long long>(c->nnz, 1));
  double *R = (double *)take(sizeof(double) * n), *b = (double *)take(sizeof(double) * n);
  double *du = (double *)take(sizeof(double) * n), *dval = (double *)take(sizeof(double) * n);
  double *part = (double *)take(sizeof(double) * NBR);
  NewtonState *st = (NewtonState *)take(sizeof(NewtonState));
  if (!c->newton_dval_ready) {  // Dirichlet values as a dense dof vector (the last entry of a repeated dof wins)
    std::vector<double> hv((size_t)n, 0.0);
    for (size_t k = 0; k < c->h_dir_idx.size(); ++k) hv[(size_t)c->h_dir_idx[k]] = c->h_dir_val[k];
    FH_HIP(hipStreamSynchronize(s));
    FH_HIP(hipMemcpy(dval, hv.data(), sizeof(double) * n, hipMemcpyHostToDevice));
    c->newton_dval_ready = true;
  }
  const long long n_hist = 2LL * load_steps * max_iter;
  const dim3 grid((unsigned)((n + 255) / 256));
  const long long n_init = std::max<long long>({n_hist, (long long)load_steps, 1});
  hipLaunchKernelGGL(k_newton_init, dim3((unsigned)((n_init + 255) / 256)), dim3(256), 0, s, st, history, n_hist,
                     steps, load_steps);
  const int *halt = &st->halt, *failed = &st->failed;
  for (int step = 1; step <= load_steps; ++step) {
    hipLaunchKernelGGL(k_apply_dirichlet_if, grid, dim3(256), 0, s, n, u, c->d_dmask, dval, (double)step / load_steps,
                       failed);
    hipLaunchKernelGGL(k_newton_begin_step, dim3(1), dim3(1), 0, s, st);
    int *step_iters = steps ? steps + (step - 1) : nullptr;
    for (int it = 1; it <= max_iter; ++it) {
      double *hist = history ? history + 2LL * ((long long)(step - 1) * max_iter + (it - 1)) : nullptr;
      // the tangent and residual are assembled whatever the flags say (their outputs are this call's
      // workspace; u only changes through the guarded update below)
      int rc = femhot_tangent_csr(h, FEMHOT_F64, field, u, vals, R, flags, stream);
      if (rc) return rc;
      hipLaunchKernelGGL(k_dot64, dim3(NBR), dim3(256), 0, s, n, R, R, part);
      hipLaunchKernelGGL(k_newton_res, dim3(1), dim3(1024), 0, s, part, st, hist, step_iters, it);
      hipLaunchKernelGGL(k_newton_rhs_if, grid, dim3(256), 0, s, n, R, b, du, halt);
      const PcgState *ps = nullptr;
      rc = pcg_async_core(c, true, c->d_row_ptr, c->d_col_idx, vals, b, du, lin_tol, lin_iterations, halt, &ps,
                          nullptr, nullptr, nullptr, s);
      if (rc) return rc;
      hipLaunchKernelGGL(k_dot64, dim3(NBR), dim3(256), 0, s, n, du, du, part);
      hipLaunchKernelGGL(k_newton_du, dim3(1), dim3(1024), 0, s, part, st, ps, hist, step_iters, it, max_iter, abs_tol,
                         rel_tol);
      hipLaunchKernelGGL(k_add64_if, grid, dim3(256), 0, s, n, u, du, halt);
    }
  }
  hipLaunchKernelGGL(k_newton_finish, dim3(1), dim3(64), 0, s, st, status);
  FH_HIP(hipGetLastError());
  return FEMHOT_OK;
}
