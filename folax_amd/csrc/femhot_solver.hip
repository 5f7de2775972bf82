// Jacobi-preconditioned conjugate gradients on the assembled CSR tangent.
//
// Reference: fol/solvers/fe_solver.py:60-78 (LinearSolve: BiCGSTAB / direct solve of K du = -R)
// and fe_linear_residual_based_solver.py:15-22.  The reference has no CG; with the reference's
// BC treatment (Dirichlet rows zeroed except the diagonal, R_D = 0) the Dirichlet unknowns stay
// zero and the iteration runs on the SPD free block, so CG applies to the matrix as assembled.
//
// Design: all CG scalars live in device memory (double, double-buffered by iteration parity), so
// an iteration is three stream-ordered kernels with no host round trip; the host reads the
// residual norm only every `check` iterations.  FP32 mode keeps the matrix and vectors in FP32,
// accumulates every dot product in FP64, and wraps the FP32 solves in FP64 iterative refinement
// r = b - A x with the FP64 matrix.
#include <cmath>

#include "femhot_kernels.h"

namespace femhot {

// y = A x with LPR lanes per row; optionally dot += x.y (FP64) into *dot.
template <class T, int LPR>
__global__ __launch_bounds__(256) void k_spmv(long long nrows, const long long *__restrict__ rp,
                                              const int *__restrict__ ci, const T *__restrict__ v,
                                              const T *__restrict__ x, T *__restrict__ y, double *dot) {
  const long long tid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long row = tid / LPR;
  const int lane = (int)(tid % LPR);
  double s = 0.0;
  if (row < nrows) {
    const long long b = rp[row], e = rp[row + 1];
    for (long long k = b + lane; k < e; k += LPR) s += (double)v[k] * (double)x[ci[k]];
  }
#pragma unroll
  for (int off = LPR / 2; off >= 1; off >>= 1) s += __shfl_xor(s, off);
  double d = 0.0;
  if (row < nrows && lane == 0) {
    y[row] = (T)s;
    d = (double)x[row] * (double)(T)s;
  }
  if (dot) {
    d = wave_sum(d);
    if ((threadIdx.x & 63) == 0) unsafeAtomicAdd(dot, d);
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

template <class T> __global__ void k_cast(long long n, const double *a, T *b) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long st = (long long)gridDim.x * blockDim.x;
  for (long long k = i; k < n; k += st) b[k] = (T)a[k];
}

// r = b - r (r holds A x on entry); p = Dinv r; sc[0] += r.Dinv r ; sc[1] += r.r
template <class T>
__global__ __launch_bounds__(256) void k_init(long long n, const double *b, T *r, T *p, const T *dinv, double *sc) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  double rz = 0.0, rr = 0.0;
  if (i < n) {
    const double ri = b[i] - (double)r[i];
    const T rt = (T)ri;
    r[i] = rt;
    const T z = rt * dinv[i];
    p[i] = z;
    rz = (double)rt * (double)z;
    rr = (double)rt * (double)rt;
  }
  rz = wave_sum(rz);
  rr = wave_sum(rr);
  if ((threadIdx.x & 63) == 0) {
    unsafeAtomicAdd(&sc[0], rz);
    unsafeAtomicAdd(&sc[1], rr);
  }
}

// scalars layout: rz[2] at sc[0..1], rr[2] at sc[2..3], pAp[2] at sc[4..5]
template <class T>
__global__ __launch_bounds__(256) void k_update_xr(long long n, int par, T *x, T *r, const T *p, const T *q,
                                                   const T *dinv, double *sc) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const double alpha = sc[par] / sc[4 + par];
  double rz = 0.0, rr = 0.0;
  if (i < n) {
    x[i] = (T)((double)x[i] + alpha * (double)p[i]);
    const T ri = (T)((double)r[i] - alpha * (double)q[i]);
    r[i] = ri;
    const T z = ri * dinv[i];
    rz = (double)ri * (double)z;
    rr = (double)ri * (double)ri;
  }
  rz = wave_sum(rz);
  rr = wave_sum(rr);
  if ((threadIdx.x & 63) == 0) {
    unsafeAtomicAdd(&sc[1 - par], rz);
    unsafeAtomicAdd(&sc[2 + 1 - par], rr);
  }
}

template <class T>
__global__ __launch_bounds__(256) void k_update_p(long long n, int par, T *p, const T *r, const T *dinv,
                                                  const double *sc) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const double beta = sc[1 - par] / sc[par];
  if (i < n) p[i] = (T)((double)(r[i] * dinv[i]) + beta * (double)p[i]);
}

__global__ void k_zero_next(double *sc, int par) {
  // zero the accumulators of the next iteration: rz[1-par], rr[1-par], pAp[1-par]
  sc[1 - par] = 0.0;
  sc[2 + 1 - par] = 0.0;
  sc[4 + 1 - par] = 0.0;
}

__global__ void k_zero_n(double *sc, int n) {
  if (threadIdx.x < n) sc[threadIdx.x] = 0.0;
}

template <class T> __global__ void k_axpy_to64(long long n, const T *d, double *x) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] += (double)d[i];
}

static int lanes_per_row(long long nnz, long long n) {
  const double avg = n ? (double)nnz / (double)n : 1.0;
  int l = 4;
  while (l < 64 && l * 2 <= avg) l *= 2;
  return l;
}

template <class T>
static void spmv_launch(long long n, const long long *rp, const int *ci, const T *v, const T *x, T *y, double *dot,
                        int lpr, hipStream_t s) {
  const long long threads = n * lpr;
  const dim3 grid((unsigned)((threads + 255) / 256));
  switch (lpr) {
    case 4: hipLaunchKernelGGL((k_spmv<T, 4>), grid, dim3(256), 0, s, n, rp, ci, v, x, y, dot); break;
    case 8: hipLaunchKernelGGL((k_spmv<T, 8>), grid, dim3(256), 0, s, n, rp, ci, v, x, y, dot); break;
    case 16: hipLaunchKernelGGL((k_spmv<T, 16>), grid, dim3(256), 0, s, n, rp, ci, v, x, y, dot); break;
    case 32: hipLaunchKernelGGL((k_spmv<T, 32>), grid, dim3(256), 0, s, n, rp, ci, v, x, y, dot); break;
    default: hipLaunchKernelGGL((k_spmv<T, 64>), grid, dim3(256), 0, s, n, rp, ci, v, x, y, dot); break;
  }
}

// Jacobi-PCG in type T on (rp, ci, v) for A x = b (b in FP64), x in T.  Scalars in sc (8 doubles).
// Returns iterations; *relres = ||r||/||b|| at exit (as tracked by the recurrence).
template <class T>
static int pcg_core(Ctx *c, long long n, const long long *rp, const int *ci, const T *v, const double *b, T *x,
                    T *r, T *p, T *q, T *dinv, double *sc, double bnorm, double tol, int maxiter, int check, int lpr,
                    hipStream_t s, int *iters_out, double *relres, bool *converged) {
  const dim3 grid((unsigned)((n + 255) / 256));
  hipLaunchKernelGGL(k_zero_n, dim3(1), dim3(64), 0, s, sc, 8);
  spmv_launch<T>(n, rp, ci, v, x, r, nullptr, lpr, s);  // r = A x
  hipLaunchKernelGGL((k_init<T>), grid, dim3(256), 0, s, n, b, r, p, dinv, sc);  // rz -> sc[0], rr -> sc[1]
  // move rr to its slot sc[2]
  double *hp = c->h_pinned;
  FH_HIP(hipMemcpyAsync(hp, sc, 2 * sizeof(double), hipMemcpyDeviceToHost, s));
  FH_HIP(hipStreamSynchronize(s));
  double rr = hp[1];
  FH_HIP(hipMemcpyAsync(sc + 2, sc + 1, sizeof(double), hipMemcpyDeviceToDevice, s));
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
    hipLaunchKernelGGL(k_zero_next, dim3(1), dim3(1), 0, s, sc, par);
    spmv_launch<T>(n, rp, ci, v, p, q, sc + 4 + par, lpr, s);
    hipLaunchKernelGGL((k_update_xr<T>), grid, dim3(256), 0, s, n, par, x, r, p, q, dinv, sc);
    hipLaunchKernelGGL((k_update_p<T>), grid, dim3(256), 0, s, n, par, p, r, dinv, sc);
    ++k;
    if (k % check == 0 || k == maxiter) {
      FH_HIP(hipMemcpyAsync(hp, sc + 2 + (1 - par), sizeof(double), hipMemcpyDeviceToHost, s));
      FH_HIP(hipStreamSynchronize(s));
      rr = hp[0];
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

__global__ void k_dot64(long long n, const double *a, const double *b, double *out) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  double v = i < n ? a[i] * b[i] : 0.0;
  v = wave_sum(v);
  if ((threadIdx.x & 63) == 0) unsafeAtomicAdd(out, v);
}

// r = b - A x (FP64), returns ||r||^2 into *out (device)
__global__ __launch_bounds__(256) void k_resid64(long long n, const double *b, const double *ax, double *r,
                                                 double *out) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  double v = 0.0;
  if (i < n) {
    const double ri = b[i] - ax[i];
    r[i] = ri;
    v = ri * ri;
  }
  v = wave_sum(v);
  if ((threadIdx.x & 63) == 0) unsafeAtomicAdd(out, v);
}

}  // namespace femhot

using namespace femhot;

extern "C" int femhot_spmv(femhot_ctx *h, const long long *row_ptr, const int *col_idx, const double *vals,
                           const double *x, double *y, void *stream) {
  FH_CHECK_ARG(h && row_ptr && col_idx && vals && x && y, FEMHOT_ERR_ARG, "null argument");
  Ctx *c = &h->c;
  int rc = ensure_pattern(c);
  if (rc) return rc;
  spmv_launch<double>(c->Ndof, row_ptr, col_idx, vals, x, y, nullptr, lanes_per_row(c->nnz, c->Ndof),
                      (hipStream_t)stream);
  FH_HIP(hipGetLastError());
  return FEMHOT_OK;
}

extern "C" int femhot_pcg(femhot_ctx *h, int mode, const long long *row_ptr, const int *col_idx, const double *vals,
                          const double *b, double *x, double tol, int maxiter, int *iters, double *rel_res,
                          void *stream) {
  FH_CHECK_ARG(h && row_ptr && col_idx && vals && b && x, FEMHOT_ERR_ARG, "null argument");
  FH_CHECK_ARG(mode == FEMHOT_F64 || mode == FEMHOT_F32, FEMHOT_ERR_ARG, "mode must be FEMHOT_F64 or FEMHOT_F32");
  FH_CHECK_ARG(tol > 0 && maxiter > 0, FEMHOT_ERR_ARG, "tol and maxiter must be positive");
  Ctx *c = &h->c;
  int rc = ensure_pattern(c);
  if (rc) return rc;
  hipStream_t s = (hipStream_t)stream;
  const long long n = c->Ndof, nnz = c->nnz;
  const int lpr = lanes_per_row(nnz, n);
  const dim3 grid((unsigned)((n + 255) / 256));
  const int check = 10;
  // workspace: scalars (16 doubles) + 5 FP64 vectors + (FP32 mode) 5 FP32 vectors + FP32 matrix
  const size_t vec64 = sizeof(double) * n, vec32 = sizeof(float) * n;
  size_t need = 256 + 5 * vec64 + 1024;
  if (mode == FEMHOT_F32) need += 5 * vec32 + sizeof(float) * nnz + 1024;
  rc = ensure_workspace(c, need);
  if (rc) return rc;
  char *ws = (char *)c->d_ws;
  auto take = [&](size_t bytes) {
    char *p = ws;
    ws += (bytes + 255) & ~size_t(255);
    return (void *)p;
  };
  double *sc = (double *)take(16 * sizeof(double));
  double *r64 = (double *)take(vec64), *p64 = (double *)take(vec64), *q64 = (double *)take(vec64);
  double *dinv64 = (double *)take(vec64), *ax64 = (double *)take(vec64);
  // ||b||
  hipLaunchKernelGGL(k_zero_n, dim3(1), dim3(64), 0, s, sc, 16);
  hipLaunchKernelGGL(k_dot64, grid, dim3(256), 0, s, n, b, b, sc + 8);
  FH_HIP(hipMemcpyAsync(c->h_pinned, sc + 8, sizeof(double), hipMemcpyDeviceToHost, s));
  FH_HIP(hipStreamSynchronize(s));
  const double bnorm = std::sqrt(c->h_pinned[0]);
  hipLaunchKernelGGL((k_diag_inv<double>), grid, dim3(256), 0, s, n, row_ptr, col_idx, vals, dinv64);
  int it = 0;
  double rel = 0.0;
  bool conv = false;
  if (mode == FEMHOT_F64) {
    rc = pcg_core<double>(c, n, row_ptr, col_idx, vals, b, x, r64, p64, q64, dinv64, sc, bnorm, tol, maxiter, check,
                          lpr, s, &it, &rel, &conv);
    if (rc) return rc;
  } else {
    float *v32 = (float *)take(sizeof(float) * nnz);
    float *x32 = (float *)take(vec32), *r32 = (float *)take(vec32), *p32 = (float *)take(vec32);
    float *q32 = (float *)take(vec32), *dinv32 = (float *)take(vec32);
    const unsigned cg = (unsigned)std::min<long long>((nnz + 255) / 256, 65536);
    hipLaunchKernelGGL((k_cast<float>), dim3(cg), dim3(256), 0, s, nnz, vals, v32);
    hipLaunchKernelGGL((k_diag_inv<float>), grid, dim3(256), 0, s, n, row_ptr, col_idx, vals, dinv32);
    // outer FP64 refinement: r = b - A x; solve A d = r in FP32 to inner_tol; x += d
    const double inner_tol = 1e-4;
    int outer = 0;
    while (true) {
      hipLaunchKernelGGL(k_zero_n, dim3(1), dim3(64), 0, s, sc + 8, 1);
      spmv_launch<double>(n, row_ptr, col_idx, vals, x, ax64, nullptr, lpr, s);
      hipLaunchKernelGGL(k_resid64, grid, dim3(256), 0, s, n, b, ax64, r64, sc + 8);
      FH_HIP(hipMemcpyAsync(c->h_pinned, sc + 8, sizeof(double), hipMemcpyDeviceToHost, s));
      FH_HIP(hipStreamSynchronize(s));
      rel = bnorm > 0 ? std::sqrt(c->h_pinned[0]) / bnorm : 0.0;
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
      rc = pcg_core<float>(c, n, row_ptr, col_idx, v32, r64, x32, r32, p32, q32, dinv32, sc, bnorm * rel,
                           std::max(inner_tol, 0.5 * tol / std::max(rel, 1e-300)), maxiter - it, check, lpr, s,
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
