// Shape sensitivities: g_x = (dR/dX)^T w, the VJP of the assembled residual with respect to the node
// coordinates, as the reference's adjoint shape derivative contracts it
// (fol/responses/fe_response.py:312-332: jacrev(ComputeElement[1], argnums=0) with the element adjoint,
// assembled per node component at :365-406).
//
// One lane per element.  The element residual routine is run on forward-mode dual numbers once per
// element coordinate (N x D directions, Gauss loop and physics exactly as in the residual kernels),
// each direction contracted with the element's cotangent and added into g_x.  Not a hot path: it is
// evaluated once per design iteration.
#include "femhot_dual.h"
#include "femhot_launch.h"

namespace femhot {

struct ShapeArgs {
  const double *coords;
  const int *conn;
  long long Ne;
  const void *field, *u, *w;
  void *gx;
  unsigned flags;
  PhysParams pp;
};

template <class R, int ET, int PH, int NGP>
__global__ __launch_bounds__(128) void k_shape_vjp(ShapeArgs A) {
  using DR = Dual<R>;
  using E = Elem<DR, ET, PH>;
  constexpr int N = E::N, D = E::D, P = E::P, NL = E::NL;
  const long long e = (long long)blockIdx.x * 128 + threadIdx.x;
  if (e >= A.Ne) return;
  int nodes[N];
  load_conn<N>(A.conn, e, nodes);
  R xr[N][D];
  {
    double x0[D];
#pragma unroll
    for (int i = 0; i < D; ++i) x0[i] = A.coords[(size_t)nodes[0] * 3 + i];
#pragma unroll
    for (int a = 0; a < N; ++a)
#pragma unroll
      for (int i = 0; i < D; ++i) xr[a][i] = R(A.coords[(size_t)nodes[a] * 3 + i] - x0[i]);
  }
  E el;
  const R *F = reinterpret_cast<const R *>(A.field);
  const R *U = reinterpret_cast<const R *>(A.u);
  const R *W = reinterpret_cast<const R *>(A.w);
  el.elemental = (A.flags & FEMHOT_ELEMENT_FIELD) != 0;
  if (el.elemental) {
    el.de = DR(F[e]);
#pragma unroll
    for (int a = 0; a < N; ++a) el.d[a] = el.de;
  } else {
    el.de = DR(0.0);
#pragma unroll
    for (int a = 0; a < N; ++a) el.d[a] = DR(F[nodes[a]]);
  }
  R wv[NL];
#pragma unroll
  for (int a = 0; a < N; ++a)
#pragma unroll
    for (int c = 0; c < P; ++c) {
      el.u[a * P + c] = DR(U[(size_t)nodes[a] * P + c]);
      wv[a * P + c] = W[(size_t)nodes[a] * P + c];
    }
  R *G = reinterpret_cast<R *>(A.gx);
#pragma unroll 1
  for (int dir = 0; dir < N * D; ++dir) {
    const int a = dir / D, i = dir - a * D;
    // seed d(x_b)/d(X_a,i): coordinates are relative to local node 0, x_b = X_b - X_0
#pragma unroll
    for (int b = 0; b < N; ++b)
#pragma unroll
      for (int j = 0; j < D; ++j) {
        const R seed = (j != i) ? R(0) : (a == 0 ? (b >= 1 ? R(-1) : R(0)) : (b == a ? R(1) : R(0)));
        el.x[b][j] = DR(xr[b][j], seed, 0);
      }
    DR r[NL], en;
    elem_residual<DR, ET, PH, NGP>(el, A.pp, r, en);
    R g = R(-0.0);
#pragma unroll
    for (int k = 0; k < NL; ++k) g += wv[k] * r[k].d;
    int na = nodes[0];
#pragma unroll
    for (int b = 1; b < N; ++b) na = (b == a) ? nodes[b] : na;
    unsafeAtomicAdd(G + (size_t)na * 3 + i, g);
  }
}

template <int ET> static int launch_shape(int dtype, int ph, int ngp, dim3 grid, hipStream_t s, const ShapeArgs &A) {
  return dispatch_r_ph_ngp(dtype, ph, ngp, [&](auto rp, auto phc, auto ngpc) -> int {
    using R = std::remove_pointer_t<decltype(rp)>;
    constexpr int PH = decltype(phc)::value, NGP = decltype(ngpc)::value;
    hipLaunchKernelGGL((k_shape_vjp<R, ET, PH, NGP>), grid, dim3(128), 0, s, A);
    FH_HIP(hipGetLastError());
    return FEMHOT_OK;
  });
}

}  // namespace femhot

using namespace femhot;

extern "C" int femhot_residual_shape_vjp(femhot_ctx *h, int dtype, const void *field, const void *u, const void *w,
                                         void *g_x, unsigned flags, void *stream) {
  FH_CHECK_ARG(h && field && u && w && g_x, FEMHOT_ERR_ARG, "null argument");
  FH_CHECK_ARG(dtype == FEMHOT_F32 || dtype == FEMHOT_F64, FEMHOT_ERR_ARG, "dtype must be FEMHOT_F32 or FEMHOT_F64");
  Ctx *c = &h->c;
  hipStream_t s = (hipStream_t)stream;
  int rc = dtype == FEMHOT_F32 ? fill<float>((float *)g_x, 3 * c->Nn, 3 * c->Nn, 1, 0.f, s)
                               : fill<double>((double *)g_x, 3 * c->Nn, 3 * c->Nn, 1, 0.0, s);
  if (rc) return rc;
  ShapeArgs A{};
  A.coords = c->d_coords64;
  A.conn = c->d_conn;
  A.Ne = c->Ne;
  A.field = field;
  A.u = u;
  A.w = w;
  A.gx = g_x;
  A.flags = flags & FEMHOT_ELEMENT_FIELD;
  A.pp = c->pp;
  const dim3 grid((unsigned)((c->Ne + 127) / 128));
  switch (c->et) {
    case QUAD4: return launch_shape<QUAD4>(dtype, c->ph, c->ngp, grid, s, A);
    case HEX8: return launch_shape<HEX8>(dtype, c->ph, c->ngp, grid, s, A);
    default: return launch_shape<TET4>(dtype, c->ph, c->ngp, grid, s, A);
  }
}

// ---------------------------------------------------------------------------------------------
// Gauss-point geometry of every element: the integration data a response integrand needs
// (fol/responses/fe_response.py:92-113 evaluates its formula at each Gauss point with weight w detJ,
// field N.d and state N.u; its shape gradient needs d(detJ)/dX = detJ dN/dx).
// ---------------------------------------------------------------------------------------------
namespace femhot {

struct GPDataArgs {
  const double *coords;
  const int *conn;
  long long Ne;
  void *N, *wdet, *dndx;
};

template <class R, int ET, int NGP>
__global__ __launch_bounds__(128) void k_gp_data(GPDataArgs A) {
  constexpr int N = ElemT<ET>::N, D = ElemT<ET>::D, G = RuleSize<ET, NGP>::G;
  const long long e = (long long)blockIdx.x * 128 + threadIdx.x;
  if (e >= A.Ne) return;
  int nodes[N];
  load_conn<N>(A.conn, e, nodes);
  R x[N][D];
  double x0[D];
#pragma unroll
  for (int i = 0; i < D; ++i) x0[i] = A.coords[(size_t)nodes[0] * 3 + i];
#pragma unroll
  for (int a = 0; a < N; ++a)
#pragma unroll
    for (int i = 0; i < D; ++i) x[a][i] = R(A.coords[(size_t)nodes[a] * 3 + i] - x0[i]);
  R *Wd = reinterpret_cast<R *>(A.wdet) + e * G;
  R *Dn = reinterpret_cast<R *>(A.dndx) + e * G * N * D;
  R *Nt = reinterpret_cast<R *>(A.N);
#pragma unroll 1
  for (int g = 0; g < G; ++g) {
    GP<R, ET> gp;
    gp_geometry<R, ET, NGP>(g, x, gp);
    if (Wd) Wd[g] = gp.wdet;
    if (Dn)
#pragma unroll
      for (int a = 0; a < N; ++a)
#pragma unroll
        for (int j = 0; j < D; ++j) Dn[(g * N + a) * D + j] = gp.DN[a][j];
    if (Nt && e == 0)
#pragma unroll
      for (int a = 0; a < N; ++a) Nt[g * N + a] = gp.Nv[a];
  }
}

template <int ET> static int launch_gp_data(int dtype, int ngp, dim3 grid, hipStream_t s, const GPDataArgs &A) {
  auto go = [&](auto rp) -> int {
    using R = std::remove_pointer_t<decltype(rp)>;
    if (ngp == 1)
      hipLaunchKernelGGL((k_gp_data<R, ET, 1>), grid, dim3(128), 0, s, A);
    else if (ngp == 2)
      hipLaunchKernelGGL((k_gp_data<R, ET, 2>), grid, dim3(128), 0, s, A);
    else
      hipLaunchKernelGGL((k_gp_data<R, ET, 3>), grid, dim3(128), 0, s, A);
    FH_HIP(hipGetLastError());
    return FEMHOT_OK;
  };
  return dtype == FEMHOT_F32 ? go((float *)nullptr) : go((double *)nullptr);
}

}  // namespace femhot

extern "C" int femhot_gauss_point_data(femhot_ctx *h, int dtype, void *N, void *wdetJ, void *dNdx, void *stream) {
  FH_CHECK_ARG(h, FEMHOT_ERR_ARG, "null context");
  FH_CHECK_ARG(dtype == FEMHOT_F32 || dtype == FEMHOT_F64, FEMHOT_ERR_ARG, "dtype must be FEMHOT_F32 or FEMHOT_F64");
  FH_CHECK_ARG(N || wdetJ || dNdx, FEMHOT_ERR_ARG, "no output requested");
  Ctx *c = &h->c;
  GPDataArgs A{};
  A.coords = c->d_coords64;
  A.conn = c->d_conn;
  A.Ne = c->Ne;
  A.N = N;
  A.wdet = wdetJ;
  A.dndx = dNdx;
  const dim3 grid((unsigned)((c->Ne + 127) / 128));
  hipStream_t s = (hipStream_t)stream;
  switch (c->et) {
    case QUAD4: return launch_gp_data<QUAD4>(dtype, c->ngp, grid, s, A);
    case HEX8: return launch_gp_data<HEX8>(dtype, c->ngp, grid, s, A);
    default: return launch_gp_data<TET4>(dtype, c->ngp, grid, s, A);
  }
}
