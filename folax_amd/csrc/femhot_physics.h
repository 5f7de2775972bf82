// Element routines of the three physics on register-resident element data.
//   residual      r_e and the reference element energy (thermal T.r, mechanical u.r,
//                 Neo-Hookean Ee)                           thermal.py:28-49, mechanical.py:98-117,
//                                                           mechanical_neohooke.py:243-275
//   vjp           g_u = (dr_e/du)^T w, g_d = (dr_e/dd)^T w  (fe_response.py:434-453 contracts the
//                 same jacobian), recomputing Gauss-point quantities
//   egrad         potential Pi_e with g_u = dPi/du (= r_e) and g_d = dPi/dd
//   tangent_row   row (a, c) of the element tangent (Se / Ke of the reference)
// All element matrices of these physics are symmetric (thermal Se is the Picard matrix), so a row
// is also a column.
#pragma once
#include "femhot_element.h"

namespace femhot {

template <int PH, int D> struct DofsPerNode { static constexpr int P = (PH == THERMAL) ? 1 : D; };

template <class R, int ET, int PH> struct Elem {
  static constexpr int N = ElemT<ET>::N, D = ElemT<ET>::D, P = DofsPerNode<PH, ElemT<ET>::D>::P, NL = N * P;
  R x[N][D];  // node coordinates relative to local node 0
  R d[N];     // nodal field values (unused when elemental)
  R de;       // per-element field value
  bool elemental;
  R u[NL];
};

template <class R, int N> __device__ __forceinline__ R field_gp(const R (&d)[N], R de, bool elemental, const R (&Nv)[N]) {
  if (elemental) return de;
  R s = R(-0.0);
#pragma unroll
  for (int a = 0; a < N; ++a) s += Nv[a] * d[a];
  return s;
}

// G[i][j] = sum_a v_(a,i) dN_a/dx_j
template <class R, int N, int D>
__device__ __forceinline__ void nodal_grad(const R (&v)[N * D], const R (&DN)[N][D], R (&G)[D][D]) {
#pragma unroll
  for (int i = 0; i < D; ++i)
#pragma unroll
    for (int j = 0; j < D; ++j) {
      R s = R(-0.0);
#pragma unroll
      for (int a = 0; a < N; ++a) s += v[a * D + i] * DN[a][j];
      G[i][j] = s;
    }
}

// Linear-elastic stress tensor (unit field value) from a displacement gradient.
// 3-D Voigt [xx,yy,zz,xy,yz,xz] with engineering shears (mechanical.py:46-58, 63-82);
// 2-D plane stress [xx,yy,xy] (mechanical.py:37-44, 60-61).
template <class R, int D>
__device__ __forceinline__ void lin_stress(const R (&G)[D][D], const PhysParams &pp, R (&S)[D][D]) {
  if constexpr (D == 3) {
    const R e0 = G[0][0], e1 = G[1][1], e2 = G[2][2];
    const R c2 = R(pp.c2), c3 = R(pp.c3), c4 = R(pp.c4);
    S[0][0] = c2 * e0 + c3 * (e1 + e2);
    S[1][1] = c2 * e1 + c3 * (e0 + e2);
    S[2][2] = c2 * e2 + c3 * (e0 + e1);
    S[0][1] = S[1][0] = c4 * (G[0][1] + G[1][0]);
    S[1][2] = S[2][1] = c4 * (G[1][2] + G[2][1]);
    S[0][2] = S[2][0] = c4 * (G[0][2] + G[2][0]);
  } else {
    const R e0 = G[0][0], e1 = G[1][1];
    S[0][0] = R(pp.f2) * e0 + R(pp.f2nu) * e1;
    S[1][1] = R(pp.f2nu) * e0 + R(pp.f2) * e1;
    S[0][1] = S[1][0] = R(pp.f2sh) * (G[0][1] + G[1][0]);
  }
}

template <class R, int D> __device__ __forceinline__ R ddot(const R (&A)[D][D], const R (&B)[D][D]) {
  R s = R(-0.0);
#pragma unroll
  for (int i = 0; i < D; ++i)
#pragma unroll
    for (int j = 0; j < D; ++j) s += A[i][j] * B[i][j];
  return s;
}

// r_(a,i) += f * sum_j S_ij dN_a,j
template <class R, int N, int D>
__device__ __forceinline__ void add_nodal_force(const R (&S)[D][D], const R (&DN)[N][D], R f, R (&r)[N * D]) {
#pragma unroll
  for (int a = 0; a < N; ++a)
#pragma unroll
    for (int i = 0; i < D; ++i) {
      R s = R(-0.0);
#pragma unroll
      for (int j = 0; j < D; ++j) s += S[i][j] * DN[a][j];
      r[a * D + i] += f * s;
    }
}

// Gradient G[i][j] = sum_a v_(a,i) dN_a/dx_j.  For TET4 dN_0 = -(dN_1 + dN_2 + dN_3), so G is formed
// from the differences v_a - v_0 of nodes 1..3 and node 0's gradient is never needed.
template <class R, int ET, int N, int D>
__device__ __forceinline__ void elem_grad(const R (&v)[N * D], const R (&DN)[N][D], R (&G)[D][D]) {
  if constexpr (ET == TET4) {
    R dv[3][D];
#pragma unroll
    for (int a = 1; a < 4; ++a)
#pragma unroll
      for (int i = 0; i < D; ++i) dv[a - 1][i] = v[a * D + i] - v[i];
#pragma unroll
    for (int i = 0; i < D; ++i)
#pragma unroll
      for (int j = 0; j < D; ++j) G[i][j] = (dv[0][i] * DN[1][j] + dv[1][i] * DN[2][j]) + dv[2][i] * DN[3][j];
  } else {
    nodal_grad<R, N, D>(v, DN, G);
  }
}

// r_(a,i) += sum_j Sf_ij dN_a,j for a stress already multiplied by its Gauss factor.  For TET4 node 0
// receives minus the sum of nodes 1..3 (dN_0 = -(dN_1 + dN_2 + dN_3)).
template <class R, int ET, int N, int D>
__device__ __forceinline__ void add_scaled_nodal_force(const R (&Sf)[D][D], const R (&DN)[N][D], R (&r)[N * D]) {
  if constexpr (ET == TET4) {
    R t[3][D];
#pragma unroll
    for (int a = 1; a < 4; ++a)
#pragma unroll
      for (int i = 0; i < D; ++i) {
        R s = Sf[i][0] * DN[a][0];
#pragma unroll
        for (int j = 1; j < D; ++j) s += Sf[i][j] * DN[a][j];
        t[a - 1][i] = s;
        r[a * D + i] += s;
      }
#pragma unroll
    for (int i = 0; i < D; ++i) r[i] -= (t[0][i] + t[1][i]) + t[2][i];
  } else {
#pragma unroll
    for (int a = 0; a < N; ++a)
#pragma unroll
      for (int i = 0; i < D; ++i) {
        R s = Sf[i][0] * DN[a][0];
#pragma unroll
        for (int j = 1; j < D; ++j) s += Sf[i][j] * DN[a][j];
        r[a * D + i] += s;
      }
  }
}

// f * S for a symmetric S (each distinct entry scaled once).
template <class R, int D> __device__ __forceinline__ void scale_sym(const R (&S)[D][D], R f, R (&Sf)[D][D]) {
#pragma unroll
  for (int i = 0; i < D; ++i)
#pragma unroll
    for (int j = i; j < D; ++j) Sf[i][j] = Sf[j][i] = f * S[i][j];
}

template <class R, int D> __device__ __forceinline__ void scale_full(const R (&S)[D][D], R f, R (&Sf)[D][D]) {
#pragma unroll
  for (int i = 0; i < D; ++i)
#pragma unroll
    for (int j = 0; j < D; ++j) Sf[i][j] = f * S[i][j];
}

// r_(a,i) += f * sum_j M_ij dN_a,j for a general (non-symmetric) M: scaled once and node 0 taken from
// the others for TET4; elsewhere the factor is applied per entry (no scaled copy of M in registers).
template <class R, int ET, int N, int D>
__device__ __forceinline__ void add_nodal_force_f(const R (&M)[D][D], const R (&DN)[N][D], R f, R (&r)[N * D]) {
  if constexpr (ET == TET4) {
    R Mf[D][D];
    scale_full<R, D>(M, f, Mf);
    add_scaled_nodal_force<R, ET, N, D>(Mf, DN, r);
  } else {
    add_nodal_force<R, N, D>(M, DN, f, r);
  }
}

// Unroll factor of the Gauss-point loop: full for rules of up to 8 points, except FP64 Neo-Hookean,
// whose unrolled HEX8 2x2x2 element routine spills heavily (rolled: 12 spilled registers instead of
// 119 in the 256-thread chunk kernel, 11 % faster); FP32 Neo-Hookean is faster unrolled.
template <class R, int PH, int G> constexpr int gp_unroll() {
  return (G <= 8 && !(PH == NEOHOOKE && sizeof(R) == 8)) ? G : 1;
}

// A : S for a symmetric S: diagonal terms plus the symmetrised off-diagonal pairs.
template <class R, int D> __device__ __forceinline__ R ddot_sym(const R (&A)[D][D], const R (&S)[D][D]) {
  R s = A[0][0] * S[0][0];
#pragma unroll
  for (int i = 1; i < D; ++i) s += A[i][i] * S[i][i];
#pragma unroll
  for (int i = 0; i < D; ++i)
#pragma unroll
    for (int j = i + 1; j < D; ++j) s += (A[i][j] + A[j][i]) * S[i][j];
  return s;
}

template <class R, int N, int D>
__device__ __forceinline__ void nh_grad_u(const R (&u)[N * D], const R (&DN)[N][D], R (&H)[D][D]) {
  nodal_grad<R, N, D>(u, DN, H);
}

// ------------------------------------------------------------------------------------------
// QUAD4 thermal in bilinear modes.  With the node signs of quadrilateral_2d_4.py:17-66 (s_xi, s_eta of
// node a), sum_a v_a dN_a/dxi = 1/4 (m_xi + m_xieta eta) and sum_a v_a dN_a/deta = 1/4 (m_eta + m_xieta
// xi) for the moments m_w = sum_a w_a v_a over the sign products (0 = 1, 1 = xi, 2 = eta, 3 = xi eta).
// The moments of the coordinates, the temperature and the cotangent are formed once per element; the
// 1/4 cancels between J and the gradients, and the nodal terms accumulate as moments spread to the
// nodes once (as for HEX8 below).
// ------------------------------------------------------------------------------------------
template <class R> __device__ __forceinline__ void quad4_moments(const R (&v)[4], R (&m)[4]) {
  // node a = (s_xi, s_eta): 0 (--), 1 (+-), 2 (++), 3 (-+)
  const R s0 = v[0] + v[1], d0 = v[1] - v[0], s1 = v[3] + v[2], d1 = v[2] - v[3];
  m[0] = s0 + s1;
  m[2] = s1 - s0;
  m[1] = d0 + d1;
  m[3] = d1 - d0;
}

// v_a = M_0 + s_xi M_1 + s_eta M_2 + s_xi s_eta M_3
template <class R> __device__ __forceinline__ void quad4_spread(const R (&M)[4], R (&v)[4]) {
  const R lo0 = M[0] - M[2], lo1 = M[1] - M[3], hi0 = M[0] + M[2], hi1 = M[1] + M[3];
  v[0] = lo0 - lo1;
  v[1] = lo0 + lo1;
  v[2] = hi0 + hi1;
  v[3] = hi0 - hi1;
}

// Gauss-point geometry of a QUAD4 from its coordinate moments: inv(4 J) and w det J.
template <class R>
__device__ __forceinline__ void quad4_modal_geometry(const R (&mx)[2][4], R X, R Y, double w, R (&inv)[2][2],
                                                     R &wdet) {
  const R j00 = mx[0][1] + mx[0][3] * Y, j01 = mx[0][2] + mx[0][3] * X;
  const R j10 = mx[1][1] + mx[1][3] * Y, j11 = mx[1][2] + mx[1][3] * X;
  const R det4 = j00 * j11 - j01 * j10;
  const R id = recip(det4);
  inv[0][0] = j11 * id;
  inv[0][1] = -j01 * id;
  inv[1][0] = -j10 * id;
  inv[1][1] = j00 * id;
  wdet = R(w * (1.0 / 16.0)) * det4;
}

// grad v at the Gauss point from the moments of v: (4 dv/dxi) . inv(4 J)
template <class R>
__device__ __forceinline__ void quad4_modal_grad(const R (&m)[4], R X, R Y, const R (&inv)[2][2], R (&g)[2]) {
  const R gx = m[1] + m[3] * Y, gy = m[2] + m[3] * X;
#pragma unroll
  for (int j = 0; j < 2; ++j) g[j] = gx * inv[0][j] + gy * inv[1][j];
}

// M += the moments of f dN_a/dx . q (the flux term of the nodal vector): Q = f inv(4 J) q
template <class R>
__device__ __forceinline__ void quad4_modal_flux(R f, const R (&q)[2], const R (&inv)[2][2], R X, R Y, R (&M)[4]) {
  const R Q0 = f * (inv[0][0] * q[0] + inv[0][1] * q[1]);
  const R Q1 = f * (inv[1][0] * q[0] + inv[1][1] * q[1]);
  M[1] += Q0;
  M[2] += Q1;
  M[3] += Q0 * Y + Q1 * X;
}

// Residual and reference energy T . r of a QUAD4 thermal element (elem_residual's quantities).
template <class R, int NGP>
__device__ __forceinline__ void elem_residual_quad4_thermal(const Elem<R, QUAD4, THERMAL> &el, const PhysParams &pp,
                                                            R (&r)[4], R &energy) {
  constexpr int N = 4, G = RuleSize<QUAD4, NGP>::G, UNR = gp_unroll<R, THERMAL, G>();
  R mx[2][4], mu[4], M[4] = {R(0), R(0), R(0), R(0)};
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    const R col[4] = {el.x[0][i], el.x[1][i], el.x[2][i], el.x[3][i]};
    quad4_moments<R>(col, mx[i]);
  }
  quad4_moments<R>(el.u, mu);
#pragma unroll UNR
  for (int g = 0; g < G; ++g) {
    double xi, eta, zeta, w;
    gauss_point<QUAD4, NGP>(g, xi, eta, zeta, w);
    R Nv[N], dNl[N][2];
    shape<QUAD4, R>(xi, eta, zeta, Nv, dNl);
    const R X = R(xi), Y = R(eta);
    const R mode[4] = {R(1), X, Y, R(xi * eta)};  // N_a = 1/4 sum_w w_a mode[w]
    R inv[2][2], wdet;
    quad4_modal_geometry<R>(mx, X, Y, w, inv, wdet);
    const R eg = field_gp<R, N>(el.d, el.de, el.elemental, Nv);
    R Tg = R(-0.0);
#pragma unroll
    for (int m = 0; m < 4; ++m) Tg += mu[m] * mode[m];
    Tg = Tg * R(0.25);
    R gT[2];
    quad4_modal_grad<R>(mu, X, Y, inv, gT);
    const R k = eg * (R(1) + R(pp.beta) * powc(Tg, R(pp.c)));
    quad4_modal_flux<R>(wdet * k, gT, inv, X, Y, M);
    if (pp.has_body) {
      const R fb = wdet * R(pp.body[0]) * R(0.25);
#pragma unroll
      for (int m = 0; m < 4; ++m) M[m] -= fb * mode[m];
    }
  }
  R en = R(-0.0);
#pragma unroll
  for (int m = 0; m < 4; ++m) en += mu[m] * M[m];
  quad4_spread<R>(M, r);
  energy = en;
}

// VJP of a QUAD4 thermal element (elem_vjp's quantities): gu = (dr/dT)^T w (Picard: the k(T) term
// dropped), gd = (dr/dk0)^T w.
template <class R, int NGP>
__device__ __forceinline__ void elem_vjp_quad4_thermal(const Elem<R, QUAD4, THERMAL> &el, const R (&wv)[4],
                                                       const PhysParams &pp, bool picard, R (&gu)[4], R (&gd)[4]) {
  constexpr int N = 4, G = RuleSize<QUAD4, NGP>::G, UNR = gp_unroll<R, THERMAL, G>();
  R mx[2][4], mu[4], mw[4], M[4] = {R(0), R(0), R(0), R(0)};
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    const R col[4] = {el.x[0][i], el.x[1][i], el.x[2][i], el.x[3][i]};
    quad4_moments<R>(col, mx[i]);
  }
  quad4_moments<R>(el.u, mu);
  quad4_moments<R>(wv, mw);
#pragma unroll
  for (int a = 0; a < N; ++a) gd[a] = R(-0.0);
  const R bt = R(pp.beta);
#pragma unroll UNR
  for (int g = 0; g < G; ++g) {
    double xi, eta, zeta, w;
    gauss_point<QUAD4, NGP>(g, xi, eta, zeta, w);
    R Nv[N], dNl[N][2];
    shape<QUAD4, R>(xi, eta, zeta, Nv, dNl);
    const R X = R(xi), Y = R(eta);
    const R mode[4] = {R(1), X, Y, R(xi * eta)};
    R inv[2][2], wdet;
    quad4_modal_geometry<R>(mx, X, Y, w, inv, wdet);
    const R eg = field_gp<R, N>(el.d, el.de, el.elemental, Nv);
    R Tg = R(-0.0);
#pragma unroll
    for (int m = 0; m < 4; ++m) Tg += mu[m] * mode[m];
    Tg = Tg * R(0.25);
    R gT[2], gW[2];
    quad4_modal_grad<R>(mu, X, Y, inv, gT);
    quad4_modal_grad<R>(mw, X, Y, inv, gW);
    const R wGT = wdet * (gW[0] * gT[0] + gW[1] * gT[1]);  // w^T G_g T
    const R fac = R(1) + bt * powc(Tg, R(pp.c));
    quad4_modal_flux<R>(wdet * eg * fac, gW, inv, X, Y, M);
    if (!picard && bt != R(0)) {  // d k / d T: gu_a += wGT dk N_a
      const R dk = eg * bt * dpowc(Tg, R(pp.c));
      const R t = wGT * dk * R(0.25);
#pragma unroll
      for (int m = 0; m < 4; ++m) M[m] += t * mode[m];
    }
    const R q = wGT * fac;
    if (el.elemental)
      gd[0] += q;
    else {
#pragma unroll
      for (int a = 0; a < N; ++a) gd[a] += q * Nv[a];
    }
  }
  quad4_spread<R>(M, gu);
}

// ------------------------------------------------------------------------------------------
// residual + reference energy
// ------------------------------------------------------------------------------------------
template <class R, int PH, int NGP, bool RES>
__device__ __forceinline__ void elem_hex8_modal(const Elem<R, HEX8, PH> &el, const PhysParams &pp,
                                                R (&r)[Elem<R, HEX8, PH>::NL], R *gd, R &Pi);

template <class R, int ET, int PH, int NGP>
__device__ __forceinline__ void elem_residual_nodal(const Elem<R, ET, PH> &el, const PhysParams &pp,
                                                    R (&r)[Elem<R, ET, PH>::NL], R &energy);

template <class R, int ET, int PH, int NGP>
__device__ __forceinline__ void elem_residual(const Elem<R, ET, PH> &el, const PhysParams &pp,
                                              R (&r)[Elem<R, ET, PH>::NL], R &energy) {
  if constexpr (ET == HEX8)
    elem_hex8_modal<R, PH, NGP, true>(el, pp, r, nullptr, energy);
  else if constexpr (ET == QUAD4 && PH == THERMAL)
    elem_residual_quad4_thermal<R, NGP>(el, pp, r, energy);
  else
    elem_residual_nodal<R, ET, PH, NGP>(el, pp, r, energy);
}

template <class R, int ET, int PH, int NGP>
__device__ __forceinline__ void elem_residual_nodal(const Elem<R, ET, PH> &el, const PhysParams &pp,
                                                    R (&r)[Elem<R, ET, PH>::NL], R &energy) {
  using E = Elem<R, ET, PH>;
  constexpr int N = E::N, D = E::D, G = RuleSize<ET, NGP>::G, UNR = gp_unroll<R, PH, G>();
#pragma unroll
  for (int i = 0; i < E::NL; ++i) r[i] = R(-0.0);
  R en = R(-0.0);
#pragma unroll UNR
  for (int g = 0; g < G; ++g) {
    GP<R, ET> gp;
    gp_geometry<R, ET, NGP>(g, el.x, gp);
    const R eg = field_gp<R, N>(el.d, el.de, el.elemental, gp.Nv);
    if constexpr (PH == THERMAL) {
      R Tg = R(-0.0), gT[D];
#pragma unroll
      for (int a = 0; a < N; ++a) Tg += gp.Nv[a] * el.u[a];
#pragma unroll
      for (int j = 0; j < D; ++j) {
        R s = R(-0.0);
#pragma unroll
        for (int a = 0; a < N; ++a) s += gp.DN[a][j] * el.u[a];
        gT[j] = s;
      }
      const R k = eg * (R(1) + R(pp.beta) * powc(Tg, R(pp.c)));
      const R f = gp.wdet * k;
      const R b = pp.has_body ? gp.wdet * R(pp.body[0]) : R(-0.0);
#pragma unroll
      for (int a = 0; a < N; ++a) {
        R s = R(-0.0);
#pragma unroll
        for (int j = 0; j < D; ++j) s += gp.DN[a][j] * gT[j];
        r[a] += f * s - b * gp.Nv[a];
      }
    } else if constexpr (PH == MECH) {
      R Gu[D][D], S[D][D], Sf[D][D];
      elem_grad<R, ET, N, D>(el.u, gp.DN, Gu);
      lin_stress<R, D>(Gu, pp, S);
      scale_sym<R, D>(S, gp.wdet * eg, Sf);
      add_scaled_nodal_force<R, ET, N, D>(Sf, gp.DN, r);
      if (pp.has_body) {
#pragma unroll
        for (int a = 0; a < N; ++a)
#pragma unroll
          for (int i = 0; i < D; ++i) r[a * D + i] -= gp.wdet * gp.Nv[a] * R(pp.body[i]);
      }
    } else {
      R H[D][D];
      elem_grad<R, ET, N, D>(el.u, gp.DN, H);
      NHState<R, D> st;
      nh_state<R, D>(H, R(pp.kap1), R(pp.mu1), st);
      // P = F S ; r_(a,i) += wdet e sum_j P_ij dN_a,j
      R Pm[D][D];
#pragma unroll
      for (int i = 0; i < D; ++i)
#pragma unroll
        for (int j = 0; j < D; ++j) {
          R s = R(-0.0);
#pragma unroll
          for (int k = 0; k < D; ++k) s += st.F[i][k] * st.S[k][j];
          Pm[i][j] = s;
        }
      add_nodal_force_f<R, ET, N, D>(Pm, gp.DN, gp.wdet * eg, r);
      if (pp.has_body) {
#pragma unroll
        for (int a = 0; a < N; ++a)
#pragma unroll
          for (int i = 0; i < D; ++i) r[a * D + i] -= gp.wdet * gp.Nv[a] * R(pp.body[i]);
      }
      en += gp.wdet * eg * st.psi;
    }
  }
  if constexpr (PH != NEOHOOKE) {
#pragma unroll
    for (int i = 0; i < E::NL; ++i) en += el.u[i] * r[i];
  }
  energy = en;
}

// ------------------------------------------------------------------------------------------
// residual VJP
// ------------------------------------------------------------------------------------------
template <class R, int PH, int NGP>
__device__ __forceinline__ void elem_vjp_hex8(const Elem<R, HEX8, PH> &el, const R (&wv)[Elem<R, HEX8, PH>::NL],
                                              const PhysParams &pp, bool picard, R (&gu)[Elem<R, HEX8, PH>::NL],
                                              R (&gd)[8]);

template <class R, int ET, int PH, int NGP>
__device__ __forceinline__ void elem_vjp(const Elem<R, ET, PH> &el, const R (&w)[Elem<R, ET, PH>::NL],
                                         const PhysParams &pp, bool picard, R (&gu)[Elem<R, ET, PH>::NL],
                                         R (&gd)[Elem<R, ET, PH>::N]) {
  if constexpr (ET == QUAD4 && PH == THERMAL) {  // bilinear modes (above)
    elem_vjp_quad4_thermal<R, NGP>(el, w, pp, picard, gu, gd);
    return;
  } else if constexpr (ET == HEX8) {  // trilinear modes (below)
    elem_vjp_hex8<R, PH, NGP>(el, w, pp, picard, gu, gd);
    return;
  }
  using E = Elem<R, ET, PH>;
  constexpr int N = E::N, D = E::D, G = RuleSize<ET, NGP>::G, UNR = gp_unroll<R, PH, G>();
#pragma unroll
  for (int i = 0; i < E::NL; ++i) gu[i] = R(-0.0);
#pragma unroll
  for (int a = 0; a < N; ++a) gd[a] = R(-0.0);
#pragma unroll UNR
  for (int g = 0; g < G; ++g) {
    GP<R, ET> gp;
    gp_geometry<R, ET, NGP>(g, el.x, gp);
    const R eg = field_gp<R, N>(el.d, el.de, el.elemental, gp.Nv);
    R q;  // d(field at gp) weight for g_d
    if constexpr (PH == THERMAL) {
      R Tg = R(-0.0), gT[D], gW[D];
#pragma unroll
      for (int a = 0; a < N; ++a) Tg += gp.Nv[a] * el.u[a];
#pragma unroll
      for (int j = 0; j < D; ++j) {
        R s = R(0), t = R(-0.0);
#pragma unroll
        for (int a = 0; a < N; ++a) {
          s += gp.DN[a][j] * el.u[a];
          t += gp.DN[a][j] * w[a];
        }
        gT[j] = s;
        gW[j] = t;
      }
      R wGT = R(-0.0);
#pragma unroll
      for (int j = 0; j < D; ++j) wGT += gW[j] * gT[j];
      wGT *= gp.wdet;  // w^T G_g T
      const R bt = R(pp.beta);
      const R fac = R(1) + bt * powc(Tg, R(pp.c));
      const R f = gp.wdet * eg * fac;
#pragma unroll
      for (int a = 0; a < N; ++a) {
        R s = R(-0.0);
#pragma unroll
        for (int j = 0; j < D; ++j) s += gp.DN[a][j] * gW[j];
        gu[a] += f * s;
      }
      if (!picard && bt != R(0)) {
        const R dk = eg * bt * dpowc(Tg, R(pp.c));
#pragma unroll
        for (int a = 0; a < N; ++a) gu[a] += wGT * dk * gp.Nv[a];
      }
      q = wGT * fac;
    } else if constexpr (PH == MECH) {
      R Gu[D][D], Gw[D][D], Su[D][D], Sw[D][D], Swf[D][D];
      elem_grad<R, ET, N, D>(el.u, gp.DN, Gu);
      elem_grad<R, ET, N, D>(w, gp.DN, Gw);
      lin_stress<R, D>(Gu, pp, Su);
      lin_stress<R, D>(Gw, pp, Sw);
      scale_sym<R, D>(Sw, gp.wdet * eg, Swf);
      add_scaled_nodal_force<R, ET, N, D>(Swf, gp.DN, gu);
      q = gp.wdet * ddot_sym<R, D>(Gw, Su);
    } else {
      R H[D][D], dF[D][D];
      elem_grad<R, ET, N, D>(el.u, gp.DN, H);
      elem_grad<R, ET, N, D>(w, gp.DN, dF);
      NHState<R, D> st;
      nh_state<R, D>(H, R(pp.kap1), R(pp.mu1), st);
      // X = sym(F^T dF)
      R X[D][D];
#pragma unroll
      for (int i = 0; i < D; ++i)
#pragma unroll
        for (int j = 0; j < D; ++j) {
          R s = R(-0.0);
#pragma unroll
          for (int k = 0; k < D; ++k) s += st.F[k][i] * dF[k][j] + st.F[k][j] * dF[k][i];
          X[i][j] = R(0.5) * s;
        }
      R dS[D][D];
      nh_tangent_apply<R, D>(st, R(pp.kap1), R(pp.mu1), X, dS);
      // M = F dS + dF S
      R M[D][D];
#pragma unroll
      for (int i = 0; i < D; ++i)
#pragma unroll
        for (int j = 0; j < D; ++j) {
          R s = R(-0.0);
#pragma unroll
          for (int k = 0; k < D; ++k) s += st.F[i][k] * dS[k][j] + dF[i][k] * st.S[k][j];
          M[i][j] = s;
        }
      add_nodal_force_f<R, ET, N, D>(M, gp.DN, gp.wdet * eg, gu);
      q = gp.wdet * ddot_sym<R, D>(X, st.S);
    }
    if (el.elemental)
      gd[0] += q;
    else {
#pragma unroll
      for (int a = 0; a < N; ++a) gd[a] += q * gp.Nv[a];
    }
  }
}

// ------------------------------------------------------------------------------------------
// potential and its gradient
// ------------------------------------------------------------------------------------------
// ------------------------------------------------------------------------------------------
// HEX8 in trilinear modes.  With the node signs of hexahedra_3d_8.py:79-88 (s_xi, s_eta, s_zeta of
// node a), dN_a/dxi = 1/8 s_xi (1 + s_eta eta)(1 + s_zeta zeta), so sum_a v_a dN_a/dxi needs only the
// eight moments m_w = sum_a w_a v_a of v over the products w of the node signs:
//   8 dv/dxi   = m_xi   + m_xieta eta + m_xizeta zeta + m_xietazeta eta zeta   (and cyclically).
// The moments of the coordinates and of the solution are formed once per element (24 adds per
// component) instead of contracting 8 nodes at every Gauss point, the 1/8 cancels between J and the
// solution gradient, and the nodal forces f (P inv(J)^T)_ik dN_a/dxi_k are accumulated as moments too
// and spread to the nodes once at the end.  Mode index: 0 = 1, 1 = xi, 2 = eta, 3 = zeta, 4 = xi eta,
// 5 = xi zeta, 6 = eta zeta, 7 = xi eta zeta.
// ------------------------------------------------------------------------------------------
template <class R> __device__ __forceinline__ void hex8_moments(const R (&v)[8], R (&m)[8]) {
  // node a = (s_xi, s_eta, s_zeta): 0 (---), 1 (+--), 2 (++-), 3 (-+-), 4 (--+), 5 (+-+), 6 (+++), 7 (-++)
  const R s00 = v[0] + v[1], d00 = v[1] - v[0], s10 = v[3] + v[2], d10 = v[2] - v[3];
  const R s01 = v[4] + v[5], d01 = v[5] - v[4], s11 = v[7] + v[6], d11 = v[6] - v[7];
  const R ss0 = s00 + s10, sd0 = s10 - s00, ss1 = s01 + s11, sd1 = s11 - s01;
  const R ds0 = d00 + d10, dd0 = d10 - d00, ds1 = d01 + d11, dd1 = d11 - d01;
  m[0] = ss0 + ss1;
  m[3] = ss1 - ss0;
  m[2] = sd0 + sd1;
  m[6] = sd1 - sd0;
  m[1] = ds0 + ds1;
  m[5] = ds1 - ds0;
  m[4] = dd0 + dd1;
  m[7] = dd1 - dd0;
}

// v_a = sum_w w_a M_w (the transpose of hex8_moments)
template <class R> __device__ __forceinline__ void hex8_spread(const R (&M)[8], R (&v)[8]) {
  R t[2][4];  // [zeta sign][1, xi, eta, xi eta]
#pragma unroll
  for (int k = 0; k < 2; ++k) {
    const R sz = k ? R(1) : R(-1);
    t[k][0] = M[0] + sz * M[3];
    t[k][1] = M[1] + sz * M[5];
    t[k][2] = M[2] + sz * M[6];
    t[k][3] = M[4] + sz * M[7];
  }
  // (xi sign, eta sign) of nodes 0..3 (bottom) and 4..7 (top)
  const int sx[4] = {-1, 1, 1, -1}, sy[4] = {-1, -1, 1, 1};
#pragma unroll
  for (int k = 0; k < 2; ++k)
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      const R c0 = sy[a] > 0 ? t[k][0] + t[k][2] : t[k][0] - t[k][2];
      const R c1 = sy[a] > 0 ? t[k][1] + t[k][3] : t[k][1] - t[k][3];
      v[4 * k + a] = sx[a] > 0 ? c0 + c1 : c0 - c1;
    }
}

// 8 d v/d(xi, eta, zeta) at (X, Y, Z) from the moments of v
template <class R>
__device__ __forceinline__ void hex8_dxi(const R (&m)[8], R X, R Y, R Z, R XY, R XZ, R YZ, R (&g)[3]) {
  g[0] = ((m[1] + m[4] * Y) + m[5] * Z) + m[7] * YZ;
  g[1] = ((m[2] + m[4] * X) + m[6] * Z) + m[7] * XZ;
  g[2] = ((m[3] + m[5] * X) + m[6] * Y) + m[7] * XY;
}

// Unroll factor of the modal Gauss-point loop, by measurement on HEX8 48^3 x 16: as gp_unroll for
// thermal and linear elasticity; two points at a time for FP32 Neo-Hookean (fully unrolled, the
// chunk kernel needs 256 VGPRs and runs one wave per SIMD; 2 points: 217 VGPRs, two waves, 4 % faster
// than the nodal form); rolled for FP64 Neo-Hookean.
template <class R, int PH, int G> constexpr int hex8_unroll() {
  return PH != NEOHOOKE ? gp_unroll<R, PH, G>() : (sizeof(R) == 4 ? 2 : 1);
}

// One HEX8 element through its modes: with RES the residual and reference energy of elem_residual
// (the thermal conductivity k0 (1 + beta T^c) included; gd unused), else the potential and gradient of
// elem_egrad (gd = dPi/dd) -- the same quantities as the nodal forms below, summed in another order.
template <class R, int PH, int NGP, bool RES>
__device__ __forceinline__ void elem_hex8_modal(const Elem<R, HEX8, PH> &el, const PhysParams &pp,
                                                R (&r)[Elem<R, HEX8, PH>::NL], R *gd, R &Pi) {
  using E = Elem<R, HEX8, PH>;
  constexpr int N = 8, D = 3, P = E::P, G = RuleSize<HEX8, NGP>::G, UNR = hex8_unroll<R, PH, G>();
  R mx[D][8], mu[P][8], M[P][8];
#pragma unroll
  for (int i = 0; i < D; ++i) {
    R col[8];
#pragma unroll
    for (int a = 0; a < N; ++a) col[a] = el.x[a][i];
    hex8_moments<R>(col, mx[i]);
  }
#pragma unroll
  for (int c = 0; c < P; ++c) {
    R col[8];
#pragma unroll
    for (int a = 0; a < N; ++a) col[a] = el.u[a * P + c];
    hex8_moments<R>(col, mu[c]);
#pragma unroll
    for (int w = 0; w < 8; ++w) M[c][w] = R(0);
  }
  if constexpr (!RES) {
#pragma unroll
    for (int a = 0; a < N; ++a) gd[a] = R(-0.0);
  }
  R pi = R(-0.0);
#pragma unroll UNR
  for (int g = 0; g < G; ++g) {
    double xi, eta, zeta, w;
    gauss_point<HEX8, NGP>(g, xi, eta, zeta, w);
    R Nv[N], dNl[N][D];
    shape<HEX8, R>(xi, eta, zeta, Nv, dNl);
    const R X = R(xi), Y = R(eta), Z = R(zeta), XY = R(xi * eta), XZ = R(xi * zeta), YZ = R(eta * zeta);
    const R mode[8] = {R(1), X, Y, Z, XY, XZ, YZ, R(xi * eta * zeta)};  // N_a = 1/8 sum_w w_a mode[w]
    // 8 J (J[i][k] = dx_i/dxi_k), its inverse and determinant
    R J[D][D];
#pragma unroll
    for (int i = 0; i < D; ++i) hex8_dxi<R>(mx[i], X, Y, Z, XY, XZ, YZ, J[i]);
    const R c00 = J[1][1] * J[2][2] - J[1][2] * J[2][1];
    const R c01 = J[1][2] * J[2][0] - J[1][0] * J[2][2];
    const R c02 = J[1][0] * J[2][1] - J[1][1] * J[2][0];
    const R det8 = J[0][0] * c00 + J[0][1] * c01 + J[0][2] * c02;
    const R id = recip(det8);
    R inv[D][D];  // inverse of 8 J
    inv[0][0] = c00 * id;
    inv[1][0] = c01 * id;
    inv[2][0] = c02 * id;
    inv[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) * id;
    inv[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) * id;
    inv[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) * id;
    inv[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) * id;
    inv[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) * id;
    inv[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) * id;
    const R wdet = R(w * (1.0 / 512.0)) * det8;
    const R eg = field_gp<R, N>(el.d, el.de, el.elemental, Nv);
    // solution gradient Gu[c][j] = (8 du_c/dxi) . inv(8 J)
    R Gu[P][D];
#pragma unroll
    for (int c = 0; c < P; ++c) {
      R hx[D];
      hex8_dxi<R>(mu[c], X, Y, Z, XY, XZ, YZ, hx);
#pragma unroll
      for (int j = 0; j < D; ++j) Gu[c][j] = (hx[0] * inv[0][j] + hx[1] * inv[1][j]) + hx[2] * inv[2][j];
    }
    // Q[c][k] = sum_j A[c][j] inv[k][j] for the flux / stress A of the physics (scaled by its factor)
    R Q[P][D], q;
    if constexpr (PH == THERMAL) {  // egrad: beta = 0 (checked on the host)
      R k = eg;
      if constexpr (RES) {
        R Tg = R(-0.0);
#pragma unroll
        for (int m = 0; m < 8; ++m) Tg += mu[0][m] * mode[m];
        k = eg * (R(1) + R(pp.beta) * powc(Tg * R(0.125), R(pp.c)));
      }
      const R f = wdet * k;
#pragma unroll
      for (int k = 0; k < D; ++k) Q[0][k] = f * ((inv[k][0] * Gu[0][0] + inv[k][1] * Gu[0][1]) + inv[k][2] * Gu[0][2]);
      q = R(0.5) * wdet * ((Gu[0][0] * Gu[0][0] + Gu[0][1] * Gu[0][1]) + Gu[0][2] * Gu[0][2]);
    } else if constexpr (PH == MECH) {
      R S[D][D], Sf[D][D];
      lin_stress<R, D>(Gu, pp, S);
      scale_sym<R, D>(S, wdet * eg, Sf);
#pragma unroll
      for (int i = 0; i < D; ++i)
#pragma unroll
        for (int k = 0; k < D; ++k) Q[i][k] = (Sf[i][0] * inv[k][0] + Sf[i][1] * inv[k][1]) + Sf[i][2] * inv[k][2];
      q = R(0.5) * wdet * ddot_sym<R, D>(Gu, S);
    } else {
      NHState<R, D> st;
      nh_state<R, D>(Gu, R(pp.kap1), R(pp.mu1), st);
      const R f = wdet * eg;
      R Pf[D][D];
#pragma unroll
      for (int i = 0; i < D; ++i)
#pragma unroll
        for (int j = 0; j < D; ++j) Pf[i][j] = f * ((st.F[i][0] * st.S[0][j] + st.F[i][1] * st.S[1][j]) + st.F[i][2] * st.S[2][j]);
#pragma unroll
      for (int i = 0; i < D; ++i)
#pragma unroll
        for (int k = 0; k < D; ++k) Q[i][k] = (Pf[i][0] * inv[k][0] + Pf[i][1] * inv[k][1]) + Pf[i][2] * inv[k][2];
      q = wdet * st.psi;
    }
#pragma unroll
    for (int c = 0; c < P; ++c) {
      M[c][1] += Q[c][0];
      M[c][2] += Q[c][1];
      M[c][3] += Q[c][2];
      M[c][4] += Q[c][0] * Y + Q[c][1] * X;
      M[c][5] += Q[c][0] * Z + Q[c][2] * X;
      M[c][6] += Q[c][1] * Z + Q[c][2] * Y;
      M[c][7] += (Q[c][0] * YZ + Q[c][1] * XZ) + Q[c][2] * XY;
    }
    R work = R(-0.0);
    if (pp.has_body) {  // body force b: r_a -= wdet N_a b, work = wdet b . u(x_g)
#pragma unroll
      for (int c = 0; c < P; ++c) {
        const R b = R(PH == THERMAL ? pp.body[0] : pp.body[c]);
        const R fb = wdet * b * R(0.125);
        R uc = R(-0.0);
#pragma unroll
        for (int k = 0; k < 8; ++k) {
          M[c][k] -= fb * mode[k];
          uc += mu[c][k] * mode[k];
        }
        work += fb * uc;
      }
    }
    if constexpr (RES) {
      if constexpr (PH == NEOHOOKE) pi += eg * q;  // the reference energy: sum of wdet e psi
    } else {
      pi += eg * q - work;
      if (el.elemental)
        gd[0] += q;
      else {
#pragma unroll
        for (int a = 0; a < N; ++a) gd[a] += q * Nv[a];
      }
    }
  }
  if constexpr (RES && PH != NEOHOOKE) {  // the reference energy u . r = sum_w (moment w of u) M_w
#pragma unroll
    for (int c = 0; c < P; ++c)
#pragma unroll
      for (int m = 0; m < 8; ++m) pi += mu[c][m] * M[c][m];
  }
#pragma unroll
  for (int c = 0; c < P; ++c) {
    R v[8];
    hex8_spread<R>(M[c], v);
#pragma unroll
    for (int a = 0; a < N; ++a) r[a * P + c] = v[a];
  }
  Pi = pi;
}

// VJP of one HEX8 element through its modes (elem_vjp's quantities, summed in another order): the
// moments of the cotangent give its gradient as those of the solution give H, and every term of the
// form A_ij dN_a/dx_j accumulates as moments spread to the nodes once.
template <class R, int PH, int NGP>
__device__ __forceinline__ void elem_vjp_hex8(const Elem<R, HEX8, PH> &el, const R (&wv)[Elem<R, HEX8, PH>::NL],
                                              const PhysParams &pp, bool picard, R (&gu)[Elem<R, HEX8, PH>::NL],
                                              R (&gd)[8]) {
  using E = Elem<R, HEX8, PH>;
  constexpr int N = 8, D = 3, P = E::P, G = RuleSize<HEX8, NGP>::G, UNR = hex8_unroll<R, PH, G>();
  R mx[D][8], mu[P][8], mw[P][8], M[P][8];
#pragma unroll
  for (int i = 0; i < D; ++i) {
    R col[8];
#pragma unroll
    for (int a = 0; a < N; ++a) col[a] = el.x[a][i];
    hex8_moments<R>(col, mx[i]);
  }
#pragma unroll
  for (int c = 0; c < P; ++c) {
    R cu[8], cw[8];
#pragma unroll
    for (int a = 0; a < N; ++a) {
      cu[a] = el.u[a * P + c];
      cw[a] = wv[a * P + c];
    }
    hex8_moments<R>(cu, mu[c]);
    hex8_moments<R>(cw, mw[c]);
#pragma unroll
    for (int m = 0; m < 8; ++m) M[c][m] = R(0);
  }
#pragma unroll
  for (int a = 0; a < N; ++a) gd[a] = R(-0.0);
#pragma unroll UNR
  for (int g = 0; g < G; ++g) {
    double xi, eta, zeta, w;
    gauss_point<HEX8, NGP>(g, xi, eta, zeta, w);
    R Nv[N], dNl[N][D];
    shape<HEX8, R>(xi, eta, zeta, Nv, dNl);
    const R X = R(xi), Y = R(eta), Z = R(zeta), XY = R(xi * eta), XZ = R(xi * zeta), YZ = R(eta * zeta);
    const R mode[8] = {R(1), X, Y, Z, XY, XZ, YZ, R(xi * eta * zeta)};
    R J[D][D];
#pragma unroll
    for (int i = 0; i < D; ++i) hex8_dxi<R>(mx[i], X, Y, Z, XY, XZ, YZ, J[i]);
    const R c00 = J[1][1] * J[2][2] - J[1][2] * J[2][1];
    const R c01 = J[1][2] * J[2][0] - J[1][0] * J[2][2];
    const R c02 = J[1][0] * J[2][1] - J[1][1] * J[2][0];
    const R det8 = J[0][0] * c00 + J[0][1] * c01 + J[0][2] * c02;
    const R id = recip(det8);
    R inv[D][D];  // inverse of 8 J
    inv[0][0] = c00 * id;
    inv[1][0] = c01 * id;
    inv[2][0] = c02 * id;
    inv[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) * id;
    inv[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) * id;
    inv[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) * id;
    inv[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) * id;
    inv[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) * id;
    inv[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) * id;
    const R wdet = R(w * (1.0 / 512.0)) * det8;
    const R eg = field_gp<R, N>(el.d, el.de, el.elemental, Nv);
    R Gu[P][D], Gw[P][D];
#pragma unroll
    for (int c = 0; c < P; ++c) {
      R hu[D], hw[D];
      hex8_dxi<R>(mu[c], X, Y, Z, XY, XZ, YZ, hu);
      hex8_dxi<R>(mw[c], X, Y, Z, XY, XZ, YZ, hw);
#pragma unroll
      for (int j = 0; j < D; ++j) {
        Gu[c][j] = (hu[0] * inv[0][j] + hu[1] * inv[1][j]) + hu[2] * inv[2][j];
        Gw[c][j] = (hw[0] * inv[0][j] + hw[1] * inv[1][j]) + hw[2] * inv[2][j];
      }
    }
    R A[P][D], q;  // gu_(a,c) += sum_j A[c][j] dN_a/dx_j
    if constexpr (PH == THERMAL) {
      R Tg = R(-0.0);
#pragma unroll
      for (int m = 0; m < 8; ++m) Tg += mu[0][m] * mode[m];
      Tg = Tg * R(0.125);
      const R wGT = wdet * ((Gw[0][0] * Gu[0][0] + Gw[0][1] * Gu[0][1]) + Gw[0][2] * Gu[0][2]);
      const R bt = R(pp.beta);
      const R fac = R(1) + bt * powc(Tg, R(pp.c));
      const R f = wdet * eg * fac;
#pragma unroll
      for (int j = 0; j < D; ++j) A[0][j] = f * Gw[0][j];
      if (!picard && bt != R(0)) {  // d k / d T: gu_a += wGT dk N_a
        const R t = wGT * (eg * bt * dpowc(Tg, R(pp.c))) * R(0.125);
#pragma unroll
        for (int m = 0; m < 8; ++m) M[0][m] += t * mode[m];
      }
      q = wGT * fac;
    } else if constexpr (PH == MECH) {
      R Su[D][D], Sw[D][D];
      lin_stress<R, D>(Gu, pp, Su);
      lin_stress<R, D>(Gw, pp, Sw);
      scale_sym<R, D>(Sw, wdet * eg, A);
      q = wdet * ddot_sym<R, D>(Gw, Su);
    } else {
      NHState<R, D> st;
      nh_state<R, D>(Gu, R(pp.kap1), R(pp.mu1), st);
      R Xs[D][D];  // sym(F^T dF)
#pragma unroll
      for (int i = 0; i < D; ++i)
#pragma unroll
        for (int j = 0; j < D; ++j) {
          R s2 = R(-0.0);
#pragma unroll
          for (int k = 0; k < D; ++k) s2 += st.F[k][i] * Gw[k][j] + st.F[k][j] * Gw[k][i];
          Xs[i][j] = R(0.5) * s2;
        }
      R dS[D][D];
      nh_tangent_apply<R, D>(st, R(pp.kap1), R(pp.mu1), Xs, dS);
      const R f = wdet * eg;
#pragma unroll
      for (int i = 0; i < D; ++i)
#pragma unroll
        for (int j = 0; j < D; ++j) {
          R s2 = R(-0.0);
#pragma unroll
          for (int k = 0; k < D; ++k) s2 += st.F[i][k] * dS[k][j] + Gw[i][k] * st.S[k][j];
          A[i][j] = f * s2;
        }
      q = wdet * ddot_sym<R, D>(Xs, st.S);
    }
#pragma unroll
    for (int c = 0; c < P; ++c) {
      R Q[D];
#pragma unroll
      for (int k = 0; k < D; ++k) Q[k] = (A[c][0] * inv[k][0] + A[c][1] * inv[k][1]) + A[c][2] * inv[k][2];
      M[c][1] += Q[0];
      M[c][2] += Q[1];
      M[c][3] += Q[2];
      M[c][4] += Q[0] * Y + Q[1] * X;
      M[c][5] += Q[0] * Z + Q[2] * X;
      M[c][6] += Q[1] * Z + Q[2] * Y;
      M[c][7] += (Q[0] * YZ + Q[1] * XZ) + Q[2] * XY;
    }
    if (el.elemental)
      gd[0] += q;
    else {
#pragma unroll
      for (int a = 0; a < N; ++a) gd[a] += q * Nv[a];
    }
  }
#pragma unroll
  for (int c = 0; c < P; ++c) {
    R v[8];
    hex8_spread<R>(M[c], v);
#pragma unroll
    for (int a = 0; a < N; ++a) gu[a * P + c] = v[a];
  }
}

template <class R, int ET, int PH, int NGP>
__device__ __forceinline__ void elem_egrad_generic(const Elem<R, ET, PH> &el, const PhysParams &pp,
                                                   R (&r)[Elem<R, ET, PH>::NL], R (&gd)[Elem<R, ET, PH>::N], R &Pi);

template <class R, int ET, int PH, int NGP>
__device__ __forceinline__ void elem_egrad(const Elem<R, ET, PH> &el, const PhysParams &pp,
                                           R (&r)[Elem<R, ET, PH>::NL], R (&gd)[Elem<R, ET, PH>::N], R &Pi) {
  if constexpr (ET == HEX8)
    elem_hex8_modal<R, PH, NGP, false>(el, pp, r, gd, Pi);
  else
    elem_egrad_generic<R, ET, PH, NGP>(el, pp, r, gd, Pi);
}

template <class R, int ET, int PH, int NGP>
__device__ __forceinline__ void elem_egrad_generic(const Elem<R, ET, PH> &el, const PhysParams &pp,
                                                   R (&r)[Elem<R, ET, PH>::NL], R (&gd)[Elem<R, ET, PH>::N], R &Pi) {
  using E = Elem<R, ET, PH>;
  constexpr int N = E::N, D = E::D, G = RuleSize<ET, NGP>::G, UNR = gp_unroll<R, PH, G>();
#pragma unroll
  for (int i = 0; i < E::NL; ++i) r[i] = R(-0.0);
#pragma unroll
  for (int a = 0; a < N; ++a) gd[a] = R(-0.0);
  R pi = R(-0.0);
#pragma unroll UNR
  for (int g = 0; g < G; ++g) {
    GP<R, ET> gp;
    gp_geometry<R, ET, NGP>(g, el.x, gp);
    const R eg = field_gp<R, N>(el.d, el.de, el.elemental, gp.Nv);
    R q;
    if constexpr (PH == THERMAL) {  // beta = 0 (checked on the host)
      R gT[D];
#pragma unroll
      for (int j = 0; j < D; ++j) {
        R s = R(-0.0);
#pragma unroll
        for (int a = 0; a < N; ++a) s += gp.DN[a][j] * el.u[a];
        gT[j] = s;
      }
      R g2 = R(-0.0);
#pragma unroll
      for (int j = 0; j < D; ++j) g2 += gT[j] * gT[j];
      const R f = gp.wdet * eg;
      const R b = pp.has_body ? gp.wdet * R(pp.body[0]) : R(-0.0);
      R Tg = R(-0.0);
#pragma unroll
      for (int a = 0; a < N; ++a) {
        R s = R(-0.0);
#pragma unroll
        for (int j = 0; j < D; ++j) s += gp.DN[a][j] * gT[j];
        r[a] += f * s - b * gp.Nv[a];
        Tg += gp.Nv[a] * el.u[a];
      }
      q = R(0.5) * gp.wdet * g2;
      pi += eg * q - b * Tg;
    } else if constexpr (PH == MECH) {
      R Gu[D][D], S[D][D], Sf[D][D];
      elem_grad<R, ET, N, D>(el.u, gp.DN, Gu);
      lin_stress<R, D>(Gu, pp, S);
      scale_sym<R, D>(S, gp.wdet * eg, Sf);
      add_scaled_nodal_force<R, ET, N, D>(Sf, gp.DN, r);
      q = R(0.5) * gp.wdet * ddot_sym<R, D>(Gu, S);
      R work = R(-0.0);
      if (pp.has_body) {
#pragma unroll
        for (int a = 0; a < N; ++a)
#pragma unroll
          for (int i = 0; i < D; ++i) {
            const R fb = gp.wdet * gp.Nv[a] * R(pp.body[i]);
            r[a * D + i] -= fb;
            work += fb * el.u[a * D + i];
          }
      }
      pi += eg * q - work;
    } else {
      R H[D][D];
      elem_grad<R, ET, N, D>(el.u, gp.DN, H);
      NHState<R, D> st;
      nh_state<R, D>(H, R(pp.kap1), R(pp.mu1), st);
      R Pm[D][D];
#pragma unroll
      for (int i = 0; i < D; ++i)
#pragma unroll
        for (int j = 0; j < D; ++j) {
          R s = R(-0.0);
#pragma unroll
          for (int k = 0; k < D; ++k) s += st.F[i][k] * st.S[k][j];
          Pm[i][j] = s;
        }
      add_nodal_force_f<R, ET, N, D>(Pm, gp.DN, gp.wdet * eg, r);
      q = gp.wdet * st.psi;
      R work = R(-0.0);
      if (pp.has_body) {
#pragma unroll
        for (int a = 0; a < N; ++a)
#pragma unroll
          for (int i = 0; i < D; ++i) {
            const R fb = gp.wdet * gp.Nv[a] * R(pp.body[i]);
            r[a * D + i] -= fb;
            work += fb * el.u[a * D + i];
          }
      }
      pi += eg * q - work;
    }
    if (el.elemental)
      gd[0] += q;
    else {
#pragma unroll
      for (int a = 0; a < N; ++a) gd[a] += q * gp.Nv[a];
    }
  }
  Pi = pi;
}

// ------------------------------------------------------------------------------------------
// tangent rows of local node a: K[c][b*P+d] = Ke[(a,c),(b,d)]
// ------------------------------------------------------------------------------------------
// Row (a, c) of a HEX8 element tangent through the trilinear modes (the quantities of the nodal form
// below): node a's gradient is formed from its signs, the solution gradient from the moments of u,
// and every term A_dj dN_b/dx_j of the row accumulates as moments spread to the 8 columns once.
template <class R, int PH, int NGP>
__device__ __forceinline__ void elem_tangent_row_hex8(const Elem<R, HEX8, PH> &el, const PhysParams &pp, int a, int c,
                                                      R (&K)[Elem<R, HEX8, PH>::NL]) {
  using E = Elem<R, HEX8, PH>;
  constexpr int N = 8, D = 3, P = E::P, G = RuleSize<HEX8, NGP>::G;
  R mx[D][8], mu[P][8], M[P][8];
#pragma unroll
  for (int i = 0; i < D; ++i) {
    R col[8];
#pragma unroll
    for (int b = 0; b < N; ++b) col[b] = el.x[b][i];
    hex8_moments<R>(col, mx[i]);
  }
#pragma unroll
  for (int d = 0; d < P; ++d) {
    if constexpr (PH != MECH) {
      R col[8];
#pragma unroll
      for (int b = 0; b < N; ++b) col[b] = el.u[b * P + d];
      hex8_moments<R>(col, mu[d]);
    }
#pragma unroll
    for (int m = 0; m < 8; ++m) M[d][m] = R(0);
  }
  // node a's signs (node order of hexahedra_3d_8.py:79-88)
  const R sa0 = ((a & 3) == 1 || (a & 3) == 2) ? R(1) : R(-1);
  const R sa1 = (a & 3) >= 2 ? R(1) : R(-1);
  const R sa2 = a >= 4 ? R(1) : R(-1);
#pragma unroll 1
  for (int g = 0; g < G; ++g) {
    double xi, eta, zeta, w;
    gauss_point<HEX8, NGP>(g, xi, eta, zeta, w);
    R Nv[N], dNl[N][D];
    shape<HEX8, R>(xi, eta, zeta, Nv, dNl);
    const R X = R(xi), Y = R(eta), Z = R(zeta), XY = R(xi * eta), XZ = R(xi * zeta), YZ = R(eta * zeta);
    R J[D][D];
#pragma unroll
    for (int i = 0; i < D; ++i) hex8_dxi<R>(mx[i], X, Y, Z, XY, XZ, YZ, J[i]);
    const R c00 = J[1][1] * J[2][2] - J[1][2] * J[2][1];
    const R c01 = J[1][2] * J[2][0] - J[1][0] * J[2][2];
    const R c02 = J[1][0] * J[2][1] - J[1][1] * J[2][0];
    const R det8 = J[0][0] * c00 + J[0][1] * c01 + J[0][2] * c02;
    const R id = recip(det8);
    R inv[D][D];  // inverse of 8 J
    inv[0][0] = c00 * id;
    inv[1][0] = c01 * id;
    inv[2][0] = c02 * id;
    inv[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) * id;
    inv[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) * id;
    inv[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) * id;
    inv[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) * id;
    inv[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) * id;
    inv[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) * id;
    const R wdet = R(w * (1.0 / 512.0)) * det8;
    const R eg = field_gp<R, N>(el.d, el.de, el.elemental, Nv);
    // dN_a/dx = (8 dN_a/dxi) . inv(8 J)
    const R px = R(1) + sa0 * X, py = R(1) + sa1 * Y, pz = R(1) + sa2 * Z;
    const R sh[3] = {sa0 * py * pz, sa1 * px * pz, sa2 * px * py};
    R dNa[D];
#pragma unroll
    for (int j = 0; j < D; ++j) dNa[j] = (sh[0] * inv[0][j] + sh[1] * inv[1][j]) + sh[2] * inv[2][j];
    R A[P][D];  // K_(b,d) += sum_j A[d][j] dN_b/dx_j
    if constexpr (PH == THERMAL) {
      R Tg = R(-0.0);
      const R mode[8] = {R(1), X, Y, Z, XY, XZ, YZ, R(xi * eta * zeta)};
#pragma unroll
      for (int m = 0; m < 8; ++m) Tg += mu[0][m] * mode[m];
      const R f = wdet * eg * (R(1) + R(pp.beta) * powc(Tg * R(0.125), R(pp.c)));
#pragma unroll
      for (int j = 0; j < D; ++j) A[0][j] = f * dNa[j];
    } else if constexpr (PH == MECH) {
      R Gc[D][D], S[D][D];
#pragma unroll
      for (int i = 0; i < D; ++i)
#pragma unroll
        for (int j = 0; j < D; ++j) Gc[i][j] = (i == c) ? dNa[j] : R(0);
      lin_stress<R, D>(Gc, pp, S);
      const R f = wdet * eg;
#pragma unroll
      for (int i = 0; i < D; ++i)
#pragma unroll
        for (int j = 0; j < D; ++j) A[i][j] = f * S[i][j];
    } else {
      R H[D][D];
#pragma unroll
      for (int i = 0; i < D; ++i) {
        R hu[D];
        hex8_dxi<R>(mu[i], X, Y, Z, XY, XZ, YZ, hu);
#pragma unroll
        for (int j = 0; j < D; ++j) H[i][j] = (hu[0] * inv[0][j] + hu[1] * inv[1][j]) + hu[2] * inv[2][j];
      }
      NHState<R, D> st;
      nh_state<R, D>(H, R(pp.kap1), R(pp.mu1), st);
      R SdNa[D];
#pragma unroll
      for (int i = 0; i < D; ++i) {
        R s2 = R(-0.0);
#pragma unroll
        for (int k = 0; k < D; ++k) s2 += st.S[i][k] * dNa[k];
        SdNa[i] = s2;
      }
      R Fc[D];
#pragma unroll
      for (int i = 0; i < D; ++i) {
        R s2 = st.F[0][i];
#pragma unroll
        for (int k = 1; k < D; ++k) s2 = (k == c) ? st.F[k][i] : s2;
        Fc[i] = s2;
      }
      R Xs[D][D];
#pragma unroll
      for (int i = 0; i < D; ++i)
#pragma unroll
        for (int j = 0; j < D; ++j) Xs[i][j] = R(0.5) * (Fc[i] * dNa[j] + Fc[j] * dNa[i]);
      R dS[D][D];
      nh_tangent_apply<R, D>(st, R(pp.kap1), R(pp.mu1), Xs, dS);
      const R f = wdet * eg;
#pragma unroll
      for (int i = 0; i < D; ++i)
#pragma unroll
        for (int j = 0; j < D; ++j) {
          R s2 = (i == c) ? SdNa[j] : R(-0.0);
#pragma unroll
          for (int k = 0; k < D; ++k) s2 += st.F[i][k] * dS[k][j];
          A[i][j] = f * s2;
        }
    }
#pragma unroll
    for (int d = 0; d < P; ++d) {
      R Q[D];
#pragma unroll
      for (int k = 0; k < D; ++k) Q[k] = (A[d][0] * inv[k][0] + A[d][1] * inv[k][1]) + A[d][2] * inv[k][2];
      M[d][1] += Q[0];
      M[d][2] += Q[1];
      M[d][3] += Q[2];
      M[d][4] += Q[0] * Y + Q[1] * X;
      M[d][5] += Q[0] * Z + Q[2] * X;
      M[d][6] += Q[1] * Z + Q[2] * Y;
      M[d][7] += (Q[0] * YZ + Q[1] * XZ) + Q[2] * XY;
    }
  }
#pragma unroll
  for (int d = 0; d < P; ++d) {
    R v[8];
    hex8_spread<R>(M[d], v);
#pragma unroll
    for (int b = 0; b < N; ++b) K[b * P + d] = v[b];
  }
}

template <class R, int ET, int PH, int NGP>
__device__ __forceinline__ void elem_tangent_row(const Elem<R, ET, PH> &el, const PhysParams &pp, int a, int c,
                                                 R (&K)[Elem<R, ET, PH>::NL]) {
  // Row (a, c) of the element tangent.  The Gauss loop is rolled and only one row is accumulated, so
  // the FP64 HEX8 case stays within the register file (24 accumulators instead of 72).
  if constexpr (ET == HEX8) {
    elem_tangent_row_hex8<R, PH, NGP>(el, pp, a, c, K);
    return;
  }
  using E = Elem<R, ET, PH>;
  constexpr int N = E::N, D = E::D, G = RuleSize<ET, NGP>::G;
#pragma unroll
  for (int i = 0; i < E::NL; ++i) K[i] = R(-0.0);
#pragma unroll 1
  for (int g = 0; g < G; ++g) {
    GP<R, ET> gp;
    gp_geometry<R, ET, NGP>(g, el.x, gp);
    const R eg = field_gp<R, N>(el.d, el.de, el.elemental, gp.Nv);
    R dNa[D];
#pragma unroll
    for (int j = 0; j < D; ++j) {
      R s = gp.DN[0][j];
#pragma unroll
      for (int b = 1; b < N; ++b) s = (b == a) ? gp.DN[b][j] : s;
      dNa[j] = s;
    }
    if constexpr (PH == THERMAL) {
      R Tg = R(-0.0);
#pragma unroll
      for (int b = 0; b < N; ++b) Tg += gp.Nv[b] * el.u[b];
      const R f = gp.wdet * eg * (R(1) + R(pp.beta) * powc(Tg, R(pp.c)));
#pragma unroll
      for (int b = 0; b < N; ++b) {
        R s = R(-0.0);
#pragma unroll
        for (int j = 0; j < D; ++j) s += dNa[j] * gp.DN[b][j];
        K[b] += f * s;
      }
    } else if constexpr (PH == MECH) {
      // stress of the unit displacement gradient e_c (x) dN_a
      R Gc[D][D], S[D][D];
#pragma unroll
      for (int i = 0; i < D; ++i)
#pragma unroll
        for (int j = 0; j < D; ++j) Gc[i][j] = (i == c) ? dNa[j] : R(0);
      lin_stress<R, D>(Gc, pp, S);
      add_nodal_force<R, N, D>(S, gp.DN, gp.wdet * eg, K);
    } else {
      R H[D][D];
      nodal_grad<R, N, D>(el.u, gp.DN, H);
      NHState<R, D> st;
      nh_state<R, D>(H, R(pp.kap1), R(pp.mu1), st);
      R SdNa[D];  // S dN_a
#pragma unroll
      for (int i = 0; i < D; ++i) {
        R s = R(-0.0);
#pragma unroll
        for (int k = 0; k < D; ++k) s += st.S[i][k] * dNa[k];
        SdNa[i] = s;
      }
      // unit displacement of (a,c): dF = e_c (x) dN_a ; X = sym(F^T dF)
      R Fc[D];
#pragma unroll
      for (int i = 0; i < D; ++i) {
        R s = st.F[0][i];
#pragma unroll
        for (int k = 1; k < D; ++k) s = (k == c) ? st.F[k][i] : s;
        Fc[i] = s;
      }
      R X[D][D];
#pragma unroll
      for (int i = 0; i < D; ++i)
#pragma unroll
        for (int j = 0; j < D; ++j) X[i][j] = R(0.5) * (Fc[i] * dNa[j] + Fc[j] * dNa[i]);
      R dS[D][D];
      nh_tangent_apply<R, D>(st, R(pp.kap1), R(pp.mu1), X, dS);
      R M[D][D];  // F dS + dF S, (dF S)_ij = delta_ic (S dN_a)_j
#pragma unroll
      for (int i = 0; i < D; ++i)
#pragma unroll
        for (int j = 0; j < D; ++j) {
          R s = (i == c) ? SdNa[j] : R(-0.0);
#pragma unroll
          for (int k = 0; k < D; ++k) s += st.F[i][k] * dS[k][j];
          M[i][j] = s;
        }
      add_nodal_force<R, N, D>(M, gp.DN, gp.wdet * eg, K);
    }
  }
}

}  // namespace femhot
