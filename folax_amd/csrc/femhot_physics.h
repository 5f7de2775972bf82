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
// residual + reference energy
// ------------------------------------------------------------------------------------------
template <class R, int ET, int PH, int NGP>
__device__ __forceinline__ void elem_residual(const Elem<R, ET, PH> &el, const PhysParams &pp,
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
template <class R, int ET, int PH, int NGP>
__device__ __forceinline__ void elem_vjp(const Elem<R, ET, PH> &el, const R (&w)[Elem<R, ET, PH>::NL],
                                         const PhysParams &pp, bool picard, R (&gu)[Elem<R, ET, PH>::NL],
                                         R (&gd)[Elem<R, ET, PH>::N]) {
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
template <class R, int ET, int PH, int NGP>
__device__ __forceinline__ void elem_egrad(const Elem<R, ET, PH> &el, const PhysParams &pp,
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
template <class R, int ET, int PH, int NGP>
__device__ __forceinline__ void elem_tangent_row(const Elem<R, ET, PH> &el, const PhysParams &pp, int a, int c,
                                                 R (&K)[Elem<R, ET, PH>::NL]) {
  // Row (a, c) of the element tangent.  The Gauss loop is rolled and only one row is accumulated, so
  // the FP64 HEX8 case stays within the register file (24 accumulators instead of 72).
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
