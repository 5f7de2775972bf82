// Element-level arithmetic of the FE residual hot path, device side (gfx950).
//
// One element is processed by one lane; everything here is register-resident and fully
// unrolled over nodes and Gauss points (element type, rule and physics are template
// parameters, so shape functions and Gauss points fold to constants).
//
// Reference behaviour followed (folax, paths relative to the repository root):
//   geometry      fol/geometries/geometry.py:87-97, quadrilateral_2d_4.py:17-66,
//                 hexahedra_3d_8.py:17-114, tetrahedra_3d_4.py:17-60
//   thermal       fol/loss_functions/thermal.py:28-49
//   linear elast. fol/loss_functions/mechanical.py:14-117
//   Neo-Hookean   fol/loss_functions/mechanical_neohooke.py:49-91,171-275 and
//                 fol/constitutive_material_models/neo_hooke.py:15-109
#pragma once
#include <hip/hip_runtime.h>

namespace femhot {

enum { QUAD4 = 0, HEX8 = 1, TET4 = 2 };
enum { THERMAL = 0, MECH = 1, NEOHOOKE = 2 };

template <int ET> struct ElemT;
template <> struct ElemT<QUAD4> { static constexpr int N = 4, D = 2; };
template <> struct ElemT<HEX8> { static constexpr int N = 8, D = 3; };
template <> struct ElemT<TET4> { static constexpr int N = 4, D = 3; };

template <int ET, int NGP> struct RuleSize;
template <> struct RuleSize<QUAD4, 1> { static constexpr int G = 1; };
template <> struct RuleSize<QUAD4, 2> { static constexpr int G = 4; };
template <> struct RuleSize<QUAD4, 3> { static constexpr int G = 9; };
template <> struct RuleSize<HEX8, 1> { static constexpr int G = 1; };
template <> struct RuleSize<HEX8, 2> { static constexpr int G = 8; };
template <> struct RuleSize<HEX8, 3> { static constexpr int G = 27; };
template <> struct RuleSize<TET4, 1> { static constexpr int G = 1; };
template <> struct RuleSize<TET4, 2> { static constexpr int G = 4; };
template <> struct RuleSize<TET4, 3> { static constexpr int G = 8; };

// Gauss point g of rule (ET, NGP): local coordinates and weight.
template <int ET, int NGP>
__host__ __device__ __forceinline__ constexpr void gauss_point(int g, double &x, double &y, double &z, double &w) {
  constexpr double s3 = 0.57735026918962576451;  // 1/sqrt(3)
  constexpr double s35 = 0.77459666924148337704; // sqrt(3/5)
  z = 0.0;
  if constexpr (ET == QUAD4 || ET == HEX8) {
    if constexpr (NGP == 1) {
      x = y = z = 0.0;
      w = (ET == QUAD4) ? 4.0 : 8.0;
    } else if constexpr (NGP == 2) {
      const int q = g & 3;  // (-,-),(+,-),(+,+),(-,+) then the upper layer for HEX8
      x = (q == 1 || q == 2) ? s3 : -s3;
      y = (q >= 2) ? s3 : -s3;
      if constexpr (ET == HEX8) z = (g >= 4) ? s3 : -s3;
      w = 1.0;
    } else {
      const double xs[3] = {-s35, 0.0, s35};
      const double ws[3] = {5.0 / 9.0, 8.0 / 9.0, 5.0 / 9.0};
      const int i = g % 3, j = (g / 3) % 3, k = g / 9;
      x = xs[i];
      y = xs[j];
      w = ws[i] * ws[j];
      if constexpr (ET == HEX8) {
        z = xs[k];
        w *= ws[k];
      }
    }
  } else {  // TET4
    if constexpr (NGP == 1) {
      x = y = z = 0.25;
      w = 1.0 / 6.0;
    } else if constexpr (NGP == 2) {
      constexpr double a = 0.58541020, b = 0.13819660;  // truncated as in tetrahedra_3d_4.py:26-29
      x = (g == 0) ? a : b;
      y = (g == 1) ? a : b;
      z = (g == 2) ? a : b;
      w = 1.0 / 24.0;
    } else {
      constexpr double a = 0.015835909865720057993, b = 0.32805469671142664734;
      constexpr double c = 0.67914317820120795168, d = 0.10695227393293068277;
      const int h = g & 3;
      const double p = (g < 4) ? a : c, q = (g < 4) ? b : d;
      x = (h == 0) ? p : q;
      y = (h == 1) ? p : q;
      z = (h == 2) ? p : q;
      w = (g < 4) ? 0.02308799441864369039 : 0.01857867224802297628;
    }
  }
}

// Shape function values and local gradients dN/dxi.
template <int ET, class R>
__device__ __forceinline__ void shape(double x, double y, double z, R (&Nv)[ElemT<ET>::N],
                                      R (&dN)[ElemT<ET>::N][ElemT<ET>::D]) {
  if constexpr (ET == QUAD4) {
    constexpr int sx[4] = {-1, 1, 1, -1}, sy[4] = {-1, -1, 1, 1};
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      Nv[a] = R(0.25 * (1.0 + sx[a] * x) * (1.0 + sy[a] * y));
      dN[a][0] = R(0.25 * sx[a] * (1.0 + sy[a] * y));
      dN[a][1] = R(0.25 * sy[a] * (1.0 + sx[a] * x));
    }
  } else if constexpr (ET == HEX8) {
    constexpr int sx[8] = {-1, 1, 1, -1, -1, 1, 1, -1};
    constexpr int sy[8] = {-1, -1, 1, 1, -1, -1, 1, 1};
    constexpr int sz[8] = {-1, -1, -1, -1, 1, 1, 1, 1};
#pragma unroll
    for (int a = 0; a < 8; ++a) {
      const double px = 1.0 + sx[a] * x, py = 1.0 + sy[a] * y, pz = 1.0 + sz[a] * z;
      Nv[a] = R(0.125 * px * py * pz);
      dN[a][0] = R(0.125 * sx[a] * py * pz);
      dN[a][1] = R(0.125 * sy[a] * px * pz);
      dN[a][2] = R(0.125 * sz[a] * px * py);
    }
  } else {
    Nv[0] = R(1.0 - (x + y + z));
    Nv[1] = R(x);
    Nv[2] = R(y);
    Nv[3] = R(z);
    dN[0][0] = dN[0][1] = dN[0][2] = R(-1);
    dN[1][0] = R(1); dN[1][1] = R(0); dN[1][2] = R(0);
    dN[2][0] = R(0); dN[2][1] = R(1); dN[2][2] = R(0);
    dN[3][0] = R(0); dN[3][1] = R(0); dN[3][2] = R(1);
  }
}

// 1/x: FP64 divides exactly; FP32 takes the hardware reciprocal (1 ulp) and one Newton step.
__device__ __forceinline__ double recip(double x) { return 1.0 / x; }
__device__ __forceinline__ float recip(float x) {
  const float r = __builtin_amdgcn_rcpf(x);
  return __builtin_fmaf(__builtin_fmaf(-x, r, 1.0f), r, r);
}

template <class R, int ET> struct GP {
  static constexpr int N = ElemT<ET>::N, D = ElemT<ET>::D;
  R Nv[N];
  R DN[N][D];  // global gradients dN_a/dx_j
  R wdet;      // weight * signed detJ
};

// J[i][k] = sum_a x_a,i dN_a/dxi_k (geometry.py:94-97); DN = dN/dxi . inv(J) (:87-92).
template <class R, int ET, int NGP>
__device__ __forceinline__ void gp_geometry(int g, const R (&x)[ElemT<ET>::N][ElemT<ET>::D], GP<R, ET> &G) {
  constexpr int N = ElemT<ET>::N, D = ElemT<ET>::D;
  double xi, eta, zeta, w;
  gauss_point<ET, NGP>(g, xi, eta, zeta, w);
  R dNl[N][D];
  shape<ET, R>(xi, eta, zeta, G.Nv, dNl);
  // x is relative to local node 0 (x[0] == 0), so node 0 drops out of J; for TET4 dN/dxi is
  // constant and J is the matrix of edge vectors x1, x2, x3.
  R J[D][D];
#pragma unroll
  for (int i = 0; i < D; ++i)
#pragma unroll
    for (int k = 0; k < D; ++k) {
      if constexpr (ET == TET4) {
        J[i][k] = x[k + 1][i];
      } else {
        R s = x[1][i] * dNl[1][k];
#pragma unroll
        for (int a = 2; a < N; ++a) s += x[a][i] * dNl[a][k];
        J[i][k] = s;
      }
    }
  R inv[D][D];
  R det;
  if constexpr (D == 2) {
    det = J[0][0] * J[1][1] - J[0][1] * J[1][0];
    const R id = recip(det);
    inv[0][0] = J[1][1] * id;
    inv[0][1] = -J[0][1] * id;
    inv[1][0] = -J[1][0] * id;
    inv[1][1] = J[0][0] * id;
  } else {
    const R c00 = J[1][1] * J[2][2] - J[1][2] * J[2][1];
    const R c01 = J[1][2] * J[2][0] - J[1][0] * J[2][2];
    const R c02 = J[1][0] * J[2][1] - J[1][1] * J[2][0];
    det = J[0][0] * c00 + J[0][1] * c01 + J[0][2] * c02;
    const R id = recip(det);
    inv[0][0] = c00 * id;
    inv[1][0] = c01 * id;
    inv[2][0] = c02 * id;
    inv[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) * id;
    inv[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) * id;
    inv[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) * id;
    inv[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) * id;
    inv[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) * id;
    inv[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) * id;
  }
  if constexpr (ET == TET4) {  // dN_a/dx = row a-1 of inv(J) for a > 0; node 0 takes minus their sum
#pragma unroll
    for (int j = 0; j < D; ++j) {
      G.DN[1][j] = inv[0][j];
      G.DN[2][j] = inv[1][j];
      G.DN[3][j] = inv[2][j];
      G.DN[0][j] = -((inv[0][j] + inv[1][j]) + inv[2][j]);
    }
  } else {
#pragma unroll
    for (int a = 0; a < N; ++a)
#pragma unroll
      for (int j = 0; j < D; ++j) {
        R s = dNl[a][0] * inv[0][j];
#pragma unroll
        for (int k = 1; k < D; ++k) s += dNl[a][k] * inv[k][j];
        G.DN[a][j] = s;
      }
  }
  G.wdet = R(w) * det;
}

// Base arithmetic type of R (R itself; Dual<T> maps to T): series lengths follow the base precision.
template <class R> struct BaseType { using T = R; };

// ---------------------------------------------------------------------------------------------
// small math helpers
// ---------------------------------------------------------------------------------------------
__device__ __forceinline__ float dlog1p(float x) { return log1pf(x); }
__device__ __forceinline__ double dlog1p(double x) { return log1p(x); }
__device__ __forceinline__ float dexpm1(float x) { return expm1f(x); }
__device__ __forceinline__ double dexpm1(double x) { return expm1(x); }
__device__ __forceinline__ float dexp(float x) { return expf(x); }
__device__ __forceinline__ double dexp(double x) { return exp(x); }
__device__ __forceinline__ float dpow(float x, float y) { return powf(x, y); }
__device__ __forceinline__ double dpow(double x, double y) { return pow(x, y); }

// x^c with the integer exponents the losses use evaluated exactly (thermal.py:33-34)
template <class R> __device__ __forceinline__ R powc(R x, R c) {
  if (c == R(1)) return x;
  if (c == R(2)) return x * x;
  if (c == R(0)) return R(1);
  if (c == R(3)) return x * x * x;
  return dpow(x, c);
}
// c * x^(c-1)
template <class R> __device__ __forceinline__ R dpowc(R x, R c) {
  if (c == R(1)) return R(1);
  if (c == R(2)) return R(2) * x;
  if (c == R(0)) return R(0);
  if (c == R(3)) return R(3) * x * x;
  return c * dpow(x, c - R(1));
}

// x - log1p(x) without cancellation: with s = x/(2+x), log1p(x) = 2 atanh(s), so
// x - log1p(x) = x^2/(2+x) - 2 (s^3/3 + s^5/5 + ...).
template <class R> __device__ __forceinline__ R xm_log1p(R x) {
  if (fabs(x) < R(0.5)) {
    const R s = x / (R(2) + x);
    const R s2 = s * s;
    constexpr int K = sizeof(typename BaseType<R>::T) == 8 ? 16 : 8;
    R t = R(-0.0);
#pragma unroll
    for (int k = K; k >= 1; --k) t = t * s2 + R(1) / R(2 * k + 1);
    return x * x / (R(2) + x) - R(2) * s * s2 * t;
  }
  return x - dlog1p(x);
}

// expm1(x) - x without cancellation (Taylor series for |x| < 0.5).
template <class R> __device__ __forceinline__ R expm1_mx(R x) {
  if (fabs(x) < R(0.5)) {
    constexpr int K = sizeof(typename BaseType<R>::T) == 8 ? 17 : 10;
    R t = R(-0.0);
    // sum_{k>=2} x^k / k!  = x^2 (1/2! + x/3! + ...)
#pragma unroll
    for (int k = K; k >= 2; --k) {
      R f = R(1);
#pragma unroll
      for (int m = 2; m <= k; ++m) f *= R(m);
      t = t * x + R(1) / f;
    }
    return x * x * t;
  }
  return dexpm1(x) - x;
}

// ---------------------------------------------------------------------------------------------
// Physics parameters passed by value to every kernel
// ---------------------------------------------------------------------------------------------
struct PhysParams {
  double E, nu, beta, c;
  double body[3];
  int has_body;  // any body[k] != 0 (the body-force terms are skipped otherwise)
  // derived on the host
  double c2, c3, c4;      // 3-D D matrix entries for young_modulus E (mechanical.py:63-82)
  double f2, f2nu, f2sh;  // 2-D plane stress entries (mechanical.py:60-61)
  double kap1, mu1;       // Neo-Hookean moduli per unit field value (mechanical_neohooke.py:253-255)
};

// ---------------------------------------------------------------------------------------------
// Neo-Hookean constitutive evaluation from H = grad u, per unit field value (psi, S and the
// tangent are linear in (kappa, mu), both proportional to the field).
// All terms are formed without first-order cancellation so that FP32 keeps relative accuracy
// at small strain: J-1 from the invariants of H, log1p/expm1 series for the volumetric and
// isochoric energies, inv(C) - I = -inv(C) (C - I).
// ---------------------------------------------------------------------------------------------
template <class R, int D> struct NHState {
  R F[D][D];
  R invC[D][D];
  R Siso[D][D];  // isochoric stress (unit moduli)
  R S[D][D];     // total stress (unit moduli)
  R J2;          // J^2
  R Jp;          // J p (unit moduli)
  R a;           // J^(-2/D)
  R trC;
  R psi;         // energy density (unit moduli)
};

template <class R, int D>
__device__ __forceinline__ void nh_state(const R (&H)[D][D], R kap, R mu, NHState<R, D> &st) {
#pragma unroll
  for (int i = 0; i < D; ++i)
#pragma unroll
    for (int j = 0; j < D; ++j) st.F[i][j] = H[i][j] + (i == j ? R(1) : R(0));
  R trH, q;
  if constexpr (D == 3) {
    trH = H[0][0] + H[1][1] + H[2][2];
    const R i2 = H[0][0] * H[1][1] - H[0][1] * H[1][0] + H[0][0] * H[2][2] - H[0][2] * H[2][0] +
                 H[1][1] * H[2][2] - H[1][2] * H[2][1];
    const R i3 = H[0][0] * (H[1][1] * H[2][2] - H[1][2] * H[2][1]) -
                 H[0][1] * (H[1][0] * H[2][2] - H[1][2] * H[2][0]) +
                 H[0][2] * (H[1][0] * H[2][1] - H[1][1] * H[2][0]);
    q = i2 + i3;
  } else {
    trH = H[0][0] + H[1][1];
    q = H[0][0] * H[1][1] - H[0][1] * H[1][0];
  }
  const R delta = trH + q;  // J - 1
  const R L = dlog1p(delta);
  // E2 = C - I = H + H^T + H^T H
  R E2[D][D];
  R hh = R(-0.0);
#pragma unroll
  for (int i = 0; i < D; ++i)
#pragma unroll
    for (int j = 0; j < D; ++j) {
      R s = H[i][j] + H[j][i];
#pragma unroll
      for (int k = 0; k < D; ++k) s += H[k][i] * H[k][j];
      E2[i][j] = s;
      hh += H[i][j] * H[i][j];
    }
  R t = R(-0.0);
#pragma unroll
  for (int i = 0; i < D; ++i) t += E2[i][i];
  const R J = R(1) + delta;
  const R J2 = J * J;
  // inv(C) = adj(C) / J^2
  R C[D][D];
#pragma unroll
  for (int i = 0; i < D; ++i)
#pragma unroll
    for (int j = 0; j < D; ++j) C[i][j] = E2[i][j] + (i == j ? R(1) : R(0));
  const R iJ2 = R(1) / J2;
  if constexpr (D == 3) {
    st.invC[0][0] = (C[1][1] * C[2][2] - C[1][2] * C[2][1]) * iJ2;
    st.invC[0][1] = (C[0][2] * C[2][1] - C[0][1] * C[2][2]) * iJ2;
    st.invC[0][2] = (C[0][1] * C[1][2] - C[0][2] * C[1][1]) * iJ2;
    st.invC[1][1] = (C[0][0] * C[2][2] - C[0][2] * C[2][0]) * iJ2;
    st.invC[1][2] = (C[0][2] * C[1][0] - C[0][0] * C[1][2]) * iJ2;
    st.invC[2][2] = (C[0][0] * C[1][1] - C[0][1] * C[1][0]) * iJ2;
    st.invC[1][0] = st.invC[0][1];
    st.invC[2][0] = st.invC[0][2];
    st.invC[2][1] = st.invC[1][2];
  } else {
    st.invC[0][0] = C[1][1] * iJ2;
    st.invC[1][1] = C[0][0] * iJ2;
    st.invC[0][1] = st.invC[1][0] = -C[0][1] * iJ2;
  }
  // Y = inv(C) - I = -inv(C) E2
  R Y[D][D];
#pragma unroll
  for (int i = 0; i < D; ++i)
#pragma unroll
    for (int j = 0; j < D; ++j) {
      R s = R(-0.0);
#pragma unroll
      for (int k = 0; k < D; ++k) s -= st.invC[i][k] * E2[k][j];
      Y[i][j] = s;
    }
  const R x = -(R(2) / R(D)) * L;
  const R am1 = dexpm1(x);
  const R a = R(1) + am1;
  const R xl = xm_log1p(delta);
  // psi_vol = k/4 (J^2 - 2 lnJ - 1) = k/4 (delta^2 + 2 (delta - L))
  const R psi_vol = R(0.25) * kap * (delta * delta + R(2) * xl);
  // Ibar1 - D = D phi(x) + (a-1) t + 2 (delta - L) + |H|^2 - 2 q
  const R i1m = R(D) * expm1_mx(x) + am1 * t + R(2) * xl + (hh - R(2) * q);
  st.psi = psi_vol + R(0.5) * mu * i1m;
  const R Jp = R(0.5) * kap * delta * (R(2) + delta);  // J p = k/2 (J^2 - 1)
  const R tD = t / R(D);
#pragma unroll
  for (int i = 0; i < D; ++i)
#pragma unroll
    for (int j = 0; j < D; ++j) {
      const R siso = a * mu * (-(i == j ? tD : R(0)) - (R(1) + tD) * Y[i][j]);
      st.Siso[i][j] = siso;
      st.S[i][j] = Jp * st.invC[i][j] + siso;
    }
  st.J2 = J2;
  st.Jp = Jp;
  st.a = a;
  st.trC = R(D) + t;
}

// dS = C : X for symmetric X (the 4th-order tangent of neo_hooke.py:104-110 applied to X):
// dS = k J^2 invC tr(invC X) - 2 Jp invC X invC + (2/D) a mu trC (invC X invC - invC tr(invC X)/D)
//      - (2/D) (invC (Siso:X) + Siso tr(invC X))
template <class R, int D>
__device__ __forceinline__ void nh_tangent_apply(const NHState<R, D> &st, R kap, R mu, const R (&X)[D][D],
                                                 R (&dS)[D][D]) {
  R IX[D][D];
#pragma unroll
  for (int i = 0; i < D; ++i)
#pragma unroll
    for (int j = 0; j < D; ++j) {
      R s = R(-0.0);
#pragma unroll
      for (int k = 0; k < D; ++k) s += st.invC[i][k] * X[k][j];
      IX[i][j] = s;
    }
  R tr = R(-0.0), sx = R(-0.0);
#pragma unroll
  for (int i = 0; i < D; ++i) tr += IX[i][i];
#pragma unroll
  for (int i = 0; i < D; ++i)
#pragma unroll
    for (int j = 0; j < D; ++j) sx += st.Siso[i][j] * X[i][j];
  const R c1 = kap * st.J2;
  const R c2 = R(2) * st.Jp;
  const R c3 = (R(2) / R(D)) * st.a * mu * st.trC;
  const R c4 = R(2) / R(D);
#pragma unroll
  for (int i = 0; i < D; ++i)
#pragma unroll
    for (int j = 0; j < D; ++j) {
      R A = R(-0.0);  // (invC X invC)_ij
#pragma unroll
      for (int k = 0; k < D; ++k) A += IX[i][k] * st.invC[k][j];
      dS[i][j] = c1 * st.invC[i][j] * tr - c2 * A + c3 * (A - st.invC[i][j] * tr / R(D)) -
                 c4 * (st.invC[i][j] * sx + st.Siso[i][j] * tr);
    }
}

}  // namespace femhot
