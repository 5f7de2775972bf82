// Row-owner assembly of the tangent on the CSR pattern, thermal and linear-elastic physics
// (the element tangents of fe_loss.py:264-318, BC-masked as fe_loss.py:191-207).
//
// One 64-lane workgroup owns node i: the P rows of i (P * P * deg consecutive values, deg = number
// of neighbours of i including i) are summed from i's incident elements in a fixed order and each
// value is written once, so the assembly needs no zero fill, no atomics and no colouring, and reads
// and writes every CSR value exactly once.  Per batch of up to EB incident elements:
//   1. lane (element s, Gauss point g) computes the geometry of g and stages dN/dx of all N nodes
//      and the Gauss factor f_g in LDS;
//   2. lane (element s, node b) forms block (a, b) of the element tangent (a = local index of i in
//      s), Dirichlet-masked, stores it in LDS and sets its bit in the mask of its column k (the
//      position of node b among i's neighbours, from the slot table);
//   3. lane per output value (row c, column k, component d) adds the blocks of the lanes in column
//      k's mask, in lane order (element order, then b), to its register accumulator.
// Columns are taken KW at a time; one window covers every node of hexahedral and tetrahedral meshes
// of ordinary valence, and a wider node repeats the batches once per window.
//
// Blocks.  Thermal: K_ab = sum_g f_g dN_a . dN_b, f_g = wdetJ k0_g (1 + beta T_g^c), the Picard
// matrix of thermal.py:31-40.  Linear elastic: with M_ab = sum_g f_g dN_a (x) dN_b and f_g = wdetJ E_g,
//   K_ab[c][c] = c2 M[c][c] + c4 sum_{j != c} M[j][j],   K_ab[c][d] = c3 M[c][d] + c4 M[d][c] (c != d),
// which is B_a^T D B_b for the isotropic D of mechanical.py:60-82 (c2 = D_00, c3 = D_01, c4 = D_33 in
// 3-D; f2, f2nu, f2sh of plane stress in 2-D).  Both element tangents are symmetric, so
// FEMHOT_TRANSPOSE leaves them unchanged.
#pragma once
#include "femhot_kernels.h"

namespace femhot {

// Shape function value and local gradient of node a (unrolled callers make a a constant).
template <int ET, class R>
__device__ __forceinline__ void shape_node(int a, double x, double y, double z, R &Na, R (&dn)[ElemT<ET>::D]) {
  if constexpr (ET == QUAD4) {
    const double sx = (a == 1 || a == 2) ? 1.0 : -1.0, sy = (a >= 2) ? 1.0 : -1.0;
    Na = R(0.25 * (1.0 + sx * x) * (1.0 + sy * y));
    dn[0] = R(0.25 * sx * (1.0 + sy * y));
    dn[1] = R(0.25 * sy * (1.0 + sx * x));
  } else if constexpr (ET == HEX8) {
    const int q = a & 3;
    const double sx = (q == 1 || q == 2) ? 1.0 : -1.0, sy = (q >= 2) ? 1.0 : -1.0, sz = (a >= 4) ? 1.0 : -1.0;
    const double px = 1.0 + sx * x, py = 1.0 + sy * y, pz = 1.0 + sz * z;
    Na = R(0.125 * px * py * pz);
    dn[0] = R(0.125 * sx * py * pz);
    dn[1] = R(0.125 * sy * px * pz);
    dn[2] = R(0.125 * sz * px * py);
  } else {
    Na = R(a == 0 ? 1.0 - (x + y + z) : a == 1 ? x : a == 2 ? y : z);
    dn[0] = R(a == 0 ? -1.0 : a == 1 ? 1.0 : 0.0);
    dn[1] = R(a == 0 ? -1.0 : a == 2 ? 1.0 : 0.0);
    dn[2] = R(a == 0 ? -1.0 : a == 3 ? 1.0 : 0.0);
  }
}

// Inverse Jacobian of Gauss point g of one element for the row-owner staging, with the shape function
// values there: the arithmetic of gp_geometry (coordinates relative to node 0, J = sum_a x_a (x)
// dN_a/dxi, inverse by cofactors), returning wdetJ.  xs: the element's node coordinates (N x 3, FP64).
// HEX8 takes the trilinear modes of femhot_physics.h instead: J from the moments of the coordinates
// (adds only, then 4 terms per entry), and the pair (8 J)^-1 with stage_gradients' 8 dN/dxi.
template <class R, int ET, int NGP>
__device__ __forceinline__ R gp_inverse(int g, const double *xs, R (&Nv)[ElemT<ET>::N],
                                        R (&inv)[ElemT<ET>::D][ElemT<ET>::D]) {
  constexpr int N = ElemT<ET>::N, D = ElemT<ET>::D;
  double xi, eta, zeta, w;
  gauss_point<ET, NGP>(g, xi, eta, zeta, w);
  if constexpr (ET == HEX8) {
    R J[3][3];  // 8 J
    const R X = R(xi), Y = R(eta), Z = R(zeta), XY = R(xi * eta), XZ = R(xi * zeta), YZ = R(eta * zeta);
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      R col[8], m[8];
      col[0] = R(0);
#pragma unroll
      for (int a = 1; a < 8; ++a) col[a] = R(xs[a * 3 + i] - xs[i]);
      hex8_moments<R>(col, m);
      hex8_dxi<R>(m, X, Y, Z, XY, XZ, YZ, J[i]);
    }
#pragma unroll
    for (int a = 0; a < 8; ++a) {
      R dn[3];
      shape_node<ET, R>(a, xi, eta, zeta, Nv[a], dn);
    }
    const R c00 = J[1][1] * J[2][2] - J[1][2] * J[2][1];
    const R c01 = J[1][2] * J[2][0] - J[1][0] * J[2][2];
    const R c02 = J[1][0] * J[2][1] - J[1][1] * J[2][0];
    const R det8 = J[0][0] * c00 + J[0][1] * c01 + J[0][2] * c02;
    const R id = recip(det8);
    inv[0][0] = c00 * id;
    inv[1][0] = c01 * id;
    inv[2][0] = c02 * id;
    inv[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) * id;
    inv[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) * id;
    inv[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) * id;
    inv[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) * id;
    inv[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) * id;
    inv[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) * id;
    return R(w * (1.0 / 512.0)) * det8;
  }
  double x0[D];
#pragma unroll
  for (int i = 0; i < D; ++i) x0[i] = xs[i];
  R J[D][D];
  {
    R dn[D];
    shape_node<ET, R>(0, xi, eta, zeta, Nv[0], dn);
  }
#pragma unroll
  for (int a = 1; a < N; ++a) {
    R xa[D], dn[D];
#pragma unroll
    for (int i = 0; i < D; ++i) xa[i] = R(xs[a * 3 + i] - x0[i]);
    shape_node<ET, R>(a, xi, eta, zeta, Nv[a], dn);
    if constexpr (ET == TET4) {
#pragma unroll
      for (int i = 0; i < D; ++i) J[i][a - 1] = xa[i];
    } else {
#pragma unroll
      for (int i = 0; i < D; ++i)
#pragma unroll
        for (int k = 0; k < D; ++k) J[i][k] = (a == 1) ? xa[i] * dn[k] : J[i][k] + xa[i] * dn[k];
    }
  }
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
  return R(w) * det;
}

// dN_b/dx = dN_b/dxi . inv(J) of every node at Gauss point g, stored at st[b * D + j] (TET4: the
// rows of inv(J), node 0 minus their sum, as gp_geometry; HEX8: (8 dN_b/dxi) . inv(8 J), from the
// node signs).
template <class R, int ET, int NGP>
__device__ __forceinline__ void stage_gradients(int g, const R (&inv)[ElemT<ET>::D][ElemT<ET>::D], R *st) {
  constexpr int N = ElemT<ET>::N, D = ElemT<ET>::D;
  if constexpr (ET == HEX8) {
    double xi, eta, zeta, w;
    gauss_point<ET, NGP>(g, xi, eta, zeta, w);
    const R X = R(xi), Y = R(eta), Z = R(zeta);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      const int q = b & 3;
      const R sx = (q == 1 || q == 2) ? R(1) : R(-1), sy = q >= 2 ? R(1) : R(-1), sz = b >= 4 ? R(1) : R(-1);
      const R px = R(1) + sx * X, py = R(1) + sy * Y, pz = R(1) + sz * Z;
      const R sh[3] = {sx * (py * pz), sy * (px * pz), sz * (px * py)};
#pragma unroll
      for (int j = 0; j < 3; ++j) st[b * 3 + j] = (sh[0] * inv[0][j] + sh[1] * inv[1][j]) + sh[2] * inv[2][j];
    }
  } else if constexpr (ET == TET4) {
#pragma unroll
    for (int j = 0; j < D; ++j) {
      st[1 * D + j] = inv[0][j];
      st[2 * D + j] = inv[1][j];
      st[3 * D + j] = inv[2][j];
      st[j] = -((inv[0][j] + inv[1][j]) + inv[2][j]);
    }
  } else {
    double xi, eta, zeta, w;
    gauss_point<ET, NGP>(g, xi, eta, zeta, w);
#pragma unroll
    for (int b = 0; b < N; ++b) {
      R Nb, dn[D];
      shape_node<ET, R>(b, xi, eta, zeta, Nb, dn);
#pragma unroll
      for (int j = 0; j < D; ++j) {
        R s = dn[0] * inv[0][j];
#pragma unroll
        for (int k = 1; k < D; ++k) s += dn[k] * inv[k][j];
        st[b * D + j] = s;
      }
    }
  }
}

template <int ET, int NGP, int N, int P> struct RowsShape {
  static constexpr int G = RuleSize<ET, NGP>::G;
  static constexpr int EB = (64 / G < 64 / N) ? 64 / G : 64 / N;  // incident elements per batch
  static constexpr int KW = (P == 3) ? 28 : 64;                    // neighbour columns per window
  static constexpr int OR = (P * P * KW + 63) / 64;                // output values per lane per window
};

// Waves per SIMD the compiler must fit: 3 for linear elasticity (168 VGPRs in FP64 HEX8 2x2x2, no spills);
// the thermal instantiations keep the compiler's own choice (the powc of k(T) would spill under 3).
template <int PH> constexpr int rows_min_waves() { return PH == MECH ? 3 : 1; }

template <class R, int ET, int PH, int NGP>
__global__ __launch_bounds__(64, rows_min_waves<PH>()) void k_tangent_rows(TArgs A) {
  static_assert(PH != NEOHOOKE, "row-owner assembly covers thermal and linear elasticity");
  using E = Elem<R, ET, PH>;
  constexpr int N = E::N, D = E::D, P = E::P;
  using SH = RowsShape<ET, NGP, N, P>;
  constexpr int G = SH::G, EB = SH::EB, KW = SH::KW, OR = SH::OR;
  constexpr int UNR = G <= 8 ? G : 1;
  constexpr int ST = N * D + 1;  // staged values per (element, Gauss point): dN/dx of N nodes, f_g
  constexpr int SW = (EB * G * ST > 64 * P * P) ? EB * G * ST : 64 * P * P;
  // LDS: S holds the staged geometry, then the 64 lanes' P x P blocks.  The node data of step 0
  // (coordinates XS, field FS, thermal solution US of lane (s, b) = node b of incident element s)
  // sits at the end of the same space: step 1 reads all of it before it writes the geometry (each
  // lane's writes depend on its reads, and the workgroup is one wave), so the two never overlap in
  // time and the workgroup needs no extra LDS.
  constexpr int XW = (64 * 3 * 8 + 2 * 64 * (int)sizeof(R)) / (int)sizeof(R);  // a multiple of 4
  constexpr int SX = ((SW > XW ? SW : XW) + 3) & ~3;  // so XS starts 16-byte aligned
  __shared__ __align__(16) R S[SX];
  __shared__ unsigned long long MK[KW];
  __shared__ int CS[KW + 1];
  double *XS = reinterpret_cast<double *>(S + SX - XW);
  R *FS = reinterpret_cast<R *>(XS + 64 * 3);
  R *US = FS + 64;
  const int lane = threadIdx.x;
  // Workgroups go to the 8 XCDs in turn, and nodes with nearby ids share elements: XCD x takes the
  // x-th eighth of the nodes in order, so its workgroups find their neighbours' element data in
  // its own L2 (a bijection of workgroup ids onto nodes)
  const long long per = gridDim.x / 8, rem = gridDim.x % 8;
  const long long xcd = blockIdx.x % 8, q = blockIdx.x / 8;
  const long long node = xcd * per + (xcd < rem ? xcd : rem) + q;
  const int off0 = A.ne_off[node], cnt = A.ne_off[node + 1] - off0;
  const long long rp = A.row_ptr[node * P];
  const int deg = (int)((A.row_ptr[node * P + 1] - rp) / P);
  bool dm[P];
#pragma unroll
  for (int c = 0; c < P; ++c) dm[c] = A.dmask[node * P + c] != 0;
  bool any_dm = false;
#pragma unroll
  for (int c = 0; c < P; ++c) any_dm = any_dm || dm[c];
  const bool elemental = (A.flags & FEMHOT_ELEMENT_FIELD) != 0;
  const bool accumulate = (A.flags & FEMHOT_ACCUMULATE) != 0;
  const R *F = reinterpret_cast<const R *>(A.field);
  const R *U = reinterpret_cast<const R *>(A.u);
  R *V = reinterpret_cast<R *>(A.vals) + rp;
  const int s2 = lane / N, b2 = lane - s2 * N;  // lane (element, node) of steps 0 and 2
  const int s1 = lane / G, g1 = lane - s1 * G;  // lane (element, Gauss point) of step 1
  for (int k0 = 0; k0 < deg; k0 += KW) {
    const int kw = min(KW, deg - k0), nout = P * P * kw;
    // this lane's outputs in the window: t = r * 64 + lane -> row c, column k0 + k, component d, at
    // offset oo in the node's P rows (row c starts at c * P * deg); no integer division by kw
    int ok[OR], ocd[OR], oo[OR];
    const int pk = P * kw, pdeg = P * deg;
#pragma unroll
    for (int r = 0; r < OR; ++r) {
      const int t = r * 64 + lane;
      int c = 0, cs = 0, cr = 0;
#pragma unroll
      for (int q = 1; q < P; ++q) {
        const bool ge = t >= cs + pk;
        c += ge ? 1 : 0;
        cs += ge ? pk : 0;
        cr += ge ? pdeg : 0;
      }
      const int rem = t - cs, k = rem / P;
      ok[r] = t < nout ? k : -1;
      ocd[r] = c * P + (rem - k * P);
      oo[r] = cr + k0 * P + rem;
    }
    R acc[OR];
#pragma unroll
    for (int r = 0; r < OR; ++r) acc[r] = R(0);
    for (int base = 0; base < cnt; base += EB) {
      const int nb = min(EB, cnt - base);
      // 0. lane (s, b) fetches node b of incident element s: coordinates, field and (thermal)
      //    solution to LDS; its column among the owned node's neighbours stays in a register
      int a2 = 0, col = -1;
      if (s2 < nb) {
        const int ea = A.ne_list[off0 + base + s2];
        const int e = ea / N;
        a2 = ea - e * N;
        col = A.slot[((size_t)e * N + a2) * N + b2];
        const int nd = A.conn[(size_t)e * N + b2];
#pragma unroll
        for (int i = 0; i < 3; ++i) XS[lane * 3 + i] = A.coords[(size_t)nd * 3 + i];
        FS[lane] = elemental ? F[e] : F[nd];
        if constexpr (PH == THERMAL) US[lane] = U[nd];
      }
      for (int t = lane; t < kw; t += 64) MK[t] = 0ull;
      __syncthreads();
      // the lanes of each column of the window, as a bit mask
      if (col >= k0 && col < k0 + kw) atomicOr(&MK[col - k0], 1ull << lane);
      if (s1 < nb) {  // 1. geometry of (element, Gauss point) pairs
        const int n0 = s1 * N;
        R *st = S + lane * ST;
        R Nv[N], inv[D][D];
        const R wdet = gp_inverse<R, ET, NGP>(g1, XS + n0 * 3, Nv, inv);
        R eg;
        if (elemental) {
          eg = FS[n0];
        } else {
          eg = R(-0.0);
#pragma unroll
          for (int b = 0; b < N; ++b) eg += Nv[b] * FS[n0 + b];
        }
        R f = wdet * eg;
        if constexpr (PH == THERMAL) {
          R Tg = R(-0.0);
#pragma unroll
          for (int b = 0; b < N; ++b) Tg += Nv[b] * US[n0 + b];
          f *= R(1) + R(A.pp.beta) * powc(Tg, R(A.pp.c));
        }
        // inv(J) and f make a round trip through this lane's own staging slot, and the Gauss point
        // index an opaque copy, so the compiler keeps neither the Jacobian step's values nor its
        // shape derivatives live into the gradient step (fewer registers, more waves per SIMD)
#pragma unroll
        for (int i = 0; i < D; ++i)
#pragma unroll
          for (int j = 0; j < D; ++j) st[i * D + j] = inv[i][j];
        st[N * D] = f;
        asm volatile("" ::: "memory");
        int go = g1;
        asm volatile("" : "+v"(go));
        R iv[D][D];
#pragma unroll
        for (int i = 0; i < D; ++i)
#pragma unroll
          for (int j = 0; j < D; ++j) iv[i][j] = st[i * D + j];
        stage_gradients<R, ET, NGP>(go, iv, st);
      }
      __syncthreads();
      {  // blocks are stored column by column, lanes in order: CS[k] = first block of column k
        int pc = lane < kw ? __popcll(MK[lane]) : 0;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
          const int y = __shfl_up(pc, o);
          if (lane >= o) pc += y;
        }
        if (lane < kw) CS[lane + 1] = pc;
        if (lane == 0) CS[0] = 0;
      }
      R K[P * P];
#pragma unroll
      for (int i = 0; i < P * P; ++i) K[i] = R(0);
      if (s2 < nb) {  // 2. block (a, b) of each incident element's tangent
        const int a = a2, b = b2;
        R M[D][D];
#pragma unroll
        for (int p = 0; p < D; ++p)
#pragma unroll
          for (int q = 0; q < D; ++q) M[p][q] = R(-0.0);
        R t = R(-0.0);
#pragma unroll UNR
        for (int g = 0; g < G; ++g) {
          const R *st = S + (s2 * G + g) * ST;
          const R f = st[N * D];
          R fa[D], db[D];
#pragma unroll
          for (int j = 0; j < D; ++j) {
            fa[j] = f * st[a * D + j];
            db[j] = st[b * D + j];
          }
          if constexpr (PH == THERMAL) {
#pragma unroll
            for (int j = 0; j < D; ++j) t += fa[j] * db[j];
          } else {
#pragma unroll
            for (int p = 0; p < D; ++p)
#pragma unroll
              for (int q = 0; q < D; ++q) M[p][q] += fa[p] * db[q];
          }
        }
        if constexpr (PH == THERMAL) {
          K[0] = t;
        } else {
          const R c2 = R(D == 3 ? A.pp.c2 : A.pp.f2), c3 = R(D == 3 ? A.pp.c3 : A.pp.f2nu),
                  c4 = R(D == 3 ? A.pp.c4 : A.pp.f2sh);
#pragma unroll
          for (int c = 0; c < D; ++c)
#pragma unroll
            for (int d = 0; d < D; ++d) {
              if (c == d) {
                R o = M[(c + 1) % D][(c + 1) % D];
                if constexpr (D == 3) o += M[(c + 2) % D][(c + 2) % D];
                K[c * P + d] = c2 * M[c][c] + c4 * o;
              } else {
                K[c * P + d] = c3 * M[c][d] + c4 * M[d][c];
              }
            }
        }
        // a Dirichlet row keeps only the element diagonal (fe_loss.py:191-207); dm is the same for
        // the whole workgroup, so free nodes skip this
        if (any_dm) {
#pragma unroll
          for (int c = 0; c < P; ++c)
#pragma unroll
            for (int d = 0; d < P; ++d)
              if (dm[c] && !(b == a && d == c)) K[c * P + d] = R(0);
        }
      }
      __syncthreads();  // every lane has read the staged geometry: the space takes the blocks
      if (col >= k0 && col < k0 + kw) {
        const int k = col - k0;
        const int dst = CS[k] + __popcll(MK[k] & ((1ull << lane) - 1ull));
#pragma unroll
        for (int i = 0; i < P * P; ++i) S[dst * P * P + i] = K[i];
      }
      __syncthreads();
#pragma unroll
      for (int r = 0; r < OR; ++r) {  // 3. each output sums its column's blocks in lane order
        if (ok[r] < 0) continue;
        const int lo = CS[ok[r]], hi = CS[ok[r] + 1];
        R sum = R(0);
        for (int i = lo; i < hi; i += 4) {  // four contributions per round, their loads issued together
          R x[4];
#pragma unroll
          for (int q = 0; q < 4; ++q) x[q] = S[min(i + q, hi - 1) * P * P + ocd[r]];
#pragma unroll
          for (int q = 0; q < 4; ++q) sum += (i + q < hi) ? x[q] : R(0);
        }
        acc[r] += sum;
      }
      __syncthreads();  // before the next batch overwrites the blocks and the masks
    }
#pragma unroll
    for (int r = 0; r < OR; ++r) {
      if (ok[r] < 0) continue;
      R *v = V + oo[r];
      *v = accumulate ? *v + acc[r] : acc[r];
    }
  }
}

}  // namespace femhot
