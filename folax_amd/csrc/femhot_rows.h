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

template <int ET, int NGP, int N, int P> struct RowsShape {
  static constexpr int G = RuleSize<ET, NGP>::G;
  static constexpr int EB = (64 / G < 64 / N) ? 64 / G : 64 / N;  // incident elements per batch
  static constexpr int KW = (P == 3) ? 56 : 64;                    // neighbour columns per window
  static constexpr int OR = (P * P * KW + 63) / 64;                // output values per lane per window
};

template <class R, int ET, int PH, int NGP>
__global__ __launch_bounds__(64) void k_tangent_rows(TArgs A) {
  static_assert(PH != NEOHOOKE, "row-owner assembly covers thermal and linear elasticity");
  using E = Elem<R, ET, PH>;
  constexpr int N = E::N, D = E::D, P = E::P;
  using SH = RowsShape<ET, NGP, N, P>;
  constexpr int G = SH::G, EB = SH::EB, KW = SH::KW, OR = SH::OR;
  constexpr int UNR = G <= 8 ? G : 1;
  constexpr int ST = N * D + 1;  // staged values per (element, Gauss point): dN/dx of N nodes, f_g
  constexpr int SW = (EB * G * ST > 64 * P * P) ? EB * G * ST : 64 * P * P;
  __shared__ R S[SW];  // staged geometry, then (aliased) the 64 lanes' P x P blocks
  __shared__ unsigned long long MK[KW];
  const int lane = threadIdx.x;
  const long long node = blockIdx.x;
  const int off0 = A.ne_off[node], cnt = A.ne_off[node + 1] - off0;
  const long long rp = A.row_ptr[node * P];
  const int deg = (int)((A.row_ptr[node * P + 1] - rp) / P);
  bool dm[P];
#pragma unroll
  for (int c = 0; c < P; ++c) dm[c] = A.dmask[node * P + c] != 0;
  const bool elemental = (A.flags & FEMHOT_ELEMENT_FIELD) != 0;
  const bool accumulate = (A.flags & FEMHOT_ACCUMULATE) != 0;
  const R *F = reinterpret_cast<const R *>(A.field);
  const R *U = reinterpret_cast<const R *>(A.u);
  R *V = reinterpret_cast<R *>(A.vals) + rp;
  for (int k0 = 0; k0 < deg; k0 += KW) {
    const int kw = min(KW, deg - k0), nout = P * P * kw;
    // this lane's outputs in the window: t = r * 64 + lane -> row c, column k0 + k, component d
    int ok[OR], ocd[OR];
#pragma unroll
    for (int r = 0; r < OR; ++r) {
      const int t = r * 64 + lane;
      const int c = t / (P * kw), rem = t - c * (P * kw), k = rem / P;
      ok[r] = t < nout ? k : -1;
      ocd[r] = c * P + (rem - k * P);
    }
    R acc[OR];
#pragma unroll
    for (int r = 0; r < OR; ++r) acc[r] = R(0);
    for (int base = 0; base < cnt; base += EB) {
      const int nb = min(EB, cnt - base);
      {  // 1. geometry of (element, Gauss point) pairs
        const int s = lane / G, g = lane - s * G;
        if (s < nb) {
          const int e = A.ne_list[off0 + base + s] / N;
          int nodes[N];
          load_conn<N>(A.conn, e, nodes);
          E el;
          load_geometry<R, ET, PH>(A.coords, nodes, el);
          GP<R, ET> gp;
          gp_geometry<R, ET, NGP>(g, el.x, gp);
          R eg;
          if (elemental) {
            eg = F[e];
          } else {
            eg = R(-0.0);
#pragma unroll
            for (int b = 0; b < N; ++b) eg += gp.Nv[b] * F[nodes[b]];
          }
          R f = gp.wdet * eg;
          if constexpr (PH == THERMAL) {
            R Tg = R(-0.0);
#pragma unroll
            for (int b = 0; b < N; ++b) Tg += gp.Nv[b] * U[nodes[b]];
            f *= R(1) + R(A.pp.beta) * powc(Tg, R(A.pp.c));
          }
          R *st = S + lane * ST;
#pragma unroll
          for (int b = 0; b < N; ++b)
#pragma unroll
            for (int j = 0; j < D; ++j) st[b * D + j] = gp.DN[b][j];
          st[N * D] = f;
        }
      }
      __syncthreads();
      R K[P * P];
      int col = -1;
      {  // 2. block (a, b) of each incident element's tangent
        const int s = lane / N, b = lane - s * N;
#pragma unroll
        for (int i = 0; i < P * P; ++i) K[i] = R(0);
        if (s < nb) {
          const int ea = A.ne_list[off0 + base + s];
          const int e = ea / N, a = ea - e * N;
          col = A.slot[((size_t)e * N + a) * N + b];
          R M[D][D];
#pragma unroll
          for (int p = 0; p < D; ++p)
#pragma unroll
            for (int q = 0; q < D; ++q) M[p][q] = R(-0.0);
          R t = R(-0.0);
#pragma unroll UNR
          for (int g = 0; g < G; ++g) {
            const R *st = S + (s * G + g) * ST;
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
          // a Dirichlet row keeps only the element diagonal (fe_loss.py:191-207)
#pragma unroll
          for (int c = 0; c < P; ++c)
#pragma unroll
            for (int d = 0; d < P; ++d)
              if (dm[c] && !(b == a && d == c)) K[c * P + d] = R(0);
        }
      }
      __syncthreads();  // every lane has read the staged geometry: the space takes the blocks
      for (int t = lane; t < kw; t += 64) MK[t] = 0ull;
#pragma unroll
      for (int i = 0; i < P * P; ++i) S[lane * P * P + i] = K[i];
      __syncthreads();
      if (col >= k0 && col < k0 + kw) atomicOr(&MK[col - k0], 1ull << lane);
      __syncthreads();
#pragma unroll
      for (int r = 0; r < OR; ++r) {  // 3. each output sums its column's blocks in lane order
        if (ok[r] < 0) continue;
        unsigned long long m = MK[ok[r]];
        R sum = R(0);
        while (m) {
          const int j = __builtin_ctzll(m);
          m &= m - 1;
          sum += S[j * P * P + ocd[r]];
        }
        acc[r] += sum;
      }
      __syncthreads();  // before the next batch overwrites the blocks and the masks
    }
#pragma unroll
    for (int r = 0; r < OR; ++r) {
      if (ok[r] < 0) continue;
      const int c = ocd[r] / P, d = ocd[r] - c * P;
      R *v = V + (long long)c * P * deg + (long long)(k0 + ok[r]) * P + d;
      *v = accumulate ? *v + acc[r] : acc[r];
    }
  }
}

}  // namespace femhot
