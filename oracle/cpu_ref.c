/*
 * Plain-C / OpenMP FP64 restatement of the FE hot path, used as the timed CPU baseline.
 *
 * TEST INFRASTRUCTURE ONLY: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * load this library; the product package never does.
 *
 * tet4_elastic_energy_grad: small-strain isotropic TET4 element (one-point rule, w = 1/6,
 * fol/geometries/tetrahedra_3d_4.py:17-60; D of fol/loss_functions/mechanical.py:63-82 scaled by a
 * per-element field; B rows [xx,yy,zz,xy,yz,xz] of mechanical.py:46-58), the element potential
 * Pi_e = 1/2 u_e^T K_e u_e, its gradient r_e = K_e u_e scattered into R with OpenMP atomics (the
 * scatter pattern of fol/loss_functions/ffi_functions/kr_small_displacement_element.cc:81-93) and
 * dPi/dE_e written per element; per-segment potentials summed.
 *
 * quad4_thermal_resnorm_vjp: nonlinear thermal QUAD4 (fol/loss_functions/thermal.py:28-49), the
 * batch residual-norm loss and its exact gradient with respect to the temperature and the nodal
 * conductivity field (the same quantities as oracle/fe.py residual_norm_loss_grad).
 */
#include <math.h>
#include <stdint.h>

int cpuref_num_threads(void) {
#ifdef _OPENMP
  extern int omp_get_max_threads(void);
  return omp_get_max_threads();
#else
  return 1;
#endif
}

void cpuref_set_threads(int n) {
#ifdef _OPENMP
  extern void omp_set_num_threads(int);
  omp_set_num_threads(n);
#else
  (void)n;
#endif
}

/* QUAD4 2x2-Gauss geometry at Gauss point g: N (4), global gradients DN (4x2), weight x detJ
 * (fol/geometries/quadrilateral_2d_4.py:17-66, geometry.py:87-97). */
static double quad4_gp(const double x[4][2], int g, double N[4], double DN[4][2]) {
  const double s = 0.57735026918962576451;
  const int sx[4] = {-1, 1, 1, -1}, sy[4] = {-1, -1, 1, 1};
  const double xi = sx[g] * s, eta = sy[g] * s;
  double dN[4][2], J[2][2] = {{0, 0}, {0, 0}};
  for (int a = 0; a < 4; ++a) {
    N[a] = 0.25 * (1 + sx[a] * xi) * (1 + sy[a] * eta);
    dN[a][0] = 0.25 * sx[a] * (1 + sy[a] * eta);
    dN[a][1] = 0.25 * sy[a] * (1 + sx[a] * xi);
    for (int i = 0; i < 2; ++i)
      for (int k = 0; k < 2; ++k) J[i][k] += x[a][i] * dN[a][k];
  }
  const double det = J[0][0] * J[1][1] - J[0][1] * J[1][0];
  const double inv[2][2] = {{J[1][1] / det, -J[0][1] / det}, {-J[1][0] / det, J[0][0] / det}};
  for (int a = 0; a < 4; ++a)
    for (int j = 0; j < 2; ++j) DN[a][j] = dN[a][0] * inv[0][j] + dN[a][1] * inv[1][j];
  return det; /* weight 1 */
}

static double powc(double x, double c) { return c == 2.0 ? x * x : (c == 1.0 ? x : pow(x, c)); }

/* Nonlinear thermal QUAD4 (fol/loss_functions/thermal.py:28-49, k = (N.k0)(1 + beta (N.T)^c)):
 * per sample b the Dirichlet-masked residual R_b = sum_e Se(k0_b, T_b) T_b, the loss
 * L = 1/2 sum_b |R_b|^2, and its exact gradient: g_T = (dR/dT)^T R (zero at Dirichlet dofs) and
 * g_k0 = (dR/dk0)^T R, through element residual VJPs scattered with OpenMP atomics.  T, k0, g_T,
 * g_k0: (batch, n_nodes) row-major; R: (n_nodes) work array; dmask: 1 at Dirichlet nodes. */
double quad4_thermal_resnorm_vjp(long long n_elems, const int *conn, const double *coords, long long n_nodes,
                                 int batch, const double *k0, const double *T, double beta, double c,
                                 const unsigned char *dmask, double *R, double *gT, double *gk0) {
  double L = 0.0;
  for (int b = 0; b < batch; ++b) {
    const double *kb = k0 + (long long)b * n_nodes, *Tb = T + (long long)b * n_nodes;
    double *gTb = gT + (long long)b * n_nodes, *gkb = gk0 + (long long)b * n_nodes;
#pragma omp parallel for schedule(static)
    for (long long i = 0; i < n_nodes; ++i) R[i] = gTb[i] = gkb[i] = 0.0;
#pragma omp parallel for schedule(static)
    for (long long e = 0; e < n_elems; ++e) {
      const int *nd = conn + 4 * e;
      double x[4][2], Te[4], ke[4], r[4] = {0, 0, 0, 0};
      for (int a = 0; a < 4; ++a) {
        x[a][0] = coords[3 * (long long)nd[a]];
        x[a][1] = coords[3 * (long long)nd[a] + 1];
        Te[a] = Tb[nd[a]];
        ke[a] = kb[nd[a]];
      }
      for (int g = 0; g < 4; ++g) {
        double N[4], DN[4][2];
        const double wd = quad4_gp(x, g, N, DN);
        double dg = 0, Tg = 0, gx = 0, gy = 0;
        for (int a = 0; a < 4; ++a) {
          dg += N[a] * ke[a];
          Tg += N[a] * Te[a];
          gx += DN[a][0] * Te[a];
          gy += DN[a][1] * Te[a];
        }
        const double k = dg * (1.0 + beta * powc(Tg, c));
        for (int a = 0; a < 4; ++a) r[a] += wd * k * (DN[a][0] * gx + DN[a][1] * gy);
      }
      for (int a = 0; a < 4; ++a) {
#pragma omp atomic
        R[nd[a]] += r[a];
      }
    }
    double lb = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : lb)
    for (long long i = 0; i < n_nodes; ++i) {
      if (dmask[i]) R[i] = 0.0;
      lb += R[i] * R[i];
    }
    L += 0.5 * lb;
#pragma omp parallel for schedule(static)
    for (long long e = 0; e < n_elems; ++e) {
      const int *nd = conn + 4 * e;
      double x[4][2], Te[4], ke[4], we[4], gu[4] = {0, 0, 0, 0}, gd[4] = {0, 0, 0, 0};
      for (int a = 0; a < 4; ++a) {
        x[a][0] = coords[3 * (long long)nd[a]];
        x[a][1] = coords[3 * (long long)nd[a] + 1];
        Te[a] = Tb[nd[a]];
        ke[a] = kb[nd[a]];
        we[a] = R[nd[a]];
      }
      for (int g = 0; g < 4; ++g) {
        double N[4], DN[4][2];
        const double wd = quad4_gp(x, g, N, DN);
        double dg = 0, Tg = 0, tx = 0, ty = 0, wx = 0, wy = 0;
        for (int a = 0; a < 4; ++a) {
          dg += N[a] * ke[a];
          Tg += N[a] * Te[a];
          tx += DN[a][0] * Te[a];
          ty += DN[a][1] * Te[a];
          wx += DN[a][0] * we[a];
          wy += DN[a][1] * we[a];
        }
        const double f = 1.0 + beta * powc(Tg, c);
        const double k = dg * f;
        const double q = wd * (wx * tx + wy * ty);
        const double dk = beta != 0.0 ? dg * beta * c * powc(Tg, c - 1.0) : 0.0;
        for (int a = 0; a < 4; ++a) {
          gu[a] += wd * k * (DN[a][0] * wx + DN[a][1] * wy) + q * dk * N[a];
          gd[a] += q * f * N[a];
        }
      }
      for (int a = 0; a < 4; ++a) {
#pragma omp atomic
        gTb[nd[a]] += gu[a];
#pragma omp atomic
        gkb[nd[a]] += gd[a];
      }
    }
#pragma omp parallel for schedule(static)
    for (long long i = 0; i < n_nodes; ++i)
      if (dmask[i]) gTb[i] = 0.0;
  }
  return L;
}

void tet4_elastic_energy_grad(long long n_elems, const int *conn, const double *coords, const double *E,
                              const double *u, double c2, double c3, double c4, const long long *seg_of_elem,
                              double *R, double *gE, double *Pi_seg) {
#pragma omp parallel for schedule(static)
  for (long long e = 0; e < n_elems; ++e) {
    const int *nd = conn + 4 * e;
    double x[4][3], U[4][3];
    for (int a = 0; a < 4; ++a)
      for (int i = 0; i < 3; ++i) {
        x[a][i] = coords[3 * (long long)nd[a] + i];
        U[a][i] = u[3 * (long long)nd[a] + i];
      }
    /* J[i][k] = sum_a x_a,i dN_a/dxi_k with dN = [[-1,-1,-1],[1,0,0],[0,1,0],[0,0,1]] */
    double J[3][3];
    for (int i = 0; i < 3; ++i)
      for (int k = 0; k < 3; ++k) J[i][k] = x[k + 1][i] - x[0][i];
    const double c00 = J[1][1] * J[2][2] - J[1][2] * J[2][1];
    const double c01 = J[1][2] * J[2][0] - J[1][0] * J[2][2];
    const double c02 = J[1][0] * J[2][1] - J[1][1] * J[2][0];
    const double det = J[0][0] * c00 + J[0][1] * c01 + J[0][2] * c02;
    const double id = 1.0 / det;
    double inv[3][3];
    inv[0][0] = c00 * id;
    inv[1][0] = c01 * id;
    inv[2][0] = c02 * id;
    inv[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) * id;
    inv[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) * id;
    inv[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) * id;
    inv[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) * id;
    inv[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) * id;
    inv[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) * id;
    const double dNl[4][3] = {{-1, -1, -1}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    double DN[4][3];
    for (int a = 0; a < 4; ++a)
      for (int j = 0; j < 3; ++j) DN[a][j] = dNl[a][0] * inv[0][j] + dNl[a][1] * inv[1][j] + dNl[a][2] * inv[2][j];
    double G[3][3];
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        double s = 0.0;
        for (int a = 0; a < 4; ++a) s += U[a][i] * DN[a][j];
        G[i][j] = s;
      }
    double S[3][3];
    S[0][0] = c2 * G[0][0] + c3 * (G[1][1] + G[2][2]);
    S[1][1] = c2 * G[1][1] + c3 * (G[0][0] + G[2][2]);
    S[2][2] = c2 * G[2][2] + c3 * (G[0][0] + G[1][1]);
    S[0][1] = S[1][0] = c4 * (G[0][1] + G[1][0]);
    S[1][2] = S[2][1] = c4 * (G[1][2] + G[2][1]);
    S[0][2] = S[2][0] = c4 * (G[0][2] + G[2][0]);
    const double wdet = det / 6.0;
    double q = 0.0;
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) q += G[i][j] * S[i][j];
    q *= 0.5 * wdet;
    const double Ee = E[e];
    gE[e] = q;
    for (int a = 0; a < 4; ++a)
      for (int i = 0; i < 3; ++i) {
        const double v = wdet * Ee * (S[i][0] * DN[a][0] + S[i][1] * DN[a][1] + S[i][2] * DN[a][2]);
#pragma omp atomic
        R[3 * (long long)nd[a] + i] += v;
      }
    const double pe = Ee * q;
#pragma omp atomic
    Pi_seg[seg_of_elem[e]] += pe;
  }
}
