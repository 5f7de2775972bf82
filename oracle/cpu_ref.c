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
