/*
 * pmo_mesh.c -- structured hex mesh, Q_p DoF numbering, geometric factors,
 * right-hand side and L2 error of the oracle (test infrastructure only).
 *
 * Restates what deal.II builds for the reference:
 *   - lexicographic cell DoF map with constrained DoFs -> invalid
 *     (portable_laplace_operator_bk3.h:360-437, invalid at :426-429);
 *   - G = JxW * J^{-1} J^{-T}, components (xx,xy,xz,yy,yz,zz), layout
 *     G[cell*6*nq^3 + c*nq^3 + q] (portable_laplace_operator_bk3.h:456-515);
 *     the reference's MappingQ(p) on straight-sided hexes reproduces the
 *     trilinear map of the 8 vertices, which is what is evaluated here;
 *   - b_i = int phi_i f, QGauss(p+1) (tests/poisson_cube_bk3/program.cc:517-566);
 *   - L2 error with QGauss(p+2) (source/geometric_multigrid_bk3/program.cc:391-400).
 * The global numbering is lexicographic on the (p*nx+1)(p*ny+1)(p*nz+1) grid
 * (deal.II's numbering differs by a permutation, which no result depends on).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "pmg_oracle.h"

long pmo_n_cells(const pmo_mesh *m) { return (long)m->nx * m->ny * m->nz; }

long pmo_n_vertices(const pmo_mesh *m)
{
  return (long)(m->nx + 1) * (m->ny + 1) * (m->nz + 1);
}

static void vertex_coord(const pmo_mesh *m, int a, int b, int c, double *X)
{
  double x = m->lo[0] + (m->hi[0] - m->lo[0]) * a / m->nx;
  double y = m->lo[1] + (m->hi[1] - m->lo[1]) * b / m->ny;
  double z = m->lo[2] + (m->hi[2] - m->lo[2]) * c / m->nz;
  X[0] = x; X[1] = y; X[2] = z;
  if (m->deform == 1) {
    const double A = m->amplitude, tp = 2.0 * M_PI;
    X[0] = x + A * sin(tp * y) * sin(tp * z);
    X[1] = y + A * sin(tp * z) * sin(tp * x);
    X[2] = z + A * sin(tp * x) * sin(tp * y);
  } else if (m->deform == 2) {
    /* affine shear x += a y, y += a z, z += a x (parallelepiped cells with a
     * non-symmetric Jacobian) */
    const double A = m->amplitude;
    X[0] = x + A * y;
    X[1] = y + A * z;
    X[2] = z + A * x;
  }
}

void pmo_vertices(const pmo_mesh *m, double *xyz)
{
  long v = 0;
  for (int c = 0; c <= m->nz; ++c)
    for (int b = 0; b <= m->ny; ++b)
      for (int a = 0; a <= m->nx; ++a, ++v) vertex_coord(m, a, b, c, xyz + 3 * v);
}

void pmo_cell_vertices(const pmo_mesh *m, double *cv)
{
  long cell = 0;
  for (int cz = 0; cz < m->nz; ++cz)
    for (int cy = 0; cy < m->ny; ++cy)
      for (int cx = 0; cx < m->nx; ++cx, ++cell)
        for (int v = 0; v < 8; ++v)
          vertex_coord(m, cx + (v & 1), cy + ((v >> 1) & 1), cz + ((v >> 2) & 1),
                       cv + cell * 24 + 3 * v);
}

long pmo_n_dofs_cg(const pmo_mesh *m, int p)
{
  return (long)(p * m->nx + 1) * (p * m->ny + 1) * (p * m->nz + 1);
}

void pmo_dofmap_cg(const pmo_mesh *m, int p, int dirichlet, uint32_t *map)
{
  const int nm = p + 1;
  const long NX = p * m->nx + 1, NY = p * m->ny + 1, NZ = p * m->nz + 1;
  long cell = 0;
  for (int cz = 0; cz < m->nz; ++cz)
    for (int cy = 0; cy < m->ny; ++cy)
      for (int cx = 0; cx < m->nx; ++cx, ++cell)
        for (int k = 0; k < nm; ++k)
          for (int j = 0; j < nm; ++j)
            for (int i = 0; i < nm; ++i) {
              long I = (long)p * cx + i, J = (long)p * cy + j, K = (long)p * cz + k;
              int bnd = (I == 0 || I == NX - 1 || J == 0 || J == NY - 1 || K == 0 || K == NZ - 1);
              uint32_t g = (uint32_t)(I + NX * (J + NY * K));
              map[cell * nm * nm * nm + i + nm * (j + nm * k)] =
                  (dirichlet && bnd) ? PMO_INVALID : g;
            }
}

long pmo_boundary_dofs_cg(const pmo_mesh *m, int p, uint32_t *out)
{
  const long NX = p * m->nx + 1, NY = p * m->ny + 1, NZ = p * m->nz + 1;
  long n = 0;
  for (long K = 0; K < NZ; ++K)
    for (long J = 0; J < NY; ++J)
      for (long I = 0; I < NX; ++I)
        if (I == 0 || I == NX - 1 || J == 0 || J == NY - 1 || K == 0 || K == NZ - 1) {
          if (out) out[n] = (uint32_t)(I + NX * (J + NY * K));
          ++n;
        }
  return n;
}

/* trilinear map of one cell at reference point (s,t,u) */
static void trilinear(const double *cv, const double xi[3], double X[3], double J[3][3])
{
  for (int d = 0; d < 3; ++d) {
    X[d] = 0;
    for (int e = 0; e < 3; ++e) J[d][e] = 0;
  }
  for (int v = 0; v < 8; ++v) {
    int b[3] = {v & 1, (v >> 1) & 1, (v >> 2) & 1};
    double f[3], df[3];
    for (int e = 0; e < 3; ++e) {
      f[e] = b[e] ? xi[e] : 1.0 - xi[e];
      df[e] = b[e] ? 1.0 : -1.0;
    }
    double N = f[0] * f[1] * f[2];
    double dN[3] = {df[0] * f[1] * f[2], f[0] * df[1] * f[2], f[0] * f[1] * df[2]};
    for (int d = 0; d < 3; ++d) {
      X[d] += N * cv[3 * v + d];
      for (int e = 0; e < 3; ++e) J[d][e] += dN[e] * cv[3 * v + d];
    }
  }
}

void pmo_geometry(const pmo_mesh *m, int nq, double *G, double *JxW, double *X)
{
  const long C = pmo_n_cells(m);
  const int nq3 = nq * nq * nq;
  double *gx = malloc(sizeof(double) * nq), *gw = malloc(sizeof(double) * nq);
  double *cv = malloc(sizeof(double) * 24 * C);
  pmo_gauss_legendre(nq, gx, gw);
  pmo_cell_vertices(m, cv);
  for (long c = 0; c < C; ++c)
    for (int r = 0; r < nq; ++r)
      for (int qq = 0; qq < nq; ++qq)
        for (int pp = 0; pp < nq; ++pp) {
          const int q = pp + nq * (qq + nq * r);
          double xi[3] = {gx[pp], gx[qq], gx[r]}, Xq[3], J[3][3];
          trilinear(cv + 24 * c, xi, Xq, J);
          double det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) -
                       J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
                       J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
          /* inverse: Ji[d][k] = d xi_d / d x_k */
          double Ji[3][3];
          Ji[0][0] = (J[1][1] * J[2][2] - J[1][2] * J[2][1]) / det;
          Ji[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) / det;
          Ji[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) / det;
          Ji[1][0] = (J[1][2] * J[2][0] - J[1][0] * J[2][2]) / det;
          Ji[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) / det;
          Ji[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) / det;
          Ji[2][0] = (J[1][0] * J[2][1] - J[1][1] * J[2][0]) / det;
          Ji[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) / det;
          Ji[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) / det;
          const double jxw = fabs(det) * gw[pp] * gw[qq] * gw[r];
          if (G) {
            int idx = 0;
            for (int d1 = 0; d1 < 3; ++d1)
              for (int d2 = d1; d2 < 3; ++d2, ++idx) {
                double s = 0;
                for (int k = 0; k < 3; ++k) s += Ji[d1][k] * Ji[d2][k];
                G[c * 6 * nq3 + (long)idx * nq3 + q] = jxw * s;
              }
          }
          if (JxW) JxW[c * nq3 + q] = jxw;
          if (X)
            for (int d = 0; d < 3; ++d) X[(c * nq3 + q) * 3 + d] = Xq[d];
        }
  free(gx); free(gw); free(cv);
}

static double f_eval(int kind, const double *X)
{
  if (kind == 0) return 1.0;
  return 3.0 * M_PI * M_PI * sin(M_PI * X[0]) * sin(M_PI * X[1]) * sin(M_PI * X[2]);
}

static double u_exact(int kind, const double *X)
{
  if (kind == 0) return 0.0;
  return sin(M_PI * X[0]) * sin(M_PI * X[1]) * sin(M_PI * X[2]);
}

void pmo_rhs(const pmo_mesh *m, int p, int nq, int f_kind, const uint32_t *map, long n_dofs,
             double *b)
{
  const long C = pmo_n_cells(m);
  const int nm = p + 1, nm3 = nm * nm * nm, nq3 = nq * nq * nq;
  double *S = malloc(sizeof(double) * nm * nq), *D = malloc(sizeof(double) * nq * nq);
  double *JxW = malloc(sizeof(double) * C * nq3), *X = malloc(sizeof(double) * C * nq3 * 3);
  pmo_shape_matrices(p, nq, S, D);
  pmo_geometry(m, nq, NULL, JxW, X);
  memset(b, 0, sizeof(double) * n_dofs);
  for (long c = 0; c < C; ++c)
    for (int i = 0; i < nm3; ++i) {
      const uint32_t g = map[c * nm3 + i];
      if (g == PMO_INVALID) continue;
      const int i0 = i % nm, i1 = (i / nm) % nm, i2 = i / (nm * nm);
      double s = 0;
      for (int q = 0; q < nq3; ++q) {
        const int q0 = q % nq, q1 = (q / nq) % nq, q2 = q / (nq * nq);
        s += S[i0 * nq + q0] * S[i1 * nq + q1] * S[i2 * nq + q2] *
             f_eval(f_kind, X + 3 * (c * nq3 + q)) * JxW[c * nq3 + q];
      }
      b[g] += s;
    }
  free(S); free(D); free(JxW); free(X);
}

double pmo_l2_error(const pmo_mesh *m, int p, int nq, int u_kind, const uint32_t *plain_map,
                    const double *u)
{
  const long C = pmo_n_cells(m);
  const int nm = p + 1, nm3 = nm * nm * nm, nq3 = nq * nq * nq;
  double *S = malloc(sizeof(double) * nm * nq), *D = malloc(sizeof(double) * nq * nq);
  double *JxW = malloc(sizeof(double) * C * nq3), *X = malloc(sizeof(double) * C * nq3 * 3);
  pmo_shape_matrices(p, nq, S, D);
  pmo_geometry(m, nq, NULL, JxW, X);
  double err = 0;
  for (long c = 0; c < C; ++c)
    for (int q = 0; q < nq3; ++q) {
      const int q0 = q % nq, q1 = (q / nq) % nq, q2 = q / (nq * nq);
      double uh = 0;
      for (int i = 0; i < nm3; ++i) {
        const int i0 = i % nm, i1 = (i / nm) % nm, i2 = i / (nm * nm);
        uh += u[plain_map[c * nm3 + i]] * S[i0 * nq + q0] * S[i1 * nq + q1] * S[i2 * nq + q2];
      }
      const double e = uh - u_exact(u_kind, X + 3 * (c * nq3 + q));
      err += e * e * JxW[c * nq3 + q];
    }
  free(S); free(D); free(JxW); free(X);
  return sqrt(err);
}
