/*
 * pmo_dg.c -- symmetric interior penalty DG Laplacian of the oracle (test
 * infrastructure only).
 *
 * Restates LaplaceOperatorDG::vmult (include/operators/portable_laplace_operator_dg.h:252-351)
 * and its kernels BK3::DG::compute_cell / compute_inner_faces /
 * compute_boundary_faces / distribute_face_to_global
 * (include/kernels/laplace_dg_kokkos_kernels.h:23-2027), in the bilinear form
 * of the reference's CPU check operator
 * (source/plain_laplace_solver_dg/program.cc:37-200):
 *   a(u,v) = sum_K int_K grad u . grad v
 *          + sum_F int_F (sigma [u] - {d_n u}) [v] - 1/2 [u] (d_n v^- + d_n v^+)
 * with [u] = u^- - u^+, {d_n u} = (d_n u^- + d_n u^+)/2 and n the outward
 * normal of the "-" cell.  Boundary faces with id 0 mirror u^+ = -u^-,
 * d_n u^+ = d_n u^- (weak homogeneous Dirichlet); other ids mirror
 * u^+ = u^-, d_n u^+ = -d_n u^- (homogeneous Neumann, no contribution)
 * (laplace_dg_kokkos_kernels.h:1563-1576).
 * Normal derivative: d_n u = m . grad_xi u with the metric vector
 *   metric 0: m = J^{-1} n      (grad_x u . n; deal.II get_normal_derivative,
 *                                as in the reference's CPU check operator)
 *   metric 1: m = J^{-T} n      (what the reference GPU face geometry stores,
 *                                portable_laplace_operator_dg.h:781-789)
 * The two agree where J is symmetric (axis-aligned boxes) and differ on
 * deformed cells.
 * Penalty per face (pmo_dg_penalty), h = |K| / |F| of each side:
 *   kind 0: sigma = p(p+1)/2 (1/h^- + 1/h^+)   (get_penalty_factor,
 *           portable_laplace_operator_dg.h:147-151, 766-772, 817-821)
 *   kind 1: sigma = (p+1)^2 (1/h^- + 1/h^+)    (= 2(p+1)^2/h on uniform cells)
 * boundary faces use h^+ = h^-.
 * The face terms are evaluated point by point with exact Lagrange values and
 * derivatives (no sum factorisation), independent of the GPU kernels.
 * DoF layout: cell-major, lexicographic inside a cell (FE_DGQ numbering).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "pmg_oracle.h"

/* trilinear Jacobian J[d][e] = d x_d / d xi_e of a cell with vertices cv[8][3] */
static void jac(const double *cv, const double xi[3], double J[3][3])
{
  for (int d = 0; d < 3; ++d)
    for (int e = 0; e < 3; ++e) J[d][e] = 0;
  for (int v = 0; v < 8; ++v) {
    const int b[3] = {v & 1, (v >> 1) & 1, (v >> 2) & 1};
    double f[3], df[3];
    for (int e = 0; e < 3; ++e) {
      f[e] = b[e] ? xi[e] : 1.0 - xi[e];
      df[e] = b[e] ? 1.0 : -1.0;
    }
    const double dN[3] = {df[0] * f[1] * f[2], f[0] * df[1] * f[2], f[0] * f[1] * df[2]};
    for (int d = 0; d < 3; ++d)
      for (int e = 0; e < 3; ++e) J[d][e] += dN[e] * cv[3 * v + d];
  }
}

static double inv3(const double J[3][3], double Ji[3][3])
{
  const double det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) -
                     J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
                     J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
  Ji[0][0] = (J[1][1] * J[2][2] - J[1][2] * J[2][1]) / det;
  Ji[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) / det;
  Ji[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) / det;
  Ji[1][0] = (J[1][2] * J[2][0] - J[1][0] * J[2][2]) / det;
  Ji[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) / det;
  Ji[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) / det;
  Ji[2][0] = (J[1][0] * J[2][1] - J[1][1] * J[2][0]) / det;
  Ji[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) / det;
  Ji[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) / det;
  return det;
}

/* value and reference gradient of a Q_p cell field (nodal values u) at xi */
static void eval_field(int nm, const double *gll, const double *u, const double xi[3],
                       double *val, double g[3])
{
  double v[3][16], d[3][16];
  for (int e = 0; e < 3; ++e) pmo_lagrange(nm, gll, xi[e], v[e], d[e]);
  double s = 0, g0 = 0, g1 = 0, g2 = 0;
  for (int k = 0; k < nm; ++k)
    for (int j = 0; j < nm; ++j)
      for (int i = 0; i < nm; ++i) {
        const double c = u[(k * nm + j) * nm + i];
        s += c * v[0][i] * v[1][j] * v[2][k];
        g0 += c * d[0][i] * v[1][j] * v[2][k];
        g1 += c * v[0][i] * d[1][j] * v[2][k];
        g2 += c * v[0][i] * v[1][j] * d[2][k];
      }
  *val = s;
  g[0] = g0;
  g[1] = g1;
  g[2] = g2;
}

/* J e_t1 x J e_t2 for the face with normal axis dn (unnormalised, orientation +e_dn) */
static void face_cross(const double J[3][3], int dn, double N[3])
{
  const int t1 = (dn + 1) % 3, t2 = (dn + 2) % 3;
  N[0] = J[1][t1] * J[2][t2] - J[2][t1] * J[1][t2];
  N[1] = J[2][t1] * J[0][t2] - J[0][t1] * J[2][t2];
  N[2] = J[0][t1] * J[1][t2] - J[1][t1] * J[0][t2];
}

/* |K| and |F_f| (f = 0..5) of a trilinear cell, Gauss 4 (exact for trilinear maps) */
static void cell_measures(const double *cv, double *vol, double area[6])
{
  double x[4], w[4];
  pmo_gauss_legendre(4, x, w);
  *vol = 0;
  for (int f = 0; f < 6; ++f) area[f] = 0;
  for (int a = 0; a < 4; ++a)
    for (int b = 0; b < 4; ++b)
      for (int c = 0; c < 4; ++c) {
        double xi[3] = {x[a], x[b], x[c]}, J[3][3], Ji[3][3];
        jac(cv, xi, J);
        *vol += fabs(inv3(J, Ji)) * w[a] * w[b] * w[c];
      }
  for (int f = 0; f < 6; ++f) {
    const int dn = f / 2, t1 = (dn + 1) % 3, t2 = (dn + 2) % 3;
    for (int a = 0; a < 4; ++a)
      for (int b = 0; b < 4; ++b) {
        double xi[3], J[3][3], N[3];
        xi[dn] = f % 2;
        xi[t1] = x[a];
        xi[t2] = x[b];
        jac(cv, xi, J);
        face_cross(J, dn, N);
        area[f] += sqrt(N[0] * N[0] + N[1] * N[1] + N[2] * N[2]) * w[a] * w[b];
      }
  }
}

long pmo_dg_n_dofs(const pmo_mesh *m, int p)
{
  return pmo_n_cells(m) * (long)(p + 1) * (p + 1) * (p + 1);
}

/* neighbour of cell (cx,cy,cz) across face f, or -1 on the boundary */
static long nb_cell(const pmo_mesh *m, int cx, int cy, int cz, int f)
{
  int c[3] = {cx, cy, cz};
  const int n[3] = {m->nx, m->ny, m->nz}, d = f / 2;
  c[d] += (f % 2) ? 1 : -1;
  if (c[d] < 0 || c[d] >= n[d]) return -1;
  return c[0] + (long)m->nx * (c[1] + (long)m->ny * c[2]);
}

void pmo_dg_penalty(const pmo_mesh *m, int p, int kind, double *sigma)
{
  const long C = pmo_n_cells(m);
  double *cv = malloc(sizeof(double) * 24 * C), *h = malloc(sizeof(double) * 6 * C);
  pmo_cell_vertices(m, cv);
  for (long c = 0; c < C; ++c) {
    double vol, area[6];
    cell_measures(cv + 24 * c, &vol, area);
    for (int f = 0; f < 6; ++f) h[6 * c + f] = vol / area[f];
  }
  const double fac = kind == 0 ? 0.5 * p * (p + 1.0) : (p + 1.0) * (p + 1.0);
  long c = 0;
  for (int cz = 0; cz < m->nz; ++cz)
    for (int cy = 0; cy < m->ny; ++cy)
      for (int cx = 0; cx < m->nx; ++cx, ++c)
        for (int f = 0; f < 6; ++f) {
          const long nb = nb_cell(m, cx, cy, cz, f);
          const double hm = h[6 * c + f], hp = nb < 0 ? hm : h[6 * nb + (f ^ 1)];
          sigma[6 * c + f] = fac * (1.0 / hm + 1.0 / hp);
        }
  free(cv);
  free(h);
}

void pmo_dg_vmult(const pmo_mesh *m, int p, int nq, const double *sigma, int metric,
                  int neumann_mask, const double *src, double *dst)
{
  const int nm = p + 1, nm3 = nm * nm * nm;
  const long C = pmo_n_cells(m);
  double gll[16], gx[16], gw[16];
  pmo_gauss_lobatto(nm, gll);
  pmo_gauss_legendre(nq, gx, gw);
  /* cell integrals: the BK3 cell operator on the cell-major DoF layout */
  {
    double *S = malloc(sizeof(double) * nm * nq), *D = malloc(sizeof(double) * nq * nq);
    double *G = malloc(sizeof(double) * C * 6 * nq * nq * nq);
    uint32_t *map = malloc(sizeof(uint32_t) * C * nm3);
    pmo_shape_matrices(p, nq, S, D);
    pmo_geometry(m, nq, G, NULL, NULL);
    for (long i = 0; i < C * nm3; ++i) map[i] = (uint32_t)i;
    memset(dst, 0, sizeof(double) * C * nm3);
    pmo_bk3_cells(p, nq, C, S, D, G, map, src, dst, 1, NULL);
    free(S);
    free(D);
    free(G);
    free(map);
  }
  double *cv = malloc(sizeof(double) * 24 * C);
  pmo_cell_vertices(m, cv);
  /* face integrals, tested with the own-side basis functions of every cell */
  long c = 0;
  for (int cz = 0; cz < m->nz; ++cz)
    for (int cy = 0; cy < m->ny; ++cy)
      for (int cx = 0; cx < m->nx; ++cx, ++c) {
        const double *u = src + c * nm3;
        for (int f = 0; f < 6; ++f) {
          const int dn = f / 2, s = f % 2, t1 = (dn + 1) % 3, t2 = (dn + 2) % 3;
          const long nb = nb_cell(m, cx, cy, cz, f);
          const int neumann = nb < 0 && ((neumann_mask >> f) & 1);
          const double sig = sigma[6 * c + f];
          for (int a = 0; a < nq; ++a)
            for (int b = 0; b < nq; ++b) {
              double xi[3], J[3][3], Ji[3][3], N[3];
              xi[dn] = s;
              xi[t1] = gx[a];
              xi[t2] = gx[b];
              jac(cv + 24 * c, xi, J);
              inv3(J, Ji);
              face_cross(J, dn, N);
              const double sgn = s ? 1.0 : -1.0;
              const double area = sqrt(N[0] * N[0] + N[1] * N[1] + N[2] * N[2]);
              const double n[3] = {sgn * N[0] / area, sgn * N[1] / area, sgn * N[2] / area};
              const double jxw = area * gw[a] * gw[b];
              double mo[3];
              for (int d = 0; d < 3; ++d) {
                mo[d] = 0;
                for (int k = 0; k < 3; ++k) mo[d] += (metric == 0 ? Ji[d][k] : Ji[k][d]) * n[k];
              }
              double um, gm[3];
              eval_field(nm, gll, u, xi, &um, gm);
              const double dnm = mo[0] * gm[0] + mo[1] * gm[1] + mo[2] * gm[2];
              double up, dnp;
              if (nb >= 0) {
                double xin[3] = {xi[0], xi[1], xi[2]}, Jn[3][3], Jni[3][3], gp[3], mn[3];
                xin[dn] = 1 - s;
                jac(cv + 24 * nb, xin, Jn);
                inv3(Jn, Jni);
                for (int d = 0; d < 3; ++d) {
                  mn[d] = 0;
                  for (int k = 0; k < 3; ++k)
                    mn[d] += (metric == 0 ? Jni[d][k] : Jni[k][d]) * n[k];
                }
                eval_field(nm, gll, src + nb * nm3, xin, &up, gp);
                dnp = mn[0] * gp[0] + mn[1] * gp[1] + mn[2] * gp[2];
              } else if (!neumann) {
                up = -um;
                dnp = dnm;
              } else {
                up = um;
                dnp = -dnm;
              }
              const double jump = um - up, avg = 0.5 * (dnm + dnp);
              const double tv = (sig * jump - avg) * jxw, td = -0.5 * jump * jxw;
              double v[3][16], d[3][16];
              for (int e = 0; e < 3; ++e) pmo_lagrange(nm, gll, xi[e], v[e], d[e]);
              for (int k = 0; k < nm; ++k)
                for (int j = 0; j < nm; ++j)
                  for (int i = 0; i < nm; ++i) {
                    const double phi = v[0][i] * v[1][j] * v[2][k];
                    const double dphi = mo[0] * d[0][i] * v[1][j] * v[2][k] +
                                        mo[1] * v[0][i] * d[1][j] * v[2][k] +
                                        mo[2] * v[0][i] * v[1][j] * d[2][k];
                    dst[c * nm3 + (k * nm + j) * nm + i] += tv * phi + td * dphi;
                  }
            }
        }
      }
  free(cv);
}

/* DG(p) <-> CG(p) on one mesh (ContinuousTransfer,
 * include/multigrid/portable_continuous_transfer.h:112-160, 219-266): the DG
 * DoF i of cell c and the CG DoF cg_map[c][i] share a support point;
 * prolongation dst[c*nm3+i] += src[cg_map[c][i]], restriction is its
 * transpose; invalid (constrained) CG entries are skipped both ways. */
void pmo_dg_cg_prolongate(int p, long n_cells, const uint32_t *cg_map, const double *src_cg,
                          double *dst_dg)
{
  const long nm3 = (long)(p + 1) * (p + 1) * (p + 1);
  for (long i = 0; i < n_cells * nm3; ++i)
    if (cg_map[i] != 0xFFFFFFFFu) dst_dg[i] += src_cg[cg_map[i]];
}

void pmo_dg_cg_restrict(int p, long n_cells, const uint32_t *cg_map, const double *src_dg,
                        double *dst_cg)
{
  const long nm3 = (long)(p + 1) * (p + 1) * (p + 1);
  for (long i = 0; i < n_cells * nm3; ++i)
    if (cg_map[i] != 0xFFFFFFFFu) dst_cg[cg_map[i]] += src_dg[i];
}
