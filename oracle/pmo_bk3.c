/*
 * pmo_bk3.c -- continuous Q_p Laplace operator of the oracle (test
 * infrastructure only).
 *
 * pmo_bk3_cells restates BK3::Parallel::KokkosKernel
 * (include/kernels/bk3_kokkos_kernel.h:25-746) cell by cell, in the same
 * contraction order: gather (:129-193), x/y/z interpolation (:195-337),
 * collocation gradient + G (:339-461), transposed gradient (:463-527),
 * z/y/x back-interpolation (:533-669), scatter-add (:671-738).
 * pmo_bk3_vmult adds LaplaceOperatorBK3::vmult's framing
 * (portable_laplace_operator_bk3.h:128-231): dst = 0, cell loop, then
 * copy_constrained_values (dst_c = src_c).
 * pmo_bk3_element_matrices / pmo_bk3_diagonal are the dense, brute-force
 * counterparts (exact Lagrange derivatives, no sum factorisation) that pin the
 * sum-factorised path; the diagonal follows MatrixFreeTools::compute_diagonal
 * + the 1/d kernel (portable_laplace_operator_bk3.h:533-561) with 1 on
 * constrained rows.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "pmg_oracle.h"

#ifdef _OPENMP
#include <omp.h>
#endif

static void cell_apply(int nm, int nq, const double *S, const double *D, const double *Gc,
                       const double *u, double *out, double *w0, double *w1, double *rr,
                       double *rs, double *rt)
{
  const int nq2 = nq * nq, nq3 = nq2 * nq;
  /* step 2: x direction, w1[k][j][p] */
  for (int k = 0; k < nm; ++k)
    for (int j = 0; j < nm; ++j)
      for (int p = 0; p < nq; ++p) {
        double t = 0;
        for (int i = 0; i < nm; ++i) t += S[i * nq + p] * u[k * nm * nm + j * nm + i];
        w1[k * nq * nm + j * nq + p] = t;
      }
  /* step 3: y direction, w0[k][q][p] */
  for (int k = 0; k < nm; ++k)
    for (int p = 0; p < nq; ++p)
      for (int q = 0; q < nq; ++q) {
        double t = 0;
        for (int j = 0; j < nm; ++j) t += S[j * nq + q] * w1[k * nq * nm + j * nq + p];
        w0[k * nq2 + q * nq + p] = t;
      }
  /* step 4: z direction, w1[r][q][p] */
  for (int q = 0; q < nq; ++q)
    for (int p = 0; p < nq; ++p)
      for (int r = 0; r < nq; ++r) {
        double t = 0;
        for (int k = 0; k < nm; ++k) t += S[k * nq + r] * w0[k * nq2 + q * nq + p];
        w1[r * nq2 + q * nq + p] = t;
      }
  /* step 5: collocation gradients and symmetric G */
  for (int q = 0; q < nq; ++q)
    for (int p = 0; p < nq; ++p)
      for (int r = 0; r < nq; ++r) {
        double qr = 0, qs = 0, qt = 0;
        for (int n = 0; n < nq; ++n) {
          qr += D[n * nq + p] * w1[r * nq2 + q * nq + n];
          qs += D[n * nq + q] * w1[r * nq2 + n * nq + p];
          qt += D[n * nq + r] * w1[n * nq2 + q * nq + p];
        }
        const int idx = r * nq2 + q * nq + p;
        const double Grr = Gc[0 * nq3 + idx], Grs = Gc[1 * nq3 + idx], Grt = Gc[2 * nq3 + idx];
        const double Gss = Gc[3 * nq3 + idx], Gst = Gc[4 * nq3 + idx], Gtt = Gc[5 * nq3 + idx];
        rr[idx] = Grr * qr + Grs * qs + Grt * qt;
        rs[idx] = Grs * qr + Gss * qs + Gst * qt;
        rt[idx] = Grt * qr + Gst * qs + Gtt * qt;
      }
  /* step 6: integrate gradients, w1[r][q][p] */
  for (int q = 0; q < nq; ++q)
    for (int p = 0; p < nq; ++p)
      for (int r = 0; r < nq; ++r) {
        double t = 0;
        for (int n = 0; n < nq; ++n) t += rr[r * nq2 + q * nq + n] * D[p * nq + n];
        for (int n = 0; n < nq; ++n) t += rs[r * nq2 + n * nq + p] * D[q * nq + n];
        for (int n = 0; n < nq; ++n) t += rt[n * nq2 + q * nq + p] * D[r * nq + n];
        w1[r * nq2 + q * nq + p] = t;
      }
  /* step 7: z back, w0[k][q][p] */
  for (int q = 0; q < nq; ++q)
    for (int p = 0; p < nq; ++p)
      for (int k = 0; k < nm; ++k) {
        double t = 0;
        for (int r = 0; r < nq; ++r) t += S[k * nq + r] * w1[r * nq2 + q * nq + p];
        w0[k * nq2 + q * nq + p] = t;
      }
  /* step 8: y back, w1[k][j][p] */
  for (int k = 0; k < nm; ++k)
    for (int p = 0; p < nq; ++p)
      for (int j = 0; j < nm; ++j) {
        double t = 0;
        for (int q = 0; q < nq; ++q) t += S[j * nq + q] * w0[k * nq2 + q * nq + p];
        w1[k * nq * nm + j * nq + p] = t;
      }
  /* step 9: x back, out[k][j][i] */
  for (int k = 0; k < nm; ++k)
    for (int j = 0; j < nm; ++j)
      for (int i = 0; i < nm; ++i) {
        double t = 0;
        for (int p = 0; p < nq; ++p) t += S[i * nq + p] * w1[k * nq * nm + j * nq + p];
        out[k * nm * nm + j * nm + i] = t;
      }
}

static void cells_range(int p, int nq, const double *S, const double *D, const double *G,
                        const uint32_t *map, const double *x, double *y, long n_cells,
                        const int *colors, int color)
{
  const int nm = p + 1, nm3 = nm * nm * nm, nq3 = nq * nq * nq;
  const int big = nq3 > nm3 ? nq3 : nm3;
  double *buf = malloc(sizeof(double) * (7 * big));
  double *u = buf, *out = buf + big, *w0 = buf + 2 * big, *w1 = buf + 3 * big;
  double *rr = buf + 4 * big, *rs = buf + 5 * big, *rt = buf + 6 * big;
#ifdef _OPENMP
#pragma omp for schedule(static)
#endif
  for (long c = 0; c < n_cells; ++c) {
    if (colors && colors[c] != color) continue;
    const uint32_t *mc = map + c * nm3;
    for (int i = 0; i < nm3; ++i) u[i] = mc[i] == PMO_INVALID ? 0.0 : x[mc[i]];
    cell_apply(nm, nq, S, D, G + c * 6 * nq3, u, out, w0, w1, rr, rs, rt);
    for (int i = 0; i < nm3; ++i)
      if (mc[i] != PMO_INVALID) y[mc[i]] += out[i];
  }
  free(buf);
}

void pmo_bk3_cells(int p, int nq, long n_cells, const double *S, const double *D,
                   const double *G, const uint32_t *map, const double *x, double *y,
                   int n_threads, const int *colors)
{
#ifdef _OPENMP
  if (n_threads > 1 && colors) {
#pragma omp parallel num_threads(n_threads)
    for (int color = 0; color < 8; ++color) {
      cells_range(p, nq, S, D, G, map, x, y, n_cells, colors, color);
      /* implicit barrier at the end of the omp for */
    }
    return;
  }
#endif
  (void)n_threads;
  cells_range(p, nq, S, D, G, map, x, y, n_cells, NULL, 0);
}

void pmo_bk3_vmult(int p, int nq, long n_cells, const double *S, const double *D,
                   const double *G, const uint32_t *map, long n_dofs,
                   const uint32_t *constrained, long n_constrained, const double *src,
                   double *dst, int n_threads, const int *colors)
{
  memset(dst, 0, sizeof(double) * n_dofs);
  pmo_bk3_cells(p, nq, n_cells, S, D, G, map, src, dst, n_threads, colors);
  for (long i = 0; i < n_constrained; ++i) dst[constrained[i]] = src[constrained[i]];
}

/* SubdomainLaplaceOperator::vmult_neumann and ::vmult_interface_cell_range
 * (include/operators/portable_subdomain_laplace_operator.h:384-454, :456-533):
 * dst = 0; the cell loop with the plain indices, in which only the physical
 * boundary DoFs are eliminated (:1056-1061), over every cell (cells == NULL)
 * or over the listed cells (the interface cells, :1181-1203); then
 * dst_b = src_b on the physical boundary DoFs. */
void pmo_bk3_vmult_subdomain(int p, int nq, long n_cells, const double *S, const double *D,
                             const double *G, const uint32_t *plain_map, long n_dofs,
                             const uint32_t *physical, long n_physical, const uint32_t *cells,
                             long n_list, const double *src, double *dst)
{
  const int nm = p + 1, nm3 = nm * nm * nm, nq3 = nq * nq * nq;
  const long n_sel = cells ? n_list : n_cells;
  unsigned char *phys = calloc((size_t)n_dofs + 1, 1);
  uint32_t *map = calloc((size_t)(n_sel * nm3 + 1), sizeof(uint32_t));
  double *g = calloc((size_t)(n_sel * 6 * nq3 + 1), sizeof(double));
  for (long i = 0; i < n_physical; ++i) phys[physical[i]] = 1;
  for (long s = 0; s < n_sel; ++s) {
    const long c = cells ? (long)cells[s] : s;
    for (int i = 0; i < nm3; ++i) {
      const uint32_t d = plain_map[c * nm3 + i];
      map[s * nm3 + i] = (d == PMO_INVALID || phys[d]) ? PMO_INVALID : d;
    }
    memcpy(g + s * 6 * nq3, G + c * 6 * nq3, sizeof(double) * 6 * nq3);
  }
  memset(dst, 0, sizeof(double) * n_dofs);
  cells_range(p, nq, S, D, g, map, src, dst, n_sel, NULL, 0);
  for (long i = 0; i < n_physical; ++i) dst[physical[i]] = src[physical[i]];
  free(g);
  free(map);
  free(phys);
}

void pmo_cell_colors(const pmo_mesh *m, int *colors)
{
  long c = 0;
  for (int cz = 0; cz < m->nz; ++cz)
    for (int cy = 0; cy < m->ny; ++cy)
      for (int cx = 0; cx < m->nx; ++cx, ++c) colors[c] = (cx & 1) + 2 * (cy & 1) + 4 * (cz & 1);
}

/* reference-gradient of basis function i at quadrature point q (exact derivatives) */
static void basis_grad(int nm, int nq, const double *S, const double *Sd, int i, int q,
                       double g[3])
{
  const int i0 = i % nm, i1 = (i / nm) % nm, i2 = i / (nm * nm);
  const int q0 = q % nq, q1 = (q / nq) % nq, q2 = q / (nq * nq);
  g[0] = Sd[i0 * nq + q0] * S[i1 * nq + q1] * S[i2 * nq + q2];
  g[1] = S[i0 * nq + q0] * Sd[i1 * nq + q1] * S[i2 * nq + q2];
  g[2] = S[i0 * nq + q0] * S[i1 * nq + q1] * Sd[i2 * nq + q2];
}

static void gmul(const double *Gc, int nq3, int q, const double g[3], double o[3])
{
  const double a = Gc[q], b = Gc[nq3 + q], c = Gc[2 * nq3 + q];
  const double d = Gc[3 * nq3 + q], e = Gc[4 * nq3 + q], f = Gc[5 * nq3 + q];
  o[0] = a * g[0] + b * g[1] + c * g[2];
  o[1] = b * g[0] + d * g[1] + e * g[2];
  o[2] = c * g[0] + e * g[1] + f * g[2];
}

void pmo_bk3_element_matrices(int p, int nq, long n_cells, const double *G, double *K)
{
  const int nm = p + 1, nm3 = nm * nm * nm, nq3 = nq * nq * nq;
  double *S = malloc(sizeof(double) * nm * nq), *D = malloc(sizeof(double) * nq * nq);
  double *Sd = malloc(sizeof(double) * nm * nq);
  double *gr = malloc(sizeof(double) * 3 * nm3 * nq3);
  pmo_shape_matrices(p, nq, S, D);
  pmo_shape_gradients(p, nq, Sd);
  for (int i = 0; i < nm3; ++i)
    for (int q = 0; q < nq3; ++q) basis_grad(nm, nq, S, Sd, i, q, gr + 3 * (i * nq3 + q));
  for (long c = 0; c < n_cells; ++c) {
    const double *Gc = G + c * 6 * nq3;
    double *Kc = K + c * nm3 * nm3;
    for (int i = 0; i < nm3; ++i)
      for (int j = 0; j < nm3; ++j) {
        double s = 0;
        for (int q = 0; q < nq3; ++q) {
          double o[3];
          gmul(Gc, nq3, q, gr + 3 * (j * nq3 + q), o);
          const double *gi = gr + 3 * (i * nq3 + q);
          s += gi[0] * o[0] + gi[1] * o[1] + gi[2] * o[2];
        }
        Kc[i * nm3 + j] = s;
      }
  }
  free(S); free(D); free(Sd); free(gr);
}

void pmo_bk3_diagonal(int p, int nq, long n_cells, const double *G, const uint32_t *map,
                      long n_dofs, const uint32_t *constrained, long n_constrained,
                      double *diag)
{
  const int nm = p + 1, nm3 = nm * nm * nm, nq3 = nq * nq * nq;
  double *S = malloc(sizeof(double) * nm * nq), *D = malloc(sizeof(double) * nq * nq);
  double *Sd = malloc(sizeof(double) * nm * nq);
  double *gr = malloc(sizeof(double) * 3 * nm3 * nq3);
  pmo_shape_matrices(p, nq, S, D);
  pmo_shape_gradients(p, nq, Sd);
  for (int i = 0; i < nm3; ++i)
    for (int q = 0; q < nq3; ++q) basis_grad(nm, nq, S, Sd, i, q, gr + 3 * (i * nq3 + q));
  memset(diag, 0, sizeof(double) * n_dofs);
  for (long c = 0; c < n_cells; ++c) {
    const double *Gc = G + c * 6 * nq3;
    for (int i = 0; i < nm3; ++i) {
      const uint32_t g = map[c * nm3 + i];
      if (g == PMO_INVALID) continue;
      double s = 0;
      for (int q = 0; q < nq3; ++q) {
        double o[3];
        const double *gi = gr + 3 * (i * nq3 + q);
        gmul(Gc, nq3, q, gi, o);
        s += gi[0] * o[0] + gi[1] * o[1] + gi[2] * o[2];
      }
      diag[g] += s;
    }
  }
  for (long i = 0; i < n_constrained; ++i) diag[constrained[i]] = 1.0;
  free(S); free(D); free(Sd); free(gr);
}

void pmo_scale_G(long n_cells, int nq, const double *kappa, const double *G, double *Gk)
{
  const int nq3 = nq * nq * nq;
  for (long c = 0; c < n_cells; ++c)
    for (int comp = 0; comp < 6; ++comp)
      for (int q = 0; q < nq3; ++q)
        Gk[c * 6 * nq3 + comp * nq3 + q] = kappa[c * nq3 + q] * G[c * 6 * nq3 + comp * nq3 + q];
}

/* reference gradients of a Q_p field at the quadrature points of one cell:
 * steps 1-5 of the reference kernel before G is applied */
static void cell_grad(int nm, int nq, const double *S, const double *D, const double *u,
                      double *gx, double *gy, double *gz, double *w0, double *w1)
{
  const int nq2 = nq * nq;
  for (int k = 0; k < nm; ++k)
    for (int j = 0; j < nm; ++j)
      for (int p = 0; p < nq; ++p) {
        double t = 0;
        for (int i = 0; i < nm; ++i) t += S[i * nq + p] * u[k * nm * nm + j * nm + i];
        w1[k * nq * nm + j * nq + p] = t;
      }
  for (int k = 0; k < nm; ++k)
    for (int p = 0; p < nq; ++p)
      for (int q = 0; q < nq; ++q) {
        double t = 0;
        for (int j = 0; j < nm; ++j) t += S[j * nq + q] * w1[k * nq * nm + j * nq + p];
        w0[k * nq2 + q * nq + p] = t;
      }
  for (int q = 0; q < nq; ++q)
    for (int p = 0; p < nq; ++p)
      for (int r = 0; r < nq; ++r) {
        double t = 0;
        for (int k = 0; k < nm; ++k) t += S[k * nq + r] * w0[k * nq2 + q * nq + p];
        w1[r * nq2 + q * nq + p] = t;
      }
  for (int r = 0; r < nq; ++r)
    for (int q = 0; q < nq; ++q)
      for (int p = 0; p < nq; ++p) {
        double a = 0, b = 0, c = 0;
        for (int n = 0; n < nq; ++n) {
          a += D[n * nq + p] * w1[r * nq2 + q * nq + n];
          b += D[n * nq + q] * w1[r * nq2 + n * nq + p];
          c += D[n * nq + r] * w1[n * nq2 + q * nq + p];
        }
        const int idx = r * nq2 + q * nq + p;
        gx[idx] = a; gy[idx] = b; gz[idx] = c;
      }
}

void pmo_grad_kappa(int p, int nq, long n_cells, const double *S, const double *D,
                    const double *G, const double *kappa, const uint32_t *map,
                    const double *u, const double *lam, double *out)
{
  const int nm = p + 1, nm3 = nm * nm * nm, nq3 = nq * nq * nq;
  const int big = nq3 > nm3 ? nq3 : nm3;
  double *b = malloc(sizeof(double) * 12 * big);
  double *uc = b, *lc = b + big, *w0 = b + 2 * big, *w1 = b + 3 * big;
  double *ux = b + 4 * big, *uy = b + 5 * big, *uz = b + 6 * big;
  double *lx = b + 7 * big, *ly = b + 8 * big, *lz = b + 9 * big;
  for (long c = 0; c < n_cells; ++c) {
    const uint32_t *mc = map + c * nm3;
    for (int i = 0; i < nm3; ++i) {
      uc[i] = mc[i] == PMO_INVALID ? 0.0 : u[mc[i]];
      lc[i] = mc[i] == PMO_INVALID ? 0.0 : lam[mc[i]];
    }
    cell_grad(nm, nq, S, D, uc, ux, uy, uz, w0, w1);
    cell_grad(nm, nq, S, D, lc, lx, ly, lz, w0, w1);
    const double *Gc = G + c * 6 * nq3;
    for (int q = 0; q < nq3; ++q) {
      double gu[3] = {ux[q], uy[q], uz[q]}, o[3];
      gmul(Gc, nq3, q, gu, o);
      out[c * nq3 + q] = -kappa[c * nq3 + q] * (o[0] * lx[q] + o[1] * ly[q] + o[2] * lz[q]);
    }
  }
  free(b);
}
