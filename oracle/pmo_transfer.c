/*
 * pmo_transfer.c -- h- and p-multigrid transfers of the oracle (test
 * infrastructure only).
 *
 * 1D matrices:
 *   h: M[i*(2p+1) + j + c*p] = phi_i(child-c support point j), from
 *      FE_Q<1>(p).get_prolongation_matrix(c) (portable_geometric_transfer.h:1285-1337);
 *      closed form: phi_i evaluated at the GLL nodes of the two-child patch.
 *   p: M[i*(pf+1) + j] = phi_i^{pc}(xi_j^{GLL,pf}), the projection matrix of
 *      FETools::get_projection_matrix between nested Q spaces
 *      (portable_polynomial_transfer.h:1067-1111), i.e. interpolation.
 * Maps and weights: coarse map with constrained -> invalid, fine map plain,
 * weights 1/multiplicity with 0 on constrained fine DoFs
 * (portable_geometric_transfer.h:1344-1486, portable_polynomial_transfer.h:1155-1360).
 * Kernels: prolongation = gather (invalid -> 0), 3 sweeps with M, times
 * weight, scatter-add (bk1_kokkos_kernels.h:462-759); restriction is its
 * transpose with the weight on the gather (bk1_kokkos_kernels.h:761-1051).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "pmg_oracle.h"

void pmo_h_matrix(int p, double *M)
{
  const int nm = p + 1, nf = 2 * p + 1;
  double *z = malloc(sizeof(double) * nm), *v = malloc(sizeof(double) * nm);
  pmo_gauss_lobatto(nm, z);
  for (int c = 0; c < 2; ++c)
    for (int j = 0; j < nm; ++j) {
      pmo_lagrange(nm, z, 0.5 * (c + z[j]), v, NULL);
      for (int i = 0; i < nm; ++i) M[i * nf + j + c * p] = v[i];
    }
  free(z); free(v);
}

void pmo_p_matrix(int pc, int pf, double *M)
{
  const int nmc = pc + 1, nmf = pf + 1;
  double *zc = malloc(sizeof(double) * nmc), *zf = malloc(sizeof(double) * nmf);
  double *v = malloc(sizeof(double) * nmc);
  pmo_gauss_lobatto(nmc, zc);
  pmo_gauss_lobatto(nmf, zf);
  for (int j = 0; j < nmf; ++j) {
    pmo_lagrange(nmc, zc, zf[j], v, NULL);
    for (int i = 0; i < nmc; ++i) M[i * nmf + j] = v[i];
  }
  free(zc); free(zf); free(v);
}

static void finish_weights(long n_cells, int nf3, const uint32_t *map_f, long n_dofs_f,
                           long NX, long NY, long NZ, double *w)
{
  double *cnt = calloc(n_dofs_f, sizeof(double));
  for (long e = 0; e < n_cells * nf3; ++e) cnt[map_f[e]] += 1.0;
  for (long e = 0; e < n_cells * nf3; ++e) {
    const uint32_t g = map_f[e];
    const long I = g % NX, J = (g / NX) % NY, K = g / (NX * NY);
    const int bnd = (I == 0 || I == NX - 1 || J == 0 || J == NY - 1 || K == 0 || K == NZ - 1);
    w[e] = bnd ? 0.0 : 1.0 / cnt[g];
  }
  free(cnt);
}

void pmo_h_maps(const pmo_mesh *coarse, int p, uint32_t *map_c, uint32_t *map_f, double *w)
{
  const int nf = 2 * p + 1, nf3 = nf * nf * nf;
  const long C = pmo_n_cells(coarse);
  const long NX = 2L * p * coarse->nx + 1, NY = 2L * p * coarse->ny + 1,
             NZ = 2L * p * coarse->nz + 1;
  pmo_dofmap_cg(coarse, p, 1, map_c);
  long cell = 0;
  for (int cz = 0; cz < coarse->nz; ++cz)
    for (int cy = 0; cy < coarse->ny; ++cy)
      for (int cx = 0; cx < coarse->nx; ++cx, ++cell)
        for (int k = 0; k < nf; ++k)
          for (int j = 0; j < nf; ++j)
            for (int i = 0; i < nf; ++i) {
              const long I = 2L * p * cx + i, J = 2L * p * cy + j, K = 2L * p * cz + k;
              map_f[cell * nf3 + i + nf * (j + nf * k)] = (uint32_t)(I + NX * (J + NY * K));
            }
  finish_weights(C, nf3, map_f, NX * NY * NZ, NX, NY, NZ, w);
}

void pmo_p_maps(const pmo_mesh *m, int pc, int pf, uint32_t *map_c, uint32_t *map_f, double *w)
{
  const int nmf = pf + 1, nf3 = nmf * nmf * nmf;
  const long C = pmo_n_cells(m);
  const long NX = (long)pf * m->nx + 1, NY = (long)pf * m->ny + 1, NZ = (long)pf * m->nz + 1;
  pmo_dofmap_cg(m, pc, 1, map_c);
  pmo_dofmap_cg(m, pf, 0, map_f);
  finish_weights(C, nf3, map_f, NX * NY * NZ, NX, NY, NZ, w);
}

/* out[c][b][a] = sum_i M[i][a] in[c][b][i] style 3D sweeps, forward (nmc -> nmf) */
static void sweep_fwd(int nc, int nf, const double *M, const double *in, double *out,
                      double *t1, double *t2)
{
  /* x: t1[k][j][a] */
  for (int k = 0; k < nc; ++k)
    for (int j = 0; j < nc; ++j)
      for (int a = 0; a < nf; ++a) {
        double s = 0;
        for (int i = 0; i < nc; ++i) s += M[i * nf + a] * in[(k * nc + j) * nc + i];
        t1[(k * nc + j) * nf + a] = s;
      }
  /* y: t2[k][b][a] */
  for (int k = 0; k < nc; ++k)
    for (int a = 0; a < nf; ++a)
      for (int b = 0; b < nf; ++b) {
        double s = 0;
        for (int j = 0; j < nc; ++j) s += M[j * nf + b] * t1[(k * nc + j) * nf + a];
        t2[(k * nf + b) * nf + a] = s;
      }
  /* z: out[c][b][a] */
  for (int b = 0; b < nf; ++b)
    for (int a = 0; a < nf; ++a)
      for (int c = 0; c < nf; ++c) {
        double s = 0;
        for (int k = 0; k < nc; ++k) s += M[k * nf + c] * t2[(k * nf + b) * nf + a];
        out[(c * nf + b) * nf + a] = s;
      }
}

static void sweep_bwd(int nc, int nf, const double *M, const double *in, double *out,
                      double *t1, double *t2)
{
  /* x: t1[c][b][i] */
  for (int c = 0; c < nf; ++c)
    for (int b = 0; b < nf; ++b)
      for (int i = 0; i < nc; ++i) {
        double s = 0;
        for (int a = 0; a < nf; ++a) s += M[i * nf + a] * in[(c * nf + b) * nf + a];
        t1[(c * nf + b) * nc + i] = s;
      }
  /* y: t2[c][j][i] */
  for (int c = 0; c < nf; ++c)
    for (int i = 0; i < nc; ++i)
      for (int j = 0; j < nc; ++j) {
        double s = 0;
        for (int b = 0; b < nf; ++b) s += M[j * nf + b] * t1[(c * nf + b) * nc + i];
        t2[(c * nc + j) * nc + i] = s;
      }
  /* z: out[k][j][i] */
  for (int j = 0; j < nc; ++j)
    for (int i = 0; i < nc; ++i)
      for (int k = 0; k < nc; ++k) {
        double s = 0;
        for (int c = 0; c < nf; ++c) s += M[k * nf + c] * t2[(c * nc + j) * nc + i];
        out[(k * nc + j) * nc + i] = s;
      }
}

void pmo_prolongate(int nmc, int nmf, long n_cells, const double *M, const uint32_t *map_c,
                    const uint32_t *map_f, const double *w, double scale, const double *xc,
                    double *yf)
{
  const int nc3 = nmc * nmc * nmc, nf3 = nmf * nmf * nmf;
  double *in = malloc(sizeof(double) * nc3), *out = malloc(sizeof(double) * nf3);
  double *t1 = malloc(sizeof(double) * nf3), *t2 = malloc(sizeof(double) * nf3);
  for (long c = 0; c < n_cells; ++c) {
    for (int i = 0; i < nc3; ++i) {
      const uint32_t g = map_c[c * nc3 + i];
      in[i] = g == PMO_INVALID ? 0.0 : xc[g];
    }
    sweep_fwd(nmc, nmf, M, in, out, t1, t2);
    for (int i = 0; i < nf3; ++i) yf[map_f[c * nf3 + i]] += scale * w[c * nf3 + i] * out[i];
  }
  free(in); free(out); free(t1); free(t2);
}

void pmo_restrict(int nmc, int nmf, long n_cells, const double *M, const uint32_t *map_c,
                  const uint32_t *map_f, const double *w, const double *xf, double *yc)
{
  const int nc3 = nmc * nmc * nmc, nf3 = nmf * nmf * nmf;
  double *in = malloc(sizeof(double) * nf3), *out = malloc(sizeof(double) * nc3);
  double *t1 = malloc(sizeof(double) * nf3), *t2 = malloc(sizeof(double) * nf3);
  for (long c = 0; c < n_cells; ++c) {
    for (int i = 0; i < nf3; ++i) in[i] = w[c * nf3 + i] * xf[map_f[c * nf3 + i]];
    sweep_bwd(nmc, nmf, M, in, out, t1, t2);
    for (int i = 0; i < nc3; ++i) {
      const uint32_t g = map_c[c * nc3 + i];
      if (g != PMO_INVALID) yc[g] += out[i];
    }
  }
  free(in); free(out); free(t1); free(t2);
}
