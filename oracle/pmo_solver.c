/*
 * pmo_solver.c -- Chebyshev smoother, V-cycle and preconditioned CG of the
 * oracle (test infrastructure only).
 *
 * V-cycle: VCycleMultigrid::v_cycle (include/multigrid/portable_v_cycle_multigrid.h:163-206):
 *   pre-smooth vmult (zero start) -> t = d - A s -> restrict -> recurse ->
 *   prolongate_and_add -> post-smooth step; coarse level = Chebyshev with
 *   automatic degree (tests/poisson_cube_bk3/program.cc:497-502) or Jacobi-PCG.
 * Chebyshev: deal.II PreconditionChebyshev<...,DiagonalMatrix> as configured at
 *   tests/poisson_cube_bk3/program.cc:474-515 (smoothing_range 15, degree n,
 *   10 eigenvalue-CG iterations; coarse: range 1e-3, degree automatic,
 *   eig iterations = m()).  Restated algorithm: Lanczos estimate from CG on
 *   P^{-1}A started from the (i mod 11) - mean vector, lambda_max = 1.2 x the
 *   largest Ritz value, alpha = lambda_max/range (range > 1) or
 *   min(0.9 lambda_max, lambda_min); theta = (lambda_max + alpha)/2,
 *   delta = (lambda_max - alpha)/2; first-kind three-term recurrence.
 *   Pinned by the textbook Chebyshev error bound, by a dense Rayleigh-Ritz
 *   projection of P^{-1}A on the same Krylov space (same start vector and
 *   iteration count: a wrong count or start vector fails), and by the
 *   closed-form automatic degree (tests/test_oracle_pinning.py); no
 *   reference test pins deal.II's own floating-point details.
 * CG: deal.II SolverCG as used at tests/common/portable_multigrid_solver.h:147-169
 *   and source/geometric_multigrid_bk3/program.cc:343-351: convergence when
 *   ||r_k|| <= tol ||r_0||, iteration count = completed iterations.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "pmg_oracle.h"

struct pmo_level {
  pmo_mesh mesh;
  int p, nq;
  int dg;          /* 1: SIP-DG level (finest only), DoFs cell-major */
  double *sigma;   /* DG penalty [C][6] */
  long C, N;
  double *S, *D, *G, *inv_diag;
  uint32_t *map, *cons;
  long ncons;
  int *colors;
  /* smoother */
  double theta, delta, max_ev, min_ev;
  int degree;
  /* transfer (level-1) -> level */
  int xfer_h; /* 1: h, 0: p, 2: CG(p) -> DG(p) */
  int nmc, nmf;
  long xcells;
  double *M, *w;
  uint32_t *map_c, *map_f;
  /* vectors */
  double *sol, *def, *t, *old, *tmp;
};

struct pmo_mg {
  int L;
  pmo_mg_params prm;
  struct pmo_level *lv;
};

static double dot(long n, const double *a, const double *b)
{
  double s = 0;
  for (long i = 0; i < n; ++i) s += a[i] * b[i];
  return s;
}

void pmo_mg_level_vmult(pmo_mg *mg, int l, double *dst, const double *src)
{
  struct pmo_level *v = &mg->lv[l];
  if (v->dg) {
    pmo_dg_vmult(&v->mesh, v->p, v->nq, v->sigma, mg->prm.dg_metric, 0, src, dst);
    return;
  }
  pmo_bk3_vmult(v->p, v->nq, v->C, v->S, v->D, v->G, v->map, v->N, v->cons, v->ncons, src, dst,
                mg->prm.n_threads, v->colors);
}

/* Lanczos eigenvalue estimate through CG on P^{-1}A (deal.II
 * internal::estimate_eigenvalues with IterationNumberControl(n, 1e-10)) */
static void estimate_eigenvalues(pmo_mg *mg, int l, int n_it, double *mn, double *mx)
{
  struct pmo_level *v = &mg->lv[l];
  const long N = v->N;
  double *x = calloc(N, sizeof(double)), *r = malloc(sizeof(double) * N);
  double *z = malloc(sizeof(double) * N), *pv = malloc(sizeof(double) * N);
  double *q = malloc(sizeof(double) * N);
  double *al = malloc(sizeof(double) * (n_it + 1)), *be = malloc(sizeof(double) * (n_it + 1));
  double mean = 0;
  for (long i = 0; i < N; ++i) { r[i] = (double)(i % 11); mean += r[i]; }
  mean /= N;
  for (long i = 0; i < N; ++i) r[i] -= mean;
  double res = sqrt(dot(N, r, r)), rho_old = 0;
  int it = 0;
  while (it < n_it && res > 1e-10) {
    for (long i = 0; i < N; ++i) z[i] = v->inv_diag[i] * r[i];
    const double rho = dot(N, r, z);
    if (it == 0) memcpy(pv, z, sizeof(double) * N);
    else {
      const double beta = rho / rho_old;
      be[it] = beta;
      for (long i = 0; i < N; ++i) pv[i] = z[i] + beta * pv[i];
    }
    pmo_mg_level_vmult(mg, l, q, pv);
    const double alpha = rho / dot(N, pv, q);
    al[it] = alpha;
    for (long i = 0; i < N; ++i) { x[i] += alpha * pv[i]; r[i] -= alpha * q[i]; }
    res = sqrt(dot(N, r, r));
    rho_old = rho;
    ++it;
  }
  if (it == 0) { *mn = *mx = 1.0; }
  else {
    double *d = malloc(sizeof(double) * it), *o = malloc(sizeof(double) * it);
    double *e = malloc(sizeof(double) * it);
    for (int k = 0; k < it; ++k) {
      d[k] = 1.0 / al[k] + (k > 0 ? be[k] / al[k - 1] : 0.0);
      if (k + 1 < it) o[k] = sqrt(be[k + 1]) / al[k];
    }
    pmo_tridiag_eigs(it, d, o, e);
    *mn = e[0];
    *mx = e[it - 1];
    free(d); free(o); free(e);
  }
  free(x); free(r); free(z); free(pv); free(q); free(al); free(be);
}

static void cheby_setup(pmo_mg *mg, int l, int degree, double range, int n_it)
{
  struct pmo_level *v = &mg->lv[l];
  double mn, mx;
  estimate_eigenvalues(mg, l, n_it, &mn, &mx);
  v->min_ev = mn;
  v->max_ev = 1.2 * mx;
  const double alpha = range > 1.0 ? v->max_ev / range : fmin(0.9 * v->max_ev, v->min_ev);
  if (degree <= 0) {
    const double actual_range = v->max_ev / alpha;
    const double sigma = (1.0 - sqrt(1.0 / actual_range)) / (1.0 + sqrt(1.0 / actual_range));
    const double eps = range;
    degree = 1 + (int)(log(1.0 / eps + sqrt(1.0 / eps / eps - 1.0)) / log(1.0 / sigma));
  }
  v->degree = degree;
  v->delta = 0.5 * (v->max_ev - alpha);
  v->theta = 0.5 * (v->max_ev + alpha);
}

/* Chebyshev smoother: step = 0 -> vmult (zero initial guess), 1 -> step */
void pmo_mg_smooth(pmo_mg *mg, int l, double *x, const double *b, int step)
{
  struct pmo_level *v = &mg->lv[l];
  const long N = v->N;
  double *old = v->old, *tmp = v->tmp;
  if (!step) {
    for (long i = 0; i < N; ++i) x[i] = b[i] * v->inv_diag[i] / v->theta;
  } else {
    pmo_mg_level_vmult(mg, l, tmp, x);
    for (long i = 0; i < N; ++i) {
      old[i] = x[i];
      x[i] = x[i] + (b[i] - tmp[i]) * v->inv_diag[i] / v->theta;
    }
  }
  if (v->degree < 2 || fabs(v->delta) < 1e-40) return;
  double rhok = v->delta / v->theta;
  const double sigma = v->theta / v->delta;
  for (int k = 0; k < v->degree - 1; ++k) {
    pmo_mg_level_vmult(mg, l, tmp, x);
    const double rhokp = 1.0 / (2.0 * sigma - rhok);
    const double f1 = rhokp * rhok, f2 = 2.0 * rhokp / v->delta;
    rhok = rhokp;
    const int first = (!step && k == 0); /* x^{k-1} = 0 */
    for (long i = 0; i < N; ++i) {
      const double xo = first ? 0.0 : old[i];
      const double xn = (1.0 + f1) * x[i] - f1 * xo + f2 * (b[i] - tmp[i]) * v->inv_diag[i];
      old[i] = x[i];
      x[i] = xn;
    }
  }
}

void pmo_mg_smoother_info(pmo_mg *mg, int l, double *max_ev, double *min_ev, int *degree)
{
  *max_ev = mg->lv[l].max_ev;
  *min_ev = mg->lv[l].min_ev;
  *degree = mg->lv[l].degree;
}

int pmo_pcg(pmo_mg *mg, int l, int precond, const double *b, double *x, double rel_tol,
            int max_it, double *res0, double *res_out)
{
  struct pmo_level *v = &mg->lv[l];
  const long N = v->N;
  double *r = malloc(sizeof(double) * N), *z = malloc(sizeof(double) * N);
  double *pv = malloc(sizeof(double) * N), *q = malloc(sizeof(double) * N);
  pmo_mg_level_vmult(mg, l, q, x);
  for (long i = 0; i < N; ++i) r[i] = b[i] - q[i];
  const double r0 = sqrt(dot(N, r, r));
  double res = r0, rho_old = 0;
  int it = 0;
  if (res0) *res0 = r0;
  while (res > rel_tol * r0 && it < max_it) {
    if (precond == 0) memcpy(z, r, sizeof(double) * N);
    else if (precond == 1)
      for (long i = 0; i < N; ++i) z[i] = v->inv_diag[i] * r[i];
    else pmo_mg_vmult(mg, z, r);
    const double rho = dot(N, r, z);
    if (it == 0) memcpy(pv, z, sizeof(double) * N);
    else {
      const double beta = rho / rho_old;
      for (long i = 0; i < N; ++i) pv[i] = z[i] + beta * pv[i];
    }
    pmo_mg_level_vmult(mg, l, q, pv);
    const double alpha = rho / dot(N, pv, q);
    for (long i = 0; i < N; ++i) { x[i] += alpha * pv[i]; r[i] -= alpha * q[i]; }
    res = sqrt(dot(N, r, r));
    rho_old = rho;
    ++it;
  }
  if (res_out) *res_out = res;
  free(r); free(z); free(pv); free(q);
  return it;
}

static void v_cycle(pmo_mg *mg, int l)
{
  struct pmo_level *v = &mg->lv[l];
  if (l == 0) {
    if (mg->prm.coarse_kind == 1) {
      memset(v->sol, 0, sizeof(double) * v->N);
      pmo_pcg(mg, 0, 1, v->def, v->sol, mg->prm.coarse_tol, 10000, NULL, NULL);
    } else pmo_mg_smooth(mg, 0, v->sol, v->def, 0);
    return;
  }
  struct pmo_level *c = &mg->lv[l - 1];
  pmo_mg_smooth(mg, l, v->sol, v->def, 0);
  pmo_mg_level_vmult(mg, l, v->t, v->sol);
  for (long i = 0; i < v->N; ++i) v->t[i] = v->def[i] - v->t[i];
  memset(c->def, 0, sizeof(double) * c->N);
  if (v->xfer_h == 2) pmo_dg_cg_restrict(v->p, v->C, v->map_c, v->t, c->def);
  else pmo_restrict(v->nmc, v->nmf, v->xcells, v->M, v->map_c, v->map_f, v->w, v->t, c->def);
  v_cycle(mg, l - 1);
  if (v->xfer_h == 2) pmo_dg_cg_prolongate(v->p, v->C, v->map_c, c->sol, v->sol);
  else
    pmo_prolongate(v->nmc, v->nmf, v->xcells, v->M, v->map_c, v->map_f, v->w,
                   v->xfer_h ? mg->prm.h_prolong_scale : 1.0, c->sol, v->sol);
  pmo_mg_smooth(mg, l, v->sol, v->def, 1);
}

void pmo_mg_vmult(pmo_mg *mg, double *dst, const double *src)
{
  struct pmo_level *v = &mg->lv[mg->L - 1];
  memcpy(v->def, src, sizeof(double) * v->N);
  v_cycle(mg, mg->L - 1);
  memcpy(dst, v->sol, sizeof(double) * v->N);
}

long pmo_mg_n_dofs(const pmo_mg *mg, int level) { return mg->lv[level].N; }

pmo_mg *pmo_mg_create(int n_levels, const pmo_mesh *ms, const int *ps, const int *nqs,
                      const pmo_mg_params *prm)
{
  pmo_mg *mg = calloc(1, sizeof(pmo_mg));
  mg->L = n_levels;
  mg->prm = *prm;
  mg->lv = calloc(n_levels, sizeof(struct pmo_level));
  for (int l = 0; l < n_levels; ++l) {
    struct pmo_level *v = &mg->lv[l];
    v->mesh = ms[l];
    v->p = ps[l];
    v->nq = nqs ? nqs[l] : ps[l] + 1;
    v->C = pmo_n_cells(&ms[l]);
    v->dg = prm->dg_top && l == n_levels - 1 && l > 0;
    v->N = v->dg ? pmo_dg_n_dofs(&ms[l], v->p) : pmo_n_dofs_cg(&ms[l], v->p);
    const int nm = v->p + 1, nm3 = nm * nm * nm, nq3 = v->nq * v->nq * v->nq;
    v->S = malloc(sizeof(double) * nm * v->nq);
    v->D = malloc(sizeof(double) * v->nq * v->nq);
    pmo_shape_matrices(v->p, v->nq, v->S, v->D);
    v->G = malloc(sizeof(double) * v->C * 6 * nq3);
    pmo_geometry(&ms[l], v->nq, v->G, NULL, NULL);
    v->map = malloc(sizeof(uint32_t) * v->C * nm3);
    if (v->dg) {
      /* cell-major DG numbering, no constraints; cell-only diagonal
       * (LaplaceOperatorDG::compute_diagonal, portable_laplace_operator_dg.h:887-915) */
      for (long i = 0; i < v->C * nm3; ++i) v->map[i] = (uint32_t)i;
      v->ncons = 0;
      v->sigma = malloc(sizeof(double) * 6 * v->C);
      pmo_dg_penalty(&ms[l], v->p, prm->dg_penalty_kind, v->sigma);
    } else {
      pmo_dofmap_cg(&ms[l], v->p, 1, v->map);
      v->ncons = pmo_boundary_dofs_cg(&ms[l], v->p, NULL);
    }
    v->cons = malloc(sizeof(uint32_t) * (v->ncons + 1));
    if (!v->dg) pmo_boundary_dofs_cg(&ms[l], v->p, v->cons);
    v->colors = malloc(sizeof(int) * v->C);
    pmo_cell_colors(&ms[l], v->colors);
    v->inv_diag = malloc(sizeof(double) * v->N);
    pmo_bk3_diagonal(v->p, v->nq, v->C, v->G, v->map, v->N, v->cons, v->ncons, v->inv_diag);
    for (long i = 0; i < v->N; ++i) v->inv_diag[i] = 1.0 / v->inv_diag[i];
    v->sol = calloc(v->N, sizeof(double));
    v->def = calloc(v->N, sizeof(double));
    v->t = calloc(v->N, sizeof(double));
    v->old = calloc(v->N, sizeof(double));
    v->tmp = calloc(v->N, sizeof(double));
    if (l > 0) {
      const pmo_mesh *mc = &ms[l - 1];
      const int pc = ps[l - 1];
      if (v->dg) {
        /* ContinuousTransfer: CG map of the level below (same mesh, degree) */
        v->xfer_h = 2;
        v->nmc = v->nmf = nm;
        v->xcells = v->C;
        v->map_c = malloc(sizeof(uint32_t) * v->C * nm3);
        pmo_dofmap_cg(mc, pc, 1, v->map_c);
      } else if (mc->nx != ms[l].nx) {
        v->xfer_h = 1;
        v->nmc = v->p + 1;
        v->nmf = 2 * v->p + 1;
        v->xcells = pmo_n_cells(mc);
        v->M = malloc(sizeof(double) * v->nmc * v->nmf);
        pmo_h_matrix(v->p, v->M);
        const long nc3 = (long)v->nmc * v->nmc * v->nmc, nf3 = (long)v->nmf * v->nmf * v->nmf;
        v->map_c = malloc(sizeof(uint32_t) * v->xcells * nc3);
        v->map_f = malloc(sizeof(uint32_t) * v->xcells * nf3);
        v->w = malloc(sizeof(double) * v->xcells * nf3);
        pmo_h_maps(mc, v->p, v->map_c, v->map_f, v->w);
      } else {
        v->xfer_h = 0;
        v->nmc = pc + 1;
        v->nmf = v->p + 1;
        v->xcells = v->C;
        v->M = malloc(sizeof(double) * v->nmc * v->nmf);
        pmo_p_matrix(pc, v->p, v->M);
        const long nc3 = (long)v->nmc * v->nmc * v->nmc, nf3 = (long)v->nmf * v->nmf * v->nmf;
        v->map_c = malloc(sizeof(uint32_t) * v->xcells * nc3);
        v->map_f = malloc(sizeof(uint32_t) * v->xcells * nf3);
        v->w = malloc(sizeof(double) * v->xcells * nf3);
        pmo_p_maps(&ms[l], pc, v->p, v->map_c, v->map_f, v->w);
      }
    }
  }
  for (int l = 0; l < n_levels; ++l) {
    if (l > 0) cheby_setup(mg, l, prm->cheby_degree, prm->smoothing_range, prm->eig_cg_n_iterations);
    else if (prm->coarse_kind == 0) cheby_setup(mg, 0, 0, 1e-3, (int)mg->lv[0].N);
  }
  return mg;
}

void pmo_mg_destroy(pmo_mg *mg)
{
  if (!mg) return;
  for (int l = 0; l < mg->L; ++l) {
    struct pmo_level *v = &mg->lv[l];
    free(v->S); free(v->D); free(v->G); free(v->inv_diag); free(v->map); free(v->cons);
    free(v->colors); free(v->sol); free(v->def); free(v->t); free(v->old); free(v->tmp);
    free(v->M); free(v->w); free(v->map_c); free(v->map_f); free(v->sigma);
  }
  free(mg->lv);
  free(mg);
}
