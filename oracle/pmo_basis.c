/*
 * pmo_basis.c -- 1D polynomial data of the oracle (test infrastructure only).
 *
 * Restates the 1D data deal.II hands the reference kernels:
 *   shape_values[i*nq+q]       = phi_i(xi_q), FE_Q<1>(p) Lagrange basis on the
 *                                Gauss-Lobatto support points, xi = QGauss<1>(nq)
 *                                (portable_laplace_operator_bk3.h:120,162;
 *                                 bk3_kokkos_kernel.h:212-218)
 *   co_shape_gradients[n*nq+q] = l_n'(xi_q), collocation derivative of the
 *                                Lagrange basis on the Gauss points
 *                                (bk3_kokkos_kernel.h:388-393,418-448)
 * Pinned by closed forms in tests/test_oracle_basis.py.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "pmg_oracle.h"

/* Legendre P_n and P_n' at t in [-1,1] by the three-term recurrence */
static void legendre(int n, double t, double *P, double *dP)
{
  double p0 = 1.0, p1 = t;
  if (n == 0) { *P = 1.0; *dP = 0.0; return; }
  for (int k = 2; k <= n; ++k) {
    double p2 = ((2.0 * k - 1.0) * t * p1 - (k - 1.0) * p0) / k;
    p0 = p1;
    p1 = p2;
  }
  *P = p1;
  /* derivative from n (t P_n - P_{n-1}) / (t^2 - 1), away from the ends */
  *dP = n * (t * p1 - p0) / (t * t - 1.0);
}

void pmo_gauss_legendre(int n, double *x, double *w)
{
  for (int i = 0; i < n; ++i) {
    double t = cos(M_PI * (i + 0.75) / (n + 0.5));
    double P = 0, dP = 1;
    for (int it = 0; it < 100; ++it) {
      legendre(n, t, &P, &dP);
      double dt = P / dP;
      t -= dt;
      if (fabs(dt) < 1e-17) break;
    }
    legendre(n, t, &P, &dP);
    /* nodes ascending on [0,1] */
    x[n - 1 - i] = 0.5 * (1.0 + t);
    w[n - 1 - i] = 1.0 / ((1.0 - t * t) * dP * dP); /* = 2/((1-t^2)P'^2) / 2 */
  }
}

void pmo_gauss_lobatto(int n, double *x)
{
  /* interior nodes are the roots of P'_{n-1}; Newton on P'_{n-1} using
   * P'' from the Legendre equation */
  const int N = n - 1;
  x[0] = 0.0;
  x[n - 1] = 1.0;
  for (int i = 1; i < n - 1; ++i) {
    double t = -cos(M_PI * i / N);
    for (int it = 0; it < 100; ++it) {
      double P, dP;
      legendre(N, t, &P, &dP);
      double d2P = (2.0 * t * dP - N * (N + 1.0) * P) / (1.0 - t * t);
      double dt = dP / d2P;
      t -= dt;
      if (fabs(dt) < 1e-17) break;
    }
    x[i] = 0.5 * (1.0 + t);
  }
  /* symmetrize to remove round-off asymmetry */
  for (int i = 0; i < n / 2; ++i) {
    double a = 0.5 * (x[i] + (1.0 - x[n - 1 - i]));
    x[i] = a;
    x[n - 1 - i] = 1.0 - a;
  }
  if (n % 2 == 1) x[n / 2] = 0.5;
}

void pmo_lagrange(int m, const double *z, double x, double *val, double *der)
{
  for (int i = 0; i < m; ++i) {
    double v = 1.0;
    for (int j = 0; j < m; ++j)
      if (j != i) v *= (x - z[j]) / (z[i] - z[j]);
    double d = 0.0;
    for (int k = 0; k < m; ++k) {
      if (k == i) continue;
      double t = 1.0 / (z[i] - z[k]);
      for (int j = 0; j < m; ++j)
        if (j != i && j != k) t *= (x - z[j]) / (z[i] - z[j]);
      d += t;
    }
    if (val) val[i] = v;
    if (der) der[i] = d;
  }
}

void pmo_shape_matrices(int p, int nq, double *S, double *D)
{
  const int nm = p + 1;
  double *gll = malloc(sizeof(double) * nm);
  double *gx = malloc(sizeof(double) * nq), *gw = malloc(sizeof(double) * nq);
  double *v = malloc(sizeof(double) * (nm > nq ? nm : nq));
  double *d = malloc(sizeof(double) * (nm > nq ? nm : nq));
  pmo_gauss_lobatto(nm, gll);
  pmo_gauss_legendre(nq, gx, gw);
  for (int q = 0; q < nq; ++q) {
    pmo_lagrange(nm, gll, gx[q], v, NULL);
    for (int i = 0; i < nm; ++i) S[i * nq + q] = v[i];
    pmo_lagrange(nq, gx, gx[q], NULL, d);
    for (int n = 0; n < nq; ++n) D[n * nq + q] = d[n];
  }
  free(gll); free(gx); free(gw); free(v); free(d);
}

void pmo_shape_gradients(int p, int nq, double *Sd)
{
  const int nm = p + 1;
  double *gll = malloc(sizeof(double) * nm);
  double *gx = malloc(sizeof(double) * nq), *gw = malloc(sizeof(double) * nq);
  double *d = malloc(sizeof(double) * nm);
  pmo_gauss_lobatto(nm, gll);
  pmo_gauss_legendre(nq, gx, gw);
  for (int q = 0; q < nq; ++q) {
    pmo_lagrange(nm, gll, gx[q], NULL, d);
    for (int i = 0; i < nm; ++i) Sd[i * nq + q] = d[i];
  }
  free(gll); free(gx); free(gw); free(d);
}

/* Symmetric tridiagonal eigenvalues by bisection with Sturm counts
 * (used for the Lanczos estimate of the Chebyshev smoother). */
static int sturm_count(int n, const double *a, const double *b, double x)
{
  int c = 0;
  double q = a[0] - x;
  if (q < 0) ++c;
  for (int i = 1; i < n; ++i) {
    if (q == 0.0) q = 1e-300;
    q = a[i] - x - b[i - 1] * b[i - 1] / q;
    if (q < 0) ++c;
  }
  return c;
}

void pmo_tridiag_eigs(int n, const double *diag, const double *off, double *eig)
{
  if (n <= 0) return;
  double lo = diag[0], hi = diag[0];
  for (int i = 0; i < n; ++i) {
    double r = 0;
    if (i > 0) r += fabs(off[i - 1]);
    if (i < n - 1) r += fabs(off[i]);
    if (diag[i] - r < lo) lo = diag[i] - r;
    if (diag[i] + r > hi) hi = diag[i] + r;
  }
  for (int k = 0; k < n; ++k) {
    double a = lo, b = hi;
    for (int it = 0; it < 200; ++it) {
      double mid = 0.5 * (a + b);
      if (sturm_count(n, diag, off, mid) > k) b = mid; else a = mid;
      if (b - a <= 1e-15 * (fabs(a) + fabs(b)) + 1e-300) break;
    }
    eig[k] = 0.5 * (a + b);
  }
}
