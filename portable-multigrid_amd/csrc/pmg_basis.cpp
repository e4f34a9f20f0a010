// pmg_basis.cpp -- closed-form 1D data (Gauss-Legendre, Gauss-Lobatto,
// Lagrange bases, collocation derivatives, transfer matrices).
#include "pmg_basis.hpp"

#include <cmath>

#include "pmg_common.hpp"

namespace pmg {

namespace {
// Legendre P_n, P_n' on [-1,1] in extended precision
void legendre(int n, long double t, long double &P, long double &dP) {
  long double pm1 = 1.0L, p = t;
  if (n == 0) { P = 1; dP = 0; return; }
  for (int k = 1; k < n; ++k) {
    long double pn = ((2 * k + 1) * t * p - k * pm1) / (k + 1);
    pm1 = p;
    p = pn;
  }
  P = p;
  dP = n * (pm1 - t * p) / (1.0L - t * t);
}
}  // namespace

Quadrature1D gauss_legendre(int n) {
  PMG_REQUIRE(n >= 1 && n <= 32, PMG_ERR_INVALID_ARG, "Gauss-Legendre order out of range");
  Quadrature1D q;
  q.points.resize(n);
  q.weights.resize(n);
  const long double pi = 3.141592653589793238462643383279502884L;
  for (int i = 0; i < (n + 1) / 2; ++i) {
    long double t = std::cos(pi * (i + 0.75L) / (n + 0.5L)), P, dP;
    for (int it = 0; it < 60; ++it) {
      legendre(n, t, P, dP);
      long double dt = P / dP;
      t -= dt;
      if (std::fabs(static_cast<double>(dt)) < 1e-19) break;
    }
    legendre(n, t, P, dP);
    const long double w = 1.0L / ((1.0L - t * t) * dP * dP);
    // t > 0 for i < n/2: mirrored pair on [0,1]
    q.points[n - 1 - i] = static_cast<double>(0.5L + 0.5L * t);
    q.points[i] = static_cast<double>(0.5L - 0.5L * t);
    q.weights[n - 1 - i] = q.weights[i] = static_cast<double>(w);
  }
  if (n % 2 == 1) q.points[n / 2] = 0.5;
  return q;
}

std::vector<double> gauss_lobatto_points(int n) {
  PMG_REQUIRE(n >= 2 && n <= 32, PMG_ERR_INVALID_ARG, "Gauss-Lobatto order out of range");
  std::vector<double> z(n);
  z[0] = 0.0;
  z[n - 1] = 1.0;
  const int N = n - 1;
  const long double pi = 3.141592653589793238462643383279502884L;
  for (int i = 1; i <= (n - 1) / 2; ++i) {
    // roots of P_N' : Newton with P_N'' from the Legendre equation
    long double t = std::cos(pi * i / N), P, dP;
    for (int it = 0; it < 60; ++it) {
      legendre(N, t, P, dP);
      long double d2P = (2 * t * dP - N * (N + 1.0L) * P) / (1.0L - t * t);
      long double dt = dP / d2P;
      t -= dt;
      if (std::fabs(static_cast<double>(dt)) < 1e-19) break;
    }
    z[n - 1 - i] = static_cast<double>(0.5L + 0.5L * t);
    z[i] = static_cast<double>(0.5L - 0.5L * t);
  }
  if (n % 2 == 1) z[n / 2] = 0.5;
  return z;
}

void lagrange_eval(const std::vector<double> &nodes, double x, double *val, double *der) {
  const int m = static_cast<int>(nodes.size());
  // barycentric weights
  std::vector<long double> bw(m, 1.0L);
  for (int i = 0; i < m; ++i)
    for (int j = 0; j < m; ++j)
      if (j != i) bw[i] /= (static_cast<long double>(nodes[i]) - nodes[j]);
  int hit = -1;
  for (int i = 0; i < m; ++i)
    if (x == nodes[i]) hit = i;
  if (hit >= 0) {
    for (int i = 0; i < m; ++i) {
      if (val) val[i] = (i == hit) ? 1.0 : 0.0;
      if (der) {
        if (i != hit) {
          der[i] = static_cast<double>(bw[i] / bw[hit] / (static_cast<long double>(nodes[hit]) - nodes[i]));
        }
      }
    }
    if (der) {
      long double s = 0;
      for (int i = 0; i < m; ++i)
        if (i != hit) s += der[i];
      der[hit] = static_cast<double>(-s);  // rows of the derivative matrix sum to zero
    }
    return;
  }
  // l(x) = prod (x - z_j); phi_i = l(x) bw_i / (x - z_i)
  long double l = 1.0L, lp = 0.0L;
  for (int j = 0; j < m; ++j) {
    lp = lp * (x - nodes[j]) + l;
    l *= (x - nodes[j]);
  }
  for (int i = 0; i < m; ++i) {
    const long double d = x - nodes[i];
    const long double phi = l * bw[i] / d;
    if (val) val[i] = static_cast<double>(phi);
    if (der) der[i] = static_cast<double>(bw[i] * (lp * d - l) / (d * d));
  }
}

ShapeData make_shape_data(int degree, int nq) {
  PMG_REQUIRE(degree >= 1 && nq >= degree + 1, PMG_ERR_INVALID_ARG,
              "need degree >= 1 and n_q_points_1d >= degree+1");
  ShapeData s;
  s.degree = degree;
  s.n_q_points_1d = nq;
  s.quad = gauss_legendre(nq);
  const auto gll = gauss_lobatto_points(degree + 1);
  const int nm = degree + 1;
  s.S.resize(nm * nq);
  s.Sd.resize(nm * nq);
  s.D.resize(nq * nq);
  std::vector<double> v(nm), d(nm), dq(nq);
  for (int q = 0; q < nq; ++q) {
    lagrange_eval(gll, s.quad.points[q], v.data(), d.data());
    for (int i = 0; i < nm; ++i) {
      s.S[i * nq + q] = v[i];
      s.Sd[i * nq + q] = d[i];
    }
    lagrange_eval(s.quad.points, s.quad.points[q], nullptr, dq.data());
    for (int n = 0; n < nq; ++n) s.D[n * nq + q] = dq[n];
  }
  return s;
}

std::vector<double> h_transfer_matrix(int p) {
  const int nm = p + 1, nf = 2 * p + 1;
  const auto z = gauss_lobatto_points(nm);
  std::vector<double> M(nm * nf), v(nm);
  for (int j = 0; j < nf; ++j) {
    // fine patch node j: child 0 for j <= p, child 1 otherwise
    const double x = (j <= p) ? 0.5 * z[j] : 0.5 * (1.0 + z[j - p]);
    lagrange_eval(z, x, v.data(), nullptr);
    for (int i = 0; i < nm; ++i) M[i * nf + j] = v[i];
  }
  return M;
}

std::vector<double> p_transfer_matrix(int pc, int pf) {
  const int nc = pc + 1, nf = pf + 1;
  const auto zc = gauss_lobatto_points(nc), zf = gauss_lobatto_points(nf);
  std::vector<double> M(nc * nf), v(nc);
  for (int j = 0; j < nf; ++j) {
    lagrange_eval(zc, zf[j], v.data(), nullptr);
    for (int i = 0; i < nc; ++i) M[i * nf + j] = v[i];
  }
  return M;
}

}  // namespace pmg
