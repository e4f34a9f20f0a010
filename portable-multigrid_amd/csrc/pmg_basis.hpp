// pmg_basis.hpp -- 1D polynomial data for the sum-factorised kernels (host side).
//
// The reference takes these from deal.II (Portable::MatrixFree shape data,
// FE_Q<1>::get_prolongation_matrix, FETools::get_projection_matrix); here they
// are computed from their closed forms:
//   S[i*nq+q]  = phi_i(xi_q), phi_i the degree-p Lagrange basis on the p+1
//                Gauss-Lobatto points, xi_q the nq Gauss-Legendre points on [0,1]
//                (deal.II "shape_values", bk3_kokkos_kernel.h:212-218)
//   D[n*nq+q]  = l_n'(xi_q), Lagrange basis on the Gauss points
//                (deal.II "co_shape_gradients", bk3_kokkos_kernel.h:388-393)
//   Mh[i*(2p+1)+j] = phi_i at the j-th node of the two-child patch
//                (portable_geometric_transfer.h:1285-1337)
//   Mp[i*(pf+1)+j] = phi_i^{pc}(GLL_j^{pf}) (portable_polynomial_transfer.h:1067-1111)
#pragma once

#include <vector>

namespace pmg {

struct Quadrature1D {
  std::vector<double> points, weights;  // on [0,1], ascending
};

Quadrature1D gauss_legendre(int n);
std::vector<double> gauss_lobatto_points(int n);
// values and derivatives of the Lagrange basis on `nodes` at x (barycentric form)
void lagrange_eval(const std::vector<double> &nodes, double x, double *val, double *der);

struct ShapeData {
  int degree = 0, n_q_points_1d = 0;
  std::vector<double> S;   // [(p+1)*nq]
  std::vector<double> D;   // [nq*nq]
  std::vector<double> Sd;  // [(p+1)*nq] exact derivative phi_i'(xi_q)
  Quadrature1D quad;
};

ShapeData make_shape_data(int degree, int n_q_points_1d);
std::vector<double> h_transfer_matrix(int degree);
std::vector<double> p_transfer_matrix(int degree_coarse, int degree_fine);

}  // namespace pmg
