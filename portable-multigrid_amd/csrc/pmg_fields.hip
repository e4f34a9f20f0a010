// pmg_fields.hip -- right-hand side assembly and L2 error of Q_p fields (gfx950).
//
// b_i = int phi_i f dx with Gauss(nq) on the trilinear cells
//   (LaplaceProblem::compute_rhs, tests/poisson_cube_bk3/program.cc:517-566;
//    constrained entries stay 0 as with distribute_local_to_global);
// ||u_h - u||_L2 with Gauss(nq_err) (VectorTools::integrate_difference,
//   source/geometric_multigrid_bk3/program.cc:391-400).
// f / u kinds: 0 -> f = 1, u = 0;  1 -> u = sin(pi x) sin(pi y) sin(pi z),
// f = 3 pi^2 u.
#include "pmg_common.hpp"
#include "pmg_fields.hpp"

#include <algorithm>

namespace pmg {

namespace {

__device__ inline void trilinear(const double *v, const double xi[3], double X[3], double &det) {
  double J[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
  X[0] = X[1] = X[2] = 0;
  for (int vi = 0; vi < 8; ++vi) {
    const int b0 = vi & 1, b1 = (vi >> 1) & 1, b2 = (vi >> 2) & 1;
    const double f0 = b0 ? xi[0] : 1.0 - xi[0], f1 = b1 ? xi[1] : 1.0 - xi[1],
                 f2 = b2 ? xi[2] : 1.0 - xi[2];
    const double d0 = b0 ? 1.0 : -1.0, d1 = b1 ? 1.0 : -1.0, d2 = b2 ? 1.0 : -1.0;
    const double N = f0 * f1 * f2;
    const double dN[3] = {d0 * f1 * f2, f0 * d1 * f2, f0 * f1 * d2};
    for (int d = 0; d < 3; ++d) {
      X[d] += N * v[3 * vi + d];
      for (int k = 0; k < 3; ++k) J[d][k] += dN[k] * v[3 * vi + d];
    }
  }
  det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) -
        J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
        J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
}

__device__ inline double exact_u(int kind, const double X[3]) {
  if (kind == 0) return 0.0;
  return sin(M_PI * X[0]) * sin(M_PI * X[1]) * sin(M_PI * X[2]);
}
__device__ inline double rhs_f(int kind, const double X[3]) {
  if (kind == 0) return 1.0;
  return 3.0 * M_PI * M_PI * sin(M_PI * X[0]) * sin(M_PI * X[1]) * sin(M_PI * X[2]);
}

// Both kernels below run one 64-lane workgroup per cell (grid-stride over the
// cells) and sum-factorise the tensor-product contractions through LDS:
// nm^2 nq + nm nq^2 + nq^3 products of length nm (or nq) per cell instead of
// nm^3 nq^3 for a direct evaluation at every quadrature point.  LDS layout
// (doubles): A [nq^3], B [nm nq^2], C [nm^2 nq], D [nm^3].
__device__ inline void sf_layout(double *lds, int nm, int nq, double *&A, double *&B, double *&C,
                                 double *&D) {
  A = lds;
  B = A + nq * nq * nq;
  C = B + nm * nq * nq;
  D = C + nm * nm * nq;
}

// b[map[c][i]] += sum_q phi_i(q) f(X_q) |det J_q| w_q
template <typename T>
__global__ void __launch_bounds__(64)
integrate_kernel(int kind, int nm, int nq, int64_t n_cells, const double *__restrict__ cv,
                 const double *__restrict__ qp, const double *__restrict__ qw,
                 const double *__restrict__ S, const uint32_t *__restrict__ map, T *__restrict__ b) {
  extern __shared__ double lds[];
  double *F, *W1, *W2, *BL;
  sf_layout(lds, nm, nq, F, W1, W2, BL);
  const int lane = threadIdx.x, nq2 = nq * nq, nq3 = nq2 * nq, nm2 = nm * nm, nm3 = nm2 * nm;
  for (int64_t c = blockIdx.x; c < n_cells; c += gridDim.x) {
    for (int q = lane; q < nq3; q += 64) {
      const int q0 = q % nq, q1 = (q / nq) % nq, q2 = q / nq2;
      const double xi[3] = {qp[q0], qp[q1], qp[q2]};
      double X[3], det;
      trilinear(cv + c * 24, xi, X, det);
      F[q] = rhs_f(kind, X) * fabs(det) * qw[q0] * qw[q1] * qw[q2];
    }
    __syncthreads();
    for (int o = lane; o < nm * nq2; o += 64) {  // W1[q2][q1][i0]
      const int i0 = o % nm, r = o / nm;
      double v = 0;
      for (int q0 = 0; q0 < nq; ++q0) v += S[i0 * nq + q0] * F[r * nq + q0];
      W1[o] = v;
    }
    __syncthreads();
    for (int o = lane; o < nm2 * nq; o += 64) {  // W2[q2][i1][i0]
      const int i0 = o % nm, i1 = (o / nm) % nm, q2 = o / nm2;
      double v = 0;
      for (int q1 = 0; q1 < nq; ++q1) v += S[i1 * nq + q1] * W1[(q2 * nq + q1) * nm + i0];
      W2[o] = v;
    }
    __syncthreads();
    for (int o = lane; o < nm3; o += 64) {  // BL[i2][i1][i0]
      const int r = o % nm2, i2 = o / nm2;
      double v = 0;
      for (int q2 = 0; q2 < nq; ++q2) v += S[i2 * nq + q2] * W2[q2 * nm2 + r];
      const uint32_t dof = map[c * nm3 + o];
      if (dof != kInvalid) atomic_add(b + dof, static_cast<T>(v));
    }
    __syncthreads();
  }
}

// per-block partial sums of (u_h - u)^2 |det J| w over the block's cells
template <typename T>
__global__ void __launch_bounds__(64)
l2_error_kernel(int kind, int nm, int nq, int64_t n_cells, const double *__restrict__ cv,
                const double *__restrict__ qp, const double *__restrict__ qw,
                const double *__restrict__ S, const uint32_t *__restrict__ map,
                const T *__restrict__ u, double *__restrict__ partials) {
  extern __shared__ double lds[];
  double *UQ, *T2, *T1, *UL;
  sf_layout(lds, nm, nq, UQ, T2, T1, UL);
  const int lane = threadIdx.x, nq2 = nq * nq, nq3 = nq2 * nq, nm2 = nm * nm, nm3 = nm2 * nm;
  double e2 = 0;
  for (int64_t c = blockIdx.x; c < n_cells; c += gridDim.x) {
    for (int o = lane; o < nm3; o += 64) {
      const uint32_t d = map[c * nm3 + o];
      UL[o] = d == kInvalid ? 0.0 : static_cast<double>(u[d]);
    }
    __syncthreads();
    for (int o = lane; o < nm2 * nq; o += 64) {  // T1[i2][i1][q0]
      const int q0 = o % nq, r = o / nq;
      double v = 0;
      for (int i0 = 0; i0 < nm; ++i0) v += S[i0 * nq + q0] * UL[r * nm + i0];
      T1[o] = v;
    }
    __syncthreads();
    for (int o = lane; o < nm * nq2; o += 64) {  // T2[i2][q1][q0]
      const int q0 = o % nq, q1 = (o / nq) % nq, i2 = o / nq2;
      double v = 0;
      for (int i1 = 0; i1 < nm; ++i1) v += S[i1 * nq + q1] * T1[(i2 * nm + i1) * nq + q0];
      T2[o] = v;
    }
    __syncthreads();
    for (int q = lane; q < nq3; q += 64) {  // u_h at the quadrature points
      const int r = q % nq2, q2 = q / nq2, q0 = q % nq, q1 = (q / nq) % nq;
      double uh = 0;
      for (int i2 = 0; i2 < nm; ++i2) uh += S[i2 * nq + q2] * T2[i2 * nq2 + r];
      const double xi[3] = {qp[q0], qp[q1], qp[q2]};
      double X[3], det;
      trilinear(cv + c * 24, xi, X, det);
      const double d = uh - exact_u(kind, X);
      e2 += d * d * fabs(det) * qw[q0] * qw[q1] * qw[q2];
    }
    __syncthreads();
  }
  for (int off = 32; off > 0; off >>= 1) e2 += __shfl_down(e2, off, 64);
  if (lane == 0) partials[blockIdx.x] = e2;
}

size_t sf_lds_bytes(int nm, int nq) {
  return sizeof(double) * (size_t(nq) * nq * nq + size_t(nm) * nq * nq + size_t(nm) * nm * nq +
                           size_t(nm) * nm * nm);
}

}  // namespace

template <typename T>
void launch_integrate(int kind, int nm, int nq, int64_t n_cells, const double *cv,
                      const double *qp, const double *qw, const double *S, const uint32_t *map,
                      T *b, hipStream_t s) {
  if (n_cells == 0) return;
  const unsigned nb = static_cast<unsigned>(std::min<int64_t>(n_cells, 16384));
  hipLaunchKernelGGL(integrate_kernel<T>, dim3(nb), dim3(64), sf_lds_bytes(nm, nq), s, kind, nm,
                     nq, n_cells, cv, qp, qw, S, map, b);
  PMG_HIP_CHECK(hipGetLastError());
}

template <typename T>
void launch_l2_error(int kind, int nm, int nq, int64_t n_cells, const double *cv, const double *qp,
                     const double *qw, const double *S, const uint32_t *map, const T *u,
                     double *partials, int n_blocks, hipStream_t s) {
  hipLaunchKernelGGL(l2_error_kernel<T>, dim3(n_blocks), dim3(64), sf_lds_bytes(nm, nq), s, kind,
                     nm, nq, n_cells, cv, qp, qw, S, map, u, partials);
  PMG_HIP_CHECK(hipGetLastError());
}

template void launch_integrate<double>(int, int, int, int64_t, const double *, const double *,
                                       const double *, const double *, const uint32_t *, double *,
                                       hipStream_t);
template void launch_integrate<float>(int, int, int, int64_t, const double *, const double *,
                                      const double *, const double *, const uint32_t *, float *,
                                      hipStream_t);
template void launch_l2_error<double>(int, int, int, int64_t, const double *, const double *,
                                      const double *, const double *, const uint32_t *,
                                      const double *, double *, int, hipStream_t);
template void launch_l2_error<float>(int, int, int, int64_t, const double *, const double *,
                                     const double *, const double *, const uint32_t *,
                                     const float *, double *, int, hipStream_t);

}  // namespace pmg
