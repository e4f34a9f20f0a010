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

// F[c][q] = f(X_q) |det J_q| w_q
__global__ void __launch_bounds__(256)
f_at_q_kernel(int kind, int nq, int64_t n_cells, const double *__restrict__ cv,
              const double *__restrict__ qp, const double *__restrict__ qw, double *__restrict__ F) {
  const int nq3 = nq * nq * nq;
  const int64_t gid = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (gid >= n_cells * nq3) return;
  const int64_t c = gid / nq3;
  const int q = static_cast<int>(gid - c * nq3);
  const int q0 = q % nq, q1 = (q / nq) % nq, q2 = q / (nq * nq);
  const double xi[3] = {qp[q0], qp[q1], qp[q2]};
  double X[3], det;
  trilinear(cv + c * 24, xi, X, det);
  F[gid] = rhs_f(kind, X) * fabs(det) * qw[q0] * qw[q1] * qw[q2];
}

// b[map[c][i]] += sum_q phi_i(q) F[c][q]   (one lane per (cell, i))
template <typename T>
__global__ void __launch_bounds__(256)
integrate_kernel(int nm, int nq, int64_t n_cells, const double *__restrict__ S,
                 const uint32_t *__restrict__ map, const double *__restrict__ F, T *__restrict__ b) {
  const int nm3 = nm * nm * nm, nq3 = nq * nq * nq;
  const int64_t gid = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (gid >= n_cells * nm3) return;
  const uint32_t dof = map[gid];
  if (dof == kInvalid) return;
  const int64_t c = gid / nm3;
  const int i = static_cast<int>(gid - c * nm3);
  const int i0 = i % nm, i1 = (i / nm) % nm, i2 = i / (nm * nm);
  const double *Fc = F + c * nq3;
  double acc = 0;
  for (int q2 = 0; q2 < nq; ++q2)
    for (int q1 = 0; q1 < nq; ++q1) {
      const double s12 = S[i1 * nq + q1] * S[i2 * nq + q2];
      double t = 0;
      for (int q0 = 0; q0 < nq; ++q0) t += S[i0 * nq + q0] * Fc[(q2 * nq + q1) * nq + q0];
      acc += s12 * t;
    }
  atomic_add(b + dof, static_cast<T>(acc));
}

// per-block partial sums of (u_h - u)^2 JxW; one lane per (cell, q)
template <typename T>
__global__ void __launch_bounds__(256)
l2_error_kernel(int kind, int nm, int nq, int64_t n_cells, const double *__restrict__ cv,
                const double *__restrict__ qp, const double *__restrict__ qw,
                const double *__restrict__ S, const uint32_t *__restrict__ map,
                const T *__restrict__ u, double *__restrict__ partials) {
  __shared__ double red[256];
  const int nm3 = nm * nm * nm, nq3 = nq * nq * nq;
  double e2 = 0;
  for (int64_t gid = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
       gid < n_cells * nq3; gid += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const int64_t c = gid / nq3;
    const int q = static_cast<int>(gid - c * nq3);
    const int q0 = q % nq, q1 = (q / nq) % nq, q2 = q / (nq * nq);
    const double xi[3] = {qp[q0], qp[q1], qp[q2]};
    double X[3], det;
    trilinear(cv + c * 24, xi, X, det);
    double uh = 0;
    const uint32_t *m = map + c * nm3;
    for (int i2 = 0; i2 < nm; ++i2)
      for (int i1 = 0; i1 < nm; ++i1) {
        const double s12 = S[i1 * nq + q1] * S[i2 * nq + q2];
        for (int i0 = 0; i0 < nm; ++i0)
          {
          const uint32_t d = m[(i2 * nm + i1) * nm + i0];
          if (d != kInvalid) uh += s12 * S[i0 * nq + q0] * static_cast<double>(u[d]);
        }
      }
    const double d = uh - exact_u(kind, X);
    e2 += d * d * fabs(det) * qw[q0] * qw[q1] * qw[q2];
  }
  red[threadIdx.x] = e2;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) partials[blockIdx.x] = red[0];
}

}  // namespace

void launch_f_at_q(int kind, int nq, int64_t n_cells, const double *cv, const double *qp,
                   const double *qw, double *F, hipStream_t s) {
  const int64_t n = n_cells * nq * nq * nq;
  if (n == 0) return;
  hipLaunchKernelGGL(f_at_q_kernel, dim3(unsigned((n + 255) / 256)), dim3(256), 0, s, kind, nq,
                     n_cells, cv, qp, qw, F);
  PMG_HIP_CHECK(hipGetLastError());
}

template <typename T>
void launch_integrate(int nm, int nq, int64_t n_cells, const double *S, const uint32_t *map,
                      const double *F, T *b, hipStream_t s) {
  const int64_t n = n_cells * nm * nm * nm;
  if (n == 0) return;
  hipLaunchKernelGGL(integrate_kernel<T>, dim3(unsigned((n + 255) / 256)), dim3(256), 0, s, nm, nq,
                     n_cells, S, map, F, b);
  PMG_HIP_CHECK(hipGetLastError());
}

template <typename T>
void launch_l2_error(int kind, int nm, int nq, int64_t n_cells, const double *cv, const double *qp,
                     const double *qw, const double *S, const uint32_t *map, const T *u,
                     double *partials, int n_blocks, hipStream_t s) {
  hipLaunchKernelGGL(l2_error_kernel<T>, dim3(n_blocks), dim3(256), 0, s, kind, nm, nq, n_cells, cv,
                     qp, qw, S, map, u, partials);
  PMG_HIP_CHECK(hipGetLastError());
}

template void launch_integrate<double>(int, int, int64_t, const double *, const uint32_t *,
                                       const double *, double *, hipStream_t);
template void launch_integrate<float>(int, int, int64_t, const double *, const uint32_t *,
                                      const double *, float *, hipStream_t);
template void launch_l2_error<double>(int, int, int, int64_t, const double *, const double *,
                                      const double *, const double *, const uint32_t *,
                                      const double *, double *, int, hipStream_t);
template void launch_l2_error<float>(int, int, int, int64_t, const double *, const double *,
                                     const double *, const double *, const uint32_t *,
                                     const float *, double *, int, hipStream_t);

}  // namespace pmg
