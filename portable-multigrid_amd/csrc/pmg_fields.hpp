// pmg_fields.hpp -- launchers for right-hand sides and L2 errors of Q_p fields.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace pmg {
// b[map[c][i]] += sum_q S_i(q) f(X_q) JxW_q on the trilinear cells
// (cv = [C][8][3] vertices, S = [(p+1)*nq] shape values, double)
template <typename T>
void launch_integrate(int kind, int nm, int nq, int64_t n_cells, const double *cv,
                      const double *qp, const double *qw, const double *S, const uint32_t *map,
                      T *b, hipStream_t s);
// per-block partial sums of (u_h - u)^2 JxW with Gauss(nq)
template <typename T>
void launch_l2_error(int kind, int nm, int nq, int64_t n_cells, const double *cv, const double *qp,
                     const double *qw, const double *S, const uint32_t *map, const T *u,
                     double *partials, int n_blocks, hipStream_t s);
}  // namespace pmg
