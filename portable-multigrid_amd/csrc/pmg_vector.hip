// pmg_vector.hip -- streaming vector kernels of the solver stack (gfx950).
// These replace the deal.II device-vector operations that SolverCG and
// PreconditionChebyshev perform between operator applications
// (tests/common/portable_multigrid_solver.h:171-271); each is fused so that a
// CG or Chebyshev iteration touches every vector entry the minimum number of
// times.  Reductions are deterministic: a fixed grid of kDotBlocks blocks
// writes partial sums that one block adds in a fixed order.
#include "pmg_common.hpp"
#include "pmg_kernels.hpp"

#include <algorithm>
#include <type_traits>

namespace pmg {

namespace {
constexpr int kThreads = 256;
inline unsigned grid_for(int64_t n) {
  int64_t b = (n + kThreads - 1) / kThreads;
  if (b > 65536) b = 65536;
  return static_cast<unsigned>(b < 1 ? 1 : b);
}

template <typename T> __global__ void fill_k(int64_t n, T *x, T v) {
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x)
    x[i] = v;
}
template <typename T> __global__ void copy_k(int64_t n, T *d, const T *s) {
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x)
    d[i] = s[i];
}
template <typename T>
__global__ void copy_idx_k(int64_t n, const uint32_t *idx, T *d, const T *s) {
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x) {
    const uint32_t j = idx[i];
    d[j] = s[j];
  }
}
// d = 0 on free DoFs, d = s on constrained ones (mask[i] != 0)
template <typename T>
__global__ void prep_constrained_k(int64_t n, const uint8_t *mask, T *d, const T *s) {
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x)
    d[i] = mask[i] ? s[i] : T(0);
}
template <typename T>
__global__ void set_idx_k(int64_t n, const uint32_t *idx, T *d, T v) {
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x)
    d[idx[i]] = v;
}
// first[0] = min { i : !(d[i] > 0) } (first[0] = n beforehand)
template <typename T>
__global__ void first_nonpositive_k(int64_t n, const T *d, unsigned long long *first) {
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x)
    if (!(d[i] > T(0))) atomicMin(first, static_cast<unsigned long long>(i));
}
template <typename T> __global__ void invert_k(int64_t n, T *x) {
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x)
    x[i] = T(1) / x[i];
}
template <typename TD, typename TS> __global__ void convert_k(int64_t n, TD *d, const TS *s) {
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x)
    d[i] = static_cast<TD>(s[i]);
}
template <typename T> __global__ void residual_k(int64_t n, T *t, const T *b) {
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x)
    t[i] = b[i] - t[i];
}

// block-level sum of a double (wave64 shuffles, then LDS across waves)
__device__ inline double block_sum(double v) {
  __shared__ double part[kThreads / 64];
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) part[wave] = v;
  __syncthreads();
  double s = 0;
  if (threadIdx.x == 0)
    for (int w = 0; w < kThreads / 64; ++w) s += part[w];
  __syncthreads();
  return s;  // valid in thread 0
}

// stream_a: a is read non-temporally (its last use for a while)
template <typename TA, typename TB>
__global__ void __launch_bounds__(kThreads)
dot_k(int64_t n, const TA *a, const TB *b, double *partials, bool stream_a) {
  double s = 0;
  for (int64_t i = blockIdx.x * int64_t(kThreads) + threadIdx.x; i < n; i += int64_t(gridDim.x) * kThreads)
    s += static_cast<double>(stream_a ? __builtin_nontemporal_load(a + i) : a[i]) *
         static_cast<double>(b[i]);
  s = block_sum(s);
  if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

// sc[slot + k] = sum_b partials[k * n_partials + b]; nothing once *skip != 0
// (a CG iteration queued after convergence)
__global__ void __launch_bounds__(kThreads)
reduce_k(const double *partials, int n_partials, double *sc, int slot, int n_sums,
         const double *skip) {
  if (skip && *skip != 0.0) return;
  for (int k = 0; k < n_sums; ++k) {
    double s = 0;
    for (int i = threadIdx.x; i < n_partials; i += kThreads) s += partials[k * n_partials + i];
    s = block_sum(s);
    if (threadIdx.x == 0) sc[slot + k] = s;
  }
}

// prep (optional, may be t itself: each element is read before it is
// written): the output of the matvec that follows, prep = mask ? x_new : 0
template <typename T>
__global__ void cheby_k(int mode, int64_t n, T *x, T *old, const T *b, const T *t, const T *pinv,
                        T f1, T f2, bool old_zero, T *prep, const uint8_t *mask) {
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x) {
    // b, the inverse diagonal, t and old are streamed (non-temporal): only
    // x (the next matvec's input) and prep (its output) should stay in the
    // last-level cache for the matvec that follows
    T xn;
    const T pi = __builtin_nontemporal_load(pinv + i), bi = __builtin_nontemporal_load(b + i);
    if (mode == 0) {
      xn = f2 * pi * bi;
    } else if (mode == 1) {
      const T xi = x[i];
      __builtin_nontemporal_store(xi, old + i);
      xn = xi + f2 * pi * (bi - __builtin_nontemporal_load(t + i));
    } else {
      const T xi = x[i];
      const T xo = old_zero ? T(0) : __builtin_nontemporal_load(old + i);
      __builtin_nontemporal_store(xi, old + i);
      xn = (T(1) + f1) * xi - f1 * xo + f2 * pi * (bi - __builtin_nontemporal_load(t + i));
    }
    x[i] = xn;
    if (prep) prep[i] = (mask && mask[i]) ? xn : T(0);
  }
}

// CG scalar slots (device doubles):
//  0 rho = r.z      1 p.q       2 r.r       3 rho_old   4 alpha   5 beta
//  6 the iteration that converged   7 converged flag
enum { kRho = 0, kPq = 1, kRr = 2, kRhoOld = 3, kAlpha = 4, kBeta = 5, kConvIt = kCgConvItSlot,
       kConverged = kCgConvergedSlot };

template <typename T>
__global__ void __launch_bounds__(kThreads)
cg_xr_k(int64_t n, T *x, T *r, const T *p, const T *q, const double *sc, double *partials) {
  const double pq = sc[kPq];
  const double alpha = pq != 0.0 ? sc[kRho] / pq : 0.0;
  const T a = static_cast<T>(alpha);
  double s = 0;
  // x, p and q are streamed (not needed again before the next iteration's
  // update); r stays cached for the preconditioner that reads it next
  for (int64_t i = blockIdx.x * int64_t(kThreads) + threadIdx.x; i < n; i += int64_t(gridDim.x) * kThreads) {
    __builtin_nontemporal_store(__builtin_nontemporal_load(x + i) + a * __builtin_nontemporal_load(p + i),
                                x + i);
    const T ri = r[i] - a * __builtin_nontemporal_load(q + i);
    r[i] = ri;
    s += static_cast<double>(ri) * static_cast<double>(ri);
  }
  s = block_sum(s);
  if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

// The CG kernels below return at once when skip (the converged flag) is set:
// the host may queue iterations beyond the one that converges, and they
// then change nothing.
template <typename T>
__global__ void __launch_bounds__(kThreads)
cg_pz_k(int64_t n, const T *r, const T *pinv, T *z, double *partials, const double *skip) {
  if (skip && *skip != 0.0) return;
  double s = 0;
  for (int64_t i = blockIdx.x * int64_t(kThreads) + threadIdx.x; i < n; i += int64_t(gridDim.x) * kThreads) {
    const T zi = pinv ? pinv[i] * r[i] : r[i];
    z[i] = zi;
    s += static_cast<double>(zi) * static_cast<double>(r[i]);
  }
  s = block_sum(s);
  if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

template <typename T, typename TZ>
__global__ void cg_p_k(int64_t n, T *p, const TZ *z, const double *sc, int first) {
  const double ro = sc[kRhoOld];
  const T beta = (first || ro == 0.0) ? T(0) : static_cast<T>(sc[kRho] / ro);
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x) {
    const T zi = static_cast<T>(__builtin_nontemporal_load(z + i));
    p[i] = first ? zi : zi + beta * p[i];
  }
}

// The CG update with x deferred: iteration k updates r with alpha_k, and the
// p update of iteration k+1, which reads p_k anyway, also applies
// x += alpha_k p_k (cg_xp_k); a last cg_x_k after the loop.  The same
// arithmetic as updating x and r together, one pass over p less.
// With stage, the new residual is also written there in the preconditioner's
// precision (the V-cycle input of the next iteration: one pass over r less).
template <typename T, typename TS>
__global__ void __launch_bounds__(kThreads)
cg_r_k(int64_t n, T *r, const T *q, const double *sc, double *partials, TS *stage,
       const double *skip) {
  if (skip && *skip != 0.0) return;
  const double pq = sc[kPq];
  const T a = static_cast<T>(pq != 0.0 ? sc[kRho] / pq : 0.0);
  double s = 0;
  for (int64_t i = blockIdx.x * int64_t(kThreads) + threadIdx.x; i < n; i += int64_t(gridDim.x) * kThreads) {
    const T ri = r[i] - a * __builtin_nontemporal_load(q + i);
    r[i] = ri;
    if (stage) stage[i] = static_cast<TS>(ri);
    s += static_cast<double>(ri) * static_cast<double>(ri);
  }
  s = block_sum(s);
  if (threadIdx.x == 0) partials[blockIdx.x] = s;
}
template <typename T, typename TZ>
__global__ void cg_xp_k(int64_t n, T *x, T *p, const TZ *z, const double *sc, int first,
                        const double *skip) {
  if (skip && *skip != 0.0) return;
  const double ro = sc[kRhoOld];
  const T beta = (first || ro == 0.0) ? T(0) : static_cast<T>(sc[kRho] / ro);
  const T a = first ? T(0) : static_cast<T>(sc[kAlpha]);
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x) {
    const T zi = static_cast<T>(__builtin_nontemporal_load(z + i));
    if (first) {
      p[i] = zi;
    } else {
      const T pi = p[i];
      __builtin_nontemporal_store(__builtin_nontemporal_load(x + i) + a * pi, x + i);
      p[i] = zi + beta * pi;
    }
  }
}
template <typename T>
__global__ void cg_x_k(int64_t n, T *x, const T *p, const double *sc) {
  const T a = static_cast<T>(sc[kAlpha]);
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x)
    x[i] += a * p[i];
}

// end of iteration it: alpha = rho / p.q kept for the deferred x update,
// rho_old = rho; host[0] = r.r (optional, pinned); sc[kConverged] = 1 once
// r.r <= target2 (or is NaN), with sc[kConvIt] = host[1] = it, which makes
// the CG kernels of iterations queued after it return at once, and skips the
// speculatively queued input pass of the next iteration.  op 2: clear the
// flag and the iteration (start of a solve); nothing once converged.
__global__ void cg_scalars_k(double *sc, int op, double *host, double target2, int it) {
  if (threadIdx.x == 0) {
    if (op == 2) {
      sc[kConverged] = 0.0;
      sc[kConvIt] = 0.0;
      if (host) host[1] = 0.0;
      return;
    }
    if (sc[kConverged] != 0.0) return;
    if (op == 0) {
      const double pq = sc[kPq];
      sc[kAlpha] = pq != 0.0 ? sc[kRho] / pq : 0.0;
      sc[kRhoOld] = sc[kRho];
    }
    const double rr = sc[kRr];
    const bool done = !(rr > target2);
    sc[kConverged] = done ? 1.0 : 0.0;
    if (done) sc[kConvIt] = it;
    if (host) {
      host[0] = rr;
      if (done) host[1] = it;
    }
  }
}
// The end of a CG iteration in one launch: sc[kRr] = the sum of the r.r
// partials (as reduce_k), then the bookkeeping of cg_scalars_k (op 0)
__global__ void __launch_bounds__(kThreads)
cg_rr_scalars_k(const double *partials, int n_partials, double *sc, double *host, double target2,
                int it) {
  if (sc[kConverged] != 0.0) return;
  double s = 0;
  for (int i = threadIdx.x; i < n_partials; i += kThreads) s += partials[i];
  s = block_sum(s);
  if (threadIdx.x == 0) {
    sc[kRr] = s;
    const double pq = sc[kPq];
    sc[kAlpha] = pq != 0.0 ? sc[kRho] / pq : 0.0;
    sc[kRhoOld] = sc[kRho];
    const bool done = !(s > target2);
    sc[kConverged] = done ? 1.0 : 0.0;
    if (done) sc[kConvIt] = it;
    if (host) {
      host[0] = s;
      if (done) host[1] = it;
    }
  }
}
template <typename TD, typename TS>
__global__ void convert_unless_k(int64_t n, TD *d, const TS *s, const double *skip) {
  if (skip && *skip != 0.0) return;
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x)
    d[i] = static_cast<TD>(__builtin_nontemporal_load(s + i));
}
template <typename T>
__global__ void fill_mod11_k(int64_t n, T *x, double mean) {
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x)
    x[i] = static_cast<T>(double(i % 11) - mean);
}
}  // namespace

template <typename T> void launch_fill_mod11(int64_t n, T *x, double mean, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(fill_mod11_k<T>, dim3(grid_for(n)), dim3(kThreads), 0, s, n, x, mean);
  PMG_HIP_CHECK(hipGetLastError());
}
template <typename T> void launch_fill(int64_t n, T *x, T v, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(fill_k<T>, dim3(grid_for(n)), dim3(kThreads), 0, s, n, x, v);
  PMG_HIP_CHECK(hipGetLastError());
}
template <typename T> void launch_copy(int64_t n, T *d, const T *src, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(copy_k<T>, dim3(grid_for(n)), dim3(kThreads), 0, s, n, d, src);
  PMG_HIP_CHECK(hipGetLastError());
}
template <typename T>
void launch_copy_indexed(int64_t n, const uint32_t *idx, T *d, const T *src, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(copy_idx_k<T>, dim3(grid_for(n)), dim3(kThreads), 0, s, n, idx, d, src);
  PMG_HIP_CHECK(hipGetLastError());
}
template <typename T>
void launch_prep_constrained(int64_t n, const uint8_t *mask, T *d, const T *src, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(prep_constrained_k<T>, dim3(grid_for(n)), dim3(kThreads), 0, s, n, mask, d,
                     src);
  PMG_HIP_CHECK(hipGetLastError());
}
template <typename T>
void launch_set_indexed(int64_t n, const uint32_t *idx, T *d, T v, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(set_idx_k<T>, dim3(grid_for(n)), dim3(kThreads), 0, s, n, idx, d, v);
  PMG_HIP_CHECK(hipGetLastError());
}
template <typename T> void launch_invert(int64_t n, T *x, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(invert_k<T>, dim3(grid_for(n)), dim3(kThreads), 0, s, n, x);
  PMG_HIP_CHECK(hipGetLastError());
}
template <typename T>
int64_t first_nonpositive(int64_t n, const T *d, hipStream_t s) {
  if (n <= 0) return -1;
  DeviceArray<unsigned long long> first(1);
  const unsigned long long init = static_cast<unsigned long long>(n);
  PMG_HIP_CHECK(hipMemcpyAsync(first.ptr, &init, sizeof(init), hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(first_nonpositive_k<T>, dim3(grid_for(n)), dim3(kThreads), 0, s, n, d, first.ptr);
  PMG_HIP_CHECK(hipGetLastError());
  unsigned long long r = 0;
  PMG_HIP_CHECK(hipMemcpyAsync(&r, first.ptr, sizeof(r), hipMemcpyDeviceToHost, s));
  PMG_HIP_CHECK(hipStreamSynchronize(s));
  return r < static_cast<unsigned long long>(n) ? static_cast<int64_t>(r) : -1;
}
template int64_t first_nonpositive<double>(int64_t, const double *, hipStream_t);
template int64_t first_nonpositive<float>(int64_t, const float *, hipStream_t);

template <typename TD, typename TS>
void launch_convert(int64_t n, TD *d, const TS *src, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL((convert_k<TD, TS>), dim3(grid_for(n)), dim3(kThreads), 0, s, n, d, src);
  PMG_HIP_CHECK(hipGetLastError());
}
template <typename TD, typename TS>
void launch_convert_unless(int64_t n, TD *d, const TS *src, const double *skip, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL((convert_unless_k<TD, TS>), dim3(grid_for(n)), dim3(kThreads), 0, s, n, d,
                     src, skip);
  PMG_HIP_CHECK(hipGetLastError());
}
template <typename T> void launch_residual(int64_t n, T *t, const T *b, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(residual_k<T>, dim3(grid_for(n)), dim3(kThreads), 0, s, n, t, b);
  PMG_HIP_CHECK(hipGetLastError());
}
template <typename TA, typename TB>
void launch_dot(int64_t n, const TA *a, const TB *b, double *partials, double *out, hipStream_t s,
                bool stream_a, const double *skip) {
  hipLaunchKernelGGL((dot_k<TA, TB>), dim3(kDotBlocks), dim3(kThreads), 0, s, n, a, b, partials,
                     stream_a);
  PMG_HIP_CHECK(hipGetLastError());
  hipLaunchKernelGGL(reduce_k, dim3(1), dim3(kThreads), 0, s, partials, kDotBlocks, out, 0, 1,
                     skip);
  PMG_HIP_CHECK(hipGetLastError());
}
void launch_reduce_partials(const double *partials, int n_partials, double *sc, int slot,
                            int n_sums, hipStream_t s, const double *skip) {
  hipLaunchKernelGGL(reduce_k, dim3(1), dim3(kThreads), 0, s, partials, n_partials, sc, slot,
                     n_sums, skip);
  PMG_HIP_CHECK(hipGetLastError());
}
template <typename T>
void launch_cheby_update(int mode, int64_t n, T *x, T *old, const T *b, const T *t,
                         const T *pinv, double f1, double f2, bool old_is_zero, hipStream_t s,
                         T *prep, const uint8_t *mask) {
  if (n <= 0) return;
  hipLaunchKernelGGL(cheby_k<T>, dim3(grid_for(n)), dim3(kThreads), 0, s, mode, n, x, old, b, t,
                     pinv, static_cast<T>(f1), static_cast<T>(f2), old_is_zero, prep, mask);
  PMG_HIP_CHECK(hipGetLastError());
}
template <typename T>
void launch_cg_update_xr(int64_t n, T *x, T *r, const T *p, const T *q, const double *sc,
                         double *partials, hipStream_t s) {
  hipLaunchKernelGGL(cg_xr_k<T>, dim3(kDotBlocks), dim3(kThreads), 0, s, n, x, r, p, q, sc,
                     partials);
  PMG_HIP_CHECK(hipGetLastError());
}
template <typename T>
void launch_cg_precond_dot(int64_t n, const T *r, const T *pinv, T *z, double *partials,
                           hipStream_t s, const double *skip) {
  hipLaunchKernelGGL(cg_pz_k<T>, dim3(kDotBlocks), dim3(kThreads), 0, s, n, r, pinv, z, partials,
                     skip);
  PMG_HIP_CHECK(hipGetLastError());
}
template <typename T, typename TZ>
void launch_cg_update_p(int64_t n, T *p, const TZ *z, const double *sc, int first, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL((cg_p_k<T, TZ>), dim3(grid_for(n)), dim3(kThreads), 0, s, n, p, z, sc, first);
  PMG_HIP_CHECK(hipGetLastError());
}
template <typename T>
void launch_cg_update_r(int64_t n, T *r, const T *q, const double *sc, double *partials,
                        hipStream_t s, void *stage, int stage_precision, const double *skip) {
  if (stage && stage_precision == PMG_FP32)
    hipLaunchKernelGGL((cg_r_k<T, float>), dim3(kDotBlocks), dim3(kThreads), 0, s, n, r, q, sc,
                       partials, static_cast<float *>(stage), skip);
  else
    hipLaunchKernelGGL((cg_r_k<T, double>), dim3(kDotBlocks), dim3(kThreads), 0, s, n, r, q, sc,
                       partials, static_cast<double *>(stage), skip);
  PMG_HIP_CHECK(hipGetLastError());
}
template <typename T, typename TZ>
void launch_cg_update_xp(int64_t n, T *x, T *p, const TZ *z, const double *sc, int first,
                         hipStream_t s, const double *skip) {
  if (n <= 0) return;
  hipLaunchKernelGGL((cg_xp_k<T, TZ>), dim3(grid_for(n)), dim3(kThreads), 0, s, n, x, p, z, sc,
                     first, skip);
  PMG_HIP_CHECK(hipGetLastError());
}
template <typename T>
void launch_cg_update_x(int64_t n, T *x, const T *p, const double *sc, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(cg_x_k<T>, dim3(grid_for(n)), dim3(kThreads), 0, s, n, x, p, sc);
  PMG_HIP_CHECK(hipGetLastError());
}
void launch_cg_rr_scalars(const double *partials, int n_partials, double *sc, hipStream_t s,
                          double *host, double target2, int it) {
  hipLaunchKernelGGL(cg_rr_scalars_k, dim3(1), dim3(kThreads), 0, s, partials, n_partials, sc, host,
                     target2, it);
  PMG_HIP_CHECK(hipGetLastError());
}
void launch_cg_scalars(double *sc, int op, hipStream_t s, double *host, double target2, int it) {
  hipLaunchKernelGGL(cg_scalars_k, dim3(1), dim3(64), 0, s, sc, op, host, target2, it);
  PMG_HIP_CHECK(hipGetLastError());
}

#define PMG_INST(T)                                                                           \
  template void launch_fill<T>(int64_t, T *, T, hipStream_t);                                 \
  template void launch_fill_mod11<T>(int64_t, T *, double, hipStream_t);                      \
  template void launch_copy<T>(int64_t, T *, const T *, hipStream_t);                         \
  template void launch_copy_indexed<T>(int64_t, const uint32_t *, T *, const T *, hipStream_t); \
  template void launch_prep_constrained<T>(int64_t, const uint8_t *, T *, const T *, hipStream_t); \
  template void launch_set_indexed<T>(int64_t, const uint32_t *, T *, T, hipStream_t);        \
  template void launch_invert<T>(int64_t, T *, hipStream_t);                                  \
  template void launch_residual<T>(int64_t, T *, const T *, hipStream_t);                     \
  template void launch_dot<T>(int64_t, const T *, const T *, double *, double *, hipStream_t, bool, const double *); \
  template void launch_cheby_update<T>(int, int64_t, T *, T *, const T *, const T *,          \
                                       const T *, double, double, bool, hipStream_t, T *,     \
                                       const uint8_t *);                                      \
  template void launch_cg_update_xr<T>(int64_t, T *, T *, const T *, const T *,               \
                                       const double *, double *, hipStream_t);                \
  template void launch_cg_precond_dot<T>(int64_t, const T *, const T *, T *, double *,        \
                                         hipStream_t, const double *);                        \
  template void launch_cg_update_p<T>(int64_t, T *, const T *, const double *, int, hipStream_t);
PMG_INST(double)
PMG_INST(float)
// CG vectors of one precision with a preconditioned vector z of the other
template void launch_dot<double, float>(int64_t, const double *, const float *, double *, double *,
                                        hipStream_t, bool, const double *);
template void launch_dot<float, double>(int64_t, const float *, const double *, double *, double *,
                                        hipStream_t, bool, const double *);
#define PMG_CGX(T, TZ)                                                                             \
  template void launch_cg_update_xp<T, TZ>(int64_t, T *, T *, const TZ *, const double *, int,    \
                                           hipStream_t, const double *);
PMG_CGX(double, double)
PMG_CGX(double, float)
PMG_CGX(float, float)
PMG_CGX(float, double)
#undef PMG_CGX
template void launch_cg_update_r<double>(int64_t, double *, const double *, const double *,
                                         double *, hipStream_t, void *, int, const double *);
template void launch_cg_update_r<float>(int64_t, float *, const float *, const double *, double *,
                                        hipStream_t, void *, int, const double *);
template void launch_cg_update_x<double>(int64_t, double *, const double *, const double *,
                                         hipStream_t);
template void launch_cg_update_x<float>(int64_t, float *, const float *, const double *,
                                        hipStream_t);

#undef PMG_INST
template void launch_convert<double, float>(int64_t, double *, const float *, hipStream_t);
template void launch_convert<float, double>(int64_t, float *, const double *, hipStream_t);
template void launch_convert<double, double>(int64_t, double *, const double *, hipStream_t);
template void launch_convert<float, float>(int64_t, float *, const float *, hipStream_t);
#define PMG_CU(TD, TS) \
  template void launch_convert_unless<TD, TS>(int64_t, TD *, const TS *, const double *, hipStream_t);
PMG_CU(double, float)
PMG_CU(float, double)
PMG_CU(double, double)
PMG_CU(float, float)
#undef PMG_CU

// ---------------------------------------------------------------- FMA rate probe
// Measurement only: each lane runs kFmaChains independent dependency chains
// of kFmaIters fused multiply-adds (2 flops each), so the issue rate, not the
// latency, bounds it; the sum is stored so nothing is eliminated.
constexpr int kFmaChains = 8, kFmaIters = 2048;
template <typename T>
__global__ void __launch_bounds__(256) fma_probe_k(T seed, T *out) {
  T a[kFmaChains];
#pragma unroll
  for (int c = 0; c < kFmaChains; ++c) a[c] = seed + T(c) + T(threadIdx.x) * T(1e-6);
  const T m = T(0.999999), add = T(1e-7);
  for (int i = 0; i < kFmaIters; ++i) {
#pragma unroll
    for (int c = 0; c < kFmaChains; ++c) {
      if constexpr (std::is_same<T, double>::value) a[c] = __builtin_fma(a[c], m, add);
      else a[c] = __builtin_fmaf(a[c], m, add);
    }
  }
  T sum = 0;
#pragma unroll
  for (int c = 0; c < kFmaChains; ++c) sum += a[c];
  out[blockIdx.x * blockDim.x + threadIdx.x] = sum;
}

double measure_fma_rate(int precision, hipStream_t s) {
  int dev = 0, n_cu = 0;
  PMG_HIP_CHECK(hipGetDevice(&dev));
  PMG_HIP_CHECK(hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, dev));
  const unsigned blocks = static_cast<unsigned>(std::max(n_cu, 1)) * 16;
  const size_t lanes = size_t(blocks) * 256;
  DeviceArray<double> out(lanes);
  hipEvent_t e0, e1;
  PMG_HIP_CHECK(hipEventCreate(&e0));
  PMG_HIP_CHECK(hipEventCreate(&e1));
  float best = 1e30f;
  for (int rep = 0; rep < 4; ++rep) {
    PMG_HIP_CHECK(hipEventRecord(e0, s));
    if (precision == PMG_FP64)
      hipLaunchKernelGGL(fma_probe_k<double>, dim3(blocks), dim3(256), 0, s, 1.0, out.ptr);
    else
      hipLaunchKernelGGL(fma_probe_k<float>, dim3(blocks), dim3(256), 0, s, 1.0f,
                         reinterpret_cast<float *>(out.ptr));
    PMG_HIP_CHECK(hipGetLastError());
    PMG_HIP_CHECK(hipEventRecord(e1, s));
    PMG_HIP_CHECK(hipEventSynchronize(e1));
    float ms = 0.f;
    PMG_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
    if (rep > 0) best = std::min(best, ms);  // the first launch loads the code object
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return 2.0 * kFmaChains * kFmaIters * double(lanes) / (1e-3 * best);
}

}  // namespace pmg
