// pmg_transfer.hip -- tensor-product h/p transfer kernels for gfx950.
//
// Prolongation (reference BK1::Parallel::KokkosProlongationBatchedKernel,
// bk1_kokkos_kernels.h:462-759): per coarse cell gather the coarse values
// (invalid -> 0), apply M x M x M (nmc -> nmf per direction), multiply by the
// fine weights and scatter-add into the fine vector.  Restriction
// (KokkosRestrictionBatchedKernel, :761-1051) is the exact transpose with the
// weight applied on the gather.
// Layout: one nmf x nmf lane tile per cell (256/nmf^2 cells per workgroup);
// the z sweep runs on a register column, x and y sweeps on register lines
// read from LDS; M is a kernel argument held in scalar registers.
#include "pmg_common.hpp"
#include "pmg_kernels.hpp"

namespace pmg {

template <int NF>
struct XferTile {
  static constexpr int NF2 = NF * NF;
  static constexpr int cells = (256 / NF2) > 0 ? 256 / NF2 : 1;
  static constexpr int lanes = ((cells * NF2 + 63) / 64) * 64;
};

template <int NC, int NF, typename T>
__global__ void __launch_bounds__(256)
prolongate_kernel(const int64_t n_cells, const Transfer1D<T> M, const uint32_t *__restrict__ map_c,
                  const uint32_t *__restrict__ map_f, const T *__restrict__ w,
                  const T *__restrict__ xc, T *__restrict__ yf)
{
  constexpr int NF2 = NF * NF, NF3 = NF2 * NF, NC2 = NC * NC, NC3 = NC2 * NC;
  constexpr int E = XferTile<NF>::cells;
  __shared__ T lds[E * 2 * NF3];
  const unsigned lb = xcd_logical_block(blockIdx.x, gridDim.x);
  const int tid = threadIdx.x, e = tid / NF2, t = tid - e * NF2, ta = t % NF, tb = t / NF;
  const int64_t cell = static_cast<int64_t>(lb) * E + e;
  const bool active = e < E && cell < n_cells;
  T *L0 = lds + (active ? e : 0) * 2 * NF3, *L1 = L0 + NF3;

  // z: column (i,j) of coarse values -> L0[c][j][i]
  if (active && ta < NC && tb < NC) {
    T u[NC];
#pragma unroll
    for (int k = 0; k < NC; ++k) {
      const uint32_t d = map_c[cell * NC3 + k * NC2 + tb * NC + ta];
      u[k] = d == kInvalid ? T(0) : xc[d];
    }
#pragma unroll
    for (int c = 0; c < NF; ++c) {
      T v = 0;
#pragma unroll
      for (int k = 0; k < NC; ++k) v += M.M[k * NF + c] * u[k];
      L0[c * NC2 + tb * NC + ta] = v;
    }
  }
  __syncthreads();
  // y: lane (i,c) -> L1[c][b][i]
  if (active && ta < NC) {
    T a[NC];
#pragma unroll
    for (int j = 0; j < NC; ++j) a[j] = L0[tb * NC2 + j * NC + ta];
#pragma unroll
    for (int b = 0; b < NF; ++b) {
      T v = 0;
#pragma unroll
      for (int j = 0; j < NC; ++j) v += M.M[j * NF + b] * a[j];
      L1[(tb * NF + b) * NC + ta] = v;
    }
  }
  __syncthreads();
  // x: lane (b,c) -> fine values along a, weighted scatter-add
  if (active) {
    T l[NC];
#pragma unroll
    for (int i = 0; i < NC; ++i) l[i] = L1[(tb * NF + ta) * NC + i];
    const int64_t base = cell * NF3 + tb * NF2 + ta * NF;
#pragma unroll
    for (int a = 0; a < NF; ++a) {
      T v = 0;
#pragma unroll
      for (int i = 0; i < NC; ++i) v += M.M[i * NF + a] * l[i];
      const T wt = w[base + a];
      if (wt != T(0)) atomic_add(yf + map_f[base + a], wt * v);
    }
  }
}

template <int NC, int NF, typename T>
__global__ void __launch_bounds__(256)
restrict_kernel(const int64_t n_cells, const Transfer1D<T> M, const uint32_t *__restrict__ map_c,
                const uint32_t *__restrict__ map_f, const T *__restrict__ w,
                const T *__restrict__ xf, T *__restrict__ yc)
{
  constexpr int NF2 = NF * NF, NF3 = NF2 * NF, NC2 = NC * NC, NC3 = NC2 * NC;
  constexpr int E = XferTile<NF>::cells;
  __shared__ T lds[E * 2 * NF3];
  const unsigned lb = xcd_logical_block(blockIdx.x, gridDim.x);
  const int tid = threadIdx.x, e = tid / NF2, t = tid - e * NF2, ta = t % NF, tb = t / NF;
  const int64_t cell = static_cast<int64_t>(lb) * E + e;
  const bool active = e < E && cell < n_cells;
  T *L0 = lds + (active ? e : 0) * 2 * NF3, *L1 = L0 + NF3;

  // x: lane (b,c) gathers w*x along a, contracts to i -> L1[c][b][i]
  if (active) {
    T f[NF];
    const int64_t base = cell * NF3 + tb * NF2 + ta * NF;
#pragma unroll
    for (int a = 0; a < NF; ++a) {
      const T wt = w[base + a];
      f[a] = wt == T(0) ? T(0) : wt * xf[map_f[base + a]];
    }
#pragma unroll
    for (int i = 0; i < NC; ++i) {
      T v = 0;
#pragma unroll
      for (int a = 0; a < NF; ++a) v += M.M[i * NF + a] * f[a];
      L1[(tb * NF + ta) * NC + i] = v;
    }
  }
  __syncthreads();
  // y: lane (i,c) -> L0[c][j][i]
  if (active && ta < NC) {
    T a[NF];
#pragma unroll
    for (int b = 0; b < NF; ++b) a[b] = L1[(tb * NF + b) * NC + ta];
#pragma unroll
    for (int j = 0; j < NC; ++j) {
      T v = 0;
#pragma unroll
      for (int b = 0; b < NF; ++b) v += M.M[j * NF + b] * a[b];
      L0[tb * NC2 + j * NC + ta] = v;
    }
  }
  __syncthreads();
  // z: lane (i,j) -> coarse column, scatter-add (skip constrained)
  if (active && ta < NC && tb < NC) {
    T col[NF];
#pragma unroll
    for (int c = 0; c < NF; ++c) col[c] = L0[c * NC2 + tb * NC + ta];
#pragma unroll
    for (int k = 0; k < NC; ++k) {
      T v = 0;
#pragma unroll
      for (int c = 0; c < NF; ++c) v += M.M[k * NF + c] * col[c];
      const uint32_t d = map_c[cell * NC3 + k * NC2 + tb * NC + ta];
      if (d != kInvalid) atomic_add(yc + d, v);
    }
  }
}

template <int NC, int NF, typename T>
static void xfer_launch(bool prolong, int64_t n_cells, const Transfer1D<T> &M,
                        const uint32_t *map_c, const uint32_t *map_f, const T *w, const T *x,
                        T *y, hipStream_t s) {
  if (n_cells <= 0) return;
  using Tile = XferTile<NF>;
  const int64_t nb = (n_cells + Tile::cells - 1) / Tile::cells;
  PMG_REQUIRE(nb < (int64_t(1) << 31), PMG_ERR_SIZE, "too many cells for one launch");
  const unsigned grid = pad_to_xcd(static_cast<unsigned>(nb));
  if (prolong)
    hipLaunchKernelGGL((prolongate_kernel<NC, NF, T>), dim3(grid), dim3(Tile::lanes), 0, s,
                       n_cells, M, map_c, map_f, w, x, y);
  else
    hipLaunchKernelGGL((restrict_kernel<NC, NF, T>), dim3(grid), dim3(Tile::lanes), 0, s,
                       n_cells, M, map_c, map_f, w, x, y);
  PMG_HIP_CHECK(hipGetLastError());
}

template <typename T>
static void xfer_dispatch(bool prolong, int nmc, int nmf, int64_t n_cells, const Transfer1D<T> &M,
                          const uint32_t *map_c, const uint32_t *map_f, const T *w, const T *x,
                          T *y, hipStream_t s) {
#define PMG_XF(C, F)                                                            \
  if (nmc == C && nmf == F) {                                                   \
    xfer_launch<C, F, T>(prolong, n_cells, M, map_c, map_f, w, x, y, s);        \
    return;                                                                     \
  }
  // h-transfers (p+1 -> 2p+1)
  PMG_XF(2, 3) PMG_XF(3, 5) PMG_XF(4, 7) PMG_XF(5, 9) PMG_XF(6, 11) PMG_XF(7, 13)
  // p-transfers (pc+1 -> pf+1), 1 <= pc < pf <= 6
  PMG_XF(2, 4) PMG_XF(2, 5) PMG_XF(2, 6) PMG_XF(2, 7)
  PMG_XF(3, 4) PMG_XF(3, 6) PMG_XF(3, 7)
  PMG_XF(4, 5) PMG_XF(4, 6)
  PMG_XF(5, 6) PMG_XF(5, 7)
  PMG_XF(6, 7)
#undef PMG_XF
  throw Error(PMG_ERR_UNSUPPORTED, "transfer kernel: unsupported (nm_coarse, nm_fine) = (" +
                                       std::to_string(nmc) + ", " + std::to_string(nmf) + ")");
}

template <typename T>
void launch_prolongate(int nmc, int nmf, int64_t n_cells, const Transfer1D<T> &M,
                       const uint32_t *map_c, const uint32_t *map_f, const T *w, const T *xc,
                       T *yf, hipStream_t s) {
  xfer_dispatch<T>(true, nmc, nmf, n_cells, M, map_c, map_f, w, xc, yf, s);
}
template <typename T>
void launch_restrict(int nmc, int nmf, int64_t n_cells, const Transfer1D<T> &M,
                     const uint32_t *map_c, const uint32_t *map_f, const T *w, const T *xf,
                     T *yc, hipStream_t s) {
  xfer_dispatch<T>(false, nmc, nmf, n_cells, M, map_c, map_f, w, xf, yc, s);
}

template void launch_prolongate<double>(int, int, int64_t, const Transfer1D<double> &,
                                        const uint32_t *, const uint32_t *, const double *,
                                        const double *, double *, hipStream_t);
template void launch_prolongate<float>(int, int, int64_t, const Transfer1D<float> &,
                                       const uint32_t *, const uint32_t *, const float *,
                                       const float *, float *, hipStream_t);
template void launch_restrict<double>(int, int, int64_t, const Transfer1D<double> &,
                                      const uint32_t *, const uint32_t *, const double *,
                                      const double *, double *, hipStream_t);
template void launch_restrict<float>(int, int, int64_t, const Transfer1D<float> &,
                                     const uint32_t *, const uint32_t *, const float *,
                                     const float *, float *, hipStream_t);

}  // namespace pmg
