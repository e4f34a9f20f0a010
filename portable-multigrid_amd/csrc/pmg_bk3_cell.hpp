// pmg_bk3_cell.hpp -- device code shared by the continuous (BK3) and the DG
// operator kernels: the sum-factorised cell stiffness apply of one cell on an
// NQ x NQ lane tile (reference: BK3::Parallel::KokkosKernel, bk3_kokkos_kernel.h:195-669).
#pragma once

#include "pmg_common.hpp"
#include "pmg_kernels.hpp"

namespace pmg {

template <int NQ>
struct Bk3Tile {
  static constexpr int NQ2 = NQ * NQ;
  static constexpr int cells = g_batch_cells(NQ);
  static constexpr int lanes = ((cells * NQ2 + 63) / 64) * 64;
  static_assert(cells * NQ2 <= 256, "a batch must fit one 256-lane workgroup");
};

// LDS elements per cell for a per-cell need of `base` elements of type T:
// padded when base is a multiple of the 64 x 4-byte bank row, so that the
// cells sharing a wave do not start on the same bank (they would otherwise
// conflict on every access of the same local index).
template <typename T>
__host__ __device__ constexpr int lds_cell_stride(int base) {
  return base % (256 / int(sizeof(T))) == 0 ? base + 64 / int(sizeof(T)) : base;
}

template <typename T>
__device__ inline T ldg_stream(const T *p) {
  return __builtin_nontemporal_load(p);
}

// G load policy of the BK3 kernel
enum : int {
  kGEarlyNT = 0,  // non-temporal loads issued at kernel entry
  kGEarly = 1,    // ordinary loads issued at kernel entry
  kGLateNT = 2    // non-temporal loads issued where G is used
};

// Geometric factors of the lane's q-point column (ta,tb) over all planes r:
// batch-interleaved G, component c, plane r at Gc[(c * NQ + r) * LANES], so
// for a batch of consecutive cells a wave reads 64 consecutive values per load.
template <int NQ, typename T, int GMODE>
__device__ __forceinline__ void bk3_load_g(const T *Gc, T (&g)[6][NQ]) {
  constexpr int LANES = Bk3Tile<NQ>::lanes;
#pragma unroll
  for (int c = 0; c < 6; ++c)
#pragma unroll
    for (int r = 0; r < NQ; ++r)
      g[c][r] = GMODE == kGEarly ? Gc[(c * NQ + r) * LANES] : ldg_stream(Gc + (c * NQ + r) * LANES);
}

// Cell stiffness applied to one cell held by an NQ x NQ lane tile (lane
// (ta,tb)), in two halves that all lanes of the workgroup must call (they
// hold barriers).  P0..P2: three NQ^3 LDS buffers of this cell.
//
// bk3_forward: u[k] is the source column (i,j) = (ta,tb) over k (read only
// where ta,tb < NM).  Leaves the values at the Gauss points in
// P0[r][q][p] (p along x) and d/dxi_0 of them in P2.  zbuf[0..zn) (LDS,
// optional) is cleared by the whole workgroup after every lane has consumed u.
template <int NM, int NQ, typename T>
__device__ __forceinline__ void
bk3_forward(const T (&u)[NM], const Shape1D<T> &sh, T *P0, T *P1, T *P2, const bool active,
            const int ta, const int tb, T *zbuf = nullptr, const int zn = 0)
{
  constexpr int NQ2 = NQ * NQ, NM2 = NM * NM;
  // ---- phase 1: interpolate the column in z
  if (active && ta < NM && tb < NM) {
#pragma unroll
    for (int r = 0; r < NQ; ++r) {
      T v = 0;
#pragma unroll
      for (int k = 0; k < NM; ++k) v += sh.S[k * NQ + r] * u[k];
      P0[r * NM2 + tb * NM + ta] = v;  // A[r][j][i]
    }
  }
  __syncthreads();

  if (zbuf)
    for (int i = threadIdx.x; i < zn; i += blockDim.x) zbuf[i] = T(0);

  // ---- phase 2: y-interpolation, lane (i,r) = (ta,tb): B[r][q][i]
  if (active && ta < NM) {
    T a[NM];
#pragma unroll
    for (int j = 0; j < NM; ++j) a[j] = P0[tb * NM2 + j * NM + ta];
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      T v = 0;
#pragma unroll
      for (int j = 0; j < NM; ++j) v += sh.S[j * NQ + q] * a[j];
      P1[(tb * NQ + q) * NM + ta] = v;
    }
  }
  __syncthreads();

  // ---- phase 3: x-interpolation + x-gradient, lane (q,r) = (ta,tb)
  if (active) {
    T b[NM], U[NQ];
#pragma unroll
    for (int i = 0; i < NM; ++i) b[i] = P1[(tb * NQ + ta) * NM + i];
#pragma unroll
    for (int p = 0; p < NQ; ++p) {
      T v = 0;
#pragma unroll
      for (int i = 0; i < NM; ++i) v += sh.S[i * NQ + p] * b[i];
      U[p] = v;
      P0[tb * NQ2 + ta * NQ + p] = v;  // U[r][q][p]
    }
#pragma unroll
    for (int p = 0; p < NQ; ++p) {
      T v = 0;
#pragma unroll
      for (int n = 0; n < NQ; ++n) v += sh.D[n * NQ + p] * U[n];
      P2[tb * NQ2 + ta * NQ + p] = v;  // du/dxi_0
    }
  }
  __syncthreads();
}

// bk3_backward: from the state bk3_forward leaves, the remaining gradient,
// the flux G grad u (g holds the column's G, loaded here when GMODE is late)
// and the test with grad phi; with HASV also sum_q phi(q) V[q] for a value
// coefficient V[r][q][p] at the Gauss points (LDS).  out[k] receives the
// result for column (i,j) = (ta,tb) over k.
template <int NM, int NQ, typename T, int GMODE, bool HASV>
__device__ __forceinline__ void
bk3_backward(T (&out)[NM], T (&g)[6][NQ], const T *Gc, const Shape1D<T> &sh, T *P0, T *P1,
             T *P2, const T *V, const bool active, const int ta, const int tb)
{
  constexpr int NQ2 = NQ * NQ, NM2 = NM * NM;
  // ---- phase 4: y-gradient, lane (p,r) = (ta,tb)
  if (active) {
    T c[NQ];
#pragma unroll
    for (int n = 0; n < NQ; ++n) c[n] = P0[tb * NQ2 + n * NQ + ta];
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      T v = 0;
#pragma unroll
      for (int n = 0; n < NQ; ++n) v += sh.D[n * NQ + q] * c[n];
      P1[tb * NQ2 + q * NQ + ta] = v;  // du/dxi_1
    }
  }
  __syncthreads();

  // ---- phase 5: z-gradient, flux G grad u, z-integration; lane (p,q) = (ta,tb)
  if (active) {
    if (GMODE == kGLateNT) bk3_load_g<NQ, T, GMODE>(Gc, g);
    T col[NQ], fz[NQ];
#pragma unroll
    for (int n = 0; n < NQ; ++n) col[n] = P0[n * NQ2 + tb * NQ + ta];
#pragma unroll
    for (int r = 0; r < NQ; ++r) {
      T uz = 0;
#pragma unroll
      for (int n = 0; n < NQ; ++n) uz += sh.D[n * NQ + r] * col[n];
      const int id = r * NQ2 + tb * NQ + ta;
      const T ux = P2[id], uy = P1[id];
      P2[id] = g[0][r] * ux + g[1][r] * uy + g[2][r] * uz;
      P1[id] = g[1][r] * ux + g[3][r] * uy + g[4][r] * uz;
      fz[r] = g[2][r] * ux + g[4][r] * uy + g[5][r] * uz;
    }
#pragma unroll
    for (int r = 0; r < NQ; ++r) {
      T v = 0;
#pragma unroll
      for (int n = 0; n < NQ; ++n) v += sh.D[r * NQ + n] * fz[n];
      P0[r * NQ2 + tb * NQ + ta] = v;
    }
  }
  __syncthreads();

  // ---- phase 6: x-integration accumulated onto the z part, lane (q,r)
  if (active) {
    T f[NQ];
#pragma unroll
    for (int n = 0; n < NQ; ++n) f[n] = P2[tb * NQ2 + ta * NQ + n];
#pragma unroll
    for (int p = 0; p < NQ; ++p) {
      T v = P0[tb * NQ2 + ta * NQ + p];
#pragma unroll
      for (int n = 0; n < NQ; ++n) v += sh.D[p * NQ + n] * f[n];
      P0[tb * NQ2 + ta * NQ + p] = v;
    }
  }
  __syncthreads();

  // ---- phase 7: y-integration + y back-interpolation, lane (p,r)
  if (active) {
    T f[NQ], acc[NQ];
#pragma unroll
    for (int n = 0; n < NQ; ++n) f[n] = P1[tb * NQ2 + n * NQ + ta];
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      T v = P0[tb * NQ2 + q * NQ + ta];
      if (HASV) v += V[tb * NQ2 + q * NQ + ta];
#pragma unroll
      for (int n = 0; n < NQ; ++n) v += sh.D[q * NQ + n] * f[n];
      acc[q] = v;
    }
#pragma unroll
    for (int j = 0; j < NM; ++j) {
      T v = 0;
#pragma unroll
      for (int q = 0; q < NQ; ++q) v += sh.S[j * NQ + q] * acc[q];
      P1[tb * NQ2 + j * NQ + ta] = v;  // W[r][j][p]
    }
  }
  __syncthreads();

  // ---- phase 8: x back-interpolation, lane (j,r) = (ta,tb)
  if (active && ta < NM) {
    T l[NQ];
#pragma unroll
    for (int p = 0; p < NQ; ++p) l[p] = P1[tb * NQ2 + ta * NQ + p];
#pragma unroll
    for (int i = 0; i < NM; ++i) {
      T v = 0;
#pragma unroll
      for (int p = 0; p < NQ; ++p) v += sh.S[i * NQ + p] * l[p];
      P2[tb * NM2 + ta * NM + i] = v;  // Z[r][j][i]
    }
  }
  __syncthreads();

  // ---- phase 9: z back-interpolation, lane (i,j) = (ta,tb)
  if (active && ta < NM && tb < NM) {
    T z[NQ];
#pragma unroll
    for (int r = 0; r < NQ; ++r) z[r] = P2[r * NM2 + tb * NM + ta];
#pragma unroll
    for (int k = 0; k < NM; ++k) {
      T v = 0;
#pragma unroll
      for (int r = 0; r < NQ; ++r) v += sh.S[k * NQ + r] * z[r];
      out[k] = v;
    }
  }
}

// Both halves: out = A_e u (see bk3_forward / bk3_backward).
template <int NM, int NQ, typename T, int GMODE>
__device__ __forceinline__ void
bk3_cell_apply(const T (&u)[NM], T (&out)[NM], T (&g)[6][NQ], const T *Gc, const Shape1D<T> &sh,
               T *P0, T *P1, T *P2, const bool active, const int ta, const int tb,
               T *zbuf = nullptr, const int zn = 0)
{
  bk3_forward<NM, NQ, T>(u, sh, P0, P1, P2, active, ta, tb, zbuf, zn);
  bk3_backward<NM, NQ, T, GMODE, false>(out, g, Gc, sh, P0, P1, P2, nullptr, active, ta, tb);
}

}  // namespace pmg
