// pmg_bk3_cell.hpp -- device code shared by the continuous (BK3) and the DG
// operator kernels: the sum-factorised cell stiffness apply of one cell on an
// NQ x NQ lane tile (reference: BK3::Parallel::KokkosKernel, bk3_kokkos_kernel.h:195-669).
#pragma once

#include <type_traits>

#include "pmg_common.hpp"
#include "pmg_kernels.hpp"

namespace pmg {

template <int NQ>
struct Bk3Tile {
  static constexpr int NQ2 = NQ * NQ;
  static constexpr int cells = g_batch_cells(NQ);
  // wave-aligned layout: per_wave whole cells in each 64-lane wave (lanes
  // beyond per_wave nq^2 idle), the batch's cells over wave_lanes lanes
  static constexpr int per_wave = 64 / NQ2 > 0 ? 64 / NQ2 : 1;
  static constexpr int wave_lanes = ((cells + per_wave - 1) / per_wave) * 64;
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
  kGLateNT = 2,   // non-temporal loads issued where G is used
  kGAffine = 3    // affine cells: G_q = A w_p w_q w_r from the cell's 6 values
};

// The column's G of an affine cell: lane (ta,tb) = (p,q) takes
// g[c][r] = A_c w_p w_q w_r from the cell's A (Ac: 6 values, the same for
// the cell's lanes) and the 1D weights.
template <int NQ, typename T>
__device__ __forceinline__ void bk3_affine_g(const T *Ac, const Shape1D<T> &sh, int ta, int tb,
                                             T (&g)[6][NQ]) {
  T a[6];
#pragma unroll
  for (int c = 0; c < 6; ++c) a[c] = Ac[c];
  // w_ta and w_tb by selects over constant indices: a lane-dependent index
  // into the kernel argument made the compiler copy it to private memory
  T wa = T(0), wb = T(0);
#pragma unroll
  for (int i = 0; i < NQ; ++i) {
    wa = ta == i ? sh.W[i] : wa;
    wb = tb == i ? sh.W[i] : wb;
  }
  const T wab = wa * wb;
#pragma unroll
  for (int c = 0; c < 6; ++c) {
    const T ac = a[c] * wab;
#pragma unroll
    for (int r = 0; r < NQ; ++r) g[c][r] = ac * sh.W[r];
  }
}

// Geometric factors of the lane's q-point column (ta,tb) over all planes r,
// Gc = G + g_index<T>(nq, cell, 0, q of plane 0): batch-interleaved G,
// component c, plane r at Gc[(c * NQ + r) * LANES] (FP32) or the component
// pair c/2 at Gc[((c / 2) * NQ + r) * 2 LANES] (FP64), so for a batch of
// consecutive cells every load instruction of a wave reads 64 consecutive
// values (pairs).
template <int NQ, typename T, int GMODE>
__device__ __forceinline__ void bk3_load_g(const T *Gc, T (&g)[6][NQ]) {
  constexpr int LANES = Bk3Tile<NQ>::lanes;
  if constexpr (sizeof(T) == 8) {
    // component pairs (0,1), (2,3), (4,5): one 16-byte load each
    typedef T T2 __attribute__((ext_vector_type(2)));
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int r = 0; r < NQ; ++r) {
        const T2 *p = reinterpret_cast<const T2 *>(Gc + (c * NQ + r) * 2 * LANES);
        const T2 v = GMODE == kGEarly ? *p : __builtin_nontemporal_load(p);
        g[2 * c][r] = v.x;
        g[2 * c + 1][r] = v.y;
      }
  } else {
#pragma unroll
    for (int c = 0; c < 6; ++c)
#pragma unroll
      for (int r = 0; r < NQ; ++r)
        g[c][r] = GMODE == kGEarly ? Gc[(c * NQ + r) * LANES] : ldg_stream(Gc + (c * NQ + r) * LANES);
  }
}

// Cell stiffness applied to one cell held by an NQ x NQ lane tile (lane
// (ta,tb)), in two halves that all lanes of the workgroup must call (they
// hold barriers).  P0..P2: three NQ^3 LDS buffers of this cell.
//
// Barrier between the phases of the cell sweeps.  A cell's nq x nq lane tile
// lies within one 64-lane wave when 64 % nq^2 == 0 (nq = 2, 4, 8), and the
// phases exchange data only through the cell's own LDS region, so only lanes
// of that wave take part: LDS instructions of a wave are processed in order,
// and a wavefront-scope fence keeps the compiler from reordering them.  So
// are the tiles of the wave-aligned layout (WAVE = true, Bk3Tile::per_wave).
// Other tiles straddle waves and take a workgroup barrier.
template <int NQ, bool WAVE = (64 % (NQ * NQ) == 0)>
__device__ __forceinline__ void cell_sync() {
  if constexpr (WAVE) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  } else {
    __syncthreads();
  }
}

// ---- even-odd 1D contractions (tables: Shape1D, fill_even_odd)
// out[r] = sum_k S[k][r] x[k]   (values at the Gauss points)
template <int NM, int NQ, typename T>
__device__ __forceinline__ void eo_interp(const Shape1D<T> &sh, const T (&x)[NM], T (&out)[NQ]) {
  constexpr int HM = NM / 2, HQ = NQ / 2, HM1 = HM > 0 ? HM : 1;
  T e[HM1], o[HM1];
#pragma unroll
  for (int k = 0; k < HM; ++k) {
    e[k] = x[k] + x[NM - 1 - k];
    o[k] = x[k] - x[NM - 1 - k];
  }
#pragma unroll
  for (int r = 0; r < (NQ + 1) / 2; ++r) {
    T a = (NM % 2) ? sh.SM[r] * x[HM] : T(0);
#pragma unroll
    for (int k = 0; k < HM; ++k) a += sh.SE[k * kEO + r] * e[k];
    if (r < HQ) {
      T b = 0;
#pragma unroll
      for (int k = 0; k < HM; ++k) b += sh.SO[k * kEO + r] * o[k];
      out[r] = a + b;
      out[NQ - 1 - r] = a - b;
    } else {
      out[r] = a;  // middle point (odd nq)
    }
  }
}
// out[k] = sum_r S[k][r] x[r]   (test with the basis)
template <int NM, int NQ, typename T>
__device__ __forceinline__ void eo_interp_t(const Shape1D<T> &sh, const T (&x)[NQ], T (&out)[NM]) {
  constexpr int HM = NM / 2, HQ = NQ / 2, HQ1 = HQ > 0 ? HQ : 1;
  T e[HQ1], o[HQ1];
#pragma unroll
  for (int r = 0; r < HQ; ++r) {
    e[r] = x[r] + x[NQ - 1 - r];
    o[r] = x[r] - x[NQ - 1 - r];
  }
#pragma unroll
  for (int k = 0; k < HM; ++k) {
    T a = (NQ % 2) ? sh.SE[k * kEO + HQ] * x[HQ] : T(0), b = 0;
#pragma unroll
    for (int r = 0; r < HQ; ++r) {
      a += sh.SE[k * kEO + r] * e[r];
      b += sh.SO[k * kEO + r] * o[r];
    }
    out[k] = a + b;
    out[NM - 1 - k] = a - b;
  }
  if (NM % 2) {
    T a = (NQ % 2) ? sh.SM[HQ] * x[HQ] : T(0);
#pragma unroll
    for (int r = 0; r < HQ; ++r) a += sh.SM[r] * e[r];
    out[HM] = a;
  }
}
// out[p] = sum_n D[n][p] x[n]   (derivative at the Gauss points)
template <int NQ, typename T>
__device__ __forceinline__ void eo_deriv(const Shape1D<T> &sh, const T (&x)[NQ], T (&out)[NQ]) {
  constexpr int HQ = NQ / 2, HQ1 = HQ > 0 ? HQ : 1;
  T e[HQ1], o[HQ1];
#pragma unroll
  for (int n = 0; n < HQ; ++n) {
    e[n] = x[n] + x[NQ - 1 - n];
    o[n] = x[n] - x[NQ - 1 - n];
  }
#pragma unroll
  for (int p = 0; p < HQ; ++p) {
    T a = (NQ % 2) ? sh.FM[p] * x[HQ] : T(0), b = 0;
#pragma unroll
    for (int n = 0; n < HQ; ++n) {
      a += sh.FE[n * kEO + p] * e[n];
      b += sh.FO[n * kEO + p] * o[n];
    }
    out[p] = a + b;
    out[NQ - 1 - p] = b - a;
  }
  if (NQ % 2) {
    T b = 0;
#pragma unroll
    for (int n = 0; n < HQ; ++n) b += sh.FO[n * kEO + HQ] * o[n];
    out[HQ] = b;
  }
}
// out[r] = sum_n D[r][n] x[n]   (test with the derivative)
template <int NQ, typename T>
__device__ __forceinline__ void eo_deriv_t(const Shape1D<T> &sh, const T (&x)[NQ], T (&out)[NQ]) {
  constexpr int HQ = NQ / 2, HQ1 = HQ > 0 ? HQ : 1;
  T e[HQ1], o[HQ1];
#pragma unroll
  for (int n = 0; n < HQ; ++n) {
    e[n] = x[n] + x[NQ - 1 - n];
    o[n] = x[n] - x[NQ - 1 - n];
  }
#pragma unroll
  for (int r = 0; r < HQ; ++r) {
    T a = (NQ % 2) ? sh.TC[r] * x[HQ] : T(0), b = 0;
#pragma unroll
    for (int n = 0; n < HQ; ++n) {
      a += sh.TE[r * kEO + n] * e[n];
      b += sh.TO[r * kEO + n] * o[n];
    }
    out[r] = a + b;
    out[NQ - 1 - r] = b - a;
  }
  if (NQ % 2) {
    T b = 0;
#pragma unroll
    for (int n = 0; n < HQ; ++n) b += sh.TM[n] * o[n];
    out[HQ] = b;
  }
}

// ---- batched 1D contractions on the matrix cores (FP32)
// v_mfma_f32_16x16x4f32, one 16 x 16 block with K = 4: lane l supplies
// A[l%16][l/16] and B[l/16][l%16] and receives D[4(l/16)+v][l%16], v < 4.
typedef float mx_f4 __attribute__((ext_vector_type(4)));

// For every column c < ncol (one line of one of the batch's cells):
//   d(c, m, sum_{k<K} a(m, k) b(c, k))    for m < M <= 16,
// and with TWO also d2(c, m, sum_k a2(m, k) b(c, k)) from the same operands.
// Columns go in tiles of 16, tile t to wave t mod (waves); every lane of the
// workgroup calls it (the tile loop is uniform per wave, and the matrix
// instruction runs with all 64 lanes).  The matrices are zero-padded to
// 16 x 4 ceil(K / 4); b is read only for c < ncol, k < K.  The bk3 sweeps
// take their matrices from mx_a, six zero-padded 16 x 16 tables A[m * 16 + k]
// (fill_mx_tables), read by lane-dependent index: indexing the Shape1D
// kernel argument that way would copy it to private memory.
template <int M, int K, bool TWO, typename FA, typename FA2, typename FB, typename FD, typename FD2>
__device__ __forceinline__ void mx_sweep(const int ncol, FA a, FA2 a2, FB b, FD d, FD2 d2) {
  static_assert(M <= 16, "one block of 16 rows");
  constexpr int KS = (K + 3) / 4;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
  const int r16 = lane & 15, q4 = lane >> 4;
  float av[KS], av2[KS];
#pragma unroll
  for (int s = 0; s < KS; ++s) {
    const int k = 4 * s + q4;
    av[s] = r16 < M && k < K ? a(r16, k) : 0.f;
    if constexpr (TWO) av2[s] = r16 < M && k < K ? a2(r16, k) : 0.f;
  }
  const int ntile = (ncol + 15) >> 4;
  for (int t = wave; t < ntile; t += nw) {
    const int c = 16 * t + r16;
    const bool cv = c < ncol;
    float bv[KS];
#pragma unroll
    for (int s = 0; s < KS; ++s) {
      const int k = 4 * s + q4;
      bv[s] = cv && k < K ? b(c, k) : 0.f;
    }
    mx_f4 acc = {0.f, 0.f, 0.f, 0.f}, acc2 = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int s = 0; s < KS; ++s) {
      acc = __builtin_amdgcn_mfma_f32_16x16x4f32(av[s], bv[s], acc, 0, 0, 0);
      if constexpr (TWO) acc2 = __builtin_amdgcn_mfma_f32_16x16x4f32(av2[s], bv[s], acc2, 0, 0, 0);
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int m = 4 * q4 + v;
      if (cv && m < M) {
        d(c, m, acc[v]);
        if constexpr (TWO) d2(c, m, acc2[v]);
      }
    }
  }
}

// bk3_forward: u[k] is the source column (i,j) = (ta,tb) over k (read only
// where ta,tb < NM).  Leaves the values at the Gauss points in
// P0[r][q][p] (p along x) and d/dxi_0 of them in P2.  zbuf[0..zn) (LDS,
// optional) is cleared by the whole workgroup after every lane has consumed u.
// MX: the x and y sweeps (phases 2 and 3) run as batched contractions on the
// matrix cores over the mx_cells cells of the batch, whose LDS regions start
// at mx_lds with the stride lds_cell_stride(3 NQ^3); FP32 only, and every
// phase ends with a workgroup barrier.  The x-gradient is then formed from
// the nodal values directly, with the product matrix S D, instead of from the
// rounded values at the Gauss points.
template <int NM, int NQ, typename T, bool WAVE = (64 % (NQ * NQ) == 0), bool MX = false>
__device__ __forceinline__ void
bk3_forward(const T (&u)[NM], const Shape1D<T> &sh, T *P0, T *P1, T *P2, const bool active,
            const int ta, const int tb, T *zbuf = nullptr, const int zn = 0, T *mx_lds = nullptr,
            const int mx_cells = 0, const float *mx_a = nullptr)
{
  constexpr int NQ2 = NQ * NQ, NM2 = NM * NM;
  static_assert(!MX || (std::is_same<T, float>::value && !WAVE), "MX: FP32, workgroup barriers");
  [[maybe_unused]] constexpr int NQ3 = NQ2 * NQ, CS = lds_cell_stride<T>(3 * NQ3);
  // ---- phase 1: interpolate the column in z
  if (active && ta < NM && tb < NM) {
    T w[NQ];
    eo_interp<NM, NQ, T>(sh, u, w);
#pragma unroll
    for (int r = 0; r < NQ; ++r) P0[r * NM2 + tb * NM + ta] = w[r];  // A[r][j][i]
  }
  if (zbuf) {
    // zbuf is shared by the workgroup: every lane has consumed u before it is
    // cleared, and the clear is complete before any lane adds into it
    __syncthreads();
    for (int i = threadIdx.x; i < zn; i += blockDim.x) zbuf[i] = T(0);
    __syncthreads();
  } else {
    cell_sync<NQ, WAVE>();
  }

  if constexpr (MX) {
    // phase 2: column (r,i) of a cell, B[r][q][i] = sum_j S[j][q] A[r][j][i]
    constexpr int L = NQ * NM;
    mx_sweep<NQ, NM, false>(
        mx_cells * L, [&](int m, int k) { return mx_a[m * 16 + k]; }, [](int, int) { return 0.f; },
        [&](int c, int k) {
          const int e = c / L, l = c - e * L, r = l / NM, i = l - r * NM;
          return mx_lds[e * CS + r * NM2 + k * NM + i];
        },
        [&](int c, int m, float v) {
          const int e = c / L, l = c - e * L, r = l / NM, i = l - r * NM;
          mx_lds[e * CS + NQ3 + (r * NQ + m) * NM + i] = v;
        },
        [](int, int, float) {});
    __syncthreads();
    // phase 3: column (r,q), U[r][q][p] = sum_i S[i][p] B[r][q][i] and
    // dU/dxi_0 = sum_i (S D)[i][p] B[r][q][i]
    mx_sweep<NQ, NM, true>(
        mx_cells * NQ2, [&](int m, int k) { return mx_a[m * 16 + k]; },
        [&](int m, int k) { return mx_a[256 + m * 16 + k]; },
        [&](int c, int k) {
          const int e = c / NQ2, l = c - e * NQ2;
          return mx_lds[e * CS + NQ3 + l * NM + k];
        },
        [&](int c, int m, float v) {
          const int e = c / NQ2, l = c - e * NQ2;
          mx_lds[e * CS + l * NQ + m] = v;
        },
        [&](int c, int m, float v) {
          const int e = c / NQ2, l = c - e * NQ2;
          mx_lds[e * CS + 2 * NQ3 + l * NQ + m] = v;
        });
    __syncthreads();
    return;
  }

  // ---- phase 2: y-interpolation, lane (i,r) = (ta,tb): B[r][q][i]
  if (active && ta < NM) {
    T a[NM];
#pragma unroll
    for (int j = 0; j < NM; ++j) a[j] = P0[tb * NM2 + j * NM + ta];
    T w[NQ];
    eo_interp<NM, NQ, T>(sh, a, w);
#pragma unroll
    for (int q = 0; q < NQ; ++q) P1[(tb * NQ + q) * NM + ta] = w[q];
  }
  cell_sync<NQ, WAVE>();

  // ---- phase 3: x-interpolation + x-gradient, lane (q,r) = (ta,tb)
  if (active) {
    T b[NM], U[NQ], dU[NQ];
#pragma unroll
    for (int i = 0; i < NM; ++i) b[i] = P1[(tb * NQ + ta) * NM + i];
    eo_interp<NM, NQ, T>(sh, b, U);
    eo_deriv<NQ, T>(sh, U, dU);
#pragma unroll
    for (int p = 0; p < NQ; ++p) {
      P0[tb * NQ2 + ta * NQ + p] = U[p];   // U[r][q][p]
      P2[tb * NQ2 + ta * NQ + p] = dU[p];  // du/dxi_0
    }
  }
  cell_sync<NQ, WAVE>();
}

// bk3_backward: from the state bk3_forward leaves, the remaining gradient,
// the flux G grad u (g holds the column's G, loaded here when GMODE is late)
// and the test with grad phi; with HASV also sum_q phi(q) V[q] for a value
// coefficient V[r][q][p] at the Gauss points (LDS).  out[k] receives the
// result for column (i,j) = (ta,tb) over k.
// MX: as for bk3_forward, phases 4, 6, 7 and 8 on the matrix cores (no V);
// phase 7 applies the product matrix S D^T to the y-flux and S to the rest in
// one contraction of length 2 NQ.
template <int NM, int NQ, typename T, int GMODE, bool HASV, bool WAVE = (64 % (NQ * NQ) == 0),
          bool MX = false>
__device__ __forceinline__ void
bk3_backward(T (&out)[NM], T (&g)[6][NQ], const T *Gc, const Shape1D<T> &sh, T *P0, T *P1,
             T *P2, const T *V, const bool active, const int ta, const int tb,
             T *mx_lds = nullptr, const int mx_cells = 0, const float *mx_a = nullptr)
{
  constexpr int NQ2 = NQ * NQ, NM2 = NM * NM;
  static_assert(!MX || (std::is_same<T, float>::value && !WAVE && !HASV), "MX: FP32, no V");
  [[maybe_unused]] constexpr int NQ3 = NQ2 * NQ, CS = lds_cell_stride<T>(3 * NQ3);
  auto none = [](int, int) { return 0.f; };
  auto drop = [](int, int, float) {};
  if constexpr (MX) {
    // phase 4: column (r,p), du/dxi_1 [r][q][p] = sum_n D[n][q] U[r][n][p]
    mx_sweep<NQ, NQ, false>(
        mx_cells * NQ2, [&](int m, int k) { return mx_a[2 * 256 + m * 16 + k]; }, none,
        [&](int c, int k) {
          const int e = c / NQ2, l = c - e * NQ2, r = l / NQ, p = l - r * NQ;
          return mx_lds[e * CS + r * NQ2 + k * NQ + p];
        },
        [&](int c, int m, float v) {
          const int e = c / NQ2, l = c - e * NQ2, r = l / NQ, p = l - r * NQ;
          mx_lds[e * CS + NQ3 + r * NQ2 + m * NQ + p] = v;
        },
        drop);
    __syncthreads();
  } else {
  // ---- phase 4: y-gradient, lane (p,r) = (ta,tb)
  if (active) {
    T c[NQ], w[NQ];
#pragma unroll
    for (int n = 0; n < NQ; ++n) c[n] = P0[tb * NQ2 + n * NQ + ta];
    eo_deriv<NQ, T>(sh, c, w);
#pragma unroll
    for (int q = 0; q < NQ; ++q) P1[tb * NQ2 + q * NQ + ta] = w[q];  // du/dxi_1
  }
  cell_sync<NQ, WAVE>();
  }

  // ---- phase 5: z-gradient, flux G grad u, z-integration; lane (p,q) = (ta,tb)
  if (active) {
    if (GMODE == kGLateNT) bk3_load_g<NQ, T, GMODE>(Gc, g);
    T col[NQ], uzs[NQ], fz[NQ], w[NQ];
#pragma unroll
    for (int n = 0; n < NQ; ++n) col[n] = P0[n * NQ2 + tb * NQ + ta];
    eo_deriv<NQ, T>(sh, col, uzs);
#pragma unroll
    for (int r = 0; r < NQ; ++r) {
      const T uz = uzs[r];
      const int id = r * NQ2 + tb * NQ + ta;
      const T ux = P2[id], uy = P1[id];
      P2[id] = g[0][r] * ux + g[1][r] * uy + g[2][r] * uz;
      P1[id] = g[1][r] * ux + g[3][r] * uy + g[4][r] * uz;
      fz[r] = g[2][r] * ux + g[4][r] * uy + g[5][r] * uz;
    }
    eo_deriv_t<NQ, T>(sh, fz, w);
#pragma unroll
    for (int r = 0; r < NQ; ++r) P0[r * NQ2 + tb * NQ + ta] = w[r];
  }
  cell_sync<NQ, WAVE>();

  if constexpr (MX) {
    // phase 6: column (r,q), Z[r][q][p] += sum_n D[p][n] F_x[r][q][n]
    mx_sweep<NQ, NQ, false>(
        mx_cells * NQ2, [&](int m, int k) { return mx_a[3 * 256 + m * 16 + k]; }, none,
        [&](int c, int k) {
          const int e = c / NQ2, l = c - e * NQ2;
          return mx_lds[e * CS + 2 * NQ3 + l * NQ + k];
        },
        [&](int c, int m, float v) {
          const int e = c / NQ2, l = c - e * NQ2;
          mx_lds[e * CS + l * NQ + m] += v;
        },
        drop);
    __syncthreads();
    // phase 7: column (r,p), W[r][j][p] = sum_q S[j][q] (sum_n D[q][n]
    // F_y[r][n][p] + Z[r][q][p]); the column is rewritten in place (a column
    // is read and written by the lanes of one tile only)
    mx_sweep<NM, 2 * NQ, false>(
        mx_cells * NQ2, [&](int m, int k) { return mx_a[4 * 256 + m * 16 + k]; }, none,
        [&](int c, int k) {
          const int e = c / NQ2, l = c - e * NQ2, r = l / NQ, p = l - r * NQ;
          return k < NQ ? mx_lds[e * CS + NQ3 + r * NQ2 + k * NQ + p]
                        : mx_lds[e * CS + r * NQ2 + (k - NQ) * NQ + p];
        },
        [&](int c, int m, float v) {
          const int e = c / NQ2, l = c - e * NQ2, r = l / NQ, p = l - r * NQ;
          mx_lds[e * CS + NQ3 + r * NQ2 + m * NQ + p] = v;
        },
        drop);
    __syncthreads();
    // phase 8: column (r,j), Z[r][j][i] = sum_p S[i][p] W[r][j][p]
    constexpr int L = NQ * NM;
    mx_sweep<NM, NQ, false>(
        mx_cells * L, [&](int m, int k) { return mx_a[5 * 256 + m * 16 + k]; }, none,
        [&](int c, int k) {
          const int e = c / L, l = c - e * L, r = l / NM, j = l - r * NM;
          return mx_lds[e * CS + NQ3 + r * NQ2 + j * NQ + k];
        },
        [&](int c, int m, float v) {
          const int e = c / L, l = c - e * L, r = l / NM, j = l - r * NM;
          mx_lds[e * CS + 2 * NQ3 + r * NM2 + j * NM + m] = v;
        },
        drop);
    __syncthreads();
  } else {
  // ---- phase 6: x-integration accumulated onto the z part, lane (q,r)
  if (active) {
    T f[NQ], w[NQ];
#pragma unroll
    for (int n = 0; n < NQ; ++n) f[n] = P2[tb * NQ2 + ta * NQ + n];
    eo_deriv_t<NQ, T>(sh, f, w);
#pragma unroll
    for (int p = 0; p < NQ; ++p) P0[tb * NQ2 + ta * NQ + p] += w[p];
  }
  cell_sync<NQ, WAVE>();

  // ---- phase 7: y-integration + y back-interpolation, lane (p,r)
  if (active) {
    T f[NQ], acc[NQ], w[NM];
#pragma unroll
    for (int n = 0; n < NQ; ++n) f[n] = P1[tb * NQ2 + n * NQ + ta];
    eo_deriv_t<NQ, T>(sh, f, acc);
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      T v = P0[tb * NQ2 + q * NQ + ta];
      if (HASV) v += V[tb * NQ2 + q * NQ + ta];
      acc[q] += v;
    }
    eo_interp_t<NM, NQ, T>(sh, acc, w);
#pragma unroll
    for (int j = 0; j < NM; ++j) P1[tb * NQ2 + j * NQ + ta] = w[j];  // W[r][j][p]
  }
  cell_sync<NQ, WAVE>();

  // ---- phase 8: x back-interpolation, lane (j,r) = (ta,tb)
  if (active && ta < NM) {
    T l[NQ], w[NM];
#pragma unroll
    for (int p = 0; p < NQ; ++p) l[p] = P1[tb * NQ2 + ta * NQ + p];
    eo_interp_t<NM, NQ, T>(sh, l, w);
#pragma unroll
    for (int i = 0; i < NM; ++i) P2[tb * NM2 + ta * NM + i] = w[i];  // Z[r][j][i]
  }
  cell_sync<NQ, WAVE>();
  }

  // ---- phase 9: z back-interpolation, lane (i,j) = (ta,tb)
  if (active && ta < NM && tb < NM) {
    T z[NQ];
#pragma unroll
    for (int r = 0; r < NQ; ++r) z[r] = P2[r * NM2 + tb * NM + ta];
    eo_interp_t<NM, NQ, T>(sh, z, out);
  }
}

// Both halves: out = A_e u (see bk3_forward / bk3_backward).
template <int NM, int NQ, typename T, int GMODE, bool WAVE = (64 % (NQ * NQ) == 0), bool MX = false>
__device__ __forceinline__ void
bk3_cell_apply(const T (&u)[NM], T (&out)[NM], T (&g)[6][NQ], const T *Gc, const Shape1D<T> &sh,
               T *P0, T *P1, T *P2, const bool active, const int ta, const int tb,
               T *zbuf = nullptr, const int zn = 0, T *mx_lds = nullptr, const int mx_cells = 0,
               const float *mx_a = nullptr)
{
  bk3_forward<NM, NQ, T, WAVE, MX>(u, sh, P0, P1, P2, active, ta, tb, zbuf, zn, mx_lds, mx_cells,
                                   mx_a);
  bk3_backward<NM, NQ, T, GMODE, false, WAVE, MX>(out, g, Gc, sh, P0, P1, P2, nullptr, active, ta,
                                                  tb, mx_lds, mx_cells, mx_a);
}

}  // namespace pmg
