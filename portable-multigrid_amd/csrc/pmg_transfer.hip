// pmg_transfer.hip -- tensor-product h/p transfer kernels for gfx950.
//
// Prolongation (reference BK1::Parallel::KokkosProlongationBatchedKernel,
// bk1_kokkos_kernels.h:462-759): per coarse cell gather the coarse values
// (invalid -> 0), apply M x M x M (nmc -> nmf per direction), multiply by the
// fine weights and scatter-add into the fine vector.  Restriction
// (KokkosRestrictionBatchedKernel, :761-1051) is the exact transpose with the
// weight applied on the gather.
// Layout: one nmf x nmf lane tile per cell (256/nmf^2 cells per workgroup),
// LDS for the x and y sweeps, a register column for the z sweep.  The
// prolongation does z LAST and the restriction z FIRST, so the fine-side
// scatter-add / gather (nmf^3 per cell, the dominant traffic) is issued by
// lanes (a,b) over consecutive fine DoFs along x; M is a kernel argument held
// in scalar registers.  The two LDS buffers of a cell hold at most
// nmc nmf^2 and nmc^2 nmf values (the partially swept tensors), which keeps
// the LDS per workgroup small enough for 8 workgroups per CU: both kernels are
// bound by the latency of their dependent gathers, so occupancy is what hides
// it.  The owner-mode prolongation loads its fine map entries and the fine
// values it will update at kernel entry (it is their only writer), so those
// two dependent loads overlap the coarse gather and the sweeps.
#include "pmg_common.hpp"
#include "pmg_kernels.hpp"

#include <algorithm>

namespace pmg {

template <int NF>
struct XferTile {
  static constexpr int NF2 = NF * NF;
  static constexpr int cells = (256 / NF2) > 0 ? 256 / NF2 : 1;
  static constexpr int lanes = ((cells * NF2 + 63) / 64) * 64;
};

// LDS elements per cell: buffer 0 holds NC x NF x NF (or fewer) values,
// buffer 1 NC x NC x NF
template <int NC, int NF>
constexpr int xfer_lds_cell() { return NC * NF * NF + NC * NC * NF; }

// OWNER: map_f carries bit 31 on the one (cell, local) entry that owns each
// fine DoF (no weights): the owner adds scale * value with a plain
// read-modify-write (no other cell of the launch touches that DoF), the
// restriction gathers each fine value once through its owner.  This equals
// the weighted form with weights 1/multiplicity because the prolongated value
// of a node shared by several cells is the same from every cell, and a coarse
// basis function of one cell vanishes on the faces it does not belong to.
// LINES (owner mode): the fine map is given per x-line, fl[cell][c][b] =
// {first DoF of the line, bit a set if entry (a,b,c) is the owner}, every
// x-line of every cell being a run of consecutive DoFs: the same owner-mode
// entries in 8 bytes per NF entries instead of 4 bytes per entry.
template <int NF>
__device__ __forceinline__ uint32_t owner_entry(const uint2 *__restrict__ fl, int64_t cell, int c,
                                                int tb, int ta) {
  const uint2 w = fl[(cell * NF + c) * NF + tb];
  return (w.x + ta) | (((w.y >> ta) & 1u) << 31);
}

template <int NC, int NF, typename T, bool OWNER, bool LINES>
__global__ void __launch_bounds__(256)
prolongate_kernel(const int64_t n_cells, const Transfer1D<T> M, const uint32_t *__restrict__ map_c,
                  const uint32_t *__restrict__ map_f, const uint2 *__restrict__ fl,
                  const T *__restrict__ w, const T scale, const T *__restrict__ xc,
                  T *__restrict__ yf)
{
  constexpr int NF2 = NF * NF, NF3 = NF2 * NF, NC2 = NC * NC, NC3 = NC2 * NC;
  constexpr int E = XferTile<NF>::cells, LC = xfer_lds_cell<NC, NF>();
  __shared__ T lds[E * LC];
  const unsigned lb = xcd_logical_block(blockIdx.x, gridDim.x);
  const int tid = threadIdx.x, e = tid / NF2, t = tid - e * NF2, ta = t % NF, tb = t / NF;
  const int64_t cell = static_cast<int64_t>(lb) * E + e;
  const bool active = e < E && cell < n_cells;
  T *L0 = lds + (active ? e : 0) * LC, *L1 = L0 + NC * NF2;

  // owner mode: this lane's fine column (map entries, current fine values)
  // is fetched now; nothing else in the launch writes these DoFs
  const int64_t base = cell * NF3 + tb * NF + ta;
  uint32_t mf[NF];
  T yv[NF];  // owner form: current fine values; weighted form: the weights
  if (OWNER && active) {
#pragma unroll
    for (int c = 0; c < NF; ++c)
      mf[c] = LINES ? owner_entry<NF>(fl, cell, c, tb, ta) : map_f[base + c * NF2];
#pragma unroll
    for (int c = 0; c < NF; ++c) yv[c] = (mf[c] & 0x80000000u) ? yf[mf[c] & 0x7FFFFFFFu] : T(0);
  } else if (active) {
#pragma unroll
    for (int c = 0; c < NF; ++c) {
      mf[c] = map_f[base + c * NF2];
      yv[c] = w[base + c * NF2];
    }
  }

  // gather the coarse cell values, lane (i,j): L0[k][j][i]; branch-free (an
  // invalid entry reads entry 0, zeroed after the load): a load under a lane
  // condition whose value is used in the branch makes the compiler wait for
  // it there, one memory latency per entry
  if (active && ta < NC && tb < NC) {
    uint32_t d[NC];
    T v[NC];
#pragma unroll
    for (int k = 0; k < NC; ++k) d[k] = map_c[cell * NC3 + k * NC2 + tb * NC + ta];
#pragma unroll
    for (int k = 0; k < NC; ++k) v[k] = xc[d[k] == kInvalid ? 0u : d[k]];
#pragma unroll
    for (int k = 0; k < NC; ++k) L0[k * NC2 + tb * NC + ta] = d[k] == kInvalid ? T(0) : v[k];
  }
  __syncthreads();
  // x sweep, lane (j,k): L1[k][j][a]
  if (active && ta < NC && tb < NC) {
    T l[NC];
#pragma unroll
    for (int i = 0; i < NC; ++i) l[i] = L0[tb * NC2 + ta * NC + i];
#pragma unroll
    for (int a = 0; a < NF; ++a) {
      T v = 0;
#pragma unroll
      for (int i = 0; i < NC; ++i) v += M.M[i * NF + a] * l[i];
      L1[(tb * NC + ta) * NF + a] = v;
    }
  }
  __syncthreads();
  // y sweep, lane (a,k): L0[k][b][a]
  if (active && tb < NC) {
    T l[NC];
#pragma unroll
    for (int j = 0; j < NC; ++j) l[j] = L1[(tb * NC + j) * NF + ta];
#pragma unroll
    for (int b = 0; b < NF; ++b) {
      T v = 0;
#pragma unroll
      for (int j = 0; j < NC; ++j) v += M.M[j * NF + b] * l[j];
      L0[(tb * NF + b) * NF + ta] = v;
    }
  }
  __syncthreads();
  // z sweep on a register column, lane (a,b); the weighted scatter-add runs
  // along a across lanes, i.e. over consecutive fine DoFs
  if (active) {
    T l[NC];
#pragma unroll
    for (int k = 0; k < NC; ++k) l[k] = L0[(k * NF + tb) * NF + ta];
#pragma unroll
    for (int c = 0; c < NF; ++c) {
      T v = 0;
#pragma unroll
      for (int k = 0; k < NC; ++k) v += M.M[k * NF + c] * l[k];
      if (OWNER) {
        if (mf[c] & 0x80000000u) yf[mf[c] & 0x7FFFFFFFu] = yv[c] + scale * v;
      } else {
        if (yv[c] != T(0)) atomic_add(yf + mf[c], yv[c] * v);
      }
    }
  }
}

template <int NC, int NF, typename T, bool OWNER, bool LINES>
__global__ void __launch_bounds__(256)
restrict_kernel(const int64_t n_cells, const Transfer1D<T> M, const uint32_t *__restrict__ map_c,
                const uint32_t *__restrict__ map_f, const uint2 *__restrict__ fl,
                const T *__restrict__ w, const T scale,
                const T *__restrict__ xf, const T *__restrict__ xs, T *__restrict__ yc,
                T *__restrict__ buf)
{
  // restricts xf - xs (xs optional: the V-cycle's residual d - A s, computed
  // here instead of in a pass of its own).  buf (optional): the cell's coarse
  // values are stored to buf[map_c[entry]] instead of added to yc with
  // atomics; map_c then holds buffer positions in coarse-DoF order, and a
  // contiguous sum per coarse DoF follows (restrict_sum_kernel).
  constexpr int NF2 = NF * NF, NF3 = NF2 * NF, NC2 = NC * NC, NC3 = NC2 * NC;
  constexpr int E = XferTile<NF>::cells, LC = xfer_lds_cell<NC, NF>();
  __shared__ T lds[E * LC];
  const unsigned lb = xcd_logical_block(blockIdx.x, gridDim.x);
  const int tid = threadIdx.x, e = tid / NF2, t = tid - e * NF2, ta = t % NF, tb = t / NF;
  const int64_t cell = static_cast<int64_t>(lb) * E + e;
  const bool active = e < E && cell < n_cells;
  const bool writer = active && ta < NC && tb < NC;
  T *L0 = lds + (active ? e : 0) * LC, *L1 = L0 + NC * NF2;
  // the coarse targets of lane (i,j) over k, requested at entry so that the
  // write-out does not wait for them (lanes that write nothing load entry 0)
  uint32_t dc[NC];
  {
    const int64_t cb = writer ? cell * NC3 + tb * NC + ta : 0;
#pragma unroll
    for (int k = 0; k < NC; ++k) dc[k] = map_c[cb + k * NC2];
  }

  // weighted gather of the fine column over c by lane (a,b) (consecutive fine
  // DoFs across lanes), z sweep in registers: L0[k][b][a]
  // Branch-free gathers (every fine map entry is a valid DoF): a load under a
  // lane condition makes the compiler wait for its result inside the branch,
  // one memory latency per entry.  Entries that are not gathered (no owner
  // bit, zero weight) are loaded too and zeroed after the load; the branch on
  // xs is uniform.
  if (active) {
    T f[NF];
    const int64_t base = cell * NF3 + tb * NF + ta;
    uint32_t m[NF];
    T wt[NF], v[NF];
#pragma unroll
    for (int c = 0; c < NF; ++c) {
      m[c] = OWNER ? (LINES ? owner_entry<NF>(fl, cell, c, tb, ta) : map_f[base + c * NF2])
                   : map_f[base + c * NF2];
      wt[c] = OWNER ? T(1) : w[base + c * NF2];
    }
    if (xs) {
      T a[NF], b[NF];
#pragma unroll
      for (int c = 0; c < NF; ++c) {
        a[c] = xf[m[c] & 0x7FFFFFFFu];
        b[c] = xs[m[c] & 0x7FFFFFFFu];
      }
#pragma unroll
      for (int c = 0; c < NF; ++c) v[c] = a[c] - b[c];
    } else {
#pragma unroll
      for (int c = 0; c < NF; ++c) v[c] = xf[m[c] & 0x7FFFFFFFu];
    }
#pragma unroll
    for (int c = 0; c < NF; ++c) {
      if (OWNER) f[c] = (m[c] & 0x80000000u) ? v[c] : T(0);
      else f[c] = wt[c] == T(0) ? T(0) : wt[c] * v[c];
    }
#pragma unroll
    for (int k = 0; k < NC; ++k) {
      T v = 0;
#pragma unroll
      for (int c = 0; c < NF; ++c) v += M.M[k * NF + c] * f[c];
      L0[(k * NF + tb) * NF + ta] = v;
    }
  }
  __syncthreads();
  // y sweep, lane (a,k): L1[k][j][a]
  if (active && tb < NC) {
    T l[NF];
#pragma unroll
    for (int b = 0; b < NF; ++b) l[b] = L0[(tb * NF + b) * NF + ta];
#pragma unroll
    for (int j = 0; j < NC; ++j) {
      T v = 0;
#pragma unroll
      for (int b = 0; b < NF; ++b) v += M.M[j * NF + b] * l[b];
      L1[(tb * NC + j) * NF + ta] = v;
    }
  }
  __syncthreads();
  // x sweep, lane (j,k): L0[k][j][i]
  if (active && ta < NC && tb < NC) {
    T l[NF];
#pragma unroll
    for (int a = 0; a < NF; ++a) l[a] = L1[(tb * NC + ta) * NF + a];
#pragma unroll
    for (int i = 0; i < NC; ++i) {
      T v = 0;
#pragma unroll
      for (int a = 0; a < NF; ++a) v += M.M[i * NF + a] * l[a];
      L0[tb * NC2 + ta * NC + i] = v;
    }
  }
  __syncthreads();
  // scatter-add to the coarse DoFs by lane (i,j) over k (skip constrained),
  // or store into the buffer position
  if (writer) {
#pragma unroll
    for (int k = 0; k < NC; ++k) {
      const int l = k * NC2 + tb * NC + ta;
      if (dc[k] != kInvalid) {
        if (buf) buf[dc[k]] = L0[l];
        else atomic_add(yc + dc[k], L0[l]);
      }
    }
  }
}

// Second half of a restriction through the buffer: coarse DoF i sums its
// entries buf[off[i] .. off[i+1]) in order (the order of the inverse map, so
// the result does not depend on scheduling).  Loads are issued in groups
// without dependent steps between them (indices clamped into the DoF's own
// range): entry 0, then entries 1-3 where there are more, then 4-7; only
// lanes whose DoF has that many entries issue a group, so a wave touches no
// more cache lines than its entries occupy.
template <typename T>
__global__ void __launch_bounds__(256)
restrict_sum_kernel(const int64_t n, const uint32_t *__restrict__ off, const T *__restrict__ buf,
                    T *__restrict__ dst)
{
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t k0 = off[i], k1 = off[i + 1];
  if (k0 == k1) return;
  const uint32_t ne = k1 - k0, kl = k1 - 1;
  T acc = buf[k0];
  if (ne > 1) {
    T v[3];
#pragma unroll
    for (int j = 0; j < 3; ++j) v[j] = buf[k0 + 1 + j <= kl ? k0 + 1 + j : kl];
    if (ne > 4) {
      T w[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) w[j] = buf[k0 + 4 + j <= kl ? k0 + 4 + j : kl];
#pragma unroll
      for (int j = 0; j < 3; ++j) acc += v[j];
#pragma unroll
      for (int j = 0; j < 4; ++j)
        if (4u + j < ne) acc += w[j];
      for (uint32_t k = k0 + 8; k < k1; ++k) acc += buf[k];
    } else {
#pragma unroll
      for (int j = 0; j < 3; ++j)
        if (1u + j < ne) acc += v[j];
    }
  }
  dst[i] += acc;
}

template <int NC, int NF, typename T>
static void xfer_launch(bool prolong, int64_t n_cells, const Transfer1D<T> &M,
                        const uint32_t *map_c, const uint32_t *map_f, const uint2 *fl, const T *w,
                        T scale, const T *x, const T *xs, T *y, T *buf, hipStream_t s) {
  if (n_cells <= 0) return;
  using Tile = XferTile<NF>;
  const int64_t nb = (n_cells + Tile::cells - 1) / Tile::cells;
  PMG_REQUIRE(nb < (int64_t(1) << 31), PMG_ERR_SIZE, "too many cells for one launch");
  const unsigned grid = pad_to_xcd(static_cast<unsigned>(nb));
  const dim3 g(grid), b(Tile::lanes);
  if (prolong && w)
    hipLaunchKernelGGL((prolongate_kernel<NC, NF, T, false, false>), g, b, 0, s, n_cells, M, map_c,
                       map_f, fl, w, scale, x, y);
  else if (prolong && fl)
    hipLaunchKernelGGL((prolongate_kernel<NC, NF, T, true, true>), g, b, 0, s, n_cells, M, map_c,
                       map_f, fl, w, scale, x, y);
  else if (prolong)
    hipLaunchKernelGGL((prolongate_kernel<NC, NF, T, true, false>), g, b, 0, s, n_cells, M, map_c,
                       map_f, fl, w, scale, x, y);
  else if (w)
    hipLaunchKernelGGL((restrict_kernel<NC, NF, T, false, false>), g, b, 0, s, n_cells, M, map_c,
                       map_f, fl, w, scale, x, xs, y, buf);
  else if (fl)
    hipLaunchKernelGGL((restrict_kernel<NC, NF, T, true, true>), g, b, 0, s, n_cells, M, map_c,
                       map_f, fl, w, scale, x, xs, y, buf);
  else
    hipLaunchKernelGGL((restrict_kernel<NC, NF, T, true, false>), g, b, 0, s, n_cells, M, map_c,
                       map_f, fl, w, scale, x, xs, y, buf);
  PMG_HIP_CHECK(hipGetLastError());
}

template <typename T>
static void xfer_dispatch(bool prolong, int nmc, int nmf, int64_t n_cells, const Transfer1D<T> &M,
                          const uint32_t *map_c, const uint32_t *map_f, const uint2 *fl,
                          const T *w, T scale, const T *x, const T *xs, T *y, T *buf,
                          hipStream_t s) {
#define PMG_XF(C, F)                                                                          \
  if (nmc == C && nmf == F) {                                                                 \
    xfer_launch<C, F, T>(prolong, n_cells, M, map_c, map_f, fl, w, scale, x, xs, y, buf, s);  \
    return;                                                                                   \
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
                       const uint32_t *map_c, const uint32_t *map_f, const T *w, T scale,
                       const T *xc, T *yf, hipStream_t s, const uint2 *fine_lines) {
  xfer_dispatch<T>(true, nmc, nmf, n_cells, M, map_c, map_f, w ? nullptr : fine_lines, w, scale,
                   xc, nullptr, yf, nullptr, s);
}
template <typename T>
void launch_restrict(int nmc, int nmf, int64_t n_cells, const Transfer1D<T> &M,
                     const uint32_t *map_c, const uint32_t *map_f, const T *w, const T *xf,
                     T *yc, hipStream_t s, const T *xf_sub, T *cell_buf,
                     const uint2 *fine_lines) {
  xfer_dispatch<T>(false, nmc, nmf, n_cells, M, map_c, map_f, w ? nullptr : fine_lines, w, T(1),
                   xf, xf_sub, yc, cell_buf, s);
}
template <typename T>
void launch_restrict_sum(int64_t n_c, const uint32_t *off, const T *buf, T *yc, hipStream_t s) {
  if (n_c <= 0) return;
  PMG_REQUIRE((n_c + 255) / 256 < (int64_t(1) << 31), PMG_ERR_SIZE, "too many coarse DoFs");
  hipLaunchKernelGGL(restrict_sum_kernel<T>, dim3(static_cast<unsigned>((n_c + 255) / 256)),
                     dim3(256), 0, s, n_c, off, buf, yc);
  PMG_HIP_CHECK(hipGetLastError());
}
template void launch_restrict_sum<double>(int64_t, const uint32_t *, const double *, double *,
                                          hipStream_t);
template void launch_restrict_sum<float>(int64_t, const uint32_t *, const float *, float *,
                                         hipStream_t);

template void launch_prolongate<double>(int, int, int64_t, const Transfer1D<double> &,
                                        const uint32_t *, const uint32_t *, const double *,
                                        double, const double *, double *, hipStream_t,
                                        const uint2 *);
template void launch_prolongate<float>(int, int, int64_t, const Transfer1D<float> &,
                                       const uint32_t *, const uint32_t *, const float *, float,
                                       const float *, float *, hipStream_t, const uint2 *);
template void launch_restrict<double>(int, int, int64_t, const Transfer1D<double> &,
                                      const uint32_t *, const uint32_t *, const double *,
                                      const double *, double *, hipStream_t, const double *,
                                      double *, const uint2 *);
template void launch_restrict<float>(int, int, int64_t, const Transfer1D<float> &,
                                     const uint32_t *, const uint32_t *, const float *,
                                     const float *, float *, hipStream_t, const float *, float *,
                                     const uint2 *);

// ---------------------------------------------------------------- set-up
namespace {
unsigned setup_grid(int64_t n) {
  return static_cast<unsigned>(std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 16384)));
}

__global__ void owner_first_kernel(const int64_t nw, const uint32_t *__restrict__ mf,
                                   const uint8_t *__restrict__ cons, uint32_t *__restrict__ first)
{
  for (int64_t e = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; e < nw;
       e += int64_t(gridDim.x) * blockDim.x) {
    const uint32_t d = mf[e];
    if (!cons[d]) atomicMin(first + d, static_cast<uint32_t>(e));
  }
}

__global__ void owner_mark_kernel(const int64_t nw, uint32_t *__restrict__ mf,
                                  const uint8_t *__restrict__ cons,
                                  const uint32_t *__restrict__ first)
{
  for (int64_t e = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; e < nw;
       e += int64_t(gridDim.x) * blockDim.x) {
    const uint32_t d = mf[e];
    if (!cons[d] && first[d] == static_cast<uint32_t>(e)) mf[e] = d | 0x80000000u;
  }
}

__global__ void fine_lines_kernel(const int64_t nl, const int nmf, const uint32_t *__restrict__ mf,
                                  uint2 *__restrict__ fl, uint32_t *__restrict__ bad)
{
  for (int64_t l = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; l < nl;
       l += int64_t(gridDim.x) * blockDim.x) {
    const uint32_t *e = mf + l * nmf;
    const uint32_t d0 = e[0] & 0x7FFFFFFFu;
    uint32_t bits = 0;
    bool ok = true;
    for (int a = 0; a < nmf; ++a) {
      ok = ok && (e[a] & 0x7FFFFFFFu) == d0 + static_cast<uint32_t>(a);
      bits |= (e[a] >> 31) << a;
    }
    fl[l] = make_uint2(d0, bits);
    if (!ok) atomicOr(bad, 1u);
  }
}
// pos[list[k]] = k: the buffer position of every entry of an inverse map
__global__ void csr_positions_kernel(const int64_t ne, const uint32_t *__restrict__ list,
                                     uint32_t *__restrict__ pos)
{
  for (int64_t k = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; k < ne;
       k += int64_t(gridDim.x) * blockDim.x)
    pos[list[k]] = static_cast<uint32_t>(k);
}
}  // namespace

void launch_csr_positions(int64_t n_pos, int64_t ne, const uint32_t *list, uint32_t *pos,
                          hipStream_t s) {
  if (n_pos <= 0) return;
  PMG_HIP_CHECK(hipMemsetAsync(pos, 0xFF, static_cast<size_t>(n_pos) * sizeof(uint32_t), s));
  if (ne > 0) {
    hipLaunchKernelGGL(csr_positions_kernel, dim3(setup_grid(ne)), dim3(256), 0, s, ne, list, pos);
    PMG_HIP_CHECK(hipGetLastError());
  }
  PMG_HIP_CHECK(hipStreamSynchronize(s));
}

void launch_owner_marks(int64_t nw, int64_t n_dofs, uint32_t *mf, const uint8_t *cons,
                        uint32_t *first, hipStream_t s) {
  if (nw <= 0) return;
  PMG_HIP_CHECK(hipMemsetAsync(first, 0xFF, static_cast<size_t>(n_dofs) * sizeof(uint32_t), s));
  hipLaunchKernelGGL(owner_first_kernel, dim3(setup_grid(nw)), dim3(256), 0, s, nw, mf, cons, first);
  PMG_HIP_CHECK(hipGetLastError());
  hipLaunchKernelGGL(owner_mark_kernel, dim3(setup_grid(nw)), dim3(256), 0, s, nw, mf, cons, first);
  PMG_HIP_CHECK(hipGetLastError());
  PMG_HIP_CHECK(hipStreamSynchronize(s));
}

void launch_fine_lines(int64_t n_lines, int nmf, const uint32_t *mf, uint2 *fl, uint32_t *bad,
                       hipStream_t s) {
  if (n_lines <= 0) return;
  hipLaunchKernelGGL(fine_lines_kernel, dim3(setup_grid(n_lines)), dim3(256), 0, s, n_lines, nmf, mf,
                     fl, bad);
  PMG_HIP_CHECK(hipGetLastError());
}

}  // namespace pmg
