// pmg_bk3.hip -- continuous Q_p Laplace operator kernels for gfx950.
//
// Operator (reference: BK3::Parallel::KokkosKernel, bk3_kokkos_kernel.h:25-746):
//   y += sum_cells P_e^T S^T D^T G_e D S P_e x
// Design (MI355X): a workgroup of <= 256 lanes holds E = 256/nq^2 cells; each
// cell owns an nq x nq tile of lanes.  A lane keeps one 1D line of the current
// tensor in registers and contracts along it, so every 1D sweep costs nm (or
// nq) LDS reads + nq (or nm) LDS writes per line instead of one read per FMA.
// The first z-interpolation and the last z-back-interpolation run on the
// register column loaded from / scattered to HBM, the z-gradient and its
// transpose run on a register column of the q-point tensor.  Three nq^3 LDS
// buffers per cell.  G (85 % of the algorithmic bytes) is fetched as 6*nq
// coalesced loads per lane issued at kernel entry so the HBM latency overlaps
// the interpolation sweeps; it is read once, so the loads are non-temporal
// and leave L2/MALL to x and y.  The DoF map column is loaded once and kept
// in registers for the scatter.  The 1D matrices are kernel arguments (scalar
// registers).  Blocks are remapped so each XCD (own L2) walks a contiguous
// range of cells, which keeps neighbouring cells' shared DoFs in one L2.
#include "pmg_bk3_cell.hpp"
#include "pmg_common.hpp"
#include "pmg_kernels.hpp"

#include <hipcub/hipcub.hpp>

#include <algorithm>

namespace pmg {

// One batch of E = Bk3Tile<NQ>::cells cells at positions cell_base .. cell_base+E-1
// of a range of n_range cells (cell index = cell_ids[pos] if cell_ids, else pos),
// processed by the first E*NQ^2 lanes of the workgroup (all lanes must call it:
// it holds workgroup barriers).  lds: E * 3 * NQ^3 elements.  SCATTER:
// kScatterAtomic adds with device atomics; kScatterPlain with a plain
// read-modify-write, valid only when no two cells of the range share a DoF
// (one colour); kScatterStore stores the cell's values to y[cell][local] (a
// per-cell buffer laid out like the DoF map), every entry, for a gather pass;
// kScatterGrid stores them to the exploded grid of a lexicographic box of
// gnx x gny cells per layer (each cell owns an nm^3 block of it, see
// bk3_grid_sum_kernel).
// Compressed indices (optional, lines != nullptr): for a cell with lflag set,
// every x-line (j,k) of its DoF map is a run of consecutive DoFs starting at
// lines[cell][k][j]; other cells read the full map.  The line starts and the
// flag are loaded together, so compressed cells need no dependent step.
// WT: wave-aligned lane layout (Bk3Tile::per_wave whole cells per wave, the
// workgroup has Bk3Tile::wave_lanes lanes): every phase of the cell sweeps is
// wave-local, no workgroup barrier.
enum : int { kScatterAtomic = 0, kScatterPlain = 1, kScatterStore = 2, kScatterGrid = 3 };

template <int NM, int NQ, typename T, int SCATTER = kScatterAtomic, int GMODE = kGEarlyNT,
          bool WT = false>
__device__ __forceinline__ void
bk3_batch(const T *__restrict__ x, T *__restrict__ y, const T *__restrict__ G,
          const uint32_t *__restrict__ map, const uint32_t *__restrict__ lines,
          const uint8_t *__restrict__ lflag, const uint32_t *__restrict__ cell_ids,
          const int64_t n_range, const Shape1D<T> &sh, const int64_t cell_base, T *lds,
          const int gnx = 0, const int gny = 0)
{
  constexpr int NQ2 = NQ * NQ, NQ3 = NQ2 * NQ, NM2 = NM * NM, NM3 = NM2 * NM;
  constexpr int E = Bk3Tile<NQ>::cells;

  const int tid = threadIdx.x;
  constexpr int CW = Bk3Tile<NQ>::per_wave;
  const int wl = tid & 63, wc = wl / NQ2;
  const int e = WT ? (tid >> 6) * CW + wc : tid / NQ2;
  const int t = WT ? wl - wc * NQ2 : tid - e * NQ2, ta = t % NQ, tb = t / NQ;
  const int64_t pos = cell_base + e;
  const bool active = (!WT || wc < CW) && (e < E) && (pos < n_range);
  const int64_t cell = active ? (cell_ids ? static_cast<int64_t>(cell_ids[pos]) : pos) : 0;
  T *P0 = lds + (active ? e : 0) * lds_cell_stride<T>(3 * NQ3);
  const T *Gc = G + g_index<T>(NQ, cell, 0, active ? t : 0);

  T g[6][NQ];
  if (GMODE == kGAffine) {
    if (active) bk3_affine_g<NQ, T>(G + cell * 6, sh, ta, tb, g);
  } else if (GMODE != kGLateNT && active) {
    bk3_load_g<NQ, T, GMODE>(Gc, g);
  }

  // gather column (i,j) = (ta,tb) over k; invalid (constrained) entries read 0
  uint32_t idx[NM];
  T u[NM], v[NM];
  if (active && ta < NM && tb < NM) {
    // branch-free loads (a load under a lane condition whose value is used in
    // the branch makes the compiler wait for it there): without compressed
    // lines the line and flag loads read the map and are ignored, invalid
    // entries read x[0] and are zeroed after the load
    const uint32_t *lc = (lines ? lines : map) + cell * NM2 + tb;
    uint32_t ln[NM];
#pragma unroll
    for (int k = 0; k < NM; ++k) ln[k] = lc[k * NM];
    const uint8_t lf = (lflag ? lflag : reinterpret_cast<const uint8_t *>(map))[cell];
#pragma unroll
    for (int k = 0; k < NM; ++k) idx[k] = ln[k] + ta;
    if (!lines || lf == 0) {
      const uint32_t *mc = map + cell * NM3 + tb * NM + ta;
#pragma unroll
      for (int k = 0; k < NM; ++k) idx[k] = mc[k * NM2];
    }
    T xv[NM];
#pragma unroll
    for (int k = 0; k < NM; ++k) xv[k] = x[idx[k] == kInvalid ? 0u : idx[k]];
#pragma unroll
    for (int k = 0; k < NM; ++k) u[k] = idx[k] == kInvalid ? T(0) : xv[k];
  }
  bk3_cell_apply<NM, NQ, T, GMODE, WT || (64 % NQ2 == 0)>(u, v, g, Gc, sh, P0, P0 + NQ3,
                                                         P0 + 2 * NQ3, active, ta, tb);
  // scatter-add
  if (active && ta < NM && tb < NM) {
    if (SCATTER == kScatterStore) {
      T *yc = y + cell * NM3 + tb * NM + ta;
#pragma unroll
      for (int k = 0; k < NM; ++k) yc[k * NM2] = v[k];
    } else if (SCATTER == kScatterGrid) {
      const int c32 = static_cast<int>(cell), cyz = c32 / gnx, cx = c32 - cyz * gnx;
      const int cz = cyz / gny, cy = cyz - cz * gny;
      const int64_t row = int64_t(gnx) * NM, plane = row * gny * NM;
      T *yc = y + cz * NM * plane + (cy * NM + tb) * row + cx * NM + ta;
#pragma unroll
      for (int k = 0; k < NM; ++k) yc[k * plane] = idx[k] == kInvalid ? T(0) : v[k];
    } else {
#pragma unroll
      for (int k = 0; k < NM; ++k)
        if (idx[k] != kInvalid) {
          if (SCATTER == kScatterAtomic) atomic_add(y + idx[k], v[k]);
          else y[idx[k]] += v[k];  // one colour: this cell is the only writer
        }
    }
  }
}

// Wave-aligned layout for the cells kernel where it measured faster: FP32
// at nq = 7 (Q6 48^3: 424 -> 398 us).  FP64 loses with it (nq = 5: 497 ->
// 530 us, nq = 7: 572 -> 866 us): the larger workgroups meet the FP64 kernels'
// register budget and lose occupancy, and a fifth of the lanes idle.
template <int NQ, typename T>
constexpr bool bk3_wave_tile() { return std::is_same<T, float>::value && NQ == 7; }
template <int NQ, typename T>
constexpr int bk3_cells_lanes() {
  return bk3_wave_tile<NQ, T>() ? Bk3Tile<NQ>::wave_lanes : Bk3Tile<NQ>::lanes;
}

template <int NM, int NQ, typename T, int SCATTER, int GMODE>
__global__ void __launch_bounds__((bk3_cells_lanes<NQ, T>()))
bk3_cells_kernel(const T *__restrict__ x, T *__restrict__ y, const T *__restrict__ G,
                 const uint32_t *__restrict__ map, const uint32_t *__restrict__ lines,
                 const uint8_t *__restrict__ lflag, const uint32_t *__restrict__ cell_ids,
                 const int64_t n_range, const Shape1D<T> sh, const int gnx, const int gny)
{
  constexpr int E = Bk3Tile<NQ>::cells;
  __shared__ T lds[E * lds_cell_stride<T>(3 * NQ * NQ * NQ)];
  const unsigned lb = xcd_logical_block(blockIdx.x, gridDim.x);
  bk3_batch<NM, NQ, T, SCATTER, GMODE, bk3_wave_tile<NQ, T>()>(
      x, y, G, map, lines, lflag, cell_ids, n_range, sh, static_cast<int64_t>(lb) * E, lds, gnx,
      gny);
}

// ---------------------------------------------------------------- patch schedule
// One batch of E cells per workgroup.  The batch's patch (the union of its
// cells' unconstrained DoFs, sorted) is read once into LDS, each cell reads
// its source column from the patch, and the cells' results are summed into
// the patch, colour by colour of a batch-local cell colouring (or with LDS
// atomics where no colouring is given); the patch is then added to y once per DoF, in
// sorted DoF order, so a wave's device atomics cover long runs of consecutive
// DoFs.  Every DoF is added atomically, also those no other batch touches:
// plain stores into cache lines that other batches' atomics also update were
// measured slower on MI355X (Q4 64^3 FP64 549 -> 500 us, FP32 330 -> 321 us,
// Q6 48^3 FP32 480 -> 406 us with all-atomic write-out), so pdof's bit 31
// (shared with another batch) is not used here.  MX: the x and y sweeps of
// the cell apply on the matrix cores (bk3_forward / bk3_backward).
template <int NM, int NQ, typename T, int GMODE, bool COLORED, bool MX = false>
__global__ void __launch_bounds__(256)
bk3_patch_kernel(const T *__restrict__ x, T *__restrict__ y, const T *__restrict__ G,
                 const uint32_t *__restrict__ pdof, const uint32_t *__restrict__ pcount,
                 const uint32_t *__restrict__ pchunk, const uint16_t *__restrict__ pidx, const uint16_t *__restrict__ pline,
                 const uint8_t *__restrict__ pflag, const uint8_t *__restrict__ pcolor,
                 const uint8_t *__restrict__ pncol, const int64_t n_cells, const Shape1D<T> sh,
                 const float *__restrict__ mxa)
{
  constexpr int NQ2 = NQ * NQ, NQ3 = NQ2 * NQ, NM2 = NM * NM, NM3 = NM2 * NM;
  constexpr int E = Bk3Tile<NQ>::cells, PMAX = E * NM3;
  constexpr int LANES = Bk3Tile<NQ>::lanes, PL = (PMAX + LANES - 1) / LANES;
  __shared__ T lds[E * lds_cell_stride<T>(3 * NQ3)];
  __shared__ T patch[PMAX];
  const int64_t batch = xcd_logical_block(blockIdx.x, gridDim.x);
  const int64_t cell_base = batch * E;
  if (cell_base >= n_cells) return;  // whole workgroup: padding blocks of the grid
  const int tid = threadIdx.x;
  const int e = tid / NQ2, t = tid - e * NQ2, ta = t % NQ, tb = t / NQ;
  const int64_t cell = cell_base + e;
  const bool active = (e < E) && (cell < n_cells);
  const bool col = active && ta < NM && tb < NM;

  // Every index load is issued before the first wait, and without branches:
  // a load under a lane condition whose result is used inside the branch
  // makes the compiler wait for it there, which serialised the patch entry
  // loads one memory latency each (FP64 Q4 64^3 vmult 490 -> 447 us, FP32
  // Q6 48^3 314 -> 264 us).  The chain is now: slot starts, flag, patch
  // entries and count (one step) -> source values and full slot lists (one
  // step) -> LDS patch; G is requested last, so that waiting for the source
  // values never waits for G (loads complete in issue order), and G arrives
  // during the first sweeps.
  const int64_t lcell = active ? cell : cell_base;  // a valid cell for every lane
  const int lta = ta < NM ? ta : 0, ltb = tb < NM ? tb : 0;
  // compressed slot starts and flag; without compressed slots these loads read
  // pidx and are ignored
  uint16_t pr[NM];
  const uint16_t *pl = (pline ? pline : pidx) + lcell * NM2 + ltb;
#pragma unroll
  for (int k = 0; k < NM; ++k) pr[k] = pl[k * NM];
  const uint8_t pf = (pflag ? pflag : reinterpret_cast<const uint8_t *>(pidx))[lcell];
  // the lane's patch entries (kept for the write-out); indices past PMAX are
  // clamped into the batch's own entries, values past the count are not used
  const uint32_t *pd = pdof + batch * PMAX;
  uint32_t pw[PL];
  if (pchunk) {
    // the 64 slots a wave handles in pass j form chunk j LANES/64 + wave,
    // described by at most 4 runs of consecutive DoFs: one record of 8 words
    // per wave (scalar loads) instead of 64 entries; chunks of more runs load
    // their entries.  All records are loaded before any is used, and taken
    // as wave-uniform values (readfirstlane), which keeps the compiler from
    // sinking their loads into the selects below.
    constexpr int QC = (PMAX + 63) / 64;
    const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
    const uint32_t lane = static_cast<uint32_t>(tid & 63);
    uint4 rr[PL];
    uint32_t rs[PL];
#pragma unroll
    for (int j = 0; j < PL; ++j) {
      const int q = j * (LANES / 64) + wv;
      const uint32_t *rc = pchunk + (batch * QC + (q < QC ? q : QC - 1)) * 8;
      rr[j] = *reinterpret_cast<const uint4 *>(rc);
      rs[j] = rc[4];
    }
#pragma unroll
    for (int j = 0; j < PL; ++j) {
      const int q = j * (LANES / 64) + wv;
      const int i = tid + j * LANES;
      const uint32_t r0 = __builtin_amdgcn_readfirstlane(rr[j].x);
      const uint32_t r1 = __builtin_amdgcn_readfirstlane(rr[j].y);
      const uint32_t r2 = __builtin_amdgcn_readfirstlane(rr[j].z);
      const uint32_t r3 = __builtin_amdgcn_readfirstlane(rr[j].w);
      const uint32_t sw = __builtin_amdgcn_readfirstlane(rs[j]);
      const uint32_t s1 = sw & 0xFFu, s2 = (sw >> 8) & 0xFFu, s3 = (sw >> 16) & 0xFFu;
      uint32_t d = lane >= s3   ? r3 + (lane - s3)
                   : lane >= s2 ? r2 + (lane - s2)
                   : lane >= s1 ? r1 + (lane - s1)
                                : r0 + lane;
      if (q >= QC) d = 0u;
      else if ((sw >> 24) == 0) d = pd[i < PMAX ? i : PMAX - 1];  // more than 4 runs
      pw[j] = d;
    }
  } else {
#pragma unroll
    for (int j = 0; j < PL; ++j) {
      const int i = tid + j * LANES;
      pw[j] = pd[i < PMAX ? i : PMAX - 1];
    }
  }
  const int np = static_cast<int>(pcount[batch]);
  // colours of the batch's cells (uniform over the workgroup; 0: LDS atomics)
  const int n_colors = COLORED ? static_cast<int>(pncol[batch]) : 0;
  const int color = COLORED && n_colors > 0 && active ? static_cast<int>(pcolor[cell]) : 0;
  // source values of every slot, those past the count reading x[0] (only the
  // first np are stored)
  T xv[PL];
#pragma unroll
  for (int j = 0; j < PL; ++j) xv[j] = x[tid + j * LANES < np ? pw[j] & 0x7FFFFFFFu : 0u];
  uint16_t pi[NM];
#pragma unroll
  for (int k = 0; k < NM; ++k) pi[k] = static_cast<uint16_t>(pr[k] + lta);
  if (col && (!pline || pf == 0)) {  // cells without compressed slots
    const uint16_t *pc = pidx + cell * NM3 + tb * NM + ta;
#pragma unroll
    for (int k = 0; k < NM; ++k) pi[k] = pc[k * NM2];
  }
  T *P0 = lds + (active ? e : 0) * lds_cell_stride<T>(3 * NQ3);
  const T *Gc = G + g_index<T>(NQ, active ? cell : 0, 0, t);
  T g[6][NQ];
  if (GMODE == kGAffine) {
    if (active) bk3_affine_g<NQ, T>(G + cell * 6, sh, ta, tb, g);
  } else if (GMODE != kGLateNT && active) {
    bk3_load_g<NQ, T, GMODE>(Gc, g);
  }
#pragma unroll
  for (int j = 0; j < PL; ++j)
    if (tid + j * LANES < np) patch[tid + j * LANES] = xv[j];
  __syncthreads();

  T u[NM], v[NM];
  if (col) {
#pragma unroll
    for (int k = 0; k < NM; ++k) u[k] = pi[k] == 0xFFFFu ? T(0) : patch[pi[k]];
  }
  const int n_act = static_cast<int>(n_cells - cell_base < E ? n_cells - cell_base : E);
  bk3_cell_apply<NM, NQ, T, GMODE, (64 % NQ2 == 0), MX>(u, v, g, Gc, sh, P0, P0 + NQ3, P0 + 2 * NQ3,
                                                         active, ta, tb, patch, np, lds, n_act, mxa);
  if (COLORED && n_colors > 0) {
    // colour by colour: no two cells of a colour share a slot, so plain
    // read-modify-writes replace the LDS float atomics (measured costly at
    // FP32), in a fixed order
    for (int c = 0; c < n_colors; ++c) {
      if (col && color == c) {
#pragma unroll
        for (int k = 0; k < NM; ++k)
          if (pi[k] != 0xFFFFu) patch[pi[k]] += v[k];
      }
      __syncthreads();
    }
  } else {
    if (col) {
#pragma unroll
      for (int k = 0; k < NM; ++k)
        if (pi[k] != 0xFFFFu) atomicAdd(patch + pi[k], v[k]);
    }
    __syncthreads();
  }
#pragma unroll
  for (int j = 0; j < PL; ++j) {
    const int i = tid + j * LANES;
    if (i < np) {
      atomic_add(y + (pw[j] & 0x7FFFFFFFu), patch[i]);
    }
  }
}

template <int NM, int NQ, typename T, int GMODE>
static void bk3_launch_patch(bool assign, const T *x, T *y, const T *G, const Bk3Schedule &sc,
                             int64_t n_cells, const Shape1D<T> &sh, hipStream_t s) {
  if (n_cells <= 0) return;
  using Tile = Bk3Tile<NQ>;
  const int64_t nb = (n_cells + Tile::cells - 1) / Tile::cells;
  PMG_REQUIRE(nb < (int64_t(1) << 31), PMG_ERR_SIZE, "too many cells for one launch");
  const unsigned grid = pad_to_xcd(static_cast<unsigned>(nb));
  (void)assign;  // y += patch either way: callers of vmult zero the free rows first
  // the colour-by-colour sum is compiled for the default G-load policy and
  // affine geometry only
  bool colored = false;
  if constexpr (GMODE == kGEarlyNT || GMODE == kGAffine) {
    colored = sc.pncol != nullptr && sc.pcolor != nullptr;
    // the x and y sweeps on the matrix cores: FP32 at nq >= 6, tiles that
    // straddle waves (the sweeps mix the batch's cells over all waves)
    if constexpr (std::is_same<T, float>::value && NQ >= 6 && 64 % (NQ * NQ) != 0 &&
                  GMODE == kGEarlyNT) {
      if (sc.mxa) {
        if (colored)
          hipLaunchKernelGGL((bk3_patch_kernel<NM, NQ, T, GMODE, true, true>), dim3(grid),
                             dim3(Tile::lanes), 0, s, x, y, G, sc.pdof, sc.pcount, sc.pchunk,
                             sc.pidx, sc.pline, sc.pflag, sc.pcolor, sc.pncol, n_cells, sh, sc.mxa);
        else
          hipLaunchKernelGGL((bk3_patch_kernel<NM, NQ, T, GMODE, false, true>), dim3(grid),
                             dim3(Tile::lanes), 0, s, x, y, G, sc.pdof, sc.pcount, sc.pchunk,
                             sc.pidx, sc.pline, sc.pflag, nullptr, nullptr, n_cells, sh, sc.mxa);
        PMG_HIP_CHECK(hipGetLastError());
        return;
      }
    }
    if (colored)
      hipLaunchKernelGGL((bk3_patch_kernel<NM, NQ, T, GMODE, true>), dim3(grid), dim3(Tile::lanes),
                         0, s, x, y, G, sc.pdof, sc.pcount, sc.pchunk, sc.pidx, sc.pline, sc.pflag, sc.pcolor,
                         sc.pncol, n_cells, sh, nullptr);
  }
  if (!colored)
    hipLaunchKernelGGL((bk3_patch_kernel<NM, NQ, T, GMODE, false>), dim3(grid), dim3(Tile::lanes),
                       0, s, x, y, G, sc.pdof, sc.pcount, sc.pchunk, sc.pidx, sc.pline, sc.pflag, nullptr,
                       nullptr, n_cells, sh, nullptr);
  PMG_HIP_CHECK(hipGetLastError());
}

// ---------------------------------------------------------------- coarse-grid PCG
// Jacobi-preconditioned CG on a small level run entirely inside ONE workgroup:
// operator applications loop over cell batches with bk3_batch, dot products
// are workgroup reductions, and the convergence test ||r|| <= tol ||r0|| is
// evaluated on the device, so a coarse solve costs one launch instead of
// ~10 launches and a host round trip per iteration.  Every lane derives the
// loop decisions from the same LDS-broadcast scalars, so barriers stay uniform.
//
// Workgroup sums broadcast to every lane with one barrier: the wave sums go
// to one of two LDS slots, used alternately, and every lane adds them in wave
// order (the same additions one lane made when it broadcast the total).  A
// slot is rewritten two calls later, after the barrier of the call between,
// which every lane reaches only after reading it.  red: 2 x 8 x 2 doubles.
struct WgSum {
  double *red;
  int slot;
  __device__ explicit WgSum(double *r) : red(r), slot(0) {}
  // a and b summed over the workgroup (both in one barrier)
  __device__ void sum2(double &a, double &b) {
    for (int off = 32; off > 0; off >>= 1) {
      a += __shfl_down(a, off, 64);
      b += __shfl_down(b, off, 64);
    }
    double *r = red + slot * 16;
    slot ^= 1;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) {
      r[2 * wave] = a;
      r[2 * wave + 1] = b;
    }
    __syncthreads();
    a = 0;
    b = 0;
    for (int w = 0; w < static_cast<int>(blockDim.x >> 6); ++w) {
      a += r[2 * w];
      b += r[2 * w + 1];
    }
  }
  __device__ double sum(double v) {
    double u = 0;
    sum2(v, u);
    return v;
  }
};

template <int NM, int NQ, typename T>
__global__ void __launch_bounds__(256)
coarse_pcg_kernel(const T *__restrict__ G, const uint32_t *__restrict__ map, const int64_t n_cells,
                  const Shape1D<T> sh, const uint32_t *__restrict__ cons, const int64_t n_cons,
                  const T *__restrict__ pinv, const int64_t n, const T *__restrict__ b, T *x, T *r,
                  T *z, T *p, T *q, const double rel_tol, const int max_it, double *info)
{
  constexpr int E = Bk3Tile<NQ>::cells;
  __shared__ T lds[E * lds_cell_stride<T>(3 * NQ * NQ * NQ)];
  __shared__ double red[32];
  WgSum wg(red);
  const int tid = threadIdx.x, nt = blockDim.x;

  auto apply = [&](const T *in, T *out) {
    for (int64_t i = tid; i < n; i += nt) out[i] = T(0);
    __syncthreads();
    for (int64_t base = 0; base < n_cells; base += E) {
      bk3_batch<NM, NQ, T>(in, out, G, map, nullptr, nullptr, nullptr, n_cells, sh, base, lds);
      __syncthreads();
    }
    for (int64_t i = tid; i < n_cons; i += nt) out[cons[i]] = in[cons[i]];
    __syncthreads();
  };

  apply(x, q);
  double s = 0;
  for (int64_t i = tid; i < n; i += nt) {
    const T ri = b[i] - q[i];
    r[i] = ri;
    s += double(ri) * double(ri);
  }
  const double r0 = sqrt(wg.sum(s));
  double res = r0, rho_old = 1.0;
  int it = 0;
  while (it < max_it && res > rel_tol * r0) {
    s = 0;
    for (int64_t i = tid; i < n; i += nt) {
      const T zi = pinv[i] * r[i];
      z[i] = zi;
      s += double(zi) * double(r[i]);
    }
    const double rho = wg.sum(s);
    // first iteration: p = z exactly (p holds stale memory, 0 * NaN would leak)
    const T beta = it == 0 ? T(0) : static_cast<T>(rho / rho_old);
    for (int64_t i = tid; i < n; i += nt) p[i] = it == 0 ? z[i] : z[i] + beta * p[i];
    __syncthreads();
    apply(p, q);
    s = 0;
    for (int64_t i = tid; i < n; i += nt) s += double(p[i]) * double(q[i]);
    const double pq = wg.sum(s);
    const T alpha = pq != 0.0 ? static_cast<T>(rho / pq) : T(0);
    s = 0;
    for (int64_t i = tid; i < n; i += nt) {
      x[i] += alpha * p[i];
      const T ri = r[i] - alpha * q[i];
      r[i] = ri;
      s += double(ri) * double(ri);
    }
    res = sqrt(wg.sum(s));
    rho_old = rho;
    ++it;
    if (pq == 0.0) break;
  }
  if (tid == 0) {
    info[0] = it;
    info[1] = r0;
    info[2] = res;
  }
}

// The same Jacobi-PCG with the vectors the operator touches held in LDS (the
// search direction p it gathers from and q = A p it scatters into, with LDS
// atomics) and the others in registers (lane entries i = tid + j 256, j <
// NPL): x, r, z and the inverse diagonal.  An iteration then goes to memory
// only for G and the map (L2-resident), instead of a global round trip per
// vector access.  Same arithmetic in the same order as coarse_pcg_kernel,
// with fewer barriers per iteration (the loop is latency-bound): the norm of
// the updated residual and the next rho = z.r are summed in one reduction,
// q is cleared while p is written, and a constrained row's q_i = p_i is read
// through a per-lane mask instead of being copied into q.  Requires n <= 256
// NPL.
template <int NM, int NQ, typename T, int NPL>
__global__ void __launch_bounds__(256)
coarse_pcg_lds_kernel(const T *__restrict__ G, const uint32_t *__restrict__ map,
                      const int64_t n_cells, const Shape1D<T> sh, const uint32_t *__restrict__ cons,
                      const int64_t n_cons, const T *__restrict__ pinv, const int64_t n,
                      const T *__restrict__ b, T *x, const double rel_tol, const int max_it,
                      double *info)
{
  constexpr int E = Bk3Tile<NQ>::cells, LANES = 256;
  __shared__ T lds[E * lds_cell_stride<T>(3 * NQ * NQ * NQ)];
  __shared__ T pv[NPL * LANES], qv[NPL * LANES];
  __shared__ double red[32];
  WgSum wg(red);
  const int tid = threadIdx.x;
  T xr[NPL], rr[NPL], zr[NPL], dr[NPL];
#pragma unroll
  for (int j = 0; j < NPL; ++j) {
    const int64_t i = tid + j * LANES;
    xr[j] = i < n ? x[i] : T(0);
    dr[j] = i < n ? pinv[i] : T(0);
    if (i < n) {
      pv[i] = xr[j];
      qv[i] = T(0);
    }
  }
  __syncthreads();
  // constrained rows: marked in q once, read into a per-lane bit mask
  for (int64_t c = tid; c < n_cons; c += LANES) qv[cons[c]] = T(1);
  __syncthreads();
  uint32_t cmask = 0;
#pragma unroll
  for (int j = 0; j < NPL; ++j) {
    const int64_t i = tid + j * LANES;
    if (i < n && qv[i] != T(0)) cmask |= 1u << j;
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < NPL; ++j)
    if (tid + j * LANES < n) qv[tid + j * LANES] = T(0);
  __syncthreads();

  // A level of at most E cells is one batch: its geometric factors and DoF
  // indices are loaded once into registers instead of once per iteration.
  constexpr int NQ2 = NQ * NQ, NQ3 = NQ2 * NQ, NM2 = NM * NM, NM3 = NM2 * NM;
  const bool one = n_cells <= E;
  // wave-aligned tiles (whole cells per wave) when the level's cells fit
  // them (2^3 cells at nq = 5: 2 per wave): the sweeps then sync per wave
  // instead of per workgroup (a tile that straddles waves needs barriers)
  constexpr int CW = Bk3Tile<NQ>::per_wave;
  const bool wt = (64 % NQ2 != 0) && n_cells <= (LANES / 64) * CW;
  const int wl = tid & 63, wc = wl / NQ2;
  const int e1 = wt ? (tid >> 6) * CW + wc : tid / NQ2;
  const int t1 = wt ? wl - wc * NQ2 : tid - e1 * NQ2, ta = t1 % NQ, tb = t1 / NQ;
  const bool act1 = one && (!wt || wc < CW) && e1 < E && e1 < n_cells;
  const bool col1 = act1 && ta < NM && tb < NM;
  T *P1 = lds + (act1 ? e1 : 0) * lds_cell_stride<T>(3 * NQ3);
  const T *G1 = G + g_index<T>(NQ, act1 ? e1 : 0, 0, t1);
  T g1[6][NQ];
  uint32_t idx1[NM];
  if (act1) bk3_load_g<NQ, T, kGEarly>(G1, g1);
  if (col1) {
#pragma unroll
    for (int k = 0; k < NM; ++k) idx1[k] = map[e1 * NM3 + k * NM2 + tb * NM + ta];
  }

  // qv += A pv over the unconstrained rows (qv cleared and pv complete on
  // entry); ends with a barrier
  auto apply = [&]() {
    if (one) {
      T u[NM], v[NM];
      if (col1) {
#pragma unroll
        for (int k = 0; k < NM; ++k) u[k] = idx1[k] == kInvalid ? T(0) : pv[idx1[k]];
      }
      if (wt)
        bk3_cell_apply<NM, NQ, T, kGEarly, true>(u, v, g1, G1, sh, P1, P1 + NQ3, P1 + 2 * NQ3, act1,
                                                 ta, tb);
      else
        bk3_cell_apply<NM, NQ, T, kGEarly>(u, v, g1, G1, sh, P1, P1 + NQ3, P1 + 2 * NQ3, act1, ta,
                                           tb);
      if (col1) {
#pragma unroll
        for (int k = 0; k < NM; ++k)
          if (idx1[k] != kInvalid) atomicAdd(qv + idx1[k], v[k]);
      }
      __syncthreads();
    } else {
      for (int64_t base = 0; base < n_cells; base += E) {
        bk3_batch<NM, NQ, T>(pv, qv, G, map, nullptr, nullptr, nullptr, n_cells, sh, base, lds);
        __syncthreads();
      }
    }
  };
  // entry i of q = A p (constrained rows: the identity)
  auto qval = [&](int j, int64_t i) { return (cmask >> j) & 1u ? pv[i] : qv[i]; };

  apply();
  double srr = 0, szr = 0;
#pragma unroll
  for (int j = 0; j < NPL; ++j) {
    const int64_t i = tid + j * LANES;
    const T ri = i < n ? b[i] - qval(j, i) : T(0);
    rr[j] = ri;
    srr += double(ri) * double(ri);
  }
#pragma unroll
  for (int j = 0; j < NPL; ++j) {
    zr[j] = dr[j] * rr[j];
    szr += double(zr[j]) * double(rr[j]);
  }
  wg.sum2(srr, szr);
  const double r0 = sqrt(srr);
  double res = r0, rho = szr, rho_old = 1.0;
  int it = 0;
  while (it < max_it && res > rel_tol * r0) {
    // first iteration: p = z exactly
    const T beta = it == 0 ? T(0) : static_cast<T>(rho / rho_old);
#pragma unroll
    for (int j = 0; j < NPL; ++j) {
      const int64_t i = tid + j * LANES;
      if (i < n) {
        pv[i] = it == 0 ? zr[j] : zr[j] + beta * pv[i];
        qv[i] = T(0);
      }
    }
    __syncthreads();
    apply();
    double s = 0;
#pragma unroll
    for (int j = 0; j < NPL; ++j) {
      const int64_t i = tid + j * LANES;
      if (i < n) s += double(pv[i]) * double(qval(j, i));
    }
    const double pq = wg.sum(s);
    const T alpha = pq != 0.0 ? static_cast<T>(rho / pq) : T(0);
    srr = 0;
    szr = 0;
#pragma unroll
    for (int j = 0; j < NPL; ++j) {
      const int64_t i = tid + j * LANES;
      if (i < n) {
        xr[j] += alpha * pv[i];
        rr[j] -= alpha * qval(j, i);
      }
      srr += double(rr[j]) * double(rr[j]);
    }
#pragma unroll
    for (int j = 0; j < NPL; ++j) {
      zr[j] = dr[j] * rr[j];
      szr += double(zr[j]) * double(rr[j]);
    }
    wg.sum2(srr, szr);
    res = sqrt(srr);
    rho_old = rho;
    rho = szr;
    ++it;
    if (pq == 0.0) break;
  }
#pragma unroll
  for (int j = 0; j < NPL; ++j)
    if (tid + j * LANES < n) x[tid + j * LANES] = xr[j];
  if (tid == 0) {
    info[0] = it;
    info[1] = r0;
    info[2] = res;
  }
}

template <int NM, int NQ, typename T, int NPL>
static bool coarse_pcg_lds(int64_t n_cells, const T *G, const uint32_t *map, const Shape1D<T> &sh,
                           const uint32_t *cons, int64_t n_cons, const T *pinv, int64_t n,
                           const T *b, T *x, double rel_tol, int max_it, double *info,
                           hipStream_t s) {
  constexpr size_t lds = sizeof(T) * (Bk3Tile<NQ>::cells * lds_cell_stride<T>(3 * NQ * NQ * NQ) +
                                      2 * NPL * 256) + 32 * sizeof(double);
  if (n > 256 * NPL || lds > 64 * 1024) return false;
  hipLaunchKernelGGL((coarse_pcg_lds_kernel<NM, NQ, T, NPL>), dim3(1), dim3(256), 0, s, G, map,
                     n_cells, sh, cons, n_cons, pinv, n, b, x, rel_tol, max_it, info);
  PMG_HIP_CHECK(hipGetLastError());
  return true;
}

template <typename T>
void launch_bk3_coarse_pcg(int degree, int nq, int64_t n_cells, const T *G, const uint32_t *map,
                           const Shape1D<T> &sh, const uint32_t *cons, int64_t n_cons,
                           const T *pinv, int64_t n, const T *b, T *x, T *r, T *z, T *p, T *q,
                           double rel_tol, int max_it, double *info, hipStream_t s) {
#define PMG_CPCG_CASE(P, Q)                                                                     \
  if (degree == P && nq == Q) {                                                                 \
    if (coarse_pcg_lds<P + 1, Q, T, 1>(n_cells, G, map, sh, cons, n_cons, pinv, n, b, x, rel_tol, \
                                       max_it, info, s) ||                                      \
        coarse_pcg_lds<P + 1, Q, T, 4>(n_cells, G, map, sh, cons, n_cons, pinv, n, b, x, rel_tol, \
                                       max_it, info, s) ||                                      \
        coarse_pcg_lds<P + 1, Q, T, 8>(n_cells, G, map, sh, cons, n_cons, pinv, n, b, x, rel_tol, \
                                       max_it, info, s))                                        \
      return;                                                                                   \
    hipLaunchKernelGGL((coarse_pcg_kernel<P + 1, Q, T>), dim3(1), dim3(Bk3Tile<Q>::lanes), 0, s, \
                       G, map, n_cells, sh, cons, n_cons, pinv, n, b, x, r, z, p, q, rel_tol,    \
                       max_it, info);                                                           \
    PMG_HIP_CHECK(hipGetLastError());                                                           \
    return;                                                                                     \
  }
  PMG_CPCG_CASE(1, 2) PMG_CPCG_CASE(1, 3) PMG_CPCG_CASE(2, 3) PMG_CPCG_CASE(2, 4)
  PMG_CPCG_CASE(3, 4) PMG_CPCG_CASE(3, 5) PMG_CPCG_CASE(4, 5) PMG_CPCG_CASE(4, 6)
  PMG_CPCG_CASE(5, 6) PMG_CPCG_CASE(5, 7) PMG_CPCG_CASE(6, 7) PMG_CPCG_CASE(6, 8)
#undef PMG_CPCG_CASE
  throw Error(PMG_ERR_UNSUPPORTED, "coarse PCG kernel: unsupported (degree, n_q_points_1d)");
}
template void launch_bk3_coarse_pcg<double>(int, int, int64_t, const double *, const uint32_t *,
                                            const Shape1D<double> &, const uint32_t *, int64_t,
                                            const double *, int64_t, const double *, double *,
                                            double *, double *, double *, double *, double,
                                            int, double *, hipStream_t);
template void launch_bk3_coarse_pcg<float>(int, int, int64_t, const float *, const uint32_t *,
                                           const Shape1D<float> &, const uint32_t *, int64_t,
                                           const float *, int64_t, const float *, float *, float *,
                                           float *, float *, float *, double, int, double *,
                                           hipStream_t);

// ---------------------------------------------------------------- gather schedule
// Second pass of the gather schedule: y[i] (=, or += if !assign) sum of
// buf[list[k]] over k in [off[i], off[i+1]); DoFs without entries (the
// constrained ones) are left untouched.  off and list are read once.
template <typename T>
__global__ void __launch_bounds__(256)
bk3_gather_kernel(const int64_t n, const uint32_t *__restrict__ off,
                  const uint32_t *__restrict__ list, const T *__restrict__ buf, T *__restrict__ y,
                  const int assign)
{
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t k0 = __builtin_nontemporal_load(off + i);
  const uint32_t k1 = __builtin_nontemporal_load(off + i + 1);
  if (k0 == k1) return;
  T acc = T(0);
  for (uint32_t k = k0; k < k1; ++k) acc += buf[__builtin_nontemporal_load(list + k)];
  y[i] = assign ? acc : y[i] + acc;
}

template <typename T>
static void launch_bk3_gather(int64_t n, const uint32_t *off, const uint32_t *list, const T *buf,
                              T *y, bool assign, hipStream_t s) {
  if (n <= 0) return;
  const int64_t nb = (n + 255) / 256;
  PMG_REQUIRE(nb < (int64_t(1) << 31), PMG_ERR_SIZE, "too many DoFs for one launch");
  hipLaunchKernelGGL(bk3_gather_kernel<T>, dim3(static_cast<unsigned>(nb)), dim3(256), 0, s, n,
                     off, list, buf, y, assign ? 1 : 0);
  PMG_HIP_CHECK(hipGetLastError());
}

// Second pass of the grid schedule on a lexicographic box of nx x ny x nz
// cells (degree p, nm = p+1): the first pass stored every cell's values to an
// exploded grid of nx nm x ny nm x nz nm entries, cell (cx,cy,cz) at the
// block starting at (cx nm, cy nm, cz nm).  Along each axis a DoF of grid
// index I = q p + r has one copy at q nm + r, or two adjacent ones, q nm - 1
// and q nm, when it lies on a face between two cells (r = 0, 0 < q < n), so
// y[I,J,K] is the sum of at most 2 x 2 x 2 neighbouring entries: every access
// coalesces, nothing is atomic, and the summation order is fixed.
// Constrained rows (cmask != 0) are left untouched.
template <int P>
__device__ __forceinline__ int grid_copies(int I, int n, int (&e)[2]) {
  constexpr int NM = P + 1;
  const int q = I / P, r = I - q * P;  // division by a constant
  if (q == n) {
    e[0] = q * NM - 1;
    return 1;
  }
  if (r == 0 && q > 0) {
    e[0] = q * NM - 1;
    e[1] = q * NM;
    return 2;
  }
  e[0] = q * NM + r;
  return 1;
}

// Each lane sums DPL DoFs 256 apart, with every load issued before the first
// sum (one memory latency per lane instead of DPL).  Constrained DoFs are
// summed too: their buffer entries are zero (the first pass stores zero for
// invalid map entries), and the caller restores the constrained rows.
template <int P, typename T, int DPL>
__global__ void __launch_bounds__(256)
bk3_grid_sum_kernel(const int nx, const int ny, const int nz, const T *__restrict__ buf,
                    T *__restrict__ y, const int assign)
{
  constexpr int NM = P + 1;
  const int n1x = P * nx + 1, n1y = P * ny + 1;
  const int64_t n = int64_t(n1x) * n1y * (P * nz + 1);
  const int64_t row = int64_t(nx) * NM, plane = row * ny * NM;
  const int base = static_cast<int>(blockIdx.x) * (256 * DPL) + threadIdx.x;
  T v[DPL][8];
  int cnt[DPL];
#pragma unroll
  for (int d = 0; d < DPL; ++d) {
    const int i = base + d * 256;
    cnt[d] = 0;
    if (i < n) {
      const int jk = i / n1x, I = i - jk * n1x;
      const int K = jk / n1y, J = jk - K * n1y;
      int ex[2], ey[2], ez[2];
      const int cx = grid_copies<P>(I, nx, ex), cy = grid_copies<P>(J, ny, ey),
                cz = grid_copies<P>(K, nz, ez);
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
          for (int c = 0; c < 2; ++c)
            if (a < cz && b < cy && c < cx)
              v[d][cnt[d]++] = buf[ez[a] * plane + ey[b] * row + ex[c]];
    }
  }
#pragma unroll
  for (int d = 0; d < DPL; ++d) {
    const int i = base + d * 256;
    if (i < n) {
      T acc = T(0);
#pragma unroll
      for (int k = 0; k < 8; ++k)
        if (k < cnt[d]) acc += v[d][k];
      y[i] = assign ? acc : y[i] + acc;
    }
  }
}

template <typename T>
static void launch_bk3_grid_sum(int p, const int *dims, const T *buf, T *y, bool assign,
                                hipStream_t s) {
  constexpr int DPL = 2;
  const int64_t n = (int64_t(p) * dims[0] + 1) * (int64_t(p) * dims[1] + 1) *
                    (int64_t(p) * dims[2] + 1);
  if (n <= 0) return;
  const int64_t nb = (n + 256 * DPL - 1) / (256 * DPL);
  PMG_REQUIRE(n + 256 * DPL < (int64_t(1) << 31), PMG_ERR_SIZE, "box too large for the grid schedule");
#define PMG_GRID_SUM_CASE(P)                                                                  \
  if (p == P) {                                                                               \
    hipLaunchKernelGGL((bk3_grid_sum_kernel<P, T, DPL>), dim3(static_cast<unsigned>(nb)),     \
                       dim3(256), 0, s, dims[0], dims[1], dims[2], buf, y, assign ? 1 : 0);   \
    PMG_HIP_CHECK(hipGetLastError());                                                         \
    return;                                                                                   \
  }
  PMG_GRID_SUM_CASE(1) PMG_GRID_SUM_CASE(2) PMG_GRID_SUM_CASE(3) PMG_GRID_SUM_CASE(4)
  PMG_GRID_SUM_CASE(5) PMG_GRID_SUM_CASE(6)
#undef PMG_GRID_SUM_CASE
  throw Error(PMG_ERR_UNSUPPORTED, "grid schedule: unsupported degree");
}

// 1 where the map is not the lexicographic box map: entry (cell, k, j, i) of
// cell (cx, cy, cz) must be the grid DoF (p cx + i, p cy + j, p cz + k) or,
// if that DoF is constrained (cmask, may be null), that DoF or invalid.
__global__ void __launch_bounds__(256)
box_map_check_kernel(const int p, const int nx, const int ny, const int64_t n_pos,
                     const uint32_t *__restrict__ map, const uint8_t *__restrict__ cmask,
                     uint32_t *bad)
{
  const int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= n_pos) return;
  const int nm = p + 1, nm3 = nm * nm * nm;
  const int64_t c = e / nm3;
  const int l = static_cast<int>(e - c * nm3), i = l % nm, j = (l / nm) % nm, k = l / (nm * nm);
  const int64_t cyz = c / nx, cx = c - cyz * nx, cz = cyz / ny, cy = cyz - cz * ny;
  const int64_t n1x = int64_t(p) * nx + 1, n1y = int64_t(p) * ny + 1;
  const int64_t d = (p * cx + i) + n1x * ((p * cy + j) + n1y * (p * cz + k));
  const uint32_t m = map[e];
  const bool ok = m == kInvalid ? (cmask && cmask[d]) : int64_t(m) == d;
  if (!ok) atomicOr(bad, 1u);
}

bool check_box_map(int degree, const int *dims, int64_t n_cells, const uint32_t *map,
                   const uint8_t *cmask, hipStream_t s) {
  const int64_t nm = degree + 1, n_pos = n_cells * nm * nm * nm;
  if (n_pos == 0) return true;
  DeviceArray<uint32_t> bad(1);
  PMG_HIP_CHECK(hipMemsetAsync(bad.ptr, 0, sizeof(uint32_t), s));
  hipLaunchKernelGGL(box_map_check_kernel, dim3(static_cast<unsigned>((n_pos + 255) / 256)),
                     dim3(256), 0, s, degree, dims[0], dims[1], n_pos, map, cmask, bad.ptr);
  PMG_HIP_CHECK(hipGetLastError());
  uint32_t h = 1;
  PMG_HIP_CHECK(hipMemcpyAsync(&h, bad.ptr, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  PMG_HIP_CHECK(hipStreamSynchronize(s));
  return h == 0;
}

__global__ void __launch_bounds__(256)
imap_count_kernel(const int64_t n_pos, const uint32_t *__restrict__ map, uint32_t *cnt) {
  const int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= n_pos) return;
  const uint32_t d = map[e];
  if (d != kInvalid) atomicAdd(cnt + d, 1u);
}
__global__ void __launch_bounds__(256)
imap_fill_kernel(const int64_t n_pos, const uint32_t *__restrict__ map, uint32_t *cursor,
                 uint32_t *__restrict__ list) {
  const int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= n_pos) return;
  const uint32_t d = map[e];
  if (d != kInvalid) list[atomicAdd(cursor + d, 1u)] = static_cast<uint32_t>(e);
}
// each DoF's entries in ascending order (insertion sort; hex meshes have <= 8)
__global__ void __launch_bounds__(256)
imap_sort_kernel(const int64_t n, const uint32_t *__restrict__ off, uint32_t *__restrict__ list) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t k0 = off[i], k1 = off[i + 1];
  for (uint32_t k = k0 + 1; k < k1; ++k) {
    const uint32_t v = list[k];
    uint32_t j = k;
    while (j > k0 && list[j - 1] > v) {
      list[j] = list[j - 1];
      --j;
    }
    list[j] = v;
  }
}

void build_inverse_map(int64_t n_pos, const uint32_t *map, int64_t n_dofs, uint32_t *off,
                       uint32_t **list_out, int64_t *n_entries, hipStream_t s) {
  PMG_REQUIRE(n_pos < (int64_t(1) << 32) && n_dofs < (int64_t(1) << 31), PMG_ERR_SIZE,
              "inverse map: too many entries for 32-bit positions");
  *list_out = nullptr;
  *n_entries = 0;
  DeviceArray<uint32_t> cnt(static_cast<size_t>(n_dofs + 1));
  PMG_HIP_CHECK(hipMemsetAsync(cnt.ptr, 0, cnt.bytes(), s));
  const unsigned gp = static_cast<unsigned>((n_pos + 255) / 256);
  const unsigned gd = static_cast<unsigned>((n_dofs + 255) / 256);
  if (n_pos > 0) hipLaunchKernelGGL(imap_count_kernel, dim3(gp), dim3(256), 0, s, n_pos, map, cnt.ptr);
  PMG_HIP_CHECK(hipGetLastError());
  size_t tmp_bytes = 0;
  PMG_HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp_bytes, cnt.ptr, off,
                                                 static_cast<int>(n_dofs + 1), s));
  DeviceArray<char> tmp(std::max<size_t>(tmp_bytes, 1));
  PMG_HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(tmp.ptr, tmp_bytes, cnt.ptr, off,
                                                 static_cast<int>(n_dofs + 1), s));
  uint32_t total = 0;
  PMG_HIP_CHECK(hipMemcpyAsync(&total, off + n_dofs, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  PMG_HIP_CHECK(hipStreamSynchronize(s));
  uint32_t *list = nullptr;
  if (total > 0) {
    const hipError_t e = hipMalloc(reinterpret_cast<void **>(&list), size_t(total) * sizeof(uint32_t));
    if (e != hipSuccess) throw Error(PMG_ERR_ALLOC, "hipMalloc of the inverse map failed");
    PMG_HIP_CHECK(hipMemcpyAsync(cnt.ptr, off, size_t(n_dofs) * sizeof(uint32_t),
                                 hipMemcpyDeviceToDevice, s));  // cursors
    hipLaunchKernelGGL(imap_fill_kernel, dim3(gp), dim3(256), 0, s, n_pos, map, cnt.ptr, list);
    PMG_HIP_CHECK(hipGetLastError());
    if (n_dofs > 0) hipLaunchKernelGGL(imap_sort_kernel, dim3(gd), dim3(256), 0, s, n_dofs, off, list);
    PMG_HIP_CHECK(hipGetLastError());
    PMG_HIP_CHECK(hipStreamSynchronize(s));
  }
  *list_out = list;
  *n_entries = total;
}

template <int NM, int NQ, typename T, int SCATTER, int GMODE>
static void bk3_launch_range(const T *x, T *y, const T *G, const uint32_t *map,
                             const uint32_t *lines, const uint8_t *lflag,
                             const uint32_t *cell_ids, int64_t n_range, const Shape1D<T> &sh,
                             hipStream_t s, int gnx = 0, int gny = 0) {
  if (n_range <= 0) return;
  using Tile = Bk3Tile<NQ>;
  const int64_t nb = (n_range + Tile::cells - 1) / Tile::cells;
  PMG_REQUIRE(nb < (int64_t(1) << 31), PMG_ERR_SIZE, "too many cells for one launch");
  const unsigned grid = pad_to_xcd(static_cast<unsigned>(nb));
  hipLaunchKernelGGL((bk3_cells_kernel<NM, NQ, T, SCATTER, GMODE>), dim3(grid),
                     dim3(bk3_cells_lanes<NQ, T>()), 0, s, x, y, G, map, lines, lflag, cell_ids,
                     n_range, sh, gnx, gny);
  PMG_HIP_CHECK(hipGetLastError());
}

// FULL instantiates every G-load policy (used for the production degrees).
template <int NM, int NQ, typename T, bool FULL>
static void bk3_launch(const T *x, T *y, const T *G, const uint32_t *map, int64_t n_cells,
                       const Shape1D<T> &sh, hipStream_t s, const Bk3Schedule *sc, bool assign) {
  if (sc && sc->kind == kPatch && sc->pdof) {
    if (sc->gaff)
      bk3_launch_patch<NM, NQ, T, kGAffine>(assign, x, y, static_cast<const T *>(sc->gaff), *sc,
                                            n_cells, sh, s);
    else if (sc->gmode == kGEarlyNT)
      bk3_launch_patch<NM, NQ, T, kGEarlyNT>(assign, x, y, G, *sc, n_cells, sh, s);
    else if (FULL && sc->gmode == kGEarly)
      bk3_launch_patch<NM, NQ, T, (FULL ? kGEarly : kGLateNT)>(assign, x, y, G, *sc, n_cells, sh, s);
    else
      bk3_launch_patch<NM, NQ, T, kGLateNT>(assign, x, y, G, *sc, n_cells, sh, s);
    return;
  }
  if (sc && sc->kind == kGrid && sc->gbuf) {
    // cells -> exploded grid (plain stores), then every DoF sums its at most
    // 8 adjacent copies: no atomics, no index loads in the second pass
    T *buf = static_cast<T *>(sc->gbuf);
    if (FULL && sc->gmode == kGEarlyNT)
      bk3_launch_range<NM, NQ, T, kScatterGrid, kGEarlyNT>(x, buf, G, map, sc->lines, sc->lflag,
                                                          nullptr, n_cells, sh, s, sc->grid[0],
                                                          sc->grid[1]);
    else
      bk3_launch_range<NM, NQ, T, kScatterGrid, kGLateNT>(x, buf, G, map, sc->lines, sc->lflag,
                                                         nullptr, n_cells, sh, s, sc->grid[0],
                                                         sc->grid[1]);
    launch_bk3_grid_sum<T>(NM - 1, sc->grid, buf, y, assign, s);
    return;
  }
  if (sc && sc->kind == kGather && sc->gbuf) {
    // cells -> per-cell buffer (plain coalesced stores), then every DoF sums
    // its entries: no atomics on the output, and no zero fill before a vmult
    T *buf = static_cast<T *>(sc->gbuf);
    if (FULL && sc->gmode == kGEarlyNT)
      bk3_launch_range<NM, NQ, T, kScatterStore, kGEarlyNT>(x, buf, G, map, sc->lines, sc->lflag,
                                                           nullptr, n_cells, sh, s);
    else
      bk3_launch_range<NM, NQ, T, kScatterStore, kGLateNT>(x, buf, G, map, sc->lines, sc->lflag,
                                                          nullptr, n_cells, sh, s);
    launch_bk3_gather<T>(sc->gn, sc->goff, sc->glist, buf, y, assign, s);
    return;
  }
  if (sc && sc->kind == kColored) {
    // one launch per colour, no atomics; launches are stream-ordered, so a
    // colour sees the updates of the previous ones
    for (size_t c = 0; c + 1 < sc->color_offsets.size(); ++c) {
      const int64_t b = sc->color_offsets[c], n = sc->color_offsets[c + 1] - b;
      bk3_launch_range<NM, NQ, T, kScatterPlain, kGEarlyNT>(x, y, G, map, sc->lines, sc->lflag,
                                                    sc->cell_ids + b, n, sh, s);
    }
    return;
  }
  const int gmode = sc ? sc->gmode : kGLateNT;
  const uint32_t *lines = sc ? sc->lines : nullptr;
  const uint8_t *lflag = sc ? sc->lflag : nullptr;
  if (sc && sc->gaff)
    bk3_launch_range<NM, NQ, T, kScatterAtomic, kGAffine>(x, y, static_cast<const T *>(sc->gaff),
                                                          map, lines, lflag, nullptr, n_cells, sh, s);
  else if (FULL && gmode == kGEarly)
    bk3_launch_range<NM, NQ, T, kScatterAtomic, (FULL ? kGEarly : kGLateNT)>(x, y, G, map, lines, lflag,
                                                                   nullptr, n_cells, sh, s);
  else if (FULL && gmode == kGEarlyNT)
    bk3_launch_range<NM, NQ, T, kScatterAtomic, kGEarlyNT>(x, y, G, map, lines, lflag, nullptr, n_cells, sh, s);
  else
    bk3_launch_range<NM, NQ, T, kScatterAtomic, kGLateNT>(x, y, G, map, lines, lflag, nullptr, n_cells, sh, s);
}

template <typename T>
void launch_bk3_cells(int degree, int nq, const T *x, T *y, const T *G, const uint32_t *map,
                      int64_t n_cells, const Shape1D<T> &sh, hipStream_t s,
                      const Bk3Schedule *sched, bool assign) {
#define PMG_BK3_CASE(P, Q, FULL)                                                  \
  if (degree == P && nq == Q) {                                                   \
    bk3_launch<P + 1, Q, T, FULL>(x, y, G, map, n_cells, sh, s, sched, assign);   \
    return;                                                                       \
  }
  PMG_BK3_CASE(1, 2, false) PMG_BK3_CASE(1, 3, false)
  PMG_BK3_CASE(2, 3, false) PMG_BK3_CASE(2, 4, false)
  PMG_BK3_CASE(3, 4, false) PMG_BK3_CASE(3, 5, false)
  PMG_BK3_CASE(4, 5, true)  PMG_BK3_CASE(4, 6, true)
  PMG_BK3_CASE(5, 6, false) PMG_BK3_CASE(5, 7, false)
  PMG_BK3_CASE(6, 7, false) PMG_BK3_CASE(6, 8, false)
#undef PMG_BK3_CASE
  throw Error(PMG_ERR_UNSUPPORTED, "BK3 kernel: unsupported (degree, n_q_points_1d) = (" +
                                       std::to_string(degree) + ", " + std::to_string(nq) + ")");
}

template <typename T>
void launch_bk3_cell_list(int degree, int nq, const T *x, T *y, const T *G, const uint32_t *map,
                          const uint32_t *lines, const uint8_t *lflag, const uint32_t *cell_ids,
                          int64_t n_list, const Shape1D<T> &sh, hipStream_t s) {
#define PMG_BK3_CASE(P, Q)                                                                  \
  if (degree == P && nq == Q) {                                                             \
    bk3_launch_range<P + 1, Q, T, kScatterAtomic, kGLateNT>(x, y, G, map, lines, lflag, cell_ids,     \
                                                  n_list, sh, s);                           \
    return;                                                                                 \
  }
  PMG_BK3_CASE(1, 2) PMG_BK3_CASE(1, 3) PMG_BK3_CASE(2, 3) PMG_BK3_CASE(2, 4)
  PMG_BK3_CASE(3, 4) PMG_BK3_CASE(3, 5) PMG_BK3_CASE(4, 5) PMG_BK3_CASE(4, 6)
  PMG_BK3_CASE(5, 6) PMG_BK3_CASE(5, 7) PMG_BK3_CASE(6, 7) PMG_BK3_CASE(6, 8)
#undef PMG_BK3_CASE
  throw Error(PMG_ERR_UNSUPPORTED, "BK3 kernel: unsupported (degree, n_q_points_1d) = (" +
                                       std::to_string(degree) + ", " + std::to_string(nq) + ")");
}
template void launch_bk3_cell_list<double>(int, int, const double *, double *, const double *,
                                           const uint32_t *, const uint32_t *, const uint8_t *,
                                           const uint32_t *, int64_t, const Shape1D<double> &,
                                           hipStream_t);
template void launch_bk3_cell_list<float>(int, int, const float *, float *, const float *,
                                          const uint32_t *, const uint32_t *, const uint8_t *,
                                          const uint32_t *, int64_t, const Shape1D<float> &,
                                          hipStream_t);

template void launch_bk3_cells<double>(int, int, const double *, double *, const double *,
                                       const uint32_t *, int64_t, const Shape1D<double> &,
                                       hipStream_t, const Bk3Schedule *, bool);
template void launch_bk3_cells<float>(int, int, const float *, float *, const float *,
                                      const uint32_t *, int64_t, const Shape1D<float> &,
                                      hipStream_t, const Bk3Schedule *, bool);

// ---------------------------------------------------------------- geometry
// One lane per (cell, q): trilinear Jacobian J = dX/dxi at the Gauss point,
// G = kappa * |det J| w * J^{-1} J^{-T}  (compute_G_tensors,
// portable_laplace_operator_bk3.h:484-511), stored [c][comp][q].
template <typename T>
__global__ void __launch_bounds__(256)
geometry_kernel(const int nq, const int64_t n_cells, const double *__restrict__ cv,
                const double *__restrict__ qp, const double *__restrict__ qw,
                const double *__restrict__ kappa, const uint32_t *__restrict__ perm,
                T *__restrict__ G)
{
  const int nq3 = nq * nq * nq;
  const int64_t gid = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (gid >= n_cells * nq3) return;
  const int64_t c = gid / nq3;
  const int q = static_cast<int>(gid - c * nq3);
  const int q0 = q % nq, q1 = (q / nq) % nq, q2 = q / (nq * nq);
  const double xi[3] = {qp[q0], qp[q1], qp[q2]};
  double J[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
  const double *v = cv + c * 24;
  for (int vi = 0; vi < 8; ++vi) {
    const int b0 = vi & 1, b1 = (vi >> 1) & 1, b2 = (vi >> 2) & 1;
    const double f0 = b0 ? xi[0] : 1.0 - xi[0], f1 = b1 ? xi[1] : 1.0 - xi[1],
                 f2 = b2 ? xi[2] : 1.0 - xi[2];
    const double d0 = b0 ? 1.0 : -1.0, d1 = b1 ? 1.0 : -1.0, d2 = b2 ? 1.0 : -1.0;
    const double dN[3] = {d0 * f1 * f2, f0 * d1 * f2, f0 * f1 * d2};
    for (int d = 0; d < 3; ++d)
      for (int k = 0; k < 3; ++k) J[d][k] += dN[k] * v[3 * vi + d];
  }
  // cofactors: inv(J)[a][b] = cof(J)[b][a] / det
  const double c00 = J[1][1] * J[2][2] - J[1][2] * J[2][1];
  const double c01 = J[1][2] * J[2][0] - J[1][0] * J[2][2];
  const double c02 = J[1][0] * J[2][1] - J[1][1] * J[2][0];
  const double c10 = J[0][2] * J[2][1] - J[0][1] * J[2][2];
  const double c11 = J[0][0] * J[2][2] - J[0][2] * J[2][0];
  const double c12 = J[0][1] * J[2][0] - J[0][0] * J[2][1];
  const double c20 = J[0][1] * J[1][2] - J[0][2] * J[1][1];
  const double c21 = J[0][2] * J[1][0] - J[0][0] * J[1][2];
  const double c22 = J[0][0] * J[1][1] - J[0][1] * J[1][0];
  const double det = J[0][0] * c00 + J[0][1] * c01 + J[0][2] * c02;
  const double id = 1.0 / det;
  // K[a][k] = d xi_a / d x_k = cof[k][a] / det
  const double K[3][3] = {{c00 * id, c10 * id, c20 * id},
                          {c01 * id, c11 * id, c21 * id},
                          {c02 * id, c12 * id, c22 * id}};
  const double kap = kappa ? kappa[(perm ? static_cast<int64_t>(perm[c]) : c) * nq3 + q] : 1.0;
  const double s = fabs(det) * qw[q0] * qw[q1] * qw[q2] * kap;
  int comp = 0;
  for (int a = 0; a < 3; ++a)
    for (int b = a; b < 3; ++b, ++comp)
      G[g_index<T>(nq, c, comp, q)] =
          static_cast<T>(s * (K[a][0] * K[b][0] + K[a][1] * K[b][1] + K[a][2] * K[b][2]));
}

template <typename T>
void launch_geometry(int nq, int64_t n_cells, const double *cell_vertices, const double *qpts,
                     const double *qwts, const double *kappa, const uint32_t *perm, T *G,
                     hipStream_t s) {
  const int64_t n = n_cells * nq * nq * nq;
  if (n == 0) return;
  const int64_t nb = (n + 255) / 256;
  hipLaunchKernelGGL(geometry_kernel<T>, dim3(static_cast<unsigned>(nb)), dim3(256), 0, s, nq,
                     n_cells, cell_vertices, qpts, qwts, kappa, perm, G);
  PMG_HIP_CHECK(hipGetLastError());
}
template void launch_geometry<double>(int, int64_t, const double *, const double *,
                                      const double *, const double *, const uint32_t *, double *,
                                      hipStream_t);
template void launch_geometry<float>(int, int64_t, const double *, const double *,
                                     const double *, const double *, const uint32_t *, float *,
                                     hipStream_t);

// Affine cells: the trilinear map of vertices v_0..v_7 (v_i at the corner
// (i & 1, i >> 1 & 1, i >> 2 & 1)) is affine when its twist terms vanish,
// and J is then the edge vectors v_1 - v_0, v_2 - v_0, v_4 - v_0 everywhere;
// A = |det J| J^-1 J^-T (the G of geometry_kernel without the weights).
template <typename T>
__global__ void affine_geometry_kernel(const int64_t n_cells, const double *__restrict__ cv,
                                       T *__restrict__ A, unsigned long long *__restrict__ bad)
{
  const int64_t c = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (c >= n_cells) return;
  const double *v = cv + c * 24;
  auto at = [&](int i, int d) { return v[3 * i + d]; };
  double J[3][3], tw = 0, h = 0;
  for (int d = 0; d < 3; ++d) {
    J[d][0] = at(1, d) - at(0, d);
    J[d][1] = at(2, d) - at(0, d);
    J[d][2] = at(4, d) - at(0, d);
    tw = fmax(tw, fabs(at(0, d) - at(1, d) - at(2, d) + at(3, d)));
    tw = fmax(tw, fabs(at(0, d) - at(1, d) - at(4, d) + at(5, d)));
    tw = fmax(tw, fabs(at(0, d) - at(2, d) - at(4, d) + at(6, d)));
    tw = fmax(tw, fabs(-at(0, d) + at(1, d) + at(2, d) - at(3, d) + at(4, d) - at(5, d) - at(6, d) +
                       at(7, d)));
  }
  for (int k = 0; k < 3; ++k)
    h = fmax(h, sqrt(J[0][k] * J[0][k] + J[1][k] * J[1][k] + J[2][k] * J[2][k]));
  if (!(tw <= 1e-10 * h)) atomicAdd(bad, 1ull);
  const double c00 = J[1][1] * J[2][2] - J[1][2] * J[2][1];
  const double c01 = J[1][2] * J[2][0] - J[1][0] * J[2][2];
  const double c02 = J[1][0] * J[2][1] - J[1][1] * J[2][0];
  const double c10 = J[0][2] * J[2][1] - J[0][1] * J[2][2];
  const double c11 = J[0][0] * J[2][2] - J[0][2] * J[2][0];
  const double c12 = J[0][1] * J[2][0] - J[0][0] * J[2][1];
  const double c20 = J[0][1] * J[1][2] - J[0][2] * J[1][1];
  const double c21 = J[0][2] * J[1][0] - J[0][0] * J[1][2];
  const double c22 = J[0][0] * J[1][1] - J[0][1] * J[1][0];
  const double det = J[0][0] * c00 + J[0][1] * c01 + J[0][2] * c02;
  const double id = 1.0 / det;
  const double K[3][3] = {{c00 * id, c10 * id, c20 * id},
                          {c01 * id, c11 * id, c21 * id},
                          {c02 * id, c12 * id, c22 * id}};
  int comp = 0;
  for (int a = 0; a < 3; ++a)
    for (int b = a; b < 3; ++b, ++comp)
      A[c * 6 + comp] =
          static_cast<T>(fabs(det) * (K[a][0] * K[b][0] + K[a][1] * K[b][1] + K[a][2] * K[b][2]));
}

template <typename T>
int64_t launch_affine_geometry(int64_t n_cells, const double *cell_vertices, T *A, hipStream_t s) {
  if (n_cells <= 0) return 0;
  DeviceArray<unsigned long long> bad(1);
  PMG_HIP_CHECK(hipMemsetAsync(bad.ptr, 0, sizeof(unsigned long long), s));
  hipLaunchKernelGGL(affine_geometry_kernel<T>, dim3(static_cast<unsigned>((n_cells + 255) / 256)),
                     dim3(256), 0, s, n_cells, cell_vertices, A, bad.ptr);
  PMG_HIP_CHECK(hipGetLastError());
  unsigned long long h = 0;
  PMG_HIP_CHECK(hipMemcpyAsync(&h, bad.ptr, sizeof(h), hipMemcpyDeviceToHost, s));
  PMG_HIP_CHECK(hipStreamSynchronize(s));
  return static_cast<int64_t>(h);
}
template int64_t launch_affine_geometry<double>(int64_t, const double *, double *, hipStream_t);
template int64_t launch_affine_geometry<float>(int64_t, const double *, float *, hipStream_t);

// ---------------------------------------------------------------- diagonal
// (A_e)_ii = sum_q g_i(q)^T G_q g_i(q), g_i(q) = (phi'phi phi, phi phi' phi,
// phi phi phi') from the exact 1D derivatives (sh.D holds Sd here).  Each of
// the six G components contributes a separable sum
//   sum_{q0,q1,q2} G_c(q) X(i0,q0) Y(i1,q1) Z(i2,q2),  X,Y,Z in {S^2, D^2, DS},
// evaluated by sum factorisation (three sweeps of nm nq^3, nm^2 nq^2, nm^3 nq
// products) instead of a quadrature loop per DoF.  One wave per cell, double
// accumulation.  Setup only (compute_diagonal,
// portable_laplace_operator_bk3.h:533-561).
template <typename TG>
__global__ void __launch_bounds__(64)
bk3_diagonal_kernel(const int nm, const int nq, const int64_t n_cells, const TG *__restrict__ G,
                    const uint32_t *__restrict__ map, const Shape1D<double> sh,
                    double *__restrict__ diag)
{
  __shared__ double f[3][kMaxNm * kMaxNq];  // S^2, D^2, DS
  __shared__ double g[kMaxNq * kMaxNq * kMaxNq];
  __shared__ double t1[kMaxNm * kMaxNq * kMaxNq];
  __shared__ double t2[kMaxNm * kMaxNm * kMaxNq];
  __shared__ double acc[kMaxNm * kMaxNm * kMaxNm];
  const int64_t cell = blockIdx.x;
  if (cell >= n_cells) return;  // whole workgroup
  const int lane = threadIdx.x;
  const int nm2 = nm * nm, nm3 = nm2 * nm, nq2 = nq * nq, nq3 = nq2 * nq;
  for (int o = lane; o < nm * nq; o += 64) {
    const double sv = sh.S[o], dv = sh.D[o];
    f[0][o] = sv * sv;
    f[1][o] = dv * dv;
    f[2][o] = dv * sv;
  }
  for (int o = lane; o < nm3; o += 64) acc[o] = 0.0;
  // component c = xx, xy, xz, yy, yz, zz: factors along x, y, z and weight
  constexpr int fx[6] = {1, 2, 2, 0, 0, 0}, fy[6] = {0, 2, 0, 1, 2, 0}, fz[6] = {0, 0, 2, 0, 2, 1};
  constexpr double wc[6] = {1.0, 2.0, 2.0, 1.0, 2.0, 1.0};
  for (int c = 0; c < 6; ++c) {
    __syncthreads();
    for (int o = lane; o < nq3; o += 64) g[o] = static_cast<double>(G[g_index<TG>(nq, cell, c, o)]);
    __syncthreads();
    const double *X = f[fx[c]], *Y = f[fy[c]], *Z = f[fz[c]];
    for (int o = lane; o < nm * nq2; o += 64) {  // t1[(q2 nq + q1) nm + a] = sum_q0 X(a,q0) g
      const int a = o % nm, r = o / nm;
      double v = 0.0;
      for (int q0 = 0; q0 < nq; ++q0) v += X[a * nq + q0] * g[r * nq + q0];
      t1[o] = v;
    }
    __syncthreads();
    for (int o = lane; o < nm2 * nq; o += 64) {  // t2[(q2 nm + b) nm + a] = sum_q1 Y(b,q1) t1
      const int a = o % nm, b = (o / nm) % nm, q2 = o / nm2;
      double v = 0.0;
      for (int q1 = 0; q1 < nq; ++q1) v += Y[b * nq + q1] * t1[(q2 * nq + q1) * nm + a];
      t2[o] = v;
    }
    __syncthreads();
    for (int o = lane; o < nm3; o += 64) {  // acc[(k nm + b) nm + a] += w sum_q2 Z(k,q2) t2
      const int a = o % nm, b = (o / nm) % nm, k = o / nm2;
      double v = 0.0;
      for (int q2 = 0; q2 < nq; ++q2) v += Z[k * nq + q2] * t2[(q2 * nm + b) * nm + a];
      acc[o] += wc[c] * v;
    }
  }
  for (int o = lane; o < nm3; o += 64) {  // the same lane wrote acc[o]
    const uint32_t dof = map[cell * nm3 + o];
    if (dof != kInvalid) atomic_add(diag + dof, acc[o]);
  }
}

template <typename T>
void launch_bk3_diagonal(int degree, int nq, int64_t n_cells, const T *G, const uint32_t *map,
                         const Shape1D<double> &S_and_Sd, double *diag, hipStream_t s) {
  const int nm = degree + 1;
  if (n_cells <= 0) return;
  PMG_REQUIRE(nm <= kMaxNm && nq <= kMaxNq, PMG_ERR_UNSUPPORTED, "diagonal: degree or nq too large");
  PMG_REQUIRE(n_cells < (int64_t(1) << 31), PMG_ERR_SIZE, "too many cells for one launch");
  hipLaunchKernelGGL(bk3_diagonal_kernel<T>, dim3(static_cast<unsigned>(n_cells)), dim3(64), 0, s,
                     nm, nq, n_cells, G, map, S_and_Sd, diag);
  PMG_HIP_CHECK(hipGetLastError());
}
template void launch_bk3_diagonal<double>(int, int, int64_t, const double *, const uint32_t *,
                                          const Shape1D<double> &, double *, hipStream_t);
template void launch_bk3_diagonal<float>(int, int, int64_t, const float *, const uint32_t *,
                                         const Shape1D<double> &, double *, hipStream_t);

// ---------------------------------------------------------------- adjoint gradient
// One workgroup per cell, one lane per quadrature point column (p,q); reference
// gradients of u and lambda by the forward half of the operator (phases 1-5),
// then out = -(grad u)^T G (grad lambda) at every q-point.
template <int NM, int NQ, typename T>
__device__ void cell_gradient(const T *__restrict__ vec, const uint32_t *__restrict__ map,
                              int64_t cell, const Shape1D<T> &sh, T *P0, T *P1, T *P2,
                              int ta, int tb, T (&gz)[NQ])
{
  constexpr int NQ2 = NQ * NQ, NM2 = NM * NM, NM3 = NM2 * NM;
  if (ta < NM && tb < NM) {
    T u[NM], w[NQ];
    for (int k = 0; k < NM; ++k) {
      const uint32_t d = map[cell * NM3 + k * NM2 + tb * NM + ta];
      u[k] = d == kInvalid ? T(0) : vec[d];
    }
    eo_interp<NM, NQ, T>(sh, u, w);
    for (int r = 0; r < NQ; ++r) P0[r * NM2 + tb * NM + ta] = w[r];
  }
  __syncthreads();
  if (ta < NM) {
    T a[NM], w[NQ];
    for (int j = 0; j < NM; ++j) a[j] = P0[tb * NM2 + j * NM + ta];
    eo_interp<NM, NQ, T>(sh, a, w);
    for (int q = 0; q < NQ; ++q) P1[(tb * NQ + q) * NM + ta] = w[q];
  }
  __syncthreads();
  {
    T b[NM], U[NQ], dU[NQ];
    for (int i = 0; i < NM; ++i) b[i] = P1[(tb * NQ + ta) * NM + i];
    eo_interp<NM, NQ, T>(sh, b, U);
    eo_deriv<NQ, T>(sh, U, dU);
    __syncthreads();
    for (int p = 0; p < NQ; ++p) {
      P0[tb * NQ2 + ta * NQ + p] = U[p];
      P2[tb * NQ2 + ta * NQ + p] = dU[p];
    }
  }
  __syncthreads();
  {
    T c[NQ], w[NQ];
    for (int n = 0; n < NQ; ++n) c[n] = P0[tb * NQ2 + n * NQ + ta];
    eo_deriv<NQ, T>(sh, c, w);
    for (int q = 0; q < NQ; ++q) P1[tb * NQ2 + q * NQ + ta] = w[q];
  }
  __syncthreads();
  {
    T col[NQ];
    for (int n = 0; n < NQ; ++n) col[n] = P0[n * NQ2 + tb * NQ + ta];
    eo_deriv<NQ, T>(sh, col, gz);
  }
}

template <int NM, int NQ, typename T>
__global__ void __launch_bounds__(NQ * NQ)
grad_coefficient_kernel(const int64_t n_cells, const T *__restrict__ G,
                        const uint32_t *__restrict__ map, const uint32_t *__restrict__ perm,
                        const Shape1D<T> sh,
                        const T *__restrict__ u, const T *__restrict__ lam, T *__restrict__ out)
{
  constexpr int NQ2 = NQ * NQ, NQ3 = NQ2 * NQ;
  __shared__ T lds[6 * NQ3];
  const int64_t cell = blockIdx.x;
  const int ta = threadIdx.x % NQ, tb = threadIdx.x / NQ;
  T uz[NQ], lz[NQ];
  cell_gradient<NM, NQ, T>(u, map, cell, sh, lds, lds + NQ3, lds + 2 * NQ3, ta, tb, uz);
  __syncthreads();
  cell_gradient<NM, NQ, T>(lam, map, cell, sh, lds + 3 * NQ3, lds + 4 * NQ3, lds + 5 * NQ3, ta,
                           tb, lz);
  __syncthreads();
  for (int r = 0; r < NQ; ++r) {
    const int id = r * NQ2 + tb * NQ + ta;
    T gq[6];
#pragma unroll
    for (int c = 0; c < 6; ++c) gq[c] = G[g_index<T>(NQ, cell, c, id)];
    const T ux = lds[2 * NQ3 + id], uy = lds[NQ3 + id], uzz = uz[r];
    const T lx = lds[5 * NQ3 + id], ly = lds[4 * NQ3 + id], lzz = lz[r];
    const T fx = gq[0] * ux + gq[1] * uy + gq[2] * uzz;
    const T fy = gq[1] * ux + gq[3] * uy + gq[4] * uzz;
    const T fz = gq[2] * ux + gq[4] * uy + gq[5] * uzz;
    out[(perm ? static_cast<int64_t>(perm[cell]) : cell) * NQ3 + id] = -(fx * lx + fy * ly + fz * lzz);
  }
}

template <typename T>
void launch_bk3_grad_coefficient(int degree, int nq, int64_t n_cells, const T *G,
                                 const uint32_t *map, const uint32_t *perm, const Shape1D<T> &sh,
                                 const T *u, const T *lam, T *out, hipStream_t s) {
  if (n_cells == 0) return;
  PMG_REQUIRE(n_cells < (int64_t(1) << 31), PMG_ERR_SIZE, "too many cells");
#define PMG_GRAD_CASE(P, Q)                                                                  \
  if (degree == P && nq == Q) {                                                              \
    hipLaunchKernelGGL((grad_coefficient_kernel<P + 1, Q, T>), dim3(unsigned(n_cells)),       \
                       dim3(Q * Q), 0, s, n_cells, G, map, perm, sh, u, lam, out);                  \
    PMG_HIP_CHECK(hipGetLastError());                                                        \
    return;                                                                                  \
  }
  PMG_GRAD_CASE(1, 2) PMG_GRAD_CASE(1, 3) PMG_GRAD_CASE(2, 3) PMG_GRAD_CASE(2, 4)
  PMG_GRAD_CASE(3, 4) PMG_GRAD_CASE(3, 5) PMG_GRAD_CASE(4, 5) PMG_GRAD_CASE(4, 6)
  PMG_GRAD_CASE(5, 6) PMG_GRAD_CASE(5, 7) PMG_GRAD_CASE(6, 7) PMG_GRAD_CASE(6, 8)
#undef PMG_GRAD_CASE
  throw Error(PMG_ERR_UNSUPPORTED, "gradient kernel: unsupported (degree, n_q_points_1d)");
}
template void launch_bk3_grad_coefficient<double>(int, int, int64_t, const double *,
                                                  const uint32_t *, const uint32_t *,
                                                  const Shape1D<double> &,
                                                  const double *, const double *, double *,
                                                  hipStream_t);
template void launch_bk3_grad_coefficient<float>(int, int, int64_t, const float *,
                                                 const uint32_t *, const uint32_t *,
                                                 const Shape1D<float> &,
                                                 const float *, const float *, float *,
                                                 hipStream_t);

}  // namespace pmg
