// pmg_setup.hip -- set-up of the index data on the device: structured-box DoF
// maps, constraint lists and vertices, the elimination of constrained DoFs
// from a caller's map, the x-line compression of maps, and the DoF patches of
// the patch schedule.  What deal.II's DoFHandler and MatrixFree build on the
// host for the reference (tests/poisson_cube_bk3/program.cc:281-566;
// setup_dof_indices_per_color, include/operators/portable_laplace_operator_bk3.h:360-437;
// GeometricTransfer::setup_dof_indices, include/multigrid/portable_geometric_transfer.h:1399-1486;
// PolynomialTransfer::setup_weights_and_dof_indices,
// include/multigrid/portable_polynomial_transfer.h:1247-1360), for the same
// numberings, without a host pass over the cells.
#include <hipcub/hipcub.hpp>

#include <algorithm>

#include "pmg_common.hpp"
#include "pmg_kernels.hpp"

namespace pmg {

namespace {

unsigned grid_for(int64_t n, int64_t cap = 65536) {
  return static_cast<unsigned>(std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, cap)));
}

__device__ inline bool on_boundary(int64_t I, int64_t J, int64_t K, int64_t NX, int64_t NY,
                                   int64_t NZ) {
  return I == 0 || J == 0 || K == 0 || I == NX - 1 || J == NY - 1 || K == NZ - 1;
}

// Entry (cell, k, j, i) of cell (cx, cy, cz) of an nx x ny box names the grid
// point (step cx + i, step cy + j, step cz + k) of an NX x NY x NZ grid,
// lexicographic; on the grid boundary it is invalid when dirichlet is set.
__global__ void __launch_bounds__(256)
box_map_kernel(const int64_t n_pos, const int nx, const int ny, const int step, const int nl,
               const int64_t NX, const int64_t NY, const int64_t NZ, const int dirichlet,
               uint32_t *__restrict__ map)
{
  const int nl3 = nl * nl * nl;
  for (int64_t e = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; e < n_pos;
       e += int64_t(gridDim.x) * blockDim.x) {
    const int64_t c = e / nl3;
    const int l = static_cast<int>(e - c * nl3);
    const int i = l % nl, j = (l / nl) % nl, k = l / (nl * nl);
    const int64_t cyz = c / nx, cx = c - cyz * nx, cz = cyz / ny, cy = cyz - cz * ny;
    const int64_t I = step * cx + i, J = step * cy + j, K = step * cz + k;
    map[e] = dirichlet && on_boundary(I, J, K, NX, NY, NZ)
                 ? kInvalid
                 : static_cast<uint32_t>(I + NX * (J + NY * K));
  }
}

// The t-th boundary point of the grid in ascending lexicographic order: the
// plane K = 0, then per inner plane the row J = 0, the pairs (0, J), (NX-1, J)
// of the inner rows and the row J = NY-1, then the plane K = NZ-1.
__global__ void __launch_bounds__(256)
box_boundary_list_kernel(const int64_t n_b, const int64_t NX, const int64_t NY, const int64_t NZ,
                         uint32_t *__restrict__ list)
{
  const int64_t plane = NX * NY, m = 2 * NX + 2 * (NY - 2);
  for (int64_t t = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; t < n_b;
       t += int64_t(gridDim.x) * blockDim.x) {
    int64_t I, J, K;
    if (t < plane) {
      K = 0;
      J = t / NX;
      I = t - J * NX;
    } else if (t - plane < (NZ - 2) * m) {
      const int64_t u = t - plane;
      K = 1 + u / m;
      const int64_t r = u - (K - 1) * m;
      if (r < NX) {
        J = 0;
        I = r;
      } else if (r < NX + 2 * (NY - 2)) {
        const int64_t r2 = r - NX;
        J = 1 + r2 / 2;
        I = (r2 & 1) ? NX - 1 : 0;
      } else {
        J = NY - 1;
        I = r - NX - 2 * (NY - 2);
      }
    } else {
      const int64_t r = t - plane - (NZ - 2) * m;
      K = NZ - 1;
      J = r / NX;
      I = r - J * NX;
    }
    list[t] = static_cast<uint32_t>(I + NX * (J + NY * K));
  }
}

__global__ void __launch_bounds__(256)
box_boundary_mask_kernel(const int64_t n, const int64_t NX, const int64_t NY, const int64_t NZ,
                         uint8_t *__restrict__ mask)
{
  for (int64_t d = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; d < n;
       d += int64_t(gridDim.x) * blockDim.x) {
    const int64_t K = d / (NX * NY), r = d - K * NX * NY, J = r / NX, I = r - J * NX;
    mask[d] = on_boundary(I, J, K, NX, NY, NZ) ? 1 : 0;
  }
}

// The displaced vertex (a, c1, c2) of the box (same expressions as the host
// mesh generator: deform 1 the sinusoidal displacement, 2 the shear).
__device__ void box_vertex(const pmg_box &b, int a, int c1, int c2, double *X) {
  const double x = b.lo[0] + (b.hi[0] - b.lo[0]) * a / b.nx;
  const double y = b.lo[1] + (b.hi[1] - b.lo[1]) * c1 / b.ny;
  const double z = b.lo[2] + (b.hi[2] - b.lo[2]) * c2 / b.nz;
  X[0] = x;
  X[1] = y;
  X[2] = z;
  if (b.deform == 1) {
    const double w = 2.0 * M_PI, A = b.amplitude;
    X[0] += A * sin(w * y) * sin(w * z);
    X[1] += A * sin(w * z) * sin(w * x);
    X[2] += A * sin(w * x) * sin(w * y);
  } else if (b.deform == 2) {
    const double A = b.amplitude;
    X[0] += A * y;
    X[1] += A * z;
    X[2] += A * x;
  }
}

// one thread per (cell, vertex): 3 doubles of [C][8][3]
__global__ void __launch_bounds__(256)
box_vertices_kernel(const pmg_box b, const int64_t n_cells, double *__restrict__ v)
{
  for (int64_t t = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; t < n_cells * 8;
       t += int64_t(gridDim.x) * blockDim.x) {
    const int64_t c = t >> 3;
    const int q = static_cast<int>(t & 7);
    const int64_t cyz = c / b.nx, cx = c - cyz * b.nx, cz = cyz / b.ny, cy = cyz - cz * b.ny;
    double X[3];
    box_vertex(b, static_cast<int>(cx) + (q & 1), static_cast<int>(cy) + ((q >> 1) & 1),
               static_cast<int>(cz) + ((q >> 2) & 1), X);
    v[3 * t] = X[0];
    v[3 * t + 1] = X[1];
    v[3 * t + 2] = X[2];
  }
}

__global__ void __launch_bounds__(256)
iota_kernel(const int64_t n, uint32_t *__restrict__ p)
{
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x)
    p[i] = static_cast<uint32_t>(i);
}

// mask[list[i]] = 1; entries >= n_dofs are reported (smallest position)
__global__ void __launch_bounds__(256)
mask_from_list_kernel(const int64_t n, const uint32_t *__restrict__ list, const int64_t n_dofs,
                      uint8_t *__restrict__ mask, unsigned long long *bad)
{
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x) {
    const uint32_t d = list[i];
    if (int64_t(d) < n_dofs) mask[d] = 1;
    else atomicMin(bad, static_cast<unsigned long long>(i));
  }
}

// out[c][i] = src[perm[c]][i] (perm null: identity; out may be src then),
// entries of masked DoFs made invalid; out-of-range entries are reported by
// their caller position and stored invalid
__global__ void __launch_bounds__(256)
map_eliminate_kernel(const int64_t n_pos, const int nm3, const uint32_t *src,
                     const uint32_t *__restrict__ perm, const uint8_t *__restrict__ mask,
                     const int64_t n_dofs, uint32_t *out, unsigned long long *bad)
{
  for (int64_t e = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; e < n_pos;
       e += int64_t(gridDim.x) * blockDim.x) {
    int64_t from = e;
    if (perm) {
      const int64_t c = e / nm3;
      from = int64_t(perm[c]) * nm3 + (e - c * nm3);
    }
    const uint32_t d = src[from];
    uint32_t r = d;
    if (d != kInvalid) {
      if (int64_t(d) >= n_dofs) {
        atomicMin(bad, static_cast<unsigned long long>(from));
        r = kInvalid;
      } else if (mask && mask[d]) {
        r = kInvalid;
      }
    }
    out[e] = r;
  }
}

// Compressed x-lines: a cell qualifies when each of its nm^2 x-lines is a run
// of nm consecutive valid indices ending within bound; its lines hold the run
// starts (all zero for cells that do not qualify).  One thread per cell;
// count += qualifying cells (one atomic per wave).
template <typename I>
__global__ void __launch_bounds__(256)
line_compress_kernel(const int64_t n_cells, const int nm, const I *__restrict__ map,
                     const int64_t bound, const I invalid, I *__restrict__ lines,
                     uint8_t *__restrict__ flag, unsigned long long *count)
{
  const int64_t c = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  bool ok = false;
  if (c < n_cells) {
    const int nm2 = nm * nm;
    const I *m = map + c * int64_t(nm2) * nm;
    ok = true;
    for (int kj = 0; kj < nm2 && ok; ++kj) {
      const I *ln = m + kj * nm;
      const I l0 = ln[0];
      ok = l0 != invalid && int64_t(l0) + nm <= bound;
      for (int i = 1; i < nm && ok; ++i) ok = ln[i] == static_cast<I>(l0 + i);
    }
    for (int kj = 0; kj < nm2; ++kj) lines[c * nm2 + kj] = ok ? m[kj * nm] : I(0);
    flag[c] = ok ? 1 : 0;
  }
  const unsigned long long votes = __ballot(ok);
  if ((threadIdx.x & 63) == 0 && votes)
    atomicAdd(count, static_cast<unsigned long long>(__popcll(votes)));
}

__global__ void __launch_bounds__(256)
remap_kernel(const int64_t n, const uint32_t *__restrict__ to, uint32_t *__restrict__ list)
{
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x)
    list[i] = to[list[i]];
}

__global__ void __launch_bounds__(256)
count_entries_kernel(const int64_t n_pos, const uint32_t *__restrict__ map, uint32_t *cnt)
{
  for (int64_t e = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; e < n_pos;
       e += int64_t(gridDim.x) * blockDim.x) {
    const uint32_t d = map[e];
    if (d != kInvalid) atomicAdd(cnt + d, 1u);
  }
}

constexpr int kPatchThreads = 1024;
constexpr int kPatchMaxEntries = 2 * kPatchThreads;

// DoF patch of batch b (cells b E .. b E + E - 1): the batch's valid entries
// as keys DoF << 16 | entry, bitonic-sorted in LDS (sort_n a power of two >=
// E nm3, padded with ~0); run heads give the patch slots (ascending DoFs),
// every entry's slot goes to pidx (0xFFFF for invalid entries), each slot's
// DoF to pdof with bit 31 set when the run is shorter than the DoF's
// multiplicity (a cell outside the batch has it); slots past the count are 0.
// Then a greedy colouring of the batch's cells in cell order (a cell takes
// the smallest colour no earlier cell sharing one of its slots has) into
// pcolor, and the number of colours into pncol[b]; pncol[b] = 0 when a cell
// names one DoF twice (its lanes would update one slot together).
__global__ void __launch_bounds__(kPatchThreads)
patch_build_kernel(const int64_t n_cells, const int nm3, const int E, const int sort_n,
                   const uint32_t *__restrict__ map, const uint32_t *__restrict__ mult,
                   uint32_t *__restrict__ pdof, uint32_t *__restrict__ pcount,
                   uint16_t *__restrict__ pidx, uint8_t *__restrict__ pcolor,
                   uint8_t *__restrict__ pncol)
{
  using Scan = hipcub::BlockScan<uint32_t, kPatchThreads>;
  __shared__ uint64_t key[kPatchMaxEntries];
  __shared__ uint16_t first[kPatchMaxEntries], last[kPatchMaxEntries], eslot[kPatchMaxEntries];
  __shared__ unsigned long long smask[kPatchMaxEntries];
  __shared__ unsigned long long forbid;
  __shared__ int dup;
  __shared__ typename Scan::TempStorage scan_tmp;
  constexpr uint64_t kNone = ~uint64_t(0);
  const int64_t b = blockIdx.x;
  const int64_t c0 = b * E;
  const int nce = static_cast<int>(std::min<int64_t>(E, n_cells - c0));
  const int ne = nce * nm3;
  const int pmax = E * nm3;
  const int t = threadIdx.x;
  uint16_t *bidx = pidx + c0 * nm3;
  if (t == 0) dup = 0;
  for (int k = t; k < sort_n; k += kPatchThreads) {
    uint64_t v = kNone;
    if (k < ne) {
      const uint32_t d = map[c0 * nm3 + k];
      if (d != kInvalid) {
        v = (uint64_t(d) << 16) | uint64_t(k);
      } else {
        bidx[k] = 0xFFFF;
        eslot[k] = 0xFFFF;
      }
    }
    key[k] = v;
  }
  __syncthreads();
  for (int size = 2; size <= sort_n; size <<= 1)
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      for (int i = t; i < sort_n / 2; i += kPatchThreads) {
        const int lo = 2 * i - (i & (stride - 1)), hi = lo + stride;
        const uint64_t a = key[lo], c = key[hi];
        if ((a > c) == ((lo & size) == 0)) {
          key[lo] = c;
          key[hi] = a;
        }
      }
      __syncthreads();
    }
  uint32_t head[2], slot[2], count = 0;
  for (int q = 0; q < 2; ++q) {
    const int k = 2 * t + q;
    head[q] = 0;
    if (k < sort_n && key[k] != kNone)
      head[q] = (k == 0 || (key[k - 1] >> 16) != (key[k] >> 16)) ? 1u : 0u;
  }
  Scan(scan_tmp).ExclusiveSum(head, slot, count);
  for (int q = 0; q < 2; ++q) {
    const int k = 2 * t + q;
    if (k >= sort_n || key[k] == kNone) continue;
    const uint64_t v = key[k];
    const uint32_t s = slot[q] + head[q] - 1;
    bidx[v & 0xFFFF] = static_cast<uint16_t>(s);
    eslot[v & 0xFFFF] = static_cast<uint16_t>(s);
    if (head[q]) first[s] = static_cast<uint16_t>(k);
    if (k + 1 == sort_n || (key[k + 1] >> 16) != (v >> 16)) last[s] = static_cast<uint16_t>(k + 1);
    // a run's entries are sorted by entry, so two of one cell are neighbours
    if (!head[q] && int(key[k - 1] & 0xFFFF) / nm3 == int(v & 0xFFFF) / nm3) dup = 1;
  }
  for (int n = t; n < pmax; n += kPatchThreads) smask[n] = 0ull;
  __syncthreads();
  for (int n = t; n < pmax; n += kPatchThreads) {
    uint32_t out = 0;
    if (uint32_t(n) < count) {
      const uint32_t d = static_cast<uint32_t>(key[first[n]] >> 16);
      const uint32_t len = uint32_t(last[n]) - first[n];
      out = d | (len < mult[d] ? 0x80000000u : 0u);
    }
    pdof[b * pmax + n] = out;
  }
  int n_colors = 0;
  for (int e = 0; e < nce; ++e) {
    if (t == 0) forbid = 0ull;
    __syncthreads();
    const uint16_t s = t < nm3 ? eslot[e * nm3 + t] : uint16_t(0xFFFF);
    if (s != 0xFFFF) {
      const unsigned long long f = smask[s];
      if (f) atomicOr(&forbid, f);
    }
    __syncthreads();
    const int c = __builtin_ctzll(~forbid);  // < 64: at most 63 earlier cells
    if (s != 0xFFFF) atomicOr(&smask[s], 1ull << c);
    if (t == 0) pcolor[c0 + e] = static_cast<uint8_t>(c);
    n_colors = c + 1 > n_colors ? c + 1 : n_colors;
    __syncthreads();
  }
  if (t == 0) {
    pcount[b] = count;
    pncol[b] = dup ? 0 : static_cast<uint8_t>(n_colors);
  }
}

struct BoxGrid {
  int64_t NX, NY, NZ;
  BoxGrid(const pmg_box &b, int p, int r = 1)
      : NX(int64_t(p) * r * b.nx + 1), NY(int64_t(p) * r * b.ny + 1), NZ(int64_t(p) * r * b.nz + 1) {}
  int64_t size() const { return NX * NY * NZ; }
};

unsigned long long read_flag(const unsigned long long *d, hipStream_t s) {
  unsigned long long h = 0;
  PMG_HIP_CHECK(hipMemcpyAsync(&h, d, sizeof(h), hipMemcpyDeviceToHost, s));
  PMG_HIP_CHECK(hipStreamSynchronize(s));
  return h;
}

}  // namespace

void box_map_device(const pmg_box &b, int degree, int refine, bool dirichlet, uint32_t *map,
                    hipStream_t s) {
  const int step = degree * refine, nl = step + 1;
  const int64_t n_pos = int64_t(b.nx) * b.ny * b.nz * nl * nl * nl;
  if (n_pos == 0) return;
  const BoxGrid g(b, degree, refine);
  PMG_REQUIRE(g.size() < int64_t(0xFFFFFFFFu), PMG_ERR_SIZE, "too many DoFs for 32-bit maps");
  hipLaunchKernelGGL(box_map_kernel, dim3(grid_for(n_pos)), dim3(256), 0, s, n_pos, b.nx, b.ny,
                     step, nl, g.NX, g.NY, g.NZ, dirichlet ? 1 : 0, map);
  PMG_HIP_CHECK(hipGetLastError());
}

int64_t box_boundary_count(const pmg_box &b, int degree) {
  const BoxGrid g(b, degree);
  return 2 * g.NX * g.NY + (g.NZ - 2) * (2 * g.NX + 2 * (g.NY - 2));
}

void box_boundary_device(const pmg_box &b, int degree, uint32_t *list, uint8_t *mask,
                         hipStream_t s) {
  const BoxGrid g(b, degree);
  const int64_t nb = box_boundary_count(b, degree);
  if (list && nb > 0) {
    hipLaunchKernelGGL(box_boundary_list_kernel, dim3(grid_for(nb)), dim3(256), 0, s, nb, g.NX,
                       g.NY, g.NZ, list);
    PMG_HIP_CHECK(hipGetLastError());
  }
  if (mask) {
    hipLaunchKernelGGL(box_boundary_mask_kernel, dim3(grid_for(g.size())), dim3(256), 0, s,
                       g.size(), g.NX, g.NY, g.NZ, mask);
    PMG_HIP_CHECK(hipGetLastError());
  }
}

void box_vertices_device(const pmg_box &b, double *v, hipStream_t s) {
  const int64_t n_cells = int64_t(b.nx) * b.ny * b.nz;
  if (n_cells == 0) return;
  hipLaunchKernelGGL(box_vertices_kernel, dim3(grid_for(8 * n_cells)), dim3(256), 0, s, b,
                     n_cells, v);
  PMG_HIP_CHECK(hipGetLastError());
}

void iota_device(uint32_t *p, int64_t n, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(iota_kernel, dim3(grid_for(n)), dim3(256), 0, s, n, p);
  PMG_HIP_CHECK(hipGetLastError());
}

int64_t mask_from_list_device(int64_t n, const uint32_t *list, int64_t n_dofs, uint8_t *mask,
                              hipStream_t s) {
  if (n_dofs > 0) PMG_HIP_CHECK(hipMemsetAsync(mask, 0, static_cast<size_t>(n_dofs), s));
  if (n <= 0) return -1;
  DeviceArray<unsigned long long> bad(1);
  PMG_HIP_CHECK(hipMemsetAsync(bad.ptr, 0xFF, sizeof(unsigned long long), s));
  hipLaunchKernelGGL(mask_from_list_kernel, dim3(grid_for(n)), dim3(256), 0, s, n, list, n_dofs,
                     mask, bad.ptr);
  PMG_HIP_CHECK(hipGetLastError());
  const unsigned long long h = read_flag(bad.ptr, s);
  return h == ~0ull ? -1 : static_cast<int64_t>(h);
}

int64_t map_eliminate_device(int64_t n_cells, int nm3, const uint32_t *src, const uint32_t *perm,
                             const uint8_t *mask, int64_t n_dofs, uint32_t *out, hipStream_t s) {
  const int64_t n_pos = n_cells * nm3;
  if (n_pos <= 0) return -1;
  DeviceArray<unsigned long long> bad(1);
  PMG_HIP_CHECK(hipMemsetAsync(bad.ptr, 0xFF, sizeof(unsigned long long), s));
  hipLaunchKernelGGL(map_eliminate_kernel, dim3(grid_for(n_pos)), dim3(256), 0, s, n_pos, nm3, src,
                     perm, mask, n_dofs, out, bad.ptr);
  PMG_HIP_CHECK(hipGetLastError());
  const unsigned long long h = read_flag(bad.ptr, s);
  return h == ~0ull ? -1 : static_cast<int64_t>(h);
}

int64_t compress_lines_device(int64_t n_cells, int nm, const uint32_t *map, int64_t n_dofs,
                              uint32_t *lines, uint8_t *flag, hipStream_t s) {
  if (n_cells <= 0) return 0;
  DeviceArray<unsigned long long> cnt(1);
  PMG_HIP_CHECK(hipMemsetAsync(cnt.ptr, 0, sizeof(unsigned long long), s));
  PMG_REQUIRE(n_cells < (int64_t(1) << 40), PMG_ERR_SIZE, "too many cells");
  hipLaunchKernelGGL(line_compress_kernel<uint32_t>, dim3(static_cast<unsigned>((n_cells + 255) / 256)),
                     dim3(256), 0, s, n_cells, nm, map, n_dofs, kInvalid, lines, flag, cnt.ptr);
  PMG_HIP_CHECK(hipGetLastError());
  return static_cast<int64_t>(read_flag(cnt.ptr, s));
}

int64_t compress_patch_lines_device(int64_t n_cells, int nm, const uint16_t *pidx, uint16_t *pline,
                                    uint8_t *flag, hipStream_t s) {
  if (n_cells <= 0) return 0;
  DeviceArray<unsigned long long> cnt(1);
  PMG_HIP_CHECK(hipMemsetAsync(cnt.ptr, 0, sizeof(unsigned long long), s));
  hipLaunchKernelGGL(line_compress_kernel<uint16_t>, dim3(static_cast<unsigned>((n_cells + 255) / 256)),
                     dim3(256), 0, s, n_cells, nm, pidx, int64_t(0xFFFF), uint16_t(0xFFFF), pline,
                     flag, cnt.ptr);
  PMG_HIP_CHECK(hipGetLastError());
  return static_cast<int64_t>(read_flag(cnt.ptr, s));
}

void remap_list_device(int64_t n, const uint32_t *to, uint32_t *list, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(remap_kernel, dim3(grid_for(n)), dim3(256), 0, s, n, to, list);
  PMG_HIP_CHECK(hipGetLastError());
}

// Run records of the patch entries, one per 64 consecutive slots of a batch
// (the slots one wave handles in the patch kernel): 8 words {first DoF of runs
// 0..3, s1 | s2 << 8 | s3 << 16 | nr << 24, 0, 0, 0}, runs of consecutive DoFs
// starting at slots 0 = s0 < s1 < s2 < s3 (64 for runs that do not exist), nr
// the number of runs (1..4) or 0 when the chunk's valid slots need more than
// 4 runs (the kernel then loads the chunk's entries).  Slots past the count
// are not described.
__global__ void __launch_bounds__(256)
patch_chunk_kernel(const int64_t n_chunks, const int pmax, const int qc,
                   const uint32_t *__restrict__ pdof, const uint32_t *__restrict__ pcount,
                   uint32_t *__restrict__ rec, unsigned long long *described)
{
  const int64_t c = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  bool ok = false;
  if (c < n_chunks) {
    const int64_t b = c / qc;
    const int q = static_cast<int>(c - b * qc);
    const int n = static_cast<int>(pcount[b]) - q * 64;
    const uint32_t *pd = pdof + b * pmax + q * 64;
    uint32_t d0[4] = {0u, 0u, 0u, 0u}, s[4] = {0u, 64u, 64u, 64u};
    int nr = 1;
    uint32_t prev = 0;
    for (int u = 0; u < n && u < 64 && nr > 0; ++u) {
      const uint32_t d = pd[u] & 0x7FFFFFFFu;
      if (u == 0) {
        d0[0] = d;
      } else if (d != prev + 1) {
        if (nr == 4) {
          nr = 0;  // too many runs
        } else {
          d0[nr] = d;
          s[nr] = static_cast<uint32_t>(u);
          ++nr;
        }
      }
      prev = d;
    }
    uint32_t *r = rec + c * 8;
    for (int k = 0; k < 4; ++k) r[k] = d0[k];
    r[4] = s[1] | (s[2] << 8) | (s[3] << 16) | (static_cast<uint32_t>(nr) << 24);
    r[5] = r[6] = r[7] = 0u;
    ok = nr > 0;
  }
  const unsigned long long votes = __ballot(ok);
  if ((threadIdx.x & 63) == 0 && votes)
    atomicAdd(described, static_cast<unsigned long long>(__popcll(votes)));
}

int64_t build_patch_chunks_device(int64_t nb, int pmax, const uint32_t *pdof,
                                  const uint32_t *pcount, uint32_t *rec, hipStream_t s) {
  const int qc = (pmax + 63) / 64;
  const int64_t n_chunks = nb * qc;
  if (n_chunks <= 0) return 0;
  DeviceArray<unsigned long long> cnt(1);
  PMG_HIP_CHECK(hipMemsetAsync(cnt.ptr, 0, sizeof(unsigned long long), s));
  PMG_REQUIRE(n_chunks < (int64_t(1) << 39), PMG_ERR_SIZE, "too many patch chunks");
  hipLaunchKernelGGL(patch_chunk_kernel, dim3(grid_for(n_chunks, int64_t(1) << 31)), dim3(256), 0, s, n_chunks, pmax,
                     qc, pdof, pcount, rec, cnt.ptr);
  PMG_HIP_CHECK(hipGetLastError());
  return static_cast<int64_t>(read_flag(cnt.ptr, s));
}

bool patch_schedule_fits(int64_t n_cells, int nm3, int E, int64_t n_dofs) {
  return n_cells > 0 && n_dofs < (int64_t(1) << 31) && int64_t(E) * nm3 <= kPatchMaxEntries;
}

void build_patches_device(int64_t n_cells, int nm3, int E, const uint32_t *map, int64_t n_dofs,
                          uint32_t *pdof, uint32_t *pcount, uint16_t *pidx, uint8_t *pcolor,
                          uint8_t *pncol, hipStream_t s) {
  PMG_REQUIRE(patch_schedule_fits(n_cells, nm3, E, n_dofs) && E <= 64 && nm3 <= kPatchThreads,
              PMG_ERR_SIZE, "patch schedule: batch too large");
  const int64_t nb = (n_cells + E - 1) / E;
  PMG_REQUIRE(nb < (int64_t(1) << 31), PMG_ERR_SIZE, "patch schedule: too many batches");
  int sort_n = 2;
  while (sort_n < E * nm3) sort_n <<= 1;
  DeviceArray<uint32_t> mult(static_cast<size_t>(std::max<int64_t>(n_dofs, 1)));
  PMG_HIP_CHECK(hipMemsetAsync(mult.ptr, 0, mult.bytes(), s));
  hipLaunchKernelGGL(count_entries_kernel, dim3(grid_for(n_cells * nm3)), dim3(256), 0, s,
                     n_cells * nm3, map, mult.ptr);
  PMG_HIP_CHECK(hipGetLastError());
  hipLaunchKernelGGL(patch_build_kernel, dim3(static_cast<unsigned>(nb)), dim3(kPatchThreads), 0, s,
                     n_cells, nm3, E, sort_n, map, mult.ptr, pdof, pcount, pidx, pcolor, pncol);
  PMG_HIP_CHECK(hipGetLastError());
  PMG_HIP_CHECK(hipStreamSynchronize(s));  // mult is released on return
}

}  // namespace pmg
