// pmg_kernels.hpp -- host-callable launchers of the gfx950 kernels.
// All launchers are stream-ordered and asynchronous; they throw pmg::Error on
// unsupported template parameters or launch failures.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "pmg.h"

namespace pmg {

// ---------------------------------------------------------------- BK3 operator
// 1D shape data in the operator precision, passed to kernels by value.
constexpr int kMaxNm = 8;
constexpr int kMaxNq = 8;

// Even-odd tables (stride kEO): the Gauss-Lobatto basis at the Gauss points is
// centro-symmetric, S[k][r] = S[m-k][n-r], and the collocation derivative
// centro-antisymmetric, D[n][p] = -D[N-n][N-p], so each 1D contraction splits
// into half-size even and odd parts (half the products, half the
// coefficients held in scalar registers).  m = nm-1, N = nq-1, hm = nm/2,
// hq = nq/2; see fill_even_odd.
constexpr int kEO = (kMaxNq + 1) / 2;
template <typename T>
struct Shape1D {
  T S[kMaxNm * kMaxNq];  // S[i*nq+q]
  T D[kMaxNq * kMaxNq];  // D[n*nq+q]
  T SE[kEO * kEO], SO[kEO * kEO], SM[kEO];  // (S[k][r] +- S[m-k][r]) / 2, S[hm][r]
  T FE[kEO * kEO], FO[kEO * kEO], FM[kEO];  // (D[n][p] +- D[N-n][p]) / 2, D[hq][p]
  T TE[kEO * kEO], TO[kEO * kEO], TC[kEO], TM[kEO];  // (D[r][n] -+ D[N-r][n]) / 2, D[r][hq], D[hq][n]
  T W[kMaxNq];  // 1D quadrature weights (affine geometry: G_q = A w_p w_q w_r)
};

// Fills the even-odd tables of sh from S [nm][nq] and D [nq][nq] (row-major).
template <typename T>
inline void fill_even_odd(Shape1D<T> &sh, const double *S, const double *D, int nm, int nq) {
  const int m = nm - 1, N = nq - 1, hm = nm / 2, hq = nq / 2, cq = (nq + 1) / 2;
  auto s = [&](int k, int r) { return S[k * nq + r]; };
  auto d = [&](int n, int p) { return D[n * nq + p]; };
  for (int k = 0; k < hm; ++k)
    for (int r = 0; r < cq; ++r) {
      sh.SE[k * kEO + r] = static_cast<T>(0.5 * (s(k, r) + s(m - k, r)));
      if (r < hq) sh.SO[k * kEO + r] = static_cast<T>(0.5 * (s(k, r) - s(m - k, r)));
    }
  if (nm % 2)
    for (int r = 0; r < cq; ++r) sh.SM[r] = static_cast<T>(s(hm, r));
  for (int n = 0; n < hq; ++n)
    for (int p = 0; p < cq; ++p) {
      if (p < hq) sh.FE[n * kEO + p] = static_cast<T>(0.5 * (d(n, p) + d(N - n, p)));
      sh.FO[n * kEO + p] = static_cast<T>(0.5 * (d(n, p) - d(N - n, p)));
    }
  for (int r = 0; r < hq; ++r) {
    for (int n = 0; n < hq; ++n) {
      sh.TE[r * kEO + n] = static_cast<T>(0.5 * (d(r, n) - d(N - r, n)));
      sh.TO[r * kEO + n] = static_cast<T>(0.5 * (d(r, n) + d(N - r, n)));
    }
    if (nq % 2) {
      sh.FM[r] = static_cast<T>(d(hq, r));
      sh.TC[r] = static_cast<T>(d(r, hq));
      sh.TM[r] = static_cast<T>(d(hq, r));
    }
  }
}

// Cells per workgroup ("batch") of the cell kernels: each cell owns an
// nq x nq lane tile of a 256-lane workgroup, E = floor(256 / nq^2).  (Cubes of
// 2^3 cells at nq = 5 were measured slower: fewer active lanes and shorter
// contiguous DoF runs per batch.)
__host__ __device__ constexpr int g_batch_cells(int nq) {
  return (256 / (nq * nq)) > 0 ? 256 / (nq * nq) : 1;
}
// Storage of the geometric factors G (6 symmetric components per q-point),
// batch-interleaved: with the lane l = e nq^2 + (q mod nq^2) of q-point q of
// cell e of a batch (r = q / nq^2),
//   FP32: G[((batch * 6 + comp) * nq + r) * lanes + l],
//   FP64: G[((batch * 3 + comp / 2) * nq + r) * 2 lanes + 2 l + comp % 2],
// so that every G load instruction of a wave reads 64 consecutive values
// (FP32) or pairs of components (FP64: one 16-byte load for two components,
// measured faster than two 8-byte loads; FP32 lost with 8-byte pairs).
// lanes = E nq^2 rounded up to 64: each wave's loads start on a line boundary
// (unpadded rows, E nq^2 lanes, measured slower: FP64 Q4 64^3 387 -> 392 us,
// FP32 211 -> 239 us).
__host__ __device__ inline int g_batch_lanes(int nq) {
  return ((g_batch_cells(nq) * nq * nq + 63) / 64) * 64;
}
template <typename T>
__host__ __device__ inline int64_t g_index(int nq, int64_t cell, int comp, int q) {
  const int nq2 = nq * nq, E = g_batch_cells(nq);
  const int r = q / nq2, t = q - r * nq2;
  const int64_t b = cell / E;
  const int e = static_cast<int>(cell - b * E);
  const int l = e * nq2 + t;
  if (sizeof(T) == 8)
    return ((b * 3 + comp / 2) * nq + r) * 2 * g_batch_lanes(nq) + 2 * l + (comp & 1);
  return ((b * 6 + comp) * nq + r) * g_batch_lanes(nq) + l;
}
__host__ __device__ inline int64_t g_size(int nq, int64_t n_cells) {
  const int64_t nb = (n_cells + g_batch_cells(nq) - 1) / g_batch_cells(nq);
  return nb * 6 * nq * g_batch_lanes(nq);
}

// How the cell loop is scheduled:
//  kAtomic  all cells in one launch, every cell scatter-adds its nm^3 values
//           with device-scope FP atomics;
//  kColored one launch per colour of a cell colouring, plain read-modify-write
//           (no two cells of a colour share a DoF);
//  kPatch   all cells in one launch; a batch gathers the union of its cells'
//           DoFs ("patch") into LDS, its cells accumulate into the patch with
//           LDS atomics, and the patch is added to y once per DoF with device
//           atomics in sorted DoF order (fewer, longer-run atomics);
//  kGather  two launches, no atomics on the output: the cells store their
//           nm^3 values to a per-cell buffer laid out like the DoF map, then
//           every DoF sums its buffer entries through the inverse map (CSR,
//           entries in ascending order: the result is reproducible bit for bit);
//  kGrid    the same two passes for a lexicographic box (map verified): the
//           buffer is the box's exploded grid, and each DoF sums its <= 8
//           adjacent copies (coalesced, no index loads, reproducible).
enum Bk3ScheduleKind : int { kAtomic = 0, kColored = 1, kPatch = 2, kGather = 3, kGrid = 4 };
struct Bk3Schedule {
  int kind = kAtomic;
  int gmode = 2;                        // G loads: 0 early non-temporal, 1 early, 2 late non-temporal
  const uint32_t *cell_ids = nullptr;   // device: cells sorted by colour
  std::vector<int64_t> color_offsets;   // host: n_colors + 1 offsets into cell_ids
  // patch data (kPatch): per batch b, pcount[b] patch DoFs at
  // pdof[b * E nm^3 ...] (bit 31 set: shared with another batch), and per
  // cell-local DoF its patch index pidx[cell][i] (0xFFFF: constrained)
  const uint32_t *pdof = nullptr, *pcount = nullptr;
  const uint16_t *pidx = nullptr;
  // run records of the patch entries, 8 words per 64 slots (see
  // patch_chunk_kernel); null when fewer than half of the chunks are runs
  const uint32_t *pchunk = nullptr;
  // compressed DoF indices (atomic and colored schedules): per cell
  // lines[cell][k][j] = first DoF of x-line (j,k) when lflag[cell] says every
  // x-line of the cell is a run of consecutive DoFs (no constrained entry);
  // null when too few cells qualify
  const uint32_t *lines = nullptr;
  const uint8_t *lflag = nullptr;
  // compressed patch indices (patch schedule): pline[cell][k][j] = patch slot
  // of the first DoF of x-line (j,k) when pflag[cell] says the slots of every
  // x-line are consecutive; null when too few cells qualify
  const uint16_t *pline = nullptr;
  const uint8_t *pflag = nullptr;
  // batch-local cell colouring of the patch schedule (null: the cells add
  // into the LDS patch with LDS atomics): colour by colour, the cells of one
  // colour add with plain read-modify-writes (no two share a slot);
  // pncol[b] = 0 keeps the atomics for batch b
  const uint8_t *pcolor = nullptr, *pncol = nullptr;
  // patch schedule, FP32 at nq >= 6: the x and y sweeps of the cell apply as
  // batched contractions on the matrix cores, with these operand tables
  // (6 x 16 x 16 floats; null: vector sweeps)
  const float *mxa = nullptr;
  // gather data (kGather): inverse map of the gn DoFs, goff[gn + 1] into
  // glist (entry positions cell * nm^3 + local, ascending per DoF), and the
  // per-cell buffer gbuf [n_cells][nm^3] of the operator precision
  const uint32_t *goff = nullptr, *glist = nullptr;
  void *gbuf = nullptr;
  int64_t gn = 0;
  // grid data (kGrid, gbuf is the exploded grid): box cells per direction;
  // the second pass writes the constrained rows too (zero, or y unchanged
  // when adding): callers restore them
  int grid[3] = {0, 0, 0};
  // affine geometry (atomic and patch schedules): per cell the 6 components
  // of A = |det J| J^-1 J^-T, [cell][6] in the operator precision, used
  // instead of G (G_q = A w_p w_q w_r); null: G
  const void *gaff = nullptr;
};
// true if map is the lexicographic box map of dims[0] x dims[1] x dims[2]
// cells at this degree (entries of constrained DoFs may be invalid)
bool check_box_map(int degree, const int *dims, int64_t n_cells, const uint32_t *map,
                   const uint8_t *cmask, hipStream_t s);

// Inverse map of a DoF map for the gather schedule, built on the device:
// off[n_dofs + 1] and list[off[n_dofs]] with the positions e of map[e] == d
// in ascending order for each DoF d (invalid entries are in no list).
// list_out is allocated here (it needs the entry count).  Synchronous.
void build_inverse_map(int64_t n_positions, const uint32_t *map, int64_t n_dofs, uint32_t *off,
                       uint32_t **list_out, int64_t *n_entries, hipStream_t s);

// ---- index data built on the device (pmg_setup.hip)
// Box map [C][(p r + 1)^3]: cell (cx, cy, cz) of the box, entry (k, j, i) ->
// grid point (p r cx + i, ...) of the box refined r times (r = 1: the Q_p
// DoF map, r = 2: the h-transfer patch map on the fine box), lexicographic on
// the (p r nx + 1)(p r ny + 1)(p r nz + 1) grid; boundary points invalid when
// dirichlet.
void box_map_device(const pmg_box &b, int degree, int refine, bool dirichlet, uint32_t *map,
                    hipStream_t s);
int64_t box_boundary_count(const pmg_box &b, int degree);
// ascending boundary DoF list (list may be null) and the 0/1 mask of the
// boundary DoFs (mask may be null) of the Q_p grid of the box
void box_boundary_device(const pmg_box &b, int degree, uint32_t *list, uint8_t *mask,
                         hipStream_t s);
// [C][8][3] cell vertices, the host mesh generator's displacements
void box_vertices_device(const pmg_box &b, double *v, hipStream_t s);
void iota_device(uint32_t *p, int64_t n, hipStream_t s);
// mask (n_dofs bytes) = 1 on the listed DoFs, 0 elsewhere; returns the first
// list position holding a DoF >= n_dofs (-1: none).  Synchronous.
int64_t mask_from_list_device(int64_t n, const uint32_t *list, int64_t n_dofs, uint8_t *mask,
                              hipStream_t s);
// out[c][i] = src[perm[c]][i] (perm null: identity, out == src allowed) with
// masked DoFs made invalid; returns the first caller position holding an
// entry >= n_dofs that is not invalid (-1: none).  Synchronous.
int64_t map_eliminate_device(int64_t n_cells, int nm3, const uint32_t *src, const uint32_t *perm,
                             const uint8_t *mask, int64_t n_dofs, uint32_t *out, hipStream_t s);
// x-line compression of a map (lines [C][nm^2], flag [C]; lines of cells
// that do not qualify are 0) and of patch slots; return the number of
// qualifying cells.  Synchronous.
int64_t compress_lines_device(int64_t n_cells, int nm, const uint32_t *map, int64_t n_dofs,
                              uint32_t *lines, uint8_t *flag, hipStream_t s);
int64_t compress_patch_lines_device(int64_t n_cells, int nm, const uint16_t *pidx, uint16_t *pline,
                                    uint8_t *flag, hipStream_t s);
// list[i] = to[list[i]] (positions -> the numbering a map gives them)
void remap_list_device(int64_t n, const uint32_t *to, uint32_t *list, hipStream_t s);
// patches of the batches of E cells (Bk3Schedule::pdof, pcount, pidx) and
// the batch-local cell colouring (pcolor per cell, pncol colours per batch)
bool patch_schedule_fits(int64_t n_cells, int nm3, int E, int64_t n_dofs);
// Run records of the patch entries, 8 words per 64 slots of every batch (see
// patch_chunk_kernel); returns the number of chunks described by <= 4 runs
int64_t build_patch_chunks_device(int64_t nb, int pmax, const uint32_t *pdof,
                                  const uint32_t *pcount, uint32_t *rec, hipStream_t s);
void build_patches_device(int64_t n_cells, int nm3, int E, const uint32_t *map, int64_t n_dofs,
                          uint32_t *pdof, uint32_t *pcount, uint16_t *pidx, uint8_t *pcolor,
                          uint8_t *pncol, hipStream_t s);

// y += sum_cells A_e x.  assign is accepted for the callers' intent (vmult,
// which zeroes the free rows first) and does not change the result.
template <typename T>
void launch_bk3_cells(int degree, int nq, const T *x, T *y, const T *G, const uint32_t *map,
                      int64_t n_cells, const Shape1D<T> &sh, hipStream_t s,
                      const Bk3Schedule *sched = nullptr, bool assign = false);
// G[c][6][nq^3] = kappa * JxW * J^{-1}J^{-T} of the trilinear map of the cell
// vertices; kappa (optional) is indexed by the caller's cell perm[c]
// Affine geometry: A[c][6] = |det J| (J^-1 J^-T) of every cell whose
// trilinear map is affine (twist terms of its vertices within 1e-10 of its
// longest edge); returns the number of cells that are not affine (their A is
// left undefined).  Synchronous on s.
template <typename T>
int64_t launch_affine_geometry(int64_t n_cells, const double *cell_vertices, T *A, hipStream_t s);
template <typename T>
void launch_geometry(int nq, int64_t n_cells, const double *cell_vertices, const double *qpts,
                     const double *qwts, const double *kappa, const uint32_t *perm, T *G,
                     hipStream_t s);
// y += sum A_e x over the cells cell_ids[0 .. n_list) (device positions), or
// over cells 0 .. n_list when cell_ids is null: the atomic scatter with late
// G loads; lines/lflag optional compressed indices of map (Bk3Schedule)
template <typename T>
void launch_bk3_cell_list(int degree, int nq, const T *x, T *y, const T *G, const uint32_t *map,
                          const uint32_t *lines, const uint8_t *lflag, const uint32_t *cell_ids,
                          int64_t n_list, const Shape1D<T> &sh, hipStream_t s);
// diag[map[c][i]] += (A_e)_ii, accumulated in FP64 from G of either precision
// (S_and_Sd: shape values and exact derivatives)
template <typename T>
void launch_bk3_diagonal(int degree, int nq, int64_t n_cells, const T *G, const uint32_t *map,
                         const Shape1D<double> &S_and_Sd, double *diag, hipStream_t s);
// out[perm[c]][q] = -(grad u)^T G_q (grad lambda)
template <typename T>
void launch_bk3_grad_coefficient(int degree, int nq, int64_t n_cells, const T *G,
                                 const uint32_t *map, const uint32_t *perm, const Shape1D<T> &sh,
                                 const T *u, const T *lam, T *out, hipStream_t s);

// Jacobi-PCG with the BK3 operator inside one workgroup (small coarse levels);
// info[0..2] = iterations, ||r0||, ||r|| (device doubles).  x holds the initial guess.
template <typename T>
void launch_bk3_coarse_pcg(int degree, int nq, int64_t n_cells, const T *G, const uint32_t *map,
                           const Shape1D<T> &sh, const uint32_t *cons, int64_t n_cons,
                           const T *pinv, int64_t n, const T *b, T *x, T *r, T *z, T *p, T *q,
                           double rel_tol, int max_it, double *info, hipStream_t s);

// ---------------------------------------------------------------- SIP-DG operator
// Face data of the DG kernel beyond Shape1D: exact derivatives of the nodal
// (Gauss-Lobatto) basis at the Gauss points, and at the cell ends 0 and 1.
template <typename T>
struct DgShape {
  T Sd[kMaxNm * kMaxNq];  // Sd[i*nq+q] = l_i'(xi_q)
  T dl[2 * kMaxNm];       // dl[s*nm+i] = l_i'(s), nodal (Gauss-Lobatto) basis
  T L[2 * kMaxNq];        // L[s*nq+r] = L_r(s), Lagrange basis on the Gauss points
  T Ld[2 * kMaxNq];       // Ld[s*nq+r] = L_r'(s)
};

// y (=, or += if add) = A_DG x; see pmg_dg.hip for the face data layout
template <typename T>
void launch_dg_apply(int degree, int nq, bool add, const T *x, T *y, const T *G,
                     const uint32_t *cface, const uint32_t *cnbr, const T *fdata, const T *fsig,
                     int64_t n_cells, const Shape1D<T> &sh, const DgShape<T> &fs, hipStream_t s);
// fdata[F][7][nq^2] from cell vertices; finfo[F] = (cell-, cell+ or invalid, face of cell-)
// SIP-DG box set-up on the device: connectivity cface/cnbr [C][6] and
// finfo [F][3] in the numbering of pmg_box_dg_faces (boundary id of side f:
// bit f of neumann_sides), cell h [C][6] = |K| / |F_f| by the 4-point Gauss
// rule (g4p, g4w) and the penalty sigma [F] = penalty_fac (1/h- + 1/h+)
void launch_dg_box_setup(const pmg_box &box, int neumann_sides, const double *cv,
                         const double *g4p, const double *g4w, double penalty_fac,
                         uint32_t *cface, uint32_t *cnbr, uint32_t *finfo, double *h,
                         double *sigma, int64_t n_faces, hipStream_t s);
template <typename T>
void launch_dg_face_geometry(int nq, int64_t n_faces, const uint32_t *finfo, const double *cv,
                             const double *qp, const double *qw, int metric, T *fdata,
                             hipStream_t s);
// DG -> CG restriction as a gather: dst[i] += sum of (src - sub)[list[off[i] .. off[i+1])]
// (sub optional)
template <typename T>
void launch_dg_cg_restrict_gather(int64_t n_cg, const uint32_t *off, const uint32_t *list,
                                  const T *src, T *dst, hipStream_t s, const T *sub = nullptr);
// DG(p) <-> CG(p): prolongate dst[dg(e)] += src[cg_map[e]], restrict
// dst[cg_map[e]] += (src - sub)[dg(e)] (sub optional); dg(e) = dg_map ? dg_map[e] : e;
// invalid CG entries skipped
template <typename T>
void launch_dg_cg_transfer(bool prolongate, int64_t n_entries, const uint32_t *cg_map,
                           const uint32_t *dg_map, const T *src, T *dst, hipStream_t s,
                           const T *sub = nullptr);

// ---------------------------------------------------------------- transfers
template <typename T>
struct Transfer1D {
  T M[kMaxNm * (2 * kMaxNm)];  // M[i*nmf + j]
};

// Owner form of a plain fine transfer map, built on the device: the weights
// of setup_weights (1/multiplicity, 0 on constrained fine DoFs) become bit 31
// on the first entry naming each unconstrained DoF.  cons: byte mask over the
// n_dofs fine DoFs; first: n_dofs words of scratch.  Synchronous on s.
void launch_owner_marks(int64_t nw, int64_t n_dofs, uint32_t *mf, const uint8_t *cons,
                        uint32_t *first, hipStream_t s);
// x-line words {first DoF, owner bits} of an owner-form map with lines of nmf
// entries; *bad becomes 1 when a line is not a run of consecutive DoFs
void launch_fine_lines(int64_t n_lines, int nmf, const uint32_t *mf, uint2 *fl, uint32_t *bad,
                       hipStream_t s);

// w == nullptr: owner mode (map_f bit 31 marks the owning entry of each fine
// DoF, see pmg_transfer.hip); otherwise weighted, w pre-scaled for the
// prolongation (scale is only used in owner mode).  fine_lines (owner mode,
// optional): the fine map as x-line words [cell][c][b] = {first DoF, owner
// bits over a}, used instead of map_f when every fine x-line is consecutive
template <typename T>
void launch_prolongate(int nmc, int nmf, int64_t n_cells, const Transfer1D<T> &M,
                       const uint32_t *map_c, const uint32_t *map_f, const T *w, T scale,
                       const T *xc, T *yf, hipStream_t s, const uint2 *fine_lines = nullptr);
// yc += R (xf - xf_sub); xf_sub optional.  cell_buf (optional): the per-cell
// coarse values are stored to cell_buf[map_c[entry]] instead of being added
// to yc, map_c then holding buffer positions in coarse-DoF order
// (launch_csr_positions; 0xFFFFFFFF: no store), and the caller sums each
// coarse DoF's range with launch_restrict_sum
template <typename T>
void launch_restrict(int nmc, int nmf, int64_t n_cells, const Transfer1D<T> &M,
                     const uint32_t *map_c, const uint32_t *map_f, const T *w, const T *xf,
                     T *yc, hipStream_t s, const T *xf_sub = nullptr, T *cell_buf = nullptr,
                     const uint2 *fine_lines = nullptr);
// yc[i] += sum of buf[off[i] .. off[i+1]), in order
template <typename T>
void launch_restrict_sum(int64_t n_c, const uint32_t *off, const T *buf, T *yc, hipStream_t s);
// pos[n_pos]: pos[list[k]] = k for the ne entries of an inverse map, every
// other position 0xFFFFFFFF.  Synchronous on s.
void launch_csr_positions(int64_t n_pos, int64_t ne, const uint32_t *list, uint32_t *pos,
                          hipStream_t s);

// ---------------------------------------------------------------- vectors
template <typename T> void launch_fill(int64_t n, T *x, T v, hipStream_t s);
// x[i] = (i mod 11) - mean (the eigenvalue estimate's start vector)
template <typename T> void launch_fill_mod11(int64_t n, T *x, double mean, hipStream_t s);
template <typename T> void launch_copy(int64_t n, T *dst, const T *src, hipStream_t s);
// dst[idx[i]] = src[idx[i]]
template <typename T>
void launch_copy_indexed(int64_t n, const uint32_t *idx, T *dst, const T *src, hipStream_t s);
// dst[i] = mask[i] ? src[i] : 0 (one pass: the zero fill and the copy of the
// constrained rows of a vmult)
template <typename T>
void launch_prep_constrained(int64_t n, const uint8_t *mask, T *dst, const T *src, hipStream_t s);
template <typename T>
void launch_set_indexed(int64_t n, const uint32_t *idx, T *dst, T v, hipStream_t s);
// x[i] = 1/x[i]
template <typename T> void launch_invert(int64_t n, T *x, hipStream_t s);
// smallest i with !(d[i] > 0), or -1 (synchronous)
template <typename T> int64_t first_nonpositive(int64_t n, const T *d, hipStream_t s);
template <typename TD, typename TS>
void launch_convert(int64_t n, TD *dst, const TS *src, hipStream_t s);
// the same, unless *skip != 0 on the device (skip may be null)
template <typename TD, typename TS>
void launch_convert_unless(int64_t n, TD *dst, const TS *src, const double *skip, hipStream_t s);
// partial sums of a.b into `partials` (kDotBlocks entries), then final sum into out[0]
constexpr int kDotBlocks = 4096;
// stream_a: read a non-temporally (its last use for a while)
template <typename TA, typename TB = TA>
void launch_dot(int64_t n, const TA *a, const TB *b, double *partials, double *out, hipStream_t s,
                bool stream_a = false, const double *skip = nullptr);
// t = b - t   (residual)
template <typename T> void launch_residual(int64_t n, T *t, const T *b, hipStream_t s);

// Chebyshev (first kind, Jacobi preconditioner); coefficient args are host scalars.
// mode 0: x = f2 * P b                                     (vmult start)
// mode 1: old = x; x = x + f2 * P (b - t)                   (step start)
// mode 2: xn = (1+f1) x - f1*old + f2 P (b - t); old = x; x = xn   (old = 0 if first)
// prep (optional, may equal t): prep = mask ? x_new : 0, the prepared output
// of the matvec that follows (mask null: no constrained rows)
template <typename T>
void launch_cheby_update(int mode, int64_t n, T *x, T *old, const T *b, const T *t,
                         const T *pinv, double f1, double f2, bool old_is_zero, hipStream_t s,
                         T *prep = nullptr, const uint8_t *mask = nullptr);

// CG kernels with device-resident scalars (sc = array of doubles, see pmg_solver.cpp)
template <typename T>
void launch_cg_update_xr(int64_t n, T *x, T *r, const T *p, const T *q, const double *sc,
                         double *partials, hipStream_t s);
template <typename T>
void launch_cg_precond_dot(int64_t n, const T *r, const T *pinv, T *z, double *partials,
                           hipStream_t s, const double *skip = nullptr);
// p = z (first) or z + beta p
template <typename T, typename TZ = T>
void launch_cg_update_p(int64_t n, T *p, const TZ *z, const double *sc, int first, hipStream_t s);
// CG with the x update deferred by one iteration (same arithmetic): r -= alpha q
// with r.r partials; x += alpha_prev p_prev (not if first) and p = z + beta p;
// the last x += alpha p after the loop.  alpha is kept in sc[4] by
// launch_cg_scalars(op 0).
// stage (optional): the new r also written there in stage_precision
template <typename T>
void launch_cg_update_r(int64_t n, T *r, const T *q, const double *sc, double *partials,
                        hipStream_t s, void *stage = nullptr, int stage_precision = PMG_FP64,
                        const double *skip = nullptr);
template <typename T, typename TZ = T>
void launch_cg_update_xp(int64_t n, T *x, T *p, const TZ *z, const double *sc, int first,
                         hipStream_t s, const double *skip = nullptr);
template <typename T>
void launch_cg_update_x(int64_t n, T *x, const T *p, const double *sc, hipStream_t s);
// final reduction of kDotBlocks partial sums into sc[slot] (optionally sqrt)
void launch_reduce_partials(const double *partials, int n_partials, double *sc, int slot,
                            int n_sums, hipStream_t s, const double *skip = nullptr);
// CG scalar slots: the iteration that converged and the converged flag
constexpr int kCgConvItSlot = 6, kCgConvergedSlot = 7;
// scalar bookkeeping for CG on the device: see kernel for slot layout
// also: sc[7] = 1 once sc[2] = r.r <= target2 (or NaN), sc[6] = host[1] = it
// then; host[0] = r.r (pinned, optional).  skip (the CG kernels above): the
// converged flag sc + 7; the kernels then change nothing.  op 2 clears the
// flag before a solve.
void launch_cg_scalars(double *sc, int op, hipStream_t s, double *host = nullptr,
                       double target2 = -1.0, int it = 0);
// launch_reduce_partials into sc[2] (gated by the converged flag) followed by
// launch_cg_scalars(op 0), in one launch
void launch_cg_rr_scalars(const double *partials, int n_partials, double *sc, hipStream_t s,
                          double *host, double target2, int it);

// Measured FMA issue rate of the device (flop/s) in the given precision:
// independent FMA chains in every lane of 16 workgroups per CU (measurement
// only, allocates a scratch array).  Synchronous.
double measure_fma_rate(int precision, hipStream_t s);

}  // namespace pmg
