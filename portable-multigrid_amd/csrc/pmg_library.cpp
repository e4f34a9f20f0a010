// pmg_library.cpp -- host-side C++ classes: operator, transfers, Chebyshev
// smoother, V-cycle and preconditioned CG.  Every device operation is a
// launch of a gfx950 kernel from pmg_*.hip on the caller's stream; host code
// only sequences launches and reads back the scalars a solver must test.
#include "pmg_library.hpp"

#include "pmg_cellgroups.hpp"
#include "pmg_fields.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <type_traits>

namespace pmg {

namespace {
template <typename T>
Shape1D<T> to_shape(const std::vector<double> &S, const std::vector<double> &D) {
  Shape1D<T> sh{};
  PMG_REQUIRE(S.size() <= size_t(kMaxNm * kMaxNq) && D.size() <= size_t(kMaxNq * kMaxNq),
              PMG_ERR_UNSUPPORTED, "1D shape data exceeds kernel limits");
  for (size_t i = 0; i < S.size(); ++i) sh.S[i] = static_cast<T>(S[i]);
  for (size_t i = 0; i < D.size(); ++i) sh.D[i] = static_cast<T>(D[i]);
  return sh;
}

template <typename T>
double dot_sync(int64_t n, const T *a, const T *b, DeviceArray<double> &partials,
                DeviceArray<double> &sc, hipStream_t s) {
  launch_dot<T>(n, a, b, partials.ptr, sc.ptr, s);
  double v = 0;
  PMG_HIP_CHECK(hipMemcpyAsync(&v, sc.ptr, sizeof(double), hipMemcpyDeviceToHost, s));
  PMG_HIP_CHECK(hipStreamSynchronize(s));
  return v;
}

// Symmetric tridiagonal eigenvalues by Sturm-sequence bisection.
std::vector<double> tridiag_eigenvalues(const std::vector<double> &d, const std::vector<double> &o) {
  const int n = static_cast<int>(d.size());
  std::vector<double> ev(n);
  if (n == 0) return ev;
  double lo = d[0], hi = d[0];
  for (int i = 0; i < n; ++i) {
    const double r = (i > 0 ? std::fabs(o[i - 1]) : 0.0) + (i + 1 < n ? std::fabs(o[i]) : 0.0);
    lo = std::min(lo, d[i] - r);
    hi = std::max(hi, d[i] + r);
  }
  auto count_below = [&](double x) {
    int c = 0;
    double q = d[0] - x;
    if (q < 0) ++c;
    for (int i = 1; i < n; ++i) {
      if (q == 0) q = 1e-300;
      q = d[i] - x - o[i - 1] * o[i - 1] / q;
      if (q < 0) ++c;
    }
    return c;
  };
  for (int k = 0; k < n; ++k) {
    double a = lo, b = hi;
    for (int it = 0; it < 200 && b - a > 1e-15 * (std::fabs(a) + std::fabs(b)) + 1e-300; ++it) {
      const double m = 0.5 * (a + b);
      if (count_below(m) > k) b = m; else a = m;
    }
    ev[k] = 0.5 * (a + b);
  }
  return ev;
}
}  // namespace

// ============================================================ operator
// Compressed indices: a cell whose every x-line (j,k) is a run of consecutive
// unconstrained DoFs needs only the line starts (nm^2 instead of nm^3 index
// loads).  Built on the device from the device map; kept, and true, when at
// least half of the cells qualify (on lexicographic numberings all but the
// boundary cells do).
static bool compress_lines(const uint32_t *map, int64_t n_cells, int nm, int64_t n_dofs,
                           DeviceArray<uint32_t> &lines_out, DeviceArray<uint8_t> &flag_out) {
  if (n_cells <= 0) return false;
  lines_out.resize(static_cast<size_t>(n_cells * nm * nm));
  flag_out.resize(static_cast<size_t>(n_cells));
  const int64_t n_comp =
      compress_lines_device(n_cells, nm, map, n_dofs, lines_out.ptr, flag_out.ptr, nullptr);
  if (2 * n_comp >= n_cells) return true;
  lines_out.release();
  flag_out.release();
  return false;
}

void LaplaceOperatorBK3::setup_shape() {
  check_precision(precision_);
  PMG_REQUIRE(degree_ >= 1 && degree_ <= 6, PMG_ERR_UNSUPPORTED, "degree must be in [1,6]");
  PMG_REQUIRE(nq_ == degree_ + 1 || nq_ == degree_ + 2, PMG_ERR_UNSUPPORTED,
              "n_q_points_1d must be degree+1 or degree+2");
  PMG_REQUIRE(n_cells_ >= 0 && n_dofs_ >= 0 && n_constrained_ >= 0, PMG_ERR_INVALID_ARG,
              "negative size");
  PMG_REQUIRE(n_dofs_ < int64_t(0xFFFFFFFFu), PMG_ERR_SIZE, "n_dofs exceeds 32-bit DoF maps");
  shape_ = make_shape_data(degree_, nq_);
  sh64_ = to_shape<double>(shape_.S, shape_.D);
  sh32_ = to_shape<float>(shape_.S, shape_.D);
  fill_even_odd(sh64_, shape_.S.data(), shape_.D.data(), degree_ + 1, shape_.n_q_points_1d);
  fill_even_odd(sh32_, shape_.S.data(), shape_.D.data(), degree_ + 1, shape_.n_q_points_1d);
  for (int q = 0; q < nq_; ++q) {
    sh64_.W[q] = shape_.quad.weights[q];
    sh32_.W[q] = static_cast<float>(shape_.quad.weights[q]);
  }
  shd64_ = to_shape<double>(shape_.S, shape_.Sd);  // exact derivatives in the D slot
}

// Compressed lines and the default scatter schedule of the device map.
void LaplaceOperatorBK3::setup_schedule() {
  if (compress_lines(map_.ptr, n_cells_, degree_ + 1, n_dofs_, lines_, lflag_)) {
    sched_.lines = lines_.ptr;
    sched_.lflag = lflag_.ptr;
  }
  // Default scatter schedule: the patch schedule (with early non-temporal G
  // loads) for FP64 nq <= 5 and FP32 nq <= 7 on operators of more than 4096
  // cells, the per-cell atomic scatter (late G loads) otherwise.  From vmult
  // timings on MI355X over the level types of the MG hierarchies, patch vs
  // atomic (us): FP64 Q4 nq 5 32^3 59.7 / 78.6, 64^3 489.5 / 507.3, 128^3
  // 3567 / 4140; FP64 Q3 nq 4 48^3 89.4 / 127.0; FP64 Q6 nq 7 48^3 728.8 /
  // 571.2 (atomic); FP32 Q4 nq 5 64^3 295.3 / 439.6; FP32 Q3 nq 4 48^3 70.5 /
  // 122.9; FP32 Q6 nq 7 24^3 59.1 / 66.4, 48^3 370.8 / 399.0.  The gather
  // and grid schedules were slower or within 3 % (grid FP32 Q4 64^3 291.6,
  // Q3 48^3 66.3) and need a cell buffer.  Levels of at most 4096 cells are
  // launch-bound: there the atomic kernel is as fast or faster (Q4 FP32
  // 2^3..16^3: 8.3-11.7 vs 8.8-13.9 us per launch).
  // PMG_BK3_SCHEDULE=atomic|colored|patch|gather|grid and PMG_BK3_GLOAD=0|1|2
  // override.
  const char *env = std::getenv("PMG_BK3_SCHEDULE");
  const char *genv = std::getenv("PMG_BK3_GLOAD");
  int kind = nq_ <= (precision_ == PMG_FP32 ? 7 : 5) && n_cells_ > 4096 ? kPatch : kAtomic;
  if (env)
    kind = std::string(env) == "colored"  ? kColored
           : std::string(env) == "atomic" ? kAtomic
           : std::string(env) == "gather" ? kGather
           : std::string(env) == "grid"   ? kGrid
                                          : kPatch;
  set_schedule(kind, genv ? std::atoi(genv) : -1);
}

// G of the trilinear cells from the device vertices (internal cell order).
void LaplaceOperatorBK3::setup_geometry() {
  G_.reinit(precision_, g_size(nq_, n_cells_));
  qpts_.upload(shape_.quad.points);
  qwts_.upload(shape_.quad.weights);
  if (precision_ == PMG_FP64) fill_geometry_t<double>(nullptr, nullptr);
  else fill_geometry_t<float>(nullptr, nullptr);
  PMG_HIP_CHECK(hipDeviceSynchronize());
}

LaplaceOperatorBK3::LaplaceOperatorBK3(int degree, int nq, int precision, int64_t n_cells,
                                       const uint32_t *dofmap, int64_t n_dofs,
                                       const uint32_t *constrained, int64_t n_constrained,
                                       const double *cell_vertices, const double *G)
    : degree_(degree), nq_(nq), precision_(precision), n_cells_(n_cells), n_dofs_(n_dofs),
      n_constrained_(n_constrained) {
  setup_shape();
  PMG_REQUIRE(n_cells == 0 || dofmap, PMG_ERR_INVALID_ARG, "dofmap is NULL");
  PMG_REQUIRE(n_constrained == 0 || constrained, PMG_ERR_INVALID_ARG, "constrained list is NULL");
  PMG_REQUIRE((cell_vertices != nullptr) != (G != nullptr), PMG_ERR_INVALID_ARG,
              "give exactly one of cell_vertices and G");
  const int64_t nm3 = int64_t(degree + 1) * (degree + 1) * (degree + 1);
  const int64_t nq3 = int64_t(nq) * nq * nq;
  // Internal cell order.  Default: the caller's order (on lexicographic
  // structured meshes a batch is a run of cells along x, whose DoF rows are
  // long and contiguous -- measured fastest for the atomic scatter).
  // PMG_CELL_ORDER=grouped forms compact groups of E cells that share many
  // DoFs (group_cells), =morton sorts by the Morton key of the cell centroids.
  // perm_ maps internal -> caller cell index; cell-indexed user data (G,
  // kappa, the coefficient gradient) is translated through it.
  std::vector<uint32_t> perm;
  const char *order_env = std::getenv("PMG_CELL_ORDER");
  const std::string order_kind = order_env ? order_env : "caller";
  if (order_kind == "grouped" && n_cells > 1) {
    // group_cells reads the map: check its range on the host first
    for (int64_t e = 0; e < n_cells * nm3; ++e)
      PMG_REQUIRE(dofmap[e] == kInvalid || int64_t(dofmap[e]) < n_dofs, PMG_ERR_INVALID_ARG,
                  "dofmap entry out of range at " + std::to_string(e));
    perm = group_cells(n_cells, nm3, dofmap, n_dofs, g_batch_cells(nq));
  }
  if (cell_vertices && n_cells > 1 && order_kind == "morton") {
    perm.resize(static_cast<size_t>(n_cells));
    for (int64_t c = 0; c < n_cells; ++c) perm[c] = static_cast<uint32_t>(c);
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    std::vector<double> cen(static_cast<size_t>(3 * n_cells));
    for (int64_t c = 0; c < n_cells; ++c)
      for (int d = 0; d < 3; ++d) {
        double m = 0;
        for (int v = 0; v < 8; ++v) m += cell_vertices[c * 24 + 3 * v + d];
        cen[3 * c + d] = m / 8;
        lo[d] = std::min(lo[d], m / 8);
        hi[d] = std::max(hi[d], m / 8);
      }
    auto spread = [](uint64_t v) {  // 21 bits -> every third bit
      v &= 0x1fffff;
      v = (v | v << 32) & 0x1f00000000ffffULL;
      v = (v | v << 16) & 0x1f0000ff0000ffULL;
      v = (v | v << 8) & 0x100f00f00f00f00fULL;
      v = (v | v << 4) & 0x10c30c30c30c30c3ULL;
      v = (v | v << 2) & 0x1249249249249249ULL;
      return v;
    };
    std::vector<uint64_t> key(static_cast<size_t>(n_cells));
    for (int64_t c = 0; c < n_cells; ++c) {
      uint64_t q[3];
      for (int d = 0; d < 3; ++d) {
        const double ext = hi[d] - lo[d];
        const double f = ext > 0 ? (cen[3 * c + d] - lo[d]) / ext : 0.0;
        q[d] = static_cast<uint64_t>(std::min(std::max(f, 0.0), 1.0) * ((1u << 20) - 1));
      }
      key[c] = spread(q[0]) | (spread(q[1]) << 1) | (spread(q[2]) << 2);
    }
    std::stable_sort(perm.begin(), perm.end(), [&](uint32_t a, uint32_t b) { return key[a] < key[b]; });
  }
  caller_order_ = true;
  for (size_t c = 0; c < perm.size() && caller_order_; ++c) caller_order_ = perm[c] == uint32_t(c);
  if (caller_order_) {
    perm_.resize(static_cast<size_t>(n_cells));
    iota_device(perm_.ptr, n_cells, nullptr);
  } else {
    perm_.upload(perm);
  }
  // constrained DoFs are eliminated (rows and columns), as the reference's
  // setup_dof_indices_per_color does: a map entry naming a constrained DoF
  // is treated as invalid.  Range checks, the mask and the elimination run
  // on the device.
  constrained_.upload(constrained, static_cast<size_t>(n_constrained));
  if (n_constrained > 0) {
    cmask_.resize(static_cast<size_t>(n_dofs));
    PMG_REQUIRE(mask_from_list_device(n_constrained, constrained_.ptr, n_dofs, cmask_.ptr,
                                      nullptr) < 0,
                PMG_ERR_INVALID_ARG, "constrained DoF out of range");
  }
  map_.upload(dofmap, static_cast<size_t>(n_cells * nm3));
  {
    DeviceArray<uint32_t> raw;
    if (!caller_order_) {
      raw = std::move(map_);
      map_.resize(static_cast<size_t>(n_cells * nm3));
    }
    const int64_t bad = map_eliminate_device(n_cells, static_cast<int>(nm3),
                                             caller_order_ ? map_.ptr : raw.ptr,
                                             caller_order_ ? nullptr : perm_.ptr, cmask_.ptr,
                                             n_dofs, map_.ptr, nullptr);
    PMG_REQUIRE(bad < 0, PMG_ERR_INVALID_ARG, "dofmap entry out of range at " + std::to_string(bad));
  }
  setup_schedule();
  if (cell_vertices) {
    if (caller_order_) {
      vertices_.upload(cell_vertices, static_cast<size_t>(n_cells * 24));
    } else {
      std::vector<double> iv(static_cast<size_t>(n_cells * 24));
      for (int64_t c = 0; c < n_cells; ++c)
        std::copy(cell_vertices + perm[c] * 24, cell_vertices + (perm[c] + 1) * 24, iv.begin() + c * 24);
      vertices_.upload(iv);
    }
    setup_geometry();
  } else {
    G_.reinit(precision, g_size(nq, n_cells));
    // caller layout [cell][comp][q] -> batch-interleaved device layout (padding 0)
    auto relayout = [&](auto *dst) {
      for (int64_t c = 0; c < n_cells; ++c) {
        const int64_t from = caller_order_ ? c : int64_t(perm[c]);
        for (int comp = 0; comp < 6; ++comp)
          for (int q = 0; q < nq3; ++q)
            dst[g_index<std::remove_pointer_t<decltype(dst)>>(nq, c, comp, q)] =
                static_cast<std::remove_pointer_t<decltype(dst)>>(
                G[(from * 6 + comp) * nq3 + q]);
      }
    };
    if (precision == PMG_FP64) {
      std::vector<double> g(static_cast<size_t>(g_size(nq, n_cells)), 0.0);
      relayout(g.data());
      PMG_HIP_CHECK(hipMemcpy(G_.data(), g.data(), G_.raw.bytes(), hipMemcpyHostToDevice));
    } else {
      std::vector<float> g(static_cast<size_t>(g_size(nq, n_cells)), 0.0f);
      relayout(g.data());
      PMG_HIP_CHECK(hipMemcpy(G_.data(), g.data(), G_.raw.bytes(), hipMemcpyHostToDevice));
    }
  }
}

// The operator of a structured box with every boundary DoF constrained, its
// index data generated on the device: the lexicographic map with boundary
// entries invalid, the ascending boundary list, the vertices; cells in the
// box's lexicographic order.  Declares the box structure (grid schedule)
// where its index limits allow.
LaplaceOperatorBK3::LaplaceOperatorBK3(const pmg_box &box, int degree, int nq, int precision)
    : degree_(degree), nq_(nq), precision_(precision),
      n_cells_(int64_t(box.nx) * box.ny * box.nz),
      n_dofs_((int64_t(degree) * box.nx + 1) * (int64_t(degree) * box.ny + 1) *
              (int64_t(degree) * box.nz + 1)),
      n_constrained_(0) {
  PMG_REQUIRE(box.nx > 0 && box.ny > 0 && box.nz > 0, PMG_ERR_INVALID_ARG,
              "box needs at least one cell per direction");
  PMG_REQUIRE(box.deform >= 0 && box.deform <= 2, PMG_ERR_INVALID_ARG, "unknown deformation");
  setup_shape();
  n_constrained_ = box_boundary_count(box, degree);
  const int64_t nm3 = int64_t(degree + 1) * (degree + 1) * (degree + 1);
  map_.resize(static_cast<size_t>(n_cells_ * nm3));
  box_map_device(box, degree, 1, true, map_.ptr, nullptr);
  constrained_.resize(static_cast<size_t>(n_constrained_));
  cmask_.resize(static_cast<size_t>(n_dofs_));
  box_boundary_device(box, degree, constrained_.ptr, cmask_.ptr, nullptr);
  perm_.resize(static_cast<size_t>(n_cells_));
  iota_device(perm_.ptr, n_cells_, nullptr);
  caller_order_ = true;
  vertices_.resize(static_cast<size_t>(n_cells_ * 24));
  box_vertices_device(box, vertices_.ptr, nullptr);
  const int64_t p = degree;
  if (p * box.ny + 1 < 65536 && p * box.nz + 1 < 65536 && n_cells_ < (int64_t(1) << 31))
    for (int d = 0; d < 3; ++d) sched_.grid[d] = d == 0 ? box.nx : d == 1 ? box.ny : box.nz;
  setup_schedule();
  setup_geometry();
}

// Greedy cell colouring (cells sharing a DoF get different colours), cells
// sorted by colour: the colored schedule launches one colour at a time.  Built
// on first use from the device map.
void LaplaceOperatorBK3::build_coloring() {
  const int64_t nm3 = int64_t(degree_ + 1) * (degree_ + 1) * (degree_ + 1);
  std::vector<uint32_t> imap(static_cast<size_t>(n_cells_ * nm3));
  if (!imap.empty())
    PMG_HIP_CHECK(hipMemcpy(imap.data(), map_.ptr, imap.size() * sizeof(uint32_t),
                            hipMemcpyDeviceToHost));
  const int64_t n_cells = n_cells_;
  std::vector<uint64_t> used(static_cast<size_t>(n_dofs_), 0);
  std::vector<int> color(static_cast<size_t>(n_cells), -1);
  bool ok = true;
  int n_col = 0;
  for (int64_t c = 0; c < n_cells && ok; ++c) {
    uint64_t forbidden = 0;
    for (int64_t i = 0; i < nm3; ++i) {
      const uint32_t d = imap[c * nm3 + i];
      if (d != kInvalid) forbidden |= used[d];
    }
    int col = 0;
    while (col < 64 && ((forbidden >> col) & 1u)) ++col;
    if (col == 64) { ok = false; break; }
    color[c] = col;
    n_col = std::max(n_col, col + 1);
    for (int64_t i = 0; i < nm3; ++i) {
      const uint32_t d = imap[c * nm3 + i];
      if (d != kInvalid) used[d] |= (uint64_t(1) << col);
    }
  }
  if (ok && n_cells > 0) {
    std::vector<int64_t> off(n_col + 1, 0);
    for (int64_t c = 0; c < n_cells; ++c) ++off[color[c] + 1];
    for (int k = 0; k < n_col; ++k) off[k + 1] += off[k];
    std::vector<uint32_t> ids(static_cast<size_t>(n_cells));
    std::vector<int64_t> fill(off.begin(), off.end() - 1);
    for (int64_t c = 0; c < n_cells; ++c) ids[fill[color[c]]++] = static_cast<uint32_t>(c);
    cell_ids_.upload(ids);
    sched_.cell_ids = cell_ids_.ptr;
    sched_.color_offsets = off;
  }
}

// DoF patches of the batches (patch schedule) of a device map in device cell
// order, built on the device into pa, and sc points at them; sc is left
// without patches when a batch's entries exceed the builder's limit.
// The cells of a batch sum into its LDS patch colour by colour at FP32: its
// LDS float atomics cost about a fifth of the kernel (vmult with plain LDS
// stores in their place, timing only: Q4 64^3 325 -> 257 us, Q6 48^3 385 ->
// 315 us, Q3 48^3 72 -> 53 us), while at FP64 they cost nothing measurable
// (493 -> 492 us) and keep the single barrier.  PMG_PATCH_ACCUM=atomic|colored
// overrides.
static void build_patch_schedule(const uint32_t *map, int64_t n_cells, int degree, int nq,
                                 int precision, int64_t n_dofs, PatchArrays &pa,
                                 Bk3Schedule &sc) {
  const int nm = degree + 1;
  const int nm3 = nm * nm * nm;
  const int E = g_batch_cells(nq);
  if (!patch_schedule_fits(n_cells, nm3, E, n_dofs)) return;
  const int64_t nb = (n_cells + E - 1) / E;
  pa.pdof.resize(static_cast<size_t>(nb * E * nm3));
  pa.pcount.resize(static_cast<size_t>(nb));
  pa.pidx.resize(static_cast<size_t>(n_cells * nm3));
  pa.pcolor.resize(static_cast<size_t>(n_cells));
  pa.pncol.resize(static_cast<size_t>(nb));
  build_patches_device(n_cells, nm3, E, map, n_dofs, pa.pdof.ptr, pa.pcount.ptr, pa.pidx.ptr,
                       pa.pcolor.ptr, pa.pncol.ptr, nullptr);
  sc.pdof = pa.pdof.ptr;
  sc.pcount = pa.pcount.ptr;
  sc.pidx = pa.pidx.ptr;
  // the patch entries as runs, 8 words per 64 slots instead of 256: used when
  // at least half of the chunks are at most 4 runs (box numberings: all but
  // a few)
  const int64_t pmax = int64_t(E) * nm3, n_chunks = nb * ((pmax + 63) / 64);
  pa.pchunk.resize(static_cast<size_t>(n_chunks * 8));
  const int64_t n_runs =
      build_patch_chunks_device(nb, static_cast<int>(pmax), pa.pdof.ptr, pa.pcount.ptr,
                                pa.pchunk.ptr, nullptr);
  if (2 * n_runs >= n_chunks) sc.pchunk = pa.pchunk.ptr;
  else pa.pchunk.release();
  bool colored = precision == PMG_FP32;
  if (const char *env = std::getenv("PMG_PATCH_ACCUM")) colored = std::string(env) == "colored";
  sc.pcolor = colored ? pa.pcolor.ptr : nullptr;
  sc.pncol = colored ? pa.pncol.ptr : nullptr;
  // x-lines whose patch slots are consecutive need only their first slot
  // (nm^2 instead of nm^3 16-bit loads per cell); used when at least half of
  // the cells qualify
  pa.pline.resize(static_cast<size_t>(n_cells * nm * nm));
  pa.pflag.resize(static_cast<size_t>(n_cells));
  const int64_t n_comp =
      compress_patch_lines_device(n_cells, nm, pa.pidx.ptr, pa.pline.ptr, pa.pflag.ptr, nullptr);
  if (2 * n_comp >= n_cells) {
    sc.pline = pa.pline.ptr;
    sc.pflag = pa.pflag.ptr;
  } else {
    pa.pline.release();
    pa.pflag.release();
  }
}

// The matrices of the cell apply's x and y sweeps in the operand form of the
// matrix-core contractions: six zero-padded 16 x 16 tables A[m * 16 + k],
// out[m] = sum_k A[m][k] in[k], with S[i][q] = shape.S[i nq + q] and
// D[n][q] = shape.D[n nq + q]; products formed in double and rounded once.
static std::vector<float> fill_mx_tables(const ShapeData &shape, int nm, int nq) {
  std::vector<float> a(6 * 256, 0.f);
  auto S = [&](int i, int q) { return shape.S[i * nq + q]; };
  auto D = [&](int n, int q) { return shape.D[n * nq + q]; };
  auto set = [&](int t, int m, int k, double v) { a[t * 256 + m * 16 + k] = static_cast<float>(v); };
  for (int m = 0; m < nq; ++m)
    for (int k = 0; k < nm; ++k) {
      set(0, m, k, S(k, m));  // values at the Gauss points
      double sd = 0;
      for (int n = 0; n < nq; ++n) sd += S(k, n) * D(n, m);
      set(1, m, k, sd);  // their derivative, from the nodal values
    }
  for (int m = 0; m < nq; ++m)
    for (int k = 0; k < nq; ++k) {
      set(2, m, k, D(k, m));  // derivative at the Gauss points
      set(3, m, k, D(m, k));  // test with the derivative
    }
  for (int m = 0; m < nm; ++m)
    for (int k = 0; k < nq; ++k) {
      double sd = 0;
      for (int q = 0; q < nq; ++q) sd += S(m, q) * D(q, k);
      set(4, m, k, sd);       // test with the basis after the derivative
      set(4, m, nq + k, S(m, k));  // test with the basis
      set(5, m, k, S(m, k));
    }
  return a;
}

// The main map's patches, built on first use.  PMG_BK3_MFMA=1 runs the x and
// y sweeps of the FP32 patch kernel at nq >= 6 on the matrix cores.
void LaplaceOperatorBK3::build_patches_once() {
  build_patch_schedule(map_.ptr, n_cells_, degree_, nq_, precision_, n_dofs_, patches_, sched_);
  const char *env = std::getenv("PMG_BK3_MFMA");
  if (env && std::string(env) == "1" && precision_ == PMG_FP32 && nq_ >= 6 && 2 * nq_ <= 16) {
    patches_.mxa.upload(fill_mx_tables(shape_, degree_ + 1, nq_));
    sched_.mxa = patches_.mxa.ptr;
  }
}

void LaplaceOperatorBK3::set_box_structure(int nx, int ny, int nz) {
  PMG_REQUIRE(nx > 0 && ny > 0 && nz > 0, PMG_ERR_INVALID_ARG, "box dimensions must be positive");
  PMG_REQUIRE(int64_t(nx) * ny * nz == n_cells_, PMG_ERR_SIZE,
              "box dimensions do not match the cell count");
  const int64_t p = degree_;
  PMG_REQUIRE((p * nx + 1) * (p * ny + 1) * (p * nz + 1) == n_dofs_, PMG_ERR_SIZE,
              "box dimensions do not match the DoF count");
  PMG_REQUIRE(caller_order_, PMG_ERR_UNSUPPORTED,
              "the box structure needs the caller's cell order (PMG_CELL_ORDER unset)");
  PMG_REQUIRE(p * ny + 1 < 65536 && p * nz + 1 < 65536 && n_cells_ < (int64_t(1) << 31),
              PMG_ERR_SIZE, "box too large for the grid schedule");
  const int dims[3] = {nx, ny, nz};
  PMG_REQUIRE(check_box_map(degree_, dims, n_cells_, map_.ptr, cmask_.ptr, nullptr),
              PMG_ERR_INVALID_ARG,
              "the DoF map is not the lexicographic box map of these dimensions");
  for (int d = 0; d < 3; ++d) sched_.grid[d] = dims[d];
}

// The cell buffer of the gather and grid schedules (n_cells nm^3 entries).
bool LaplaceOperatorBK3::ensure_cell_buffer() {
  const int64_t nm3 = int64_t(degree_ + 1) * (degree_ + 1) * (degree_ + 1);
  if (n_cells_ == 0 || n_cells_ * nm3 >= (int64_t(1) << 32) || n_dofs_ >= (int64_t(1) << 31))
    return false;
  gbuf_.reinit(precision_, n_cells_ * nm3);
  sched_.gbuf = gbuf_.data();
  return true;
}

// The gather schedule's inverse map (built on the device) and cell buffer.
void LaplaceOperatorBK3::build_gather_once() {
  if (sched_.goff || n_cells_ == 0) return;
  const int64_t nm3 = int64_t(degree_ + 1) * (degree_ + 1) * (degree_ + 1);
  if (!ensure_cell_buffer()) return;
  goff_.resize(static_cast<size_t>(n_dofs_ + 1));
  uint32_t *list = nullptr;
  int64_t ne = 0;
  build_inverse_map(n_cells_ * nm3, map_.ptr, n_dofs_, goff_.ptr, &list, &ne, nullptr);
  glist_.release();
  glist_.ptr = list;
  glist_.n = static_cast<size_t>(ne);
  sched_.goff = goff_.ptr;
  sched_.glist = glist_.ptr;
  sched_.gn = n_dofs_;
}

LaplaceOperatorBK3::IndexArray LaplaceOperatorBK3::index_array(int which) const {
  auto arr = [](const auto &a) {
    return IndexArray{a.ptr, a.ptr ? static_cast<int64_t>(a.n) : 0,
                      static_cast<int>(sizeof(*a.ptr))};
  };
  switch (which) {
    case PMG_INDEX_MAP: return arr(map_);
    case PMG_INDEX_CONSTRAINED: return arr(constrained_);
    case PMG_INDEX_CELL_ORDER: return arr(perm_);
    case PMG_INDEX_LINES: return arr(lines_);
    case PMG_INDEX_LINE_FLAGS: return arr(lflag_);
    case PMG_INDEX_PATCH_DOFS: return arr(patches_.pdof);
    case PMG_INDEX_PATCH_COUNTS: return arr(patches_.pcount);
    case PMG_INDEX_PATCH_SLOTS: return arr(patches_.pidx);
    case PMG_INDEX_PATCH_LINES: return arr(patches_.pline);
    case PMG_INDEX_PATCH_LINE_FLAGS: return arr(patches_.pflag);
    case PMG_INDEX_VERTICES: return arr(vertices_);
    case PMG_INDEX_PATCH_COLORS: return arr(patches_.pcolor);
    case PMG_INDEX_PATCH_NCOLORS: return arr(patches_.pncol);
    case PMG_INDEX_PATCH_CHUNKS: return arr(patches_.pchunk);
    default: throw Error(PMG_ERR_INVALID_ARG, "unknown index array");
  }
}

void LaplaceOperatorBK3::set_schedule(int kind, int gmode) {
  PMG_REQUIRE(kind >= kAtomic && kind <= kGrid, PMG_ERR_INVALID_ARG, "unknown schedule");
  if (kind == kPatch && sched_.pdof == nullptr) build_patches_once();
  if (kind == kColored && sched_.color_offsets.empty()) build_coloring();
  if (kind == kGather && sched_.goff == nullptr) build_gather_once();
  if (kind == kGrid && sched_.grid[0] > 0 && sched_.gbuf == nullptr) ensure_cell_buffer();
  if (kind == kColored && sched_.color_offsets.empty()) kind = kAtomic;
  if (kind == kPatch && sched_.pdof == nullptr) kind = kAtomic;
  if (kind == kGather && sched_.goff == nullptr) kind = kAtomic;
  if (kind == kGrid && (sched_.grid[0] == 0 || sched_.gbuf == nullptr)) kind = kAtomic;
  sched_.kind = kind;
  // -1: the schedule's measured best G-load policy (the patch kernel's first
  // phase only gathers indices and x, which the G stream then overlaps)
  sched_.gmode = gmode >= 0 ? gmode : (kind == kPatch ? 0 : 2);  // 0 early NT, 2 late NT
}

template <typename T>
void LaplaceOperatorBK3::fill_geometry_t(const double *kappa, hipStream_t s) {
  launch_geometry<T>(nq_, n_cells_, vertices_.ptr, qpts_.ptr, qwts_.ptr, kappa, perm_.ptr,
                     static_cast<T *>(G_.data()), s);
}

template <typename T>
void LaplaceOperatorBK3::vmult_t(T *dst, const T *src, hipStream_t s) const {
  const Shape1D<T> &sh = *reinterpret_cast<const Shape1D<T> *>(
      std::is_same<T, double>::value ? static_cast<const void *>(&sh64_) : static_cast<const void *>(&sh32_));
  // dst = 0 on free rows, dst = src on constrained ones; the cell kernels
  // never write constrained rows.  Small vectors: one fused pass before the
  // cells (saves a launch, which dominates there); large ones: a zero fill
  // and an indexed copy of the constrained rows (fewer bytes than the fused
  // pass's byte mask and conditional loads, measured faster at 17M DoFs; a
  // bit mask instead of the byte mask lost too: Q4 64^3 FP64 vmult 417 ->
  // 437 us, its dependent mask load makes the pure write stream wait; so did
  // finding a box operator's boundary rows from the index, 420 -> 428 us).
  // The gather and grid schedules overwrite every free row themselves: no fill.
  const bool gather = sched_.kind == kGather || sched_.kind == kGrid;
  const bool fused = !gather && n_constrained_ > 0 && n_dofs_ <= kFusedPrepMaxDofs;
  if (fused) launch_prep_constrained<T>(n_dofs_, cmask_.ptr, dst, src, s);
  else if (!gather) launch_fill<T>(n_dofs_, dst, T(0), s);
  launch_bk3_cells<T>(degree_, nq_, src, dst, static_cast<const T *>(G_.data()), map_.ptr,
                      n_cells_, sh, s, &sched_, /*assign=*/true);
  if (!fused) launch_copy_indexed<T>(n_constrained_, constrained_.ptr, dst, src, s);
}

void LaplaceOperatorBK3::vmult(void *dst, const void *src, hipStream_t s) const {
  PMG_REQUIRE(dst && src, PMG_ERR_INVALID_ARG, "NULL vector");
  PMG_REQUIRE(dst != src, PMG_ERR_INVALID_ARG, "vmult: dst and src must not alias");
  if (precision_ == PMG_FP64) vmult_t(static_cast<double *>(dst), static_cast<const double *>(src), s);
  else vmult_t(static_cast<float *>(dst), static_cast<const float *>(src), s);
}

void LaplaceOperatorBK3::set_subdomain(const uint32_t *dofmap, const uint32_t *physical,
                                       int64_t n_physical, const uint32_t *interface_cells,
                                       int64_t n_interface) {
  PMG_REQUIRE(dofmap, PMG_ERR_INVALID_ARG, "subdomain dofmap is NULL");
  PMG_REQUIRE(n_physical >= 0 && (physical || n_physical == 0), PMG_ERR_INVALID_ARG,
              "bad physical boundary DoF list");
  PMG_REQUIRE(n_interface >= 0 && (interface_cells || n_interface == 0), PMG_ERR_INVALID_ARG,
              "bad interface cell list");
  const int nm = degree_ + 1;
  const int64_t nm3 = int64_t(nm) * nm * nm;
  // the plain indices (setup of plain_dof_indices_per_color,
  // portable_subdomain_laplace_operator.h:1056-1061), in device cell order,
  // physical boundary DoFs eliminated: range checks, mask and elimination on
  // the device
  DeviceArray<uint32_t> dphys;
  dphys.upload(physical, static_cast<size_t>(n_physical));
  DeviceArray<uint8_t> pmask(static_cast<size_t>(std::max<int64_t>(n_dofs_, 1)));
  PMG_REQUIRE(mask_from_list_device(n_physical, dphys.ptr, n_dofs_, pmask.ptr, nullptr) < 0,
              PMG_ERR_INVALID_ARG, "physical boundary DoF out of range");
  DeviceArray<uint32_t> smap, raw;
  if (caller_order_) {
    smap.upload(dofmap, static_cast<size_t>(n_cells_ * nm3));
  } else {
    raw.upload(dofmap, static_cast<size_t>(n_cells_ * nm3));
    smap.resize(static_cast<size_t>(n_cells_ * nm3));
  }
  PMG_REQUIRE(map_eliminate_device(n_cells_, static_cast<int>(nm3),
                                   caller_order_ ? smap.ptr : raw.ptr,
                                   caller_order_ ? nullptr : perm_.ptr, pmask.ptr, n_dofs_,
                                   smap.ptr, nullptr) < 0,
              PMG_ERR_INVALID_ARG, "subdomain dofmap entry out of range");
  // interface cells: caller index -> device position, ascending
  std::vector<uint32_t> ids(static_cast<size_t>(n_interface));
  for (int64_t i = 0; i < n_interface; ++i)
    PMG_REQUIRE(int64_t(interface_cells[i]) < n_cells_, PMG_ERR_INVALID_ARG,
                "interface cell out of range");
  if (caller_order_) {
    std::copy(interface_cells, interface_cells + n_interface, ids.begin());
  } else {
    const std::vector<uint32_t> perm = perm_.download();
    std::vector<uint32_t> where(static_cast<size_t>(n_cells_));
    for (int64_t c = 0; c < n_cells_; ++c) where[perm[c]] = static_cast<uint32_t>(c);
    for (int64_t i = 0; i < n_interface; ++i) ids[i] = where[interface_cells[i]];
  }
  std::sort(ids.begin(), ids.end());
  PMG_REQUIRE(std::adjacent_find(ids.begin(), ids.end()) == ids.end(), PMG_ERR_INVALID_ARG,
              "interface cell listed twice");
  smap_ = std::move(smap);
  // its scatter schedule: the main map's kind (patch or atomic), same G loads
  ssched_ = Bk3Schedule{};
  if (compress_lines(smap_.ptr, n_cells_, nm, n_dofs_, slines_, sflag_)) {
    ssched_.lines = slines_.ptr;
    ssched_.lflag = sflag_.ptr;
  }
  spatches_ = PatchArrays{};
  if (sched_.kind == kPatch) {
    build_patch_schedule(smap_.ptr, n_cells_, degree_, nq_, precision_, n_dofs_, spatches_,
                         ssched_);
    if (ssched_.pdof) {
      ssched_.kind = kPatch;
      ssched_.gmode = sched_.gmode;
    }
  }
  phys_ = std::move(dphys);
  icells_.upload(ids);
  n_phys_ = n_physical;
  n_icells_ = n_interface;
  has_subdomain_ = true;
}

template <typename T>
void LaplaceOperatorBK3::vmult_subdomain_t(T *dst, const T *src, bool interface_only,
                                           hipStream_t s) const {
  const Shape1D<T> &sh = *reinterpret_cast<const Shape1D<T> *>(
      std::is_same<T, double>::value ? static_cast<const void *>(&sh64_) : static_cast<const void *>(&sh32_));
  const T *G = static_cast<const T *>(G_.data());
  launch_fill<T>(n_dofs_, dst, T(0), s);
  if (interface_only)
    launch_bk3_cell_list<T>(degree_, nq_, src, dst, G, smap_.ptr, ssched_.lines, ssched_.lflag,
                            icells_.ptr, n_icells_, sh, s);
  else
    launch_bk3_cells<T>(degree_, nq_, src, dst, G, smap_.ptr, n_cells_, sh, s, &ssched_,
                        /*assign=*/true);
  launch_copy_indexed<T>(n_phys_, phys_.ptr, dst, src, s);
}

// vmult_neumann / vmult_interface_cell_range
// (portable_subdomain_laplace_operator.h:384-454, :456-533)
void LaplaceOperatorBK3::vmult_subdomain(void *dst, const void *src, bool interface_only,
                                         hipStream_t s) const {
  PMG_REQUIRE(has_subdomain_, PMG_ERR_INVALID_ARG,
              "no subdomain index set: call pmg_operator_set_subdomain first");
  PMG_REQUIRE(dst && src, PMG_ERR_INVALID_ARG, "NULL vector");
  PMG_REQUIRE(dst != src, PMG_ERR_INVALID_ARG, "vmult: dst and src must not alias");
  if (precision_ == PMG_FP64)
    vmult_subdomain_t(static_cast<double *>(dst), static_cast<const double *>(src), interface_only, s);
  else
    vmult_subdomain_t(static_cast<float *>(dst), static_cast<const float *>(src), interface_only, s);
}

bool LaplaceOperatorBK3::supports_prepared_vmult() const {
  const int64_t elem = precision_ == PMG_FP64 ? 8 : 4;
  return n_dofs_ * elem <= kPreparedOutputMaxBytes;
}

// dst already holds src on constrained rows and 0 elsewhere
void LaplaceOperatorBK3::vmult_prepared(void *dst, const void *src, hipStream_t s) const {
  PMG_REQUIRE(dst && src && dst != src, PMG_ERR_INVALID_ARG, "bad vectors");
  if (precision_ == PMG_FP64)
    launch_bk3_cells<double>(degree_, nq_, static_cast<const double *>(src), static_cast<double *>(dst),
                             static_cast<const double *>(G_.data()), map_.ptr, n_cells_, sh64_, s,
                             &sched_, /*assign=*/true);
  else
    launch_bk3_cells<float>(degree_, nq_, static_cast<const float *>(src), static_cast<float *>(dst),
                            static_cast<const float *>(G_.data()), map_.ptr, n_cells_, sh32_, s,
                            &sched_, /*assign=*/true);
  // the grid schedule's second pass overwrites the constrained rows with 0
  if (sched_.kind == kGrid) copy_constrained(dst, src, s);
}

void LaplaceOperatorBK3::copy_constrained(void *dst, const void *src, hipStream_t s) const {
  PMG_REQUIRE(dst && src, PMG_ERR_INVALID_ARG, "NULL vector");
  if (precision_ == PMG_FP64)
    launch_copy_indexed<double>(n_constrained_, constrained_.ptr, static_cast<double *>(dst),
                                static_cast<const double *>(src), s);
  else
    launch_copy_indexed<float>(n_constrained_, constrained_.ptr, static_cast<float *>(dst),
                               static_cast<const float *>(src), s);
}

void LaplaceOperatorBK3::apply_add(void *dst, const void *src, hipStream_t s) const {
  PMG_REQUIRE(dst && src && dst != src, PMG_ERR_INVALID_ARG, "bad vectors");
  if (precision_ == PMG_FP64)
    launch_bk3_cells<double>(degree_, nq_, static_cast<const double *>(src), static_cast<double *>(dst),
                             static_cast<const double *>(G_.data()), map_.ptr, n_cells_, sh64_, s, &sched_);
  else
    launch_bk3_cells<float>(degree_, nq_, static_cast<const float *>(src), static_cast<float *>(dst),
                            static_cast<const float *>(G_.data()), map_.ptr, n_cells_, sh32_, s, &sched_);
}

// The diagonal is assembled and inverted in FP64 at every level precision
// (an FP32 level's G, FP64 arithmetic), then stored in the level precision.
template <typename T>
void LaplaceOperatorBK3::compute_diagonal_t() {
  DeviceArray<double> tmp;
  double *d = nullptr;
  if (std::is_same<T, double>::value) {
    d = static_cast<double *>(inv_diag_.data());
  } else {
    tmp.resize(static_cast<size_t>(n_dofs_));
    d = tmp.ptr;
  }
  launch_fill<double>(n_dofs_, d, 0.0, nullptr);
  launch_bk3_diagonal<T>(degree_, nq_, n_cells_, static_cast<const T *>(G_.data()), map_.ptr,
                         shd64_, d, nullptr);
  launch_set_indexed<double>(n_constrained_, constrained_.ptr, d, 1.0, nullptr);
  // DoFs touched by no cell (possible only with exotic maps) would divide by 0
  const int64_t bad = first_nonpositive<double>(n_dofs_, d, nullptr);
  PMG_REQUIRE(bad < 0, PMG_ERR_INVALID_ARG,
              "non-positive diagonal entry at DoF " + std::to_string(bad));
  launch_invert<double>(n_dofs_, d, nullptr);
  if (!std::is_same<T, double>::value)
    launch_convert<float, double>(n_dofs_, static_cast<float *>(inv_diag_.data()), d, nullptr);
  PMG_HIP_CHECK(hipDeviceSynchronize());
}

void LaplaceOperatorBK3::compute_diagonal() {
  inv_diag_.reinit(precision_, n_dofs_);
  if (precision_ == PMG_FP64) compute_diagonal_t<double>();
  else compute_diagonal_t<float>();
  has_diag_ = true;
}

const void *LaplaceOperatorBK3::inverse_diagonal() const {
  PMG_REQUIRE(has_diag_, PMG_ERR_INVALID_ARG, "compute_diagonal() has not been called");
  return inv_diag_.data();
}

void LaplaceOperatorBK3::set_coefficient(const double *kappa, hipStream_t s) {
  PMG_REQUIRE(vertices_.ptr != nullptr, PMG_ERR_INVALID_ARG,
              "set_coefficient needs an operator created from cell vertices");
  if (precision_ == PMG_FP64) fill_geometry_t<double>(kappa, s);
  else fill_geometry_t<float>(kappa, s);
  has_diag_ = false;
  has_kappa_ = kappa != nullptr;
  if (has_kappa_) sched_.gaff = nullptr;  // G_q now varies beyond the weights
}

void LaplaceOperatorBK3::use_affine_geometry(bool on) {
  if (!on) {
    sched_.gaff = nullptr;
    return;
  }
  PMG_REQUIRE(vertices_.ptr != nullptr, PMG_ERR_INVALID_ARG,
              "affine geometry needs an operator created from cell vertices");
  PMG_REQUIRE(!has_kappa_, PMG_ERR_INVALID_ARG,
              "affine geometry needs a constant coefficient (a kappa was set)");
  gaff_.reinit(precision_, std::max<int64_t>(n_cells_ * 6, 1));
  const int64_t bad =
      precision_ == PMG_FP64
          ? launch_affine_geometry<double>(n_cells_, vertices_.ptr, static_cast<double *>(gaff_.data()), nullptr)
          : launch_affine_geometry<float>(n_cells_, vertices_.ptr, static_cast<float *>(gaff_.data()), nullptr);
  PMG_REQUIRE(bad == 0, PMG_ERR_INVALID_ARG,
              std::to_string(bad) + " of " + std::to_string(n_cells_) + " cells are not affine");
  sched_.gaff = gaff_.data();
}

void LaplaceOperatorBK3::gradient_coefficient(const void *u, const void *lam, void *out,
                                              hipStream_t s) const {
  PMG_REQUIRE(u && lam && out, PMG_ERR_INVALID_ARG, "NULL vector");
  if (precision_ == PMG_FP64)
    launch_bk3_grad_coefficient<double>(degree_, nq_, n_cells_, static_cast<const double *>(G_.data()),
                                        map_.ptr, perm_.ptr, sh64_, static_cast<const double *>(u),
                                        static_cast<const double *>(lam), static_cast<double *>(out), s);
  else
    launch_bk3_grad_coefficient<float>(degree_, nq_, n_cells_, static_cast<const float *>(G_.data()),
                                       map_.ptr, perm_.ptr, sh32_, static_cast<const float *>(u),
                                       static_cast<const float *>(lam), static_cast<float *>(out), s);
}

void assemble_rhs_cells(int precision, int degree, int64_t n_cells, int64_t n_dofs,
                        const double *vertices, const uint32_t *map, int f_kind, int nq, void *b,
                        hipStream_t s) {
  PMG_REQUIRE(b, PMG_ERR_INVALID_ARG, "NULL vector");
  PMG_REQUIRE(f_kind == 0 || f_kind == 1, PMG_ERR_INVALID_ARG, "unknown right-hand side kind");
  PMG_REQUIRE(vertices != nullptr, PMG_ERR_INVALID_ARG,
              "assemble_rhs needs an operator created from cell vertices");
  if (nq <= 0) nq = degree + 1;
  PMG_REQUIRE(nq >= 1 && nq <= 12, PMG_ERR_INVALID_ARG, "n_q_points_1d out of range");
  const ShapeData sd = make_shape_data(degree, std::max(nq, degree + 1));
  PMG_REQUIRE(sd.n_q_points_1d == nq, PMG_ERR_INVALID_ARG, "need n_q_points_1d >= degree+1");
  DeviceArray<double> S, qp, qw;
  S.upload(sd.S);
  qp.upload(sd.quad.points);
  qw.upload(sd.quad.weights);
  if (precision == PMG_FP64) {
    launch_fill<double>(n_dofs, static_cast<double *>(b), 0.0, s);
    launch_integrate<double>(f_kind, degree + 1, nq, n_cells, vertices, qp.ptr, qw.ptr, S.ptr, map,
                             static_cast<double *>(b), s);
  } else {
    launch_fill<float>(n_dofs, static_cast<float *>(b), 0.0f, s);
    launch_integrate<float>(f_kind, degree + 1, nq, n_cells, vertices, qp.ptr, qw.ptr, S.ptr, map,
                            static_cast<float *>(b), s);
  }
  PMG_HIP_CHECK(hipStreamSynchronize(s));  // temporaries are freed on return
}

double l2_error_cells(int precision, int degree, int64_t n_cells, const double *vertices,
                      const uint32_t *map, const void *u, int u_kind, int nq, hipStream_t s) {
  PMG_REQUIRE(u, PMG_ERR_INVALID_ARG, "NULL vector");
  PMG_REQUIRE(u_kind == 0 || u_kind == 1, PMG_ERR_INVALID_ARG, "unknown solution kind");
  PMG_REQUIRE(vertices != nullptr, PMG_ERR_INVALID_ARG,
              "l2_error needs an operator created from cell vertices");
  if (nq <= 0) nq = degree + 2;
  PMG_REQUIRE(nq >= degree + 1 && nq <= 12, PMG_ERR_INVALID_ARG, "n_q_points_1d out of range");
  const ShapeData sd = make_shape_data(degree, nq);
  DeviceArray<double> S, qp, qw;
  S.upload(sd.S);
  qp.upload(sd.quad.points);
  qw.upload(sd.quad.weights);
  const int nb = static_cast<int>(std::max<int64_t>(1, std::min<int64_t>(n_cells, 16384)));
  DeviceArray<double> partials(nb);
  if (precision == PMG_FP64)
    launch_l2_error<double>(u_kind, degree + 1, nq, n_cells, vertices, qp.ptr, qw.ptr, S.ptr, map,
                            static_cast<const double *>(u), partials.ptr, nb, s);
  else
    launch_l2_error<float>(u_kind, degree + 1, nq, n_cells, vertices, qp.ptr, qw.ptr, S.ptr, map,
                           static_cast<const float *>(u), partials.ptr, nb, s);
  std::vector<double> h(nb);
  PMG_HIP_CHECK(hipMemcpyAsync(h.data(), partials.ptr, nb * sizeof(double), hipMemcpyDeviceToHost, s));
  PMG_HIP_CHECK(hipStreamSynchronize(s));
  double e2 = 0;
  for (double v : h) e2 += v;
  return std::sqrt(e2);
}

void LaplaceOperatorBK3::assemble_rhs(int f_kind, int nq, void *b, hipStream_t s) const {
  assemble_rhs_cells(precision_, degree_, n_cells_, n_dofs_, vertices_.ptr, map_.ptr, f_kind, nq,
                     b, s);
}

double LaplaceOperatorBK3::l2_error(const void *u, int u_kind, int nq, hipStream_t s) const {
  return l2_error_cells(precision_, degree_, n_cells_, vertices_.ptr, map_.ptr, u, u_kind, nq, s);
}

void LaplaceOperatorBK3::coarse_pcg(const void *b, void *x, void *r, void *z, void *p, void *q,
                                    double rel_tol, int max_it, double *info,
                                    hipStream_t s) const {
  PMG_REQUIRE(has_diag_, PMG_ERR_INVALID_ARG, "coarse PCG needs compute_diagonal()");
  if (precision_ == PMG_FP64)
    launch_bk3_coarse_pcg<double>(degree_, nq_, n_cells_, static_cast<const double *>(G_.data()),
                                  map_.ptr, sh64_, constrained_.ptr, n_constrained_,
                                  static_cast<const double *>(inv_diag_.data()), n_dofs_,
                                  static_cast<const double *>(b), static_cast<double *>(x),
                                  static_cast<double *>(r), static_cast<double *>(z),
                                  static_cast<double *>(p), static_cast<double *>(q), rel_tol,
                                  max_it, info, s);
  else
    launch_bk3_coarse_pcg<float>(degree_, nq_, n_cells_, static_cast<const float *>(G_.data()),
                                 map_.ptr, sh32_, constrained_.ptr, n_constrained_,
                                 static_cast<const float *>(inv_diag_.data()), n_dofs_,
                                 static_cast<const float *>(b), static_cast<float *>(x),
                                 static_cast<float *>(r), static_cast<float *>(z),
                                 static_cast<float *>(p), static_cast<float *>(q), rel_tol, max_it,
                                 info, s);
}

double LaplaceOperatorBK3::algorithmic_bytes() const {
  const double s = element_size(precision_);
  const double nq3 = double(nq_) * nq_ * nq_;
  return s * (2.0 * n_dofs_ + 6.0 * n_cells_ * nq3);
}

double LaplaceOperatorBK3::algorithmic_flops() const {
  const double nm = degree_ + 1, nq = nq_;
  return n_cells_ * 2.0 *
         (2.0 * (nm * nm * nm * nq + nm * nm * nq * nq + nm * nq * nq * nq) + 6.0 * nq * nq * nq * nq +
          9.0 * nq * nq * nq);
}

// ============================================================ transfers
// Inverse map (CSR) of a device map, built on the device: off[n_dofs + 1],
// list = the positions of each DoF's entries in ascending order.
static void inverse_csr(const uint32_t *map, int64_t n_pos, int64_t n_dofs,
                        DeviceArray<uint32_t> &off, DeviceArray<uint32_t> &list) {
  off.resize(static_cast<size_t>(n_dofs + 1));
  uint32_t *l = nullptr;
  int64_t ne = 0;
  build_inverse_map(n_pos, map, n_dofs, off.ptr, &l, &ne, nullptr);
  list.release();
  list.ptr = l;
  list.n = static_cast<size_t>(ne);
}

// Kind and degree checks; entries per cell and the 1D transfer matrix.
void MGTransfer::init_kind() {
  check_precision(precision_);
  PMG_REQUIRE(kind_ == PMG_TRANSFER_H || kind_ == PMG_TRANSFER_P || kind_ == PMG_TRANSFER_DG_CG,
              PMG_ERR_UNSUPPORTED, "transfer kind not supported");
  PMG_REQUIRE(n_cells_ >= 0 && n_dofs_c_ >= 0 && n_dofs_f_ >= 0, PMG_ERR_INVALID_ARG,
              "negative size");
  if (kind_ == PMG_TRANSFER_DG_CG) {
    // ContinuousTransfer (portable_continuous_transfer.h:18-485): CG(p) coarse,
    // DG(p) fine on the same cells
    PMG_REQUIRE(pc_ == pf_ && pc_ >= 1 && pc_ <= 6, PMG_ERR_UNSUPPORTED,
                "DG-CG transfer needs degree_coarse == degree_fine in [1,6]");
    nmc_ = nmf_ = pc_ + 1;
    return;
  }
  std::vector<double> M;
  if (kind_ == PMG_TRANSFER_H) {
    PMG_REQUIRE(pc_ == pf_ && pc_ >= 1 && pc_ <= 6, PMG_ERR_UNSUPPORTED,
                "h-transfer needs degree_coarse == degree_fine in [1,6]");
    nmc_ = pc_ + 1;
    nmf_ = 2 * pc_ + 1;
    M = h_transfer_matrix(pc_);
  } else {
    PMG_REQUIRE(pc_ >= 1 && pf_ > pc_ && pf_ <= 6, PMG_ERR_UNSUPPORTED,
                "p-transfer needs 1 <= degree_coarse < degree_fine <= 6");
    nmc_ = pc_ + 1;
    nmf_ = pf_ + 1;
    M = p_transfer_matrix(pc_, pf_);
  }
  for (size_t i = 0; i < M.size(); ++i) {
    M64_.M[i] = M[i];
    M32_.M[i] = static_cast<float>(M[i]);
  }
}

// DG-CG: the CG DoFs' DG copies (CSR) for the atomics-free restriction, from
// the device coarse map; list entries are DG positions cell * n^3 + i, or
// map_fine of them when the DG numbering is given.
void MGTransfer::setup_dg_inverse(const uint32_t *map_fine) {
  const int64_t n3 = int64_t(nmc_) * nmc_ * nmc_;
  if (n_cells_ * n3 >= (int64_t(1) << 32)) return;
  inverse_csr(map_c_.ptr, n_cells_ * n3, n_dofs_c_, inv_off_, inv_list_);
  if (map_fine && inv_list_.n > 0) remap_list_device(static_cast<int64_t>(inv_list_.n), map_fine,
                                                     inv_list_.ptr, nullptr);
}

// Restriction without coarse atomics: cells store their coarse values into a
// buffer ordered by coarse DoF (each (cell, coarse DoF) entry at its position
// in the inverse map, constrained entries nowhere), and each coarse DoF sums
// its contiguous range.  It costs a second launch and the buffer's bytes, so
// it is used where it measured faster: p-transfers of >= 2^18 entries and
// FP32 h-transfers of >= 2^20 entries.  Q4 h 64^3 -> 32^3 FP32: atomics 76
// -> buffer 55 us; Q6 -> Q3 on 48^3: FP32 89 us, FP64 119 us against 134 /
// 148 us with atomics; the small Q1 h-levels of 2^12..2^17 entries lose a
// few us each.  FP64 h-transfers keep the atomics: alone the buffer is
// faster (64^3 -> 32^3: 67 -> 63 us), but the FP64 V-cycle at 64^3 Q4 ran
// 45 us slower with it (4447 -> 4492 us).  The DoF-ordered buffer replaced a cell-ordered
// one summed through the inverse list: its stores scatter (restriction kernel
// Q6 -> Q3 FP32 60 -> 72 us) but the sum streams (37 -> 17 us), a net gain
// on every level measured (Q6 -> Q3 FP32 97 -> 89 us, FP64 121 -> 118 us;
// h 64^3 -> 32^3 FP32 61 -> 53 us).  PMG_RESTRICT_GATHER=0|1 forces either
// path.
void MGTransfer::setup_restrict_gather() {
  const int64_t nc3 = int64_t(nmc_) * nmc_ * nmc_;
  const bool big = n_cells_ * nc3 >= (int64_t(1) << (kind_ == PMG_TRANSFER_P ? 18 : 20));
  bool use_gather = big && (precision_ == PMG_FP32 || kind_ == PMG_TRANSFER_P);
  if (const char *genv = std::getenv("PMG_RESTRICT_GATHER")) use_gather = std::string(genv) == "1";
  if (use_gather && n_cells_ * nc3 < (int64_t(1) << 32) && n_dofs_c_ > 0) {
    DeviceArray<uint32_t> list;
    inverse_csr(map_c_.ptr, n_cells_ * nc3, n_dofs_c_, cinv_off_, list);
    cpos_.resize(static_cast<size_t>(n_cells_ * nc3));
    launch_csr_positions(n_cells_ * nc3, static_cast<int64_t>(list.n), list.ptr, cpos_.ptr, nullptr);
    cell_buf_.reinit(precision_, std::max<int64_t>(static_cast<int64_t>(list.n), 1));
  }
}

// setup_weights (portable_geometric_transfer.h:1344-1380) on the device: the
// weights are 1/multiplicity and 0 on constrained fine DoFs, which is the
// owner form (each unconstrained DoF's first entry marked in map_f_); fine
// x-lines that are runs of consecutive DoFs become one word each.
void MGTransfer::setup_owner(const uint8_t *dmask) {
  const size_t nw = static_cast<size_t>(n_cells_) * nmf_ * nmf_ * nmf_;
  PMG_REQUIRE(nw < (size_t(1) << 32) && n_dofs_f_ < (int64_t(1) << 31), PMG_ERR_SIZE,
              "fine map too large for owner marks");
  DeviceArray<uint32_t> first(static_cast<size_t>(std::max<int64_t>(n_dofs_f_, 1)));
  launch_owner_marks(static_cast<int64_t>(nw), n_dofs_f_, map_f_.ptr, dmask, first.ptr, nullptr);
  owner_ = true;
  scale_ = pscale_;
  if (nmf_ <= 32 && n_cells_ > 0) {
    const int64_t nl = n_cells_ * nmf_ * nmf_;
    fine_lines_.resize(static_cast<size_t>(2 * nl));
    DeviceArray<uint32_t> bad(1);
    PMG_HIP_CHECK(hipMemset(bad.ptr, 0, sizeof(uint32_t)));
    launch_fine_lines(nl, nmf_, map_f_.ptr, reinterpret_cast<uint2 *>(fine_lines_.ptr), bad.ptr,
                      nullptr);
    uint32_t hbad = 1;
    PMG_HIP_CHECK(hipMemcpy(&hbad, bad.ptr, sizeof(uint32_t), hipMemcpyDeviceToHost));
    if (hbad) fine_lines_.release();
  }
  PMG_HIP_CHECK(hipDeviceSynchronize());
}

MGTransfer::MGTransfer(int kind, int precision, int pc, int pf, int64_t n_cells,
                       const uint32_t *map_coarse, const uint32_t *map_fine, const double *weights,
                       int64_t n_dofs_coarse, int64_t n_dofs_fine, double scale,
                       const uint32_t *constrained_fine, int64_t n_constrained_fine)
    : kind_(kind), precision_(precision), pc_(pc), pf_(pf), pscale_(scale), n_cells_(n_cells),
      n_dofs_c_(n_dofs_coarse), n_dofs_f_(n_dofs_fine) {
  init_kind();
  if (kind == PMG_TRANSFER_DG_CG) {
    // fine map NULL = cell-major DG numbering
    PMG_REQUIRE(map_coarse, PMG_ERR_INVALID_ARG, "coarse (CG) map is NULL");
    const int64_t n3 = int64_t(nmc_) * nmc_ * nmc_;
    PMG_REQUIRE(map_fine || n_dofs_fine == n_cells * n3, PMG_ERR_SIZE,
                "implicit DG numbering needs n_dofs_fine = n_cells (p+1)^3");
    for (int64_t e = 0; e < n_cells * n3; ++e) {
      PMG_REQUIRE(map_coarse[e] == kInvalid || int64_t(map_coarse[e]) < n_dofs_coarse,
                  PMG_ERR_INVALID_ARG, "coarse map entry out of range");
      PMG_REQUIRE(!map_fine || int64_t(map_fine[e]) < n_dofs_fine, PMG_ERR_INVALID_ARG,
                  "fine map entry out of range");
    }
    map_c_.upload(map_coarse, static_cast<size_t>(n_cells * n3));
    if (map_fine) map_f_.upload(map_fine, static_cast<size_t>(n_cells * n3));
    setup_dg_inverse(map_f_.ptr);
    return;
  }
  const bool device_weights = weights == nullptr && n_constrained_fine >= 0;
  PMG_REQUIRE(map_coarse && map_fine && (weights || device_weights), PMG_ERR_INVALID_ARG,
              "NULL setup array");
  PMG_REQUIRE(!device_weights || constrained_fine || n_constrained_fine == 0, PMG_ERR_INVALID_ARG,
              "NULL constrained fine DoF list");
  const int64_t nc3 = int64_t(nmc_) * nmc_ * nmc_, nf3 = int64_t(nmf_) * nmf_ * nmf_;
  for (int64_t e = 0; e < n_cells * nc3; ++e)
    PMG_REQUIRE(map_coarse[e] == kInvalid || int64_t(map_coarse[e]) < n_dofs_coarse,
                PMG_ERR_INVALID_ARG, "coarse map entry out of range");
  for (int64_t e = 0; e < n_cells * nf3; ++e)
    PMG_REQUIRE(int64_t(map_fine[e]) < n_dofs_fine, PMG_ERR_INVALID_ARG,
                "fine map entry out of range (fine maps carry no invalid entries)");
  map_c_.upload(map_coarse, static_cast<size_t>(n_cells * nc3));
  setup_restrict_gather();
  const size_t nw = static_cast<size_t>(n_cells * nf3);
  if (device_weights) {
    DeviceArray<uint32_t> dcons;
    dcons.upload(constrained_fine, static_cast<size_t>(n_constrained_fine));
    DeviceArray<uint8_t> dmask(static_cast<size_t>(std::max<int64_t>(n_dofs_fine, 1)));
    PMG_REQUIRE(mask_from_list_device(n_constrained_fine, dcons.ptr, n_dofs_fine, dmask.ptr,
                                      nullptr) < 0,
                PMG_ERR_INVALID_ARG, "constrained fine DoF out of range");
    map_f_.upload(map_fine, nw);
    setup_owner(dmask.ptr);
    return;
  }
  // Owner mode when the weights are 1/multiplicity (0 on constrained DoFs,
  // consistently for every entry of a DoF), as the reference's setup_weights
  // builds them: mark the first entry of each weighted fine DoF, drop weights.
  {
    bool owner_ok = n_dofs_fine < (int64_t(1) << 31);
    std::vector<int32_t> mult(owner_ok ? static_cast<size_t>(n_dofs_fine) : 0, 0);
    std::vector<signed char> zero(mult.size(), -1);  // -1 unseen, 0 weighted, 1 zero weight
    for (size_t e = 0; e < nw && owner_ok; ++e) {
      const uint32_t d = map_fine[e];
      ++mult[d];
      const signed char z = weights[e] == 0.0 ? 1 : 0;
      if (zero[d] >= 0 && zero[d] != z) owner_ok = false;
      zero[d] = z;
    }
    for (size_t e = 0; e < nw && owner_ok; ++e)
      if (weights[e] != 0.0 && std::fabs(weights[e] * mult[map_fine[e]] - 1.0) > 1e-12)
        owner_ok = false;
    if (owner_ok) {
      std::vector<uint32_t> mf(map_fine, map_fine + nw);
      std::vector<char> taken(static_cast<size_t>(n_dofs_fine), 0);
      for (size_t e = 0; e < nw; ++e)
        if (weights[e] != 0.0 && !taken[mf[e]]) {
          taken[mf[e]] = 1;
          mf[e] |= 0x80000000u;
        }
      map_f_.upload(mf);
      owner_ = true;
      scale_ = scale;
      // every fine x-line a run of consecutive DoFs (box numberings): one
      // word per line {first DoF, owner bits} replaces nmf map entries
      if (nmf_ <= 32) {
        const int64_t nl = n_cells * nmf_ * nmf_;
        std::vector<uint32_t> fl(static_cast<size_t>(2 * nl));
        bool lines_ok = true;
        for (int64_t l = 0; l < nl && lines_ok; ++l) {
          const uint32_t *e = mf.data() + l * nmf_;
          const uint32_t first = e[0] & 0x7FFFFFFFu;
          uint32_t bits = 0;
          for (int a = 0; a < nmf_; ++a) {
            lines_ok = lines_ok && (e[a] & 0x7FFFFFFFu) == first + static_cast<uint32_t>(a);
            bits |= (e[a] >> 31) << a;
          }
          fl[2 * l] = first;
          fl[2 * l + 1] = bits;
        }
        if (lines_ok) fine_lines_.upload(fl);
      }
      return;
    }
  }
  map_f_.upload(map_fine, nw);
  w_prolong_.reinit(precision, n_cells * nf3);
  w_restrict_.reinit(precision, n_cells * nf3);
  if (precision == PMG_FP64) {
    std::vector<double> wp(weights, weights + nw);
    for (auto &v : wp) v *= scale;
    PMG_HIP_CHECK(hipMemcpy(w_prolong_.data(), wp.data(), nw * 8, hipMemcpyHostToDevice));
    PMG_HIP_CHECK(hipMemcpy(w_restrict_.data(), weights, nw * 8, hipMemcpyHostToDevice));
  } else {
    std::vector<float> wp(nw), wr(nw);
    for (size_t i = 0; i < nw; ++i) {
      wp[i] = static_cast<float>(weights[i] * scale);
      wr[i] = static_cast<float>(weights[i]);
    }
    PMG_HIP_CHECK(hipMemcpy(w_prolong_.data(), wp.data(), nw * 4, hipMemcpyHostToDevice));
    PMG_HIP_CHECK(hipMemcpy(w_restrict_.data(), wr.data(), nw * 4, hipMemcpyHostToDevice));
  }
}

// The transfer between two levels of structured boxes with every boundary DoF
// constrained, maps generated on the device: h (coarse box -> the box refined
// once, degree pc = pf; the fine map is each coarse cell's (2p+1)^3 patch),
// p (one box, pc -> pf) or DG-CG (one box, degree pc = pf, cell-major DG).
MGTransfer::MGTransfer(int kind, int precision, const pmg_box &box, int pc, int pf, double scale)
    : kind_(kind), precision_(precision), pc_(pc), pf_(pf), pscale_(scale),
      n_cells_(int64_t(box.nx) * box.ny * box.nz), n_dofs_c_(0), n_dofs_f_(0) {
  PMG_REQUIRE(box.nx > 0 && box.ny > 0 && box.nz > 0, PMG_ERR_INVALID_ARG,
              "box needs at least one cell per direction");
  auto grid_size = [&](int p, int r) {
    return (int64_t(p) * r * box.nx + 1) * (int64_t(p) * r * box.ny + 1) *
           (int64_t(p) * r * box.nz + 1);
  };
  init_kind();
  n_dofs_c_ = grid_size(pc, 1);
  const int64_t nc3 = int64_t(nmc_) * nmc_ * nmc_;
  map_c_.resize(static_cast<size_t>(n_cells_ * nc3));
  box_map_device(box, pc, 1, true, map_c_.ptr, nullptr);
  if (kind == PMG_TRANSFER_DG_CG) {
    n_dofs_f_ = n_cells_ * nc3;
    setup_dg_inverse(nullptr);
    PMG_HIP_CHECK(hipDeviceSynchronize());
    return;
  }
  const bool h = kind == PMG_TRANSFER_H;
  n_dofs_f_ = h ? grid_size(pc, 2) : grid_size(pf, 1);
  setup_restrict_gather();
  const int64_t nf3 = int64_t(nmf_) * nmf_ * nmf_;
  map_f_.resize(static_cast<size_t>(n_cells_ * nf3));
  box_map_device(box, h ? pc : pf, h ? 2 : 1, false, map_f_.ptr, nullptr);
  DeviceArray<uint8_t> dmask(static_cast<size_t>(std::max<int64_t>(n_dofs_f_, 1)));
  if (h) {
    pmg_box fine = box;
    fine.nx *= 2;
    fine.ny *= 2;
    fine.nz *= 2;
    box_boundary_device(fine, pc, nullptr, dmask.ptr, nullptr);
  } else {
    box_boundary_device(box, pf, nullptr, dmask.ptr, nullptr);
  }
  setup_owner(dmask.ptr);
}

LaplaceOperatorBK3::IndexArray MGTransfer::index_array(int which) const {
  auto arr = [](const auto &a) {
    return LaplaceOperatorBK3::IndexArray{a.ptr, a.ptr ? static_cast<int64_t>(a.n) : 0,
                                          static_cast<int>(sizeof(*a.ptr))};
  };
  switch (which) {
    case PMG_TRANSFER_INDEX_COARSE_MAP: return arr(map_c_);
    case PMG_TRANSFER_INDEX_FINE_MAP: return arr(map_f_);
    case PMG_TRANSFER_INDEX_FINE_LINES: return arr(fine_lines_);
    case PMG_TRANSFER_INDEX_RESTRICT_OFFSETS: return arr(cinv_off_);
    case PMG_TRANSFER_INDEX_RESTRICT_POSITIONS: return arr(cpos_);
    case PMG_TRANSFER_INDEX_DG_OFFSETS: return arr(inv_off_);
    case PMG_TRANSFER_INDEX_DG_LIST: return arr(inv_list_);
    default: throw Error(PMG_ERR_INVALID_ARG, "unknown index array");
  }
}

void MGTransfer::reinit(int64_t n_cells, const uint32_t *map_coarse, const uint32_t *map_fine,
                        const uint32_t *constrained_fine, int64_t n_constrained_fine,
                        int64_t n_dofs_coarse, int64_t n_dofs_fine) {
  PMG_REQUIRE(kind_ == PMG_TRANSFER_DG_CG || n_constrained_fine >= 0, PMG_ERR_INVALID_ARG,
              "negative constrained DoF count");
  MGTransfer fresh(kind_, precision_, pc_, pf_, n_cells, map_coarse, map_fine, nullptr,
                   n_dofs_coarse, n_dofs_fine, pscale_, constrained_fine,
                   kind_ == PMG_TRANSFER_DG_CG ? -1 : n_constrained_fine);
  const uint64_t gen = generation_ + 1;
  *this = std::move(fresh);  // the old device buffers are released here
  generation_ = gen;
}

void MGTransfer::prolongate_and_add(void *dst, const void *src, hipStream_t s) const {
  PMG_REQUIRE(dst && src, PMG_ERR_INVALID_ARG, "NULL vector");
  if (kind_ == PMG_TRANSFER_DG_CG) {
    const int64_t n = n_cells_ * nmc_ * nmc_ * nmc_;
    if (precision_ == PMG_FP64)
      launch_dg_cg_transfer<double>(true, n, map_c_.ptr, map_f_.ptr, static_cast<const double *>(src),
                                    static_cast<double *>(dst), s);
    else
      launch_dg_cg_transfer<float>(true, n, map_c_.ptr, map_f_.ptr, static_cast<const float *>(src),
                                   static_cast<float *>(dst), s);
    return;
  }
  const uint2 *fl = reinterpret_cast<const uint2 *>(fine_lines_.ptr);
  if (precision_ == PMG_FP64)
    launch_prolongate<double>(nmc_, nmf_, n_cells_, M64_, map_c_.ptr, map_f_.ptr,
                              owner_ ? nullptr : static_cast<const double *>(w_prolong_.data()),
                              scale_, static_cast<const double *>(src), static_cast<double *>(dst),
                              s, fl);
  else
    launch_prolongate<float>(nmc_, nmf_, n_cells_, M32_, map_c_.ptr, map_f_.ptr,
                             owner_ ? nullptr : static_cast<const float *>(w_prolong_.data()),
                             static_cast<float>(scale_), static_cast<const float *>(src),
                             static_cast<float *>(dst), s, fl);
}

void MGTransfer::restrict_and_add(void *dst, const void *src, hipStream_t s) const {
  PMG_REQUIRE(dst && src, PMG_ERR_INVALID_ARG, "NULL vector");
  restrict_sub(dst, src, nullptr, s);
}

void MGTransfer::restrict_residual_and_add(void *dst, const void *b, const void *t,
                                           hipStream_t s) const {
  PMG_REQUIRE(dst && b && t, PMG_ERR_INVALID_ARG, "NULL vector");
  restrict_sub(dst, b, t, s);
}

// dst += R (src - sub), sub optional
void MGTransfer::restrict_sub(void *dst, const void *src, const void *sub, hipStream_t s) const {
  const bool f64 = precision_ == PMG_FP64;
  auto D = [](const void *p) { return static_cast<const double *>(p); };
  auto F = [](const void *p) { return static_cast<const float *>(p); };
  if (kind_ == PMG_TRANSFER_DG_CG) {
    const int64_t n = n_cells_ * nmc_ * nmc_ * nmc_;
    if (inv_off_.ptr && f64)
      launch_dg_cg_restrict_gather<double>(n_dofs_c_, inv_off_.ptr, inv_list_.ptr, D(src),
                                           static_cast<double *>(dst), s, D(sub));
    else if (inv_off_.ptr)
      launch_dg_cg_restrict_gather<float>(n_dofs_c_, inv_off_.ptr, inv_list_.ptr, F(src),
                                          static_cast<float *>(dst), s, F(sub));
    else if (f64)
      launch_dg_cg_transfer<double>(false, n, map_c_.ptr, map_f_.ptr, D(src),
                                    static_cast<double *>(dst), s, D(sub));
    else
      launch_dg_cg_transfer<float>(false, n, map_c_.ptr, map_f_.ptr, F(src),
                                   static_cast<float *>(dst), s, F(sub));
    return;
  }
  const bool gather = cinv_off_.ptr != nullptr;
  const uint2 *fl = reinterpret_cast<const uint2 *>(fine_lines_.ptr);
  const uint32_t *targets = gather ? cpos_.ptr : map_c_.ptr;
  if (f64) {
    launch_restrict<double>(nmc_, nmf_, n_cells_, M64_, targets, map_f_.ptr,
                            owner_ ? nullptr : D(w_restrict_.data()), D(src),
                            static_cast<double *>(dst), s, D(sub),
                            gather ? static_cast<double *>(cell_buf_.data()) : nullptr, fl);
    if (gather)
      launch_restrict_sum<double>(n_dofs_c_, cinv_off_.ptr, D(cell_buf_.data()),
                                  static_cast<double *>(dst), s);
  } else {
    launch_restrict<float>(nmc_, nmf_, n_cells_, M32_, targets, map_f_.ptr,
                           owner_ ? nullptr : F(w_restrict_.data()), F(src),
                           static_cast<float *>(dst), s, F(sub),
                           gather ? static_cast<float *>(cell_buf_.data()) : nullptr, fl);
    if (gather)
      launch_restrict_sum<float>(n_dofs_c_, cinv_off_.ptr, F(cell_buf_.data()),
                                 static_cast<float *>(dst), s);
  }
}

// ============================================================ CG
void CGWork::reinit(int precision, int64_t n) {
  if (sc.n == 0) {
    sc.resize(8);
    partials.resize(kDotBlocks);
  }
  if (r.n != n || r.precision != precision) {
    r.reinit(precision, n);
    z.reinit(precision, n);
    p.reinit(precision, n);
    q.reinit(precision, n);
  }
}

CGSpace CGWork::space() {
  host.allocate(2);  // r.r and the iteration that converged
  CGSpace w;
  w.sc = sc.ptr;
  w.partials = partials.ptr;
  w.r = r.data();
  w.z = z.data();
  w.p = p.data();
  w.q = q.data();
  w.host_rr = host.ptr;
  return w;
}

namespace {
template <typename T>
CGResult solve_cg_t(const Operator &op, int precond, const VCycleMultigrid *mg,
                    const T *b, T *x, double rel_tol, int max_it, const CGSpace &w,
                    hipStream_t s, int check_every) {
  const int64_t n = op.m();
  T *r = static_cast<T *>(w.r), *z = static_cast<T *>(w.z);
  T *p = static_cast<T *>(w.p), *q = static_cast<T *>(w.q);
  const T *pinv = precond == PMG_PRECOND_JACOBI ? static_cast<const T *>(op.inverse_diagonal()) : nullptr;
  double *sc = w.sc;
  double *partials = w.partials;
  // r = b - A x
  op.vmult(q, x, s);
  launch_residual<T>(n, q, b, s);  // q = b - A x
  launch_copy<T>(n, r, q, s);
  launch_dot<T>(n, r, r, partials, sc + 2, s);
  double rr = 0;
  PMG_HIP_CHECK(hipMemcpyAsync(&rr, sc + 2, sizeof(double), hipMemcpyDeviceToHost, s));
  PMG_HIP_CHECK(hipStreamSynchronize(s));
  CGResult res;
  res.res0 = res.res = std::sqrt(rr);
  const double target = rel_tol * res.res0;
  if (res.res <= target || res.res0 == 0.0) {
    res.converged = true;
    return res;
  }
  int it = 0;
  static const bool verbose = std::getenv("PMG_CG_VERBOSE") != nullptr;
  const double target2 = target * target;
  // the converged flag: once set, the CG kernels of iterations queued after
  // the converged one change nothing, so the host may check every few
  // iterations and the result is the same as checking after each one
  const double *skip = sc + kCgConvergedSlot;
  launch_cg_scalars(sc, 2, s, w.host_rr);
  if (check_every <= 0) {
    check_every = precond == PMG_PRECOND_MG ? 1 : kCgCheckEvery;
    if (const char *e = std::getenv("PMG_CG_CHECK_EVERY")) check_every = std::max(1, std::atoi(e));
  }
  if (precond == PMG_PRECOND_MG) mg->stage(r, s, nullptr);
  while (it < max_it) {
    if (precond == PMG_PRECOND_MG) {
      // z = the V-cycle's own finest-level solution, read in its precision
      // (the same values as after a conversion, one pass less)
      const void *zl = mg->run(s);
      if (mg->level_precision() == PMG_FP64) {
        launch_dot<T, double>(n, r, static_cast<const double *>(zl), partials, sc + 0, s, true, skip);
        launch_cg_update_xp<T, double>(n, x, p, static_cast<const double *>(zl), sc, it == 0 ? 1 : 0, s,
                                       skip);
      } else {
        launch_dot<T, float>(n, r, static_cast<const float *>(zl), partials, sc + 0, s, true, skip);
        launch_cg_update_xp<T, float>(n, x, p, static_cast<const float *>(zl), sc, it == 0 ? 1 : 0, s,
                                      skip);
      }
    } else {
      launch_cg_precond_dot<T>(n, r, pinv, z, partials, s, skip);
      launch_reduce_partials(partials, kDotBlocks, sc, 0, 1, s, skip);
      launch_cg_update_xp<T>(n, x, p, z, sc, it == 0 ? 1 : 0, s, skip);
    }
    op.vmult(q, p, s);
    launch_dot<T>(n, p, q, partials, sc + 1, s, false, skip);
    // the residual update also writes the next V-cycle input (its staging),
    // so the GPU stays busy while the host waits for r.r
    if (precond == PMG_PRECOND_MG)
      launch_cg_update_r<T>(n, r, q, sc, partials, s, mg->input_buffer(), mg->level_precision(),
                            skip);
    else
      launch_cg_update_r<T>(n, r, q, sc, partials, s, nullptr, PMG_FP64, skip);
    launch_cg_rr_scalars(partials, kDotBlocks, sc, s, w.host_rr, target2, it + 1);
    ++it;
    const bool check = it % check_every == 0 || it == max_it;
    if (check) {
      double conv_it = 0;
      if (w.host_rr) {
        PMG_HIP_CHECK(hipStreamSynchronize(s));
        rr = w.host_rr[0];
        conv_it = w.host_rr[1];
      } else {
        double scv[8];
        PMG_HIP_CHECK(hipMemcpyAsync(scv, sc, sizeof(scv), hipMemcpyDeviceToHost, s));
        PMG_HIP_CHECK(hipStreamSynchronize(s));
        rr = scv[2];
        conv_it = scv[kCgConvItSlot];
      }
      res.res = std::sqrt(rr);
      if (verbose) std::fprintf(stderr, "[pmg cg] it %d  |r|/|r0| = %.3e\n", it, res.res / res.res0);
      if (conv_it > 0) {  // r.r <= target^2 at that iteration, or NaN
        it = static_cast<int>(conv_it);
        res.converged = res.res <= target;
        break;
      }
    }
  }
  if (it > 0) launch_cg_update_x<T>(n, x, p, sc, s);  // the deferred last x update
  res.iterations = it;
  return res;
}
}  // namespace

CGResult solve_cg(const Operator &op, int precond, const VCycleMultigrid *mg,
                  const void *b, void *x, double rel_tol, int max_it, CGWork &work,
                  hipStream_t s, int check_every) {
  work.reinit(op.precision(), op.m());
  return solve_cg(op, precond, mg, b, x, rel_tol, max_it, work.space(), s, check_every);
}

CGResult solve_cg(const Operator &op, int precond, const VCycleMultigrid *mg,
                  const void *b, void *x, double rel_tol, int max_it, const CGSpace &work,
                  hipStream_t s, int check_every) {
  PMG_REQUIRE(work.sc && work.partials && work.r && work.z && work.p && work.q,
              PMG_ERR_INVALID_ARG, "CG work space missing");
  PMG_REQUIRE(b && x && b != x, PMG_ERR_INVALID_ARG, "bad vectors");
  PMG_REQUIRE(precond >= 0 && precond <= 2, PMG_ERR_INVALID_ARG, "unknown preconditioner");
  PMG_REQUIRE(precond != PMG_PRECOND_MG || (mg && mg->outer_precision() == op.precision()),
              PMG_ERR_INVALID_ARG, "MG preconditioner missing or of the wrong outer precision");
  PMG_REQUIRE(precond != PMG_PRECOND_MG || mg->m() == op.m(), PMG_ERR_SIZE,
              "the MG preconditioner's finest level does not match the operator size");
  PMG_REQUIRE(rel_tol > 0 && max_it >= 0, PMG_ERR_INVALID_ARG, "bad tolerance or max_it");
  if (op.precision() == PMG_FP64)
    return solve_cg_t<double>(op, precond, mg, static_cast<const double *>(b),
                              static_cast<double *>(x), rel_tol, max_it, work, s, check_every);
  return solve_cg_t<float>(op, precond, mg, static_cast<const float *>(b), static_cast<float *>(x),
                           rel_tol, max_it, work, s, check_every);
}

// ============================================================ Chebyshev
namespace {
// Lanczos estimate of the extreme eigenvalues of P^{-1}A through CG with
// IterationNumberControl(n_it, 1e-10) semantics, right-hand side (i mod 11)
// minus its mean, zero start (deal.II internal::estimate_eigenvalues).
template <typename T>
void estimate_eigenvalues(const Operator &op, int n_it, double &mn, double &mx) {
  const int64_t n = op.m();
  DeviceVector xv, rv, zv, pv, qv;
  xv.reinit(op.precision(), n); rv.reinit(op.precision(), n); zv.reinit(op.precision(), n);
  pv.reinit(op.precision(), n); qv.reinit(op.precision(), n);
  DeviceArray<double> partials(kDotBlocks), sc(4);
  T *x = static_cast<T *>(xv.data()), *r = static_cast<T *>(rv.data()), *z = static_cast<T *>(zv.data());
  T *p = static_cast<T *>(pv.data()), *q = static_cast<T *>(qv.data());
  const T *pinv = static_cast<const T *>(op.inverse_diagonal());
  // start vector r_i = (i mod 11) - mean, filled on the device; the sum of
  // i mod 11 is an integer below 2^53, so the closed form is exact
  const int64_t full = n / 11, rest = n % 11;
  const double mean = double(55 * full + rest * (rest - 1) / 2) / double(std::max<int64_t>(n, 1));
  launch_fill_mod11<T>(n, r, mean, nullptr);
  launch_fill<T>(n, x, T(0), nullptr);
  double res = std::sqrt(dot_sync<T>(n, r, r, partials, sc, nullptr));
  std::vector<double> al, be;
  double rho_old = 0;
  int it = 0;
  while (it < n_it && res > 1e-10) {
    launch_cg_precond_dot<T>(n, r, pinv, z, partials.ptr, nullptr);
    launch_reduce_partials(partials.ptr, kDotBlocks, sc.ptr, 0, 1, nullptr);
    double rho = 0;
    PMG_HIP_CHECK(hipMemcpy(&rho, sc.ptr, sizeof(double), hipMemcpyDeviceToHost));
    const double beta = it == 0 ? 0.0 : rho / rho_old;
    if (it > 0) be.push_back(beta);
    // p = z + beta p
    PMG_HIP_CHECK(hipMemcpy(sc.ptr + 3, &rho_old, sizeof(double), hipMemcpyHostToDevice));
    launch_cg_update_p<T>(n, p, z, sc.ptr, it == 0 ? 1 : 0, nullptr);
    op.vmult(q, p, nullptr);
    const double pq = dot_sync<T>(n, p, q, partials, sc, nullptr);
    if (!(pq > 0)) break;
    const double alpha = rho / pq;
    al.push_back(alpha);
    double s2[4] = {rho, pq, 0, 0};
    PMG_HIP_CHECK(hipMemcpy(sc.ptr, s2, 2 * sizeof(double), hipMemcpyHostToDevice));
    launch_cg_update_xr<T>(n, x, r, p, q, sc.ptr, partials.ptr, nullptr);
    launch_reduce_partials(partials.ptr, kDotBlocks, sc.ptr, 2, 1, nullptr);
    double rr = 0;
    PMG_HIP_CHECK(hipMemcpy(&rr, sc.ptr + 2, sizeof(double), hipMemcpyDeviceToHost));
    res = std::sqrt(rr);
    rho_old = rho;
    ++it;
  }
  const int k = static_cast<int>(al.size());
  if (k == 0) {
    mn = mx = 1.0;
    return;
  }
  std::vector<double> d(k), o(std::max(k - 1, 0));
  for (int i = 0; i < k; ++i) d[i] = 1.0 / al[i] + (i > 0 ? be[i - 1] / al[i - 1] : 0.0);
  for (int i = 0; i + 1 < k; ++i) o[i] = std::sqrt(be[i]) / al[i];
  const auto ev = tridiag_eigenvalues(d, o);
  mn = ev.front();
  mx = ev.back();
}
}  // namespace

PreconditionChebyshev::PreconditionChebyshev(const Operator &op, int degree,
                                             double smoothing_range, int eig_cg_n_iterations)
    : op_(op), degree_(degree), range_(smoothing_range) {
  PMG_REQUIRE(op.has_diagonal(), PMG_ERR_INVALID_ARG,
              "Chebyshev needs compute_diagonal() on its operator first");
  PMG_REQUIRE(smoothing_range > 0 && eig_cg_n_iterations > 2, PMG_ERR_INVALID_ARG,
              "need smoothing_range > 0 and eig_cg_n_iterations > 2");
  PMG_REQUIRE(degree > 0 || smoothing_range < 1, PMG_ERR_INVALID_ARG,
              "automatic degree needs smoothing_range < 1");
  double mn, mx;
  if (op.precision() == PMG_FP64) estimate_eigenvalues<double>(op, eig_cg_n_iterations, mn, mx);
  else estimate_eigenvalues<float>(op, eig_cg_n_iterations, mn, mx);
  min_ev_ = mn;
  max_ev_ = 1.2 * mx;  // safety factor: CG is not converged
  const double alpha = range_ > 1.0 ? max_ev_ / range_ : std::min(0.9 * max_ev_, min_ev_);
  if (degree_ <= 0) {
    const double actual_range = max_ev_ / alpha;
    const double sigma = (1.0 - std::sqrt(1.0 / actual_range)) / (1.0 + std::sqrt(1.0 / actual_range));
    const double eps = range_;
    degree_ = 1 + static_cast<int>(std::log(1.0 / eps + std::sqrt(1.0 / eps / eps - 1.0)) /
                                   std::log(1.0 / sigma));
  }
  delta_ = 0.5 * (max_ev_ - alpha);
  theta_ = 0.5 * (max_ev_ + alpha);
  tmp_.reinit(op.precision(), op.m());
  old_.reinit(op.precision(), op.m());
}

void PreconditionChebyshev::run(void *dst, const void *src, bool step, void *next_out,
                                hipStream_t s) const {
  PMG_REQUIRE(dst && src && dst != src, PMG_ERR_INVALID_ARG, "bad vectors");
  const int64_t n = op_.m();
  const bool f64 = op_.precision() == PMG_FP64;
  const bool prep = op_.supports_prepared_vmult();
  const uint8_t *mask = op_.constrained_mask();
  // Each update also writes the prepared output of the matvec that follows
  // it (tmp_ inside the recurrence, next_out after the last update), which
  // saves the operator's own zero fill and constrained copy.
  auto upd = [&](int mode, double f1, double f2, bool old_zero, void *out) {
    void *pp = prep ? out : nullptr;
    if (f64)
      launch_cheby_update<double>(mode, n, static_cast<double *>(dst), static_cast<double *>(old_.data()),
                                  static_cast<const double *>(src), static_cast<const double *>(tmp_.data()),
                                  static_cast<const double *>(op_.inverse_diagonal()), f1, f2, old_zero, s,
                                  static_cast<double *>(pp), mask);
    else
      launch_cheby_update<float>(mode, n, static_cast<float *>(dst), static_cast<float *>(old_.data()),
                                 static_cast<const float *>(src), static_cast<const float *>(tmp_.data()),
                                 static_cast<const float *>(op_.inverse_diagonal()), f1, f2, old_zero, s,
                                 static_cast<float *>(pp), mask);
  };
  const bool more = degree_ >= 2 && std::fabs(delta_) >= 1e-40;
  if (!step) {
    upd(0, 0.0, 1.0 / theta_, false, more ? tmp_.data() : next_out);
  } else {
    op_.vmult(tmp_.data(), dst, s);
    upd(1, 0.0, 1.0 / theta_, false, more ? tmp_.data() : next_out);
  }
  if (!more) return;
  double rhok = delta_ / theta_;
  const double sigma = theta_ / delta_;
  for (int k = 0; k < degree_ - 1; ++k) {
    if (prep) op_.vmult_prepared(tmp_.data(), dst, s);
    else op_.vmult(tmp_.data(), dst, s);
    const double rhokp = 1.0 / (2.0 * sigma - rhok);
    const double f1 = rhokp * rhok, f2 = 2.0 * rhokp / delta_;
    rhok = rhokp;
    upd(2, f1, f2, !step && k == 0, k + 2 < degree_ ? tmp_.data() : next_out);
  }
}

void PreconditionChebyshev::vmult(void *dst, const void *src, hipStream_t s) const {
  run(dst, src, false, nullptr, s);
}
void PreconditionChebyshev::step(void *dst, const void *src, hipStream_t s) const {
  run(dst, src, true, nullptr, s);
}
bool PreconditionChebyshev::vmult_and_prepare(void *dst, const void *src, void *next_out,
                                              hipStream_t s) const {
  PMG_REQUIRE(next_out && next_out != dst && next_out != src, PMG_ERR_INVALID_ARG, "bad vectors");
  run(dst, src, false, next_out, s);
  return op_.supports_prepared_vmult();
}

// ============================================================ V-cycle
VCycleMultigrid::VCycleMultigrid(std::vector<const Operator *> ops,
                                 std::vector<const MGTransfer *> transfers,
                                 std::vector<const PreconditionChebyshev *> smoothers,
                                 int coarse_kind, double coarse_tol, int outer_precision)
    : ops_(std::move(ops)), xfer_(std::move(transfers)), sm_(std::move(smoothers)),
      coarse_kind_(coarse_kind), outer_precision_(outer_precision), coarse_tol_(coarse_tol) {
  const int L = static_cast<int>(ops_.size());
  PMG_REQUIRE(L >= 1, PMG_ERR_INVALID_ARG, "need at least one level");
  PMG_REQUIRE(int(xfer_.size()) == L && int(sm_.size()) == L, PMG_ERR_INVALID_ARG,
              "transfers/smoothers arrays must have n_levels entries");
  check_precision(outer_precision);
  PMG_REQUIRE(coarse_kind == PMG_COARSE_PCG || coarse_kind == PMG_COARSE_CHEBYSHEV,
              PMG_ERR_INVALID_ARG, "unknown coarse solver");
  for (int l = 0; l < L; ++l)
    PMG_REQUIRE(ops_[l], PMG_ERR_INVALID_ARG, "missing level operator");
  // one level precision; the coarsest operator may be FP64 under FP32 levels
  // (the coarse-grid solve kept in FP64)
  precision_ = ops_[L - 1]->precision();
  coarse64_ = L >= 2 && precision_ == PMG_FP32 && ops_[0]->precision() == PMG_FP64;
  for (int l = 0; l < L; ++l) {
    PMG_REQUIRE(ops_[l]->precision() == precision_ || (l == 0 && coarse64_), PMG_ERR_INVALID_ARG,
                "level operators must share one precision (the coarsest may be FP64)");
    PMG_REQUIRE(ops_[l]->has_diagonal(), PMG_ERR_INVALID_ARG,
                "level operators need compute_diagonal()");
    if (l > 0) {
      PMG_REQUIRE(xfer_[l] && sm_[l], PMG_ERR_INVALID_ARG, "missing transfer or smoother");
      PMG_REQUIRE(xfer_[l]->precision() == precision_ && &sm_[l]->op() == ops_[l],
                  PMG_ERR_INVALID_ARG, "transfer/smoother precision or operator mismatch");
      PMG_REQUIRE(xfer_[l]->n_dofs_coarse() == ops_[l - 1]->m() &&
                      xfer_[l]->n_dofs_fine() == ops_[l]->m(),
                  PMG_ERR_SIZE, "transfer sizes do not match the level operators");
    }
  }
  if (coarse_kind == PMG_COARSE_CHEBYSHEV) {
    const int n_it = static_cast<int>(std::max<int64_t>(ops_[0]->m(), 3));
    coarse_cheby_ = std::make_unique<PreconditionChebyshev>(*ops_[0], 0, 1e-3, n_it);
  }
  sol_.resize(L);
  def_.resize(L);
  t_.resize(L);
  for (int l = 0; l < L; ++l) {
    sol_[l].reinit(precision_, ops_[l]->m());
    def_[l].reinit(precision_, ops_[l]->m());
    t_[l].reinit(precision_, ops_[l]->m());
  }
  coarse_work_.reinit(ops_[0]->precision(), ops_[0]->m());
  if (coarse64_) {
    c64_def_.reinit(PMG_FP64, ops_[0]->m());
    c64_sol_.reinit(PMG_FP64, ops_[0]->m());
  }
  // The V-cycle is a fixed launch sequence on device buffers of its own; it is
  // captured once into a HIP graph and replayed, unless the coarse solve needs
  // host round trips (CG on a large coarse level).  PMG_VCYCLE_GRAPH=0 turns
  // graphs off.
  const auto *bk3c = dynamic_cast<const LaplaceOperatorBK3 *>(ops_[0]);
  const bool host_free = coarse_kind_ == PMG_COARSE_CHEBYSHEV ||
                         (bk3c && bk3c->n_cells() <= kCoarseSingleWorkgroupCells);
  const char *genv = std::getenv("PMG_VCYCLE_GRAPH");
  use_graph_ = host_free && !(genv && std::string(genv) == "0");
}

VCycleMultigrid::~VCycleMultigrid() {
  // a replay may still run on the caller's stream: let it finish before the
  // graph and the level vectors it uses go away
  if (graph_exec_) (void)hipDeviceSynchronize();
  if (graph_exec_) (void)hipGraphExecDestroy(graph_exec_);
  if (graph_) (void)hipGraphDestroy(graph_);
  if (cap_stream_) (void)hipStreamDestroy(cap_stream_);
}

// Captures the V-cycle body once.  The graph is an optimisation: if capturing
// or instantiating fails, everything built so far is released and the eager
// path is used from then on (a genuine launch error then shows there).
void VCycleMultigrid::capture() const {
  auto give_up = [this] {
    if (graph_) (void)hipGraphDestroy(graph_);
    graph_ = nullptr;
    graph_exec_ = nullptr;
    use_graph_ = false;
    (void)hipGetLastError();  // the failed capture's error is not the eager path's
  };
  if (!cap_stream_ && hipStreamCreateWithFlags(&cap_stream_, hipStreamNonBlocking) != hipSuccess) {
    cap_stream_ = nullptr;
    give_up();
    return;
  }
  if (hipStreamBeginCapture(cap_stream_, hipStreamCaptureModeThreadLocal) != hipSuccess) {
    give_up();
    return;
  }
  bool recorded = true;
  try {
    v_cycle(levels() - 1, cap_stream_);
  } catch (...) {
    recorded = false;
  }
  hipGraph_t g = nullptr;
  const hipError_t end = hipStreamEndCapture(cap_stream_, &g);
  if (!recorded || end != hipSuccess || !g) {
    if (g) (void)hipGraphDestroy(g);
    give_up();
    return;
  }
  graph_ = g;
  if (hipGraphInstantiate(&graph_exec_, graph_, nullptr, nullptr, 0) != hipSuccess) give_up();
}

void VCycleMultigrid::mark(int level, int stage, bool begin, hipStream_t s) const {
  if (!prof_) return;
  if (begin) {
    StageEvent ev{level, stage, nullptr, nullptr};
    PMG_HIP_CHECK(hipEventCreate(&ev.begin));
    PMG_HIP_CHECK(hipEventCreate(&ev.end));
    prof_->push_back(ev);
    PMG_HIP_CHECK(hipEventRecord(ev.begin, s));
  } else {
    PMG_HIP_CHECK(hipEventRecord(prof_->back().end, s));
  }
}

void VCycleMultigrid::level_times(const void *src, int n_applications, double *times,
                                  hipStream_t s) const {
  PMG_REQUIRE(src && times && n_applications > 0, PMG_ERR_INVALID_ARG, "bad arguments");
  std::vector<StageEvent> events;
  events.reserve(static_cast<size_t>(n_applications) * levels() * kStages);
  struct Reset {  // the events are destroyed and profiling is off on every exit
    const VCycleMultigrid *mg;
    std::vector<StageEvent> *ev;
    ~Reset() {
      mg->prof_ = nullptr;
      for (auto &e : *ev) {
        if (e.begin) (void)hipEventDestroy(e.begin);
        if (e.end) (void)hipEventDestroy(e.end);
      }
    }
  } reset{this, &events};
  stage(src, s, nullptr);
  prof_ = &events;
  for (int a = 0; a < n_applications; ++a) v_cycle(levels() - 1, s);
  prof_ = nullptr;
  PMG_HIP_CHECK(hipStreamSynchronize(s));
  for (int i = 0; i < levels() * kStages; ++i) times[i] = 0.0;
  for (const auto &e : events) {
    float ms = 0.f;
    PMG_HIP_CHECK(hipEventElapsedTime(&ms, e.begin, e.end));
    times[e.level * kStages + e.stage] += 1e-3 * ms / n_applications;
  }
}

void VCycleMultigrid::v_cycle(int l, hipStream_t s) const {
  const bool f64 = precision_ == PMG_FP64;
  const int64_t n = ops_[l]->m();
  if (l == 0) {
    mark(0, 3, true, s);
    const bool c64 = ops_[0]->precision() == PMG_FP64;
    void *cdef = coarse64_ ? c64_def_.data() : def_[0].data();
    void *csol = coarse64_ ? c64_sol_.data() : sol_[0].data();
    if (coarse64_)
      launch_convert<double, float>(n, static_cast<double *>(cdef),
                                    static_cast<const float *>(def_[0].data()), s);
    if (coarse_kind_ == PMG_COARSE_PCG) {
      if (c64) launch_fill<double>(n, static_cast<double *>(csol), 0.0, s);
      else launch_fill<float>(n, static_cast<float *>(csol), 0.0f, s);
      // FP32 arithmetic cannot reduce the residual below ~1e-7 relative
      const double tol = c64 ? coarse_tol_ : std::max(coarse_tol_, 1e-6);
      const int max_it = static_cast<int>(std::min<int64_t>(100000, 4 * n + 100));
      const auto *bk3 = dynamic_cast<const LaplaceOperatorBK3 *>(ops_[0]);
      if (bk3 && bk3->n_cells() <= kCoarseSingleWorkgroupCells) {
        CGWork &w = coarse_work_;
        bk3->coarse_pcg(cdef, csol, w.r.data(), w.z.data(), w.p.data(), w.q.data(), tol, max_it,
                        w.sc.ptr + 4, s);
      } else {
        solve_cg(*ops_[0], PMG_PRECOND_JACOBI, nullptr, cdef, csol, tol, max_it, coarse_work_, s,
                 4);
      }
    } else {
      coarse_cheby_->vmult(csol, cdef, s);
    }
    if (coarse64_)
      launch_convert<float, double>(n, static_cast<float *>(sol_[0].data()),
                                    static_cast<const double *>(csol), s);
    mark(0, 3, false, s);
    return;
  }
  // pre-smoothing; its last update prepares t for the residual's matvec
  mark(l, 0, true, s);
  const bool prepared = sm_[l]->vmult_and_prepare(sol_[l].data(), def_[l].data(), t_[l].data(), s);
  mark(l, 0, false, s);
  mark(l, 1, true, s);
  if (prepared) ops_[l]->vmult_prepared(t_[l].data(), sol_[l].data(), s);
  else ops_[l]->vmult(t_[l].data(), sol_[l].data(), s);
  mark(l, 1, false, s);
  mark(l, 2, true, s);
  const int64_t nc = ops_[l - 1]->m();
  if (f64) launch_fill<double>(nc, static_cast<double *>(def_[l - 1].data()), 0.0, s);
  else launch_fill<float>(nc, static_cast<float *>(def_[l - 1].data()), 0.0f, s);
  // coarse defect = R (d - A s): the residual is formed inside the restriction
  xfer_[l]->restrict_residual_and_add(def_[l - 1].data(), def_[l].data(), t_[l].data(), s);
  mark(l, 2, false, s);
  v_cycle(l - 1, s);
  mark(l, 4, true, s);
  xfer_[l]->prolongate_and_add(sol_[l].data(), sol_[l - 1].data(), s);
  mark(l, 4, false, s);
  mark(l, 5, true, s);
  sm_[l]->step(sol_[l].data(), def_[l].data(), s);
  mark(l, 5, false, s);
}

void VCycleMultigrid::stage(const void *src, hipStream_t s, const double *skip) const {
  PMG_REQUIRE(src, PMG_ERR_INVALID_ARG, "NULL vector");
  const int top = levels() - 1;
  const int64_t n = ops_[top]->m();
  const bool in64 = outer_precision_ == PMG_FP64, lv64 = precision_ == PMG_FP64;
  void *d = def_[top].data();
  if (in64 && lv64) launch_convert_unless<double, double>(n, static_cast<double *>(d), static_cast<const double *>(src), skip, s);
  else if (in64) launch_convert_unless<float, double>(n, static_cast<float *>(d), static_cast<const double *>(src), skip, s);
  else if (lv64) launch_convert_unless<double, float>(n, static_cast<double *>(d), static_cast<const float *>(src), skip, s);
  else launch_convert_unless<float, float>(n, static_cast<float *>(d), static_cast<const float *>(src), skip, s);
}

// A transfer rebuilt by reinit() has new buffers and maybe new sizes: the
// sizes must still fit the level vectors, and a graph captured with the old
// buffers is dropped (the next run captures again).
void VCycleMultigrid::check_transfers() const {
  bool changed = xfer_gen_.size() != xfer_.size();
  for (size_t l = 1; l < xfer_.size() && !changed; ++l)
    changed = xfer_[l] && xfer_[l]->generation() != xfer_gen_[l];
  if (!changed) return;
  for (size_t l = 1; l < xfer_.size(); ++l)
    PMG_REQUIRE(xfer_[l]->n_dofs_coarse() == ops_[l - 1]->m() &&
                    xfer_[l]->n_dofs_fine() == ops_[l]->m(),
                PMG_ERR_SIZE, "transfer sizes no longer match the V-cycle levels");
  if (graph_exec_) (void)hipGraphExecDestroy(graph_exec_);
  if (graph_) (void)hipGraphDestroy(graph_);
  graph_exec_ = nullptr;
  graph_ = nullptr;
  xfer_gen_.assign(xfer_.size(), 0);
  for (size_t l = 1; l < xfer_.size(); ++l)
    if (xfer_[l]) xfer_gen_[l] = xfer_[l]->generation();
}

const void *VCycleMultigrid::run(hipStream_t s) const {
  check_transfers();
  if (use_graph_ && !graph_exec_) capture();  // may fall back to the eager path
  if (use_graph_) PMG_HIP_CHECK(hipGraphLaunch(graph_exec_, s));
  else v_cycle(levels() - 1, s);
  return sol_[levels() - 1].data();
}

const void *VCycleMultigrid::apply(const void *src, hipStream_t s) const {
  stage(src, s, nullptr);
  return run(s);
}

void VCycleMultigrid::vmult(void *dst, const void *src, hipStream_t s) const {
  PMG_REQUIRE(dst && src, PMG_ERR_INVALID_ARG, "NULL vector");
  const int64_t n = ops_.back()->m();
  const bool in64 = outer_precision_ == PMG_FP64, lv64 = precision_ == PMG_FP64;
  const void *sol = apply(src, s);
  if (in64 && lv64) launch_copy<double>(n, static_cast<double *>(dst), static_cast<const double *>(sol), s);
  else if (in64) launch_convert<double, float>(n, static_cast<double *>(dst), static_cast<const float *>(sol), s);
  else if (lv64) launch_convert<float, double>(n, static_cast<float *>(dst), static_cast<const double *>(sol), s);
  else launch_copy<float>(n, static_cast<float *>(dst), static_cast<const float *>(sol), s);
}

}  // namespace pmg
