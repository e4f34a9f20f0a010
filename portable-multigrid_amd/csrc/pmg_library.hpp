// pmg_library.hpp -- host-side C++ classes behind the C ABI.  They mirror the
// reference's deal.II-side interfaces (operator, transfer, smoother, V-cycle,
// CG) with raw device pointers in place of deal.II device vectors.
#pragma once

#include <memory>
#include <string>
#include <vector>

#include "pmg_basis.hpp"
#include "pmg_common.hpp"
#include "pmg_kernels.hpp"

namespace pmg {

// vmult of BK3 operators up to this many DoFs: fused zero fill + constrained copy
constexpr int64_t kFusedPrepMaxDofs = int64_t(1) << 20;
// BK3 operators whose vectors are at most this large let the Chebyshev
// update prepare their matvec output (one launch less per matvec).  Larger
// ones keep the zero fill right before the cell kernel: it leaves the output
// resident in the 256 MB last-level cache, where the kernel's atomic
// scatter then lands (measured faster on 17M-DoF FP32 levels).
constexpr int64_t kPreparedOutputMaxBytes = int64_t(32) << 20;


// Untyped owning device buffer of either precision.  reinit keeps the
// buffer (and its address) when precision and length are unchanged: device
// pointers to it may already be baked into a captured V-cycle graph (the
// inverse diagonal a later compute_diagonal() refills, for one).
struct DeviceVector {
  DeviceArray<char> raw;
  int precision = PMG_FP64;
  int64_t n = 0;
  void reinit(int prec, int64_t count) {
    if (raw.ptr && prec == precision && count == n) return;
    precision = prec;
    n = count;
    raw.resize(static_cast<size_t>(count) * element_size(prec));
  }
  void *data() { return raw.ptr; }
  const void *data() const { return raw.ptr; }
};

// Right-hand side b = int phi_i f and L2 error over cells given by their
// vertices (device [C][8][3]) and DoF map (device [C][(p+1)^3]).
void assemble_rhs_cells(int precision, int degree, int64_t n_cells, int64_t n_dofs,
                        const double *vertices, const uint32_t *map, int f_kind, int nq, void *b,
                        hipStream_t s);
double l2_error_cells(int precision, int degree, int64_t n_cells, const double *vertices,
                      const uint32_t *map, const void *u, int u_kind, int nq, hipStream_t s);

// Level operator interface (Portable::LaplaceOperatorBase,
// include/base/portable_laplace_operator_base.h:16-61): what the Chebyshev
// smoother, the V-cycle and CG need from an operator.
class Operator {
 public:
  virtual ~Operator() = default;
  // dst = A src (constrained rows: dst_c = src_c)
  virtual void vmult(void *dst, const void *src, hipStream_t s) const = 0;
  // dst += (cell/face part of) A src, no constrained-row handling
  virtual void apply_add(void *dst, const void *src, hipStream_t s) const = 0;
  virtual void compute_diagonal() = 0;
  virtual const void *inverse_diagonal() const = 0;
  virtual bool has_diagonal() const = 0;
  // b = int phi_i f (f kind 0: 1, 1: 3 pi^2 sin sin sin), Gauss(nq), constrained rows 0
  virtual void assemble_rhs(int f_kind, int nq, void *b, hipStream_t s) const = 0;
  // ||u_h - u||_L2 with Gauss(nq) (u kind 0: 0, 1: sin sin sin); synchronous
  virtual double l2_error(const void *u, int u_kind, int nq, hipStream_t s) const = 0;
  virtual int64_t m() const = 0;
  virtual int degree() const = 0;
  virtual int n_q_points_1d() const = 0;
  virtual int precision() const = 0;
  virtual int64_t n_cells() const = 0;
  virtual double algorithmic_bytes() const = 0;
  virtual double algorithmic_flops() const = 0;
  // Prepared output: an operator whose vmult first sets dst = mask ? src : 0
  // (mask = constrained_mask(), null: no constrained rows) can let the
  // producer of src write that state, and vmult_prepared then runs only the
  // cell part.  Solver glue (the Chebyshev update) uses it to save a pass.
  virtual bool supports_prepared_vmult() const { return false; }
  virtual const uint8_t *constrained_mask() const { return nullptr; }
  virtual void vmult_prepared(void *dst, const void *src, hipStream_t s) const {
    vmult(dst, src, s);
  }
  // dst_c = src_c on the constrained rows (copy_constrained_values); none by default
  virtual void copy_constrained(void *dst, const void *src, hipStream_t s) const {
    (void)dst; (void)src; (void)s;
  }
};

// Continuous Q_p Laplace operator (LaplaceOperatorBK3).
// Device arrays of a patch schedule (Bk3Schedule::pdof .. pflag)
struct PatchArrays {
  DeviceArray<uint32_t> pdof, pcount, pchunk;
  DeviceArray<uint16_t> pidx, pline;
  DeviceArray<uint8_t> pflag, pcolor, pncol;
  DeviceArray<float> mxa;  // matrix-core operand tables (fill_mx_tables)
};

class LaplaceOperatorBK3 : public Operator {
 public:
  LaplaceOperatorBK3(int degree, int nq, int precision, int64_t n_cells, const uint32_t *dofmap,
                     int64_t n_dofs, const uint32_t *constrained, int64_t n_constrained,
                     const double *cell_vertices, const double *G);
  // Structured box, every boundary DoF constrained; map, constraints and
  // vertices generated on the device, box structure declared
  LaplaceOperatorBK3(const pmg_box &box, int degree, int nq, int precision);
  void vmult(void *dst, const void *src, hipStream_t s) const override;
  void apply_add(void *dst, const void *src, hipStream_t s) const override;
  bool supports_prepared_vmult() const override;
  const uint8_t *constrained_mask() const override { return cmask_.ptr; }
  void vmult_prepared(void *dst, const void *src, hipStream_t s) const override;
  void copy_constrained(void *dst, const void *src, hipStream_t s) const override;
  void compute_diagonal() override;
  const void *inverse_diagonal() const override;
  void set_coefficient(const double *kappa, hipStream_t s);
  // Affine geometry for the vmult's atomic and patch schedules: per cell
  // A = |det J| J^-1 J^-T (6 values) instead of G at every q-point (G_q =
  // A w_p w_q w_r).  Needs cell vertices, affine cells (throws otherwise) and
  // no coefficient; set_coefficient with a kappa turns it off.  G stays for
  // the diagonal, the coarse solve, the gradient and the other schedules.
  void use_affine_geometry(bool on);
  bool uses_affine_geometry() const { return sched_.gaff != nullptr; }
  void gradient_coefficient(const void *u, const void *lam, void *out, hipStream_t s) const;
  void assemble_rhs(int f_kind, int nq, void *b, hipStream_t s) const override;
  double l2_error(const void *u, int u_kind, int nq, hipStream_t s) const override;

  int64_t m() const override { return n_dofs_; }
  int degree() const override { return degree_; }
  int n_q_points_1d() const override { return nq_; }
  int precision() const override { return precision_; }
  int64_t n_cells() const override { return n_cells_; }
  double algorithmic_bytes() const override;
  double algorithmic_flops() const override;
  bool has_diagonal() const override { return has_diag_; }
  // Jacobi-PCG solved inside one workgroup (for small coarse levels); x is the
  // initial guess; r,z,p,q work vectors of m() entries; info = 3 device doubles
  void coarse_pcg(const void *b, void *x, void *r, void *z, void *p, void *q, double rel_tol,
                  int max_it, double *info, hipStream_t s) const;
  // Domain-decomposition subdomain operator (SubdomainLaplaceOperator,
  // include/operators/portable_subdomain_laplace_operator.h): this operator,
  // built with every subdomain constraint, is vmult_bk3; set_subdomain adds
  // the plain index set (dofmap in caller cell order, only the physical
  // boundary DoFs eliminated) and the interface cells (caller indices).
  void set_subdomain(const uint32_t *dofmap, const uint32_t *physical, int64_t n_physical,
                     const uint32_t *interface_cells, int64_t n_interface);
  bool has_subdomain() const { return has_subdomain_; }
  int64_t n_interface_cells() const { return n_icells_; }
  // dst = A_plain src over every cell (vmult_neumann) or over the interface
  // cells (vmult_interface_cell_range); dst = src on the physical DoFs
  void vmult_subdomain(void *dst, const void *src, bool interface_only, hipStream_t s) const;

 private:
  template <typename T> void vmult_t(T *dst, const T *src, hipStream_t s) const;
  template <typename T>
  void vmult_subdomain_t(T *dst, const T *src, bool interface_only, hipStream_t s) const;
  template <typename T> void compute_diagonal_t();
  template <typename T> void fill_geometry_t(const double *kappa, hipStream_t s);
  void setup_shape();
  void setup_schedule();
  void setup_geometry();
  void build_coloring();
  void build_patches_once();
  void build_gather_once();
  bool ensure_cell_buffer();
  bool caller_order_ = true;  // internal cell order = caller order

  int degree_, nq_, precision_;
  int64_t n_cells_, n_dofs_, n_constrained_;
  ShapeData shape_;
  Shape1D<double> sh64_{}, shd64_{};
  Shape1D<float> sh32_{};
  DeviceArray<uint32_t> map_, constrained_, perm_;
  DeviceArray<double> vertices_, qpts_, qwts_;
  DeviceVector G_, inv_diag_;
  DeviceVector gaff_;  // affine geometry: [cell][6]
  bool has_diag_ = false;
  bool has_kappa_ = false;
  // cell colouring for the atomics-free schedule, DoF patches for the patch schedule
  DeviceArray<uint32_t> cell_ids_;
  PatchArrays patches_;
  // gather schedule: inverse map (CSR) and the per-cell buffer
  DeviceArray<uint32_t> goff_, glist_;
  DeviceVector gbuf_;
  // compressed DoF indices (x-line starts) and their per-cell flag
  DeviceArray<uint32_t> lines_;
  DeviceArray<uint8_t> lflag_;
  DeviceArray<uint8_t> cmask_;  // 1 on constrained DoFs
  Bk3Schedule sched_;
  // subdomain index set: plain map (physical DoFs eliminated) with its own
  // scatter schedule (compressed lines; patches where the main map's default
  // is the patch schedule), the physical DoFs and the interface cells
  // (device positions)
  DeviceArray<uint32_t> smap_, slines_, phys_, icells_;
  DeviceArray<uint8_t> sflag_;
  PatchArrays spatches_;
  Bk3Schedule ssched_;
  int64_t n_phys_ = 0, n_icells_ = 0;
  bool has_subdomain_ = false;

 public:
  // kAtomic (one launch, device FP atomics per cell), kColored (one launch per
  // colour, plain scatter), kPatch (batch patches in LDS) or kGather (cell
  // buffer + inverse-map gather), see pmg_bk3.hip; a schedule whose set-up
  // data is missing falls back to kAtomic.
  // gmode: G load policy (see pmg_bk3.hip)
  void set_schedule(int kind, int gmode);
  // Declares the cells a lexicographic nx x ny x nz box with the lexicographic
  // DoF numbering (verified on the device against the map; throws if not):
  // enables the grid schedule.
  void set_box_structure(int nx, int ny, int nz);
  bool has_box_structure() const { return sched_.grid[0] > 0; }
  int n_colors() const { return static_cast<int>(sched_.color_offsets.size()) - 1; }
  const Bk3Schedule &schedule() const { return sched_; }
  // Device index data by PMG_INDEX_* (pmg.h): pointer, element count and
  // element size; {nullptr, 0} where the array is not built
  struct IndexArray {
    const void *ptr;
    int64_t n;
    int elem;
  };
  IndexArray index_array(int which) const;
};

// Symmetric interior penalty DG Laplacian (LaplaceOperatorDG,
// include/operators/portable_laplace_operator_dg.h:20-151) on trilinear cells
// with cell-major FE_DGQ numbering.  inner_faces[F] = (cell-, cell+, face-,
// face+) with face+ = face- ^ 1 (standard orientation); boundary_faces[F] =
// (cell, face, boundary_id), id 0 weak homogeneous Dirichlet, others
// homogeneous Neumann.  penalty: NULL (computed from penalty_kind, 0:
// p(p+1)/2 (1/h- + 1/h+), 1: (p+1)^2 (1/h- + 1/h+), h = |K|/|F|) or one value
// per face, inner faces first.  metric 0: d_n = (J^{-1} n).grad_xi, 1: (J^{-T} n).
class LaplaceOperatorDG : public Operator {
 public:
  LaplaceOperatorDG(int degree, int nq, int precision, int64_t n_cells,
                    const double *cell_vertices, int64_t n_inner, const uint32_t *inner_faces,
                    int64_t n_boundary, const uint32_t *boundary_faces, const double *penalty,
                    int penalty_kind, int metric);
  // a box built on the device (faces in the numbering of pmg_box_dg_faces;
  // boundary id of side f: bit f of neumann_sides)
  LaplaceOperatorDG(const pmg_box &box, int degree, int nq, int precision, int penalty_kind,
                    int metric, int neumann_sides);
  void vmult(void *dst, const void *src, hipStream_t s) const override;
  void apply_add(void *dst, const void *src, hipStream_t s) const override;
  // cell-integral diagonal only, as LaplaceOperatorDG::compute_diagonal
  // (portable_laplace_operator_dg.h:887-915)
  void compute_diagonal() override;
  const void *inverse_diagonal() const override;
  bool has_diagonal() const override { return has_diag_; }
  void assemble_rhs(int f_kind, int nq, void *b, hipStream_t s) const override;
  double l2_error(const void *u, int u_kind, int nq, hipStream_t s) const override;
  int64_t m() const override { return n_cells_ * nm3_; }
  int degree() const override { return degree_; }
  int n_q_points_1d() const override { return nq_; }
  int precision() const override { return precision_; }
  int64_t n_cells() const override { return n_cells_; }
  int64_t n_faces() const { return n_faces_; }
  double algorithmic_bytes() const override;
  double algorithmic_flops() const override;
  // penalty per face (host copy, inner faces first)
  const std::vector<double> &penalty() const { return sigma_; }

 private:
  template <typename T> void apply_t(bool add, T *dst, const T *src, hipStream_t s) const;
  const uint32_t *identity_map() const;
  void setup_1d();
  void setup_geometry(const uint32_t *finfo_device, int metric);
  int degree_, nq_, precision_;
  int64_t n_cells_, nm3_, n_faces_ = 0;
  ShapeData shape_;
  Shape1D<double> sh64_{}, shd64_{};
  Shape1D<float> sh32_{};
  DgShape<double> fs64_{};
  DgShape<float> fs32_{};
  std::vector<double> sigma_;
  DeviceArray<uint32_t> cface_, cnbr_;
  DeviceArray<double> vertices_;
  DeviceVector G_, fdata_, fsig_, inv_diag_;
  mutable DeviceArray<uint32_t> ident_;
  bool has_diag_ = false;
};

class MGTransfer {
 public:
  // weights == nullptr: the weights of setup_weights (1/multiplicity, 0 on the
  // n_constrained_fine listed fine DoFs) are built on the device, in owner form
  MGTransfer(int kind, int precision, int degree_coarse, int degree_fine, int64_t n_cells,
             const uint32_t *map_coarse, const uint32_t *map_fine, const double *weights,
             int64_t n_dofs_coarse, int64_t n_dofs_fine, double prolongation_scale,
             const uint32_t *constrained_fine = nullptr, int64_t n_constrained_fine = -1);
  // Between structured boxes, every boundary DoF constrained, maps built on
  // the device: h (box -> box refined once), p or DG-CG (same box)
  MGTransfer(int kind, int precision, const pmg_box &box, int degree_coarse, int degree_fine,
             double prolongation_scale);
  void prolongate_and_add(void *dst_fine, const void *src_coarse, hipStream_t s) const;
  void restrict_and_add(void *dst_coarse, const void *src_fine, hipStream_t s) const;
  // dst_coarse += R (b - t) (the V-cycle's defect restriction, one pass)
  void restrict_residual_and_add(void *dst_coarse, const void *b, const void *t,
                                 hipStream_t s) const;
  int precision() const { return precision_; }
  int kind() const { return kind_; }
  int degree_coarse() const { return pc_; }
  int degree_fine() const { return pf_; }
  int64_t n_dofs_coarse() const { return n_dofs_c_; }
  int64_t n_dofs_fine() const { return n_dofs_f_; }
  // fine-map storage: 0 weighted (explicit weights), 1 owner marks in the
  // full fine map, 2 owner marks as x-line words {first DoF, owner bits}
  int layout() const { return !owner_ ? 0 : (fine_lines_.ptr ? 2 : 1); }
  // MGTransferBase::reinit (include/base/portable_mg_transfer_base.h:32-35):
  // rebuild in place from new maps and fine constraints (weights built on
  // the device; DG-CG: constraints unused), keeping kind, precision, degrees
  // and prolongation scale.  generation() changes, so a V-cycle that
  // captured this transfer's buffers in its graph captures it again.
  void reinit(int64_t n_cells, const uint32_t *map_coarse, const uint32_t *map_fine,
              const uint32_t *constrained_fine, int64_t n_constrained_fine,
              int64_t n_dofs_coarse, int64_t n_dofs_fine);
  uint64_t generation() const { return generation_; }
  // Device index data by PMG_TRANSFER_INDEX_* (pmg.h)
  LaplaceOperatorBK3::IndexArray index_array(int which) const;

 private:
  void restrict_sub(void *dst, const void *src, const void *sub, hipStream_t s) const;
  void init_kind();
  void setup_dg_inverse(const uint32_t *map_fine);
  void setup_restrict_gather();
  void setup_owner(const uint8_t *dmask);
  int kind_, precision_, pc_ = 0, pf_ = 0, nmc_, nmf_;
  double pscale_ = 1.0;
  uint64_t generation_ = 0;
  int64_t n_cells_, n_dofs_c_, n_dofs_f_;
  bool owner_ = false;  // owner-computes mode (weights 1/multiplicity), see pmg_transfer.hip
  double scale_ = 1.0;
  Transfer1D<double> M64_{};
  Transfer1D<float> M32_{};
  DeviceArray<uint32_t> map_c_, map_f_;
  // DG-CG: inverse map (CSR) of the CG DoFs to their DG copies for the
  // atomics-free restriction (empty if the entry count needs 64-bit offsets)
  DeviceArray<uint32_t> inv_off_, inv_list_;
  // h/p restriction through a cell buffer: inverse coarse map (CSR over
  // positions cell * nc^3 + i) and the buffer; a transfer's restrictions
  // must therefore be ordered (they share the buffer)
  DeviceArray<uint32_t> cinv_off_, cpos_;  // restriction buffer: ranges and positions
  DeviceArray<uint32_t> fine_lines_;  // owner mode: x-line words {first DoF, owner bits}
  mutable DeviceVector cell_buf_;
  DeviceVector w_prolong_, w_restrict_;
};

class PreconditionChebyshev {
 public:
  PreconditionChebyshev(const Operator &op, int degree, double smoothing_range,
                        int eig_cg_n_iterations);
  void vmult(void *dst, const void *src, hipStream_t s) const;
  void step(void *dst, const void *src, hipStream_t s) const;
  // vmult that also writes next_out = the prepared output (see Operator) of
  // a matvec of dst; returns false (next_out untouched) if the operator has none
  bool vmult_and_prepare(void *dst, const void *src, void *next_out, hipStream_t s) const;
  double max_eigenvalue() const { return max_ev_; }
  double min_eigenvalue() const { return min_ev_; }
  int degree() const { return degree_; }
  const Operator &op() const { return op_; }

 private:
  void run(void *dst, const void *src, bool step, void *next_out, hipStream_t s) const;
  const Operator &op_;
  int degree_;
  double range_, max_ev_ = 1, min_ev_ = 1, theta_ = 1, delta_ = 0;
  mutable DeviceVector tmp_, old_;
};

// Work space of one CG run as raw device pointers: 8 scalars, kDotBlocks
// partial sums, and r, z, p, q (m() entries of the operator precision).
// host_rr: two pinned host doubles the device stores r.r and the iteration
// that converged into, or null (the scalars are then copied back at each
// convergence check).
struct CGSpace {
  double *sc = nullptr, *partials = nullptr;
  void *r = nullptr, *z = nullptr, *p = nullptr, *q = nullptr;
  double *host_rr = nullptr;
};

// Scalars and work space of a CG run on the device (owning).
struct CGWork {
  DeviceArray<double> sc, partials;
  DeviceVector r, z, p, q;
  PinnedDoubles host;  // r.r of the last iteration, stored by the device
  void reinit(int precision, int64_t n);
  CGSpace space();  // allocates host on first use
};

// Preconditioned CG: precond 0 none, 1 Jacobi, 2 callback.
struct CGResult {
  int iterations = 0;
  double res0 = 0, res = 0;
  bool converged = false;
};

class VCycleMultigrid;

// check_every: iterations queued between host checks of the residual (the
// result does not depend on it: kernels queued after convergence change
// nothing); 0 = kCgCheckEvery without, 1 with the multigrid preconditioner
// (a V-cycle queued past convergence costs more than a check).
constexpr int kCgCheckEvery = 8;
CGResult solve_cg(const Operator &op, int precond, const VCycleMultigrid *mg,
                  const void *b, void *x, double rel_tol, int max_it, CGWork &work,
                  hipStream_t s, int check_every = 0);
// The same on caller-provided work space (allocates nothing itself).
CGResult solve_cg(const Operator &op, int precond, const VCycleMultigrid *mg,
                  const void *b, void *x, double rel_tol, int max_it, const CGSpace &space,
                  hipStream_t s, int check_every = 0);

// solve_cg on the library's fixed device work array (pmg_noalloc.hip):
// allocates nothing; throws PMG_ERR_SIZE above cg_noalloc_max_dofs(precision).
CGResult solve_cg_noalloc(const Operator &op, int precond, const VCycleMultigrid *mg,
                          const void *b, void *x, double rel_tol, int max_it, hipStream_t s);
int64_t cg_noalloc_max_dofs(int precision);

// Coarse levels with at most this many cells are solved by the one-workgroup PCG kernel.
constexpr int64_t kCoarseSingleWorkgroupCells = 4096;

class VCycleMultigrid {
 public:
  VCycleMultigrid(std::vector<const Operator *> ops,
                  std::vector<const MGTransfer *> transfers,
                  std::vector<const PreconditionChebyshev *> smoothers, int coarse_kind,
                  double coarse_tol, int outer_precision);
  ~VCycleMultigrid();
  VCycleMultigrid(const VCycleMultigrid &) = delete;
  VCycleMultigrid &operator=(const VCycleMultigrid &) = delete;
  void vmult(void *dst, const void *src, hipStream_t s) const;
  // The V-cycle without the conversion of its result: returns the finest
  // level's solution (level precision, valid until the next application).
  const void *apply(const void *src, hipStream_t s) const;
  // apply() in two parts: stage converts src into the finest level's defect
  // (on the device: not if *skip != 0; skip may be null), run executes the
  // V-cycle on it and returns the finest level's solution
  void stage(const void *src, hipStream_t s, const double *skip) const;
  const void *run(hipStream_t s) const;
  // the buffer stage() writes (the finest level's defect, level precision):
  // a producer of the next input may write it directly instead
  void *input_buffer() const { return def_.back().data(); }
  int outer_precision() const { return outer_precision_; }
  int level_precision() const { return precision_; }
  int64_t m() const { return ops_.back()->m(); }  // size of the finest level
  int levels() const { return static_cast<int>(ops_.size()); }
  // the V-cycle body runs as a captured HIP graph (no host sync inside it)
  bool uses_graph() const { return use_graph_; }
  // Per-level stage times of n_applications eager V-cycles on src (the
  // reference's timing table, tests/common/portable_multigrid_solver.h:108-132):
  // times[l * kStages + stage] = average seconds per application, stages
  // pre-smoothing, residual matvec, restriction (with the defect), coarse
  // solve (level 0 only), prolongation, post-smoothing.  Synchronous.
  static constexpr int kStages = 6;
  void level_times(const void *src, int n_applications, double *times, hipStream_t s) const;

 private:
  void v_cycle(int level, hipStream_t s) const;
  void capture() const;
  // stage events of level_times (null outside it)
  struct StageEvent {
    int level, stage;
    hipEvent_t begin, end;
  };
  mutable std::vector<StageEvent> *prof_ = nullptr;
  void mark(int level, int stage, bool begin, hipStream_t s) const;
  std::vector<const Operator *> ops_;
  std::vector<const MGTransfer *> xfer_;
  std::vector<const PreconditionChebyshev *> sm_;
  int coarse_kind_, outer_precision_, precision_;
  double coarse_tol_;
  std::unique_ptr<PreconditionChebyshev> coarse_cheby_;
  mutable std::vector<DeviceVector> sol_, def_, t_;
  mutable CGWork coarse_work_;
  // FP64 coarse level under FP32 levels: the coarse solve's own FP64 defect
  // and solution, converted from and to the level-0 vectors around it
  bool coarse64_ = false;
  mutable DeviceVector c64_def_, c64_sol_;
  mutable bool use_graph_ = false;  // cleared for good if capturing fails
  mutable hipStream_t cap_stream_ = nullptr;
  mutable hipGraph_t graph_ = nullptr;
  mutable hipGraphExec_t graph_exec_ = nullptr;
  // transfers' generation() at capture: a reinit'ed transfer re-captures
  mutable std::vector<uint64_t> xfer_gen_;
  void check_transfers() const;
};

}  // namespace pmg
