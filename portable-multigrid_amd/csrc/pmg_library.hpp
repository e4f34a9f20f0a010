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

// Untyped owning device buffer of either precision.
struct DeviceVector {
  DeviceArray<char> raw;
  int precision = PMG_FP64;
  int64_t n = 0;
  void reinit(int prec, int64_t count) {
    precision = prec;
    n = count;
    raw.resize(static_cast<size_t>(count) * element_size(prec));
  }
  void *data() { return raw.ptr; }
  const void *data() const { return raw.ptr; }
};

// Continuous Q_p Laplace operator (LaplaceOperatorBK3 / LaplaceOperatorBase).
class LaplaceOperatorBK3 {
 public:
  LaplaceOperatorBK3(int degree, int nq, int precision, int64_t n_cells, const uint32_t *dofmap,
                     int64_t n_dofs, const uint32_t *constrained, int64_t n_constrained,
                     const double *cell_vertices, const double *G);
  void vmult(void *dst, const void *src, hipStream_t s) const;
  void apply_add(void *dst, const void *src, hipStream_t s) const;
  void compute_diagonal();
  const void *inverse_diagonal() const;
  void set_coefficient(const double *kappa, hipStream_t s);
  void gradient_coefficient(const void *u, const void *lam, void *out, hipStream_t s) const;
  // b = int phi_i f (f kind 0: 1, 1: 3 pi^2 sin sin sin), Gauss(nq), constrained rows 0
  void assemble_rhs(int f_kind, int nq, void *b, hipStream_t s) const;
  // ||u_h - u||_L2 with Gauss(nq) (u kind 0: 0, 1: sin sin sin); synchronous
  double l2_error(const void *u, int u_kind, int nq, hipStream_t s) const;

  int64_t m() const { return n_dofs_; }
  int degree() const { return degree_; }
  int n_q_points_1d() const { return nq_; }
  int precision() const { return precision_; }
  int64_t n_cells() const { return n_cells_; }
  double algorithmic_bytes() const;
  double algorithmic_flops() const;
  bool has_diagonal() const { return has_diag_; }
  // Jacobi-PCG solved inside one workgroup (for small coarse levels); x is the
  // initial guess; r,z,p,q work vectors of m() entries; info = 3 device doubles
  void coarse_pcg(const void *b, void *x, void *r, void *z, void *p, void *q, double rel_tol,
                  int max_it, double *info, hipStream_t s) const;

 private:
  template <typename T> void vmult_t(T *dst, const T *src, hipStream_t s) const;
  template <typename T> void compute_diagonal_t();
  template <typename T> void fill_geometry_t(const double *kappa, hipStream_t s);

  int degree_, nq_, precision_;
  int64_t n_cells_, n_dofs_, n_constrained_;
  ShapeData shape_;
  Shape1D<double> sh64_{}, shd64_{};
  Shape1D<float> sh32_{}, shd32_{};
  DeviceArray<uint32_t> map_, constrained_, perm_;
  DeviceArray<double> vertices_, qpts_, qwts_;
  DeviceVector G_, inv_diag_;
  bool has_diag_ = false;
  // cell colouring for the atomics-free schedule
  DeviceArray<uint32_t> cell_ids_;
  Bk3Schedule sched_;

 public:
  // "atomic" (one launch, hardware FP atomics) or "colored" (one launch per
  // colour, plain scatter); gmode: G load policy (see pmg_bk3.hip)
  void set_schedule(bool colored, int gmode) {
    sched_.colored = colored && !sched_.color_offsets.empty();
    sched_.gmode = gmode;
  }
  int n_colors() const { return static_cast<int>(sched_.color_offsets.size()) - 1; }
  const Bk3Schedule &schedule() const { return sched_; }
};

class MGTransfer {
 public:
  MGTransfer(int kind, int precision, int degree_coarse, int degree_fine, int64_t n_cells,
             const uint32_t *map_coarse, const uint32_t *map_fine, const double *weights,
             int64_t n_dofs_coarse, int64_t n_dofs_fine, double prolongation_scale);
  void prolongate_and_add(void *dst_fine, const void *src_coarse, hipStream_t s) const;
  void restrict_and_add(void *dst_coarse, const void *src_fine, hipStream_t s) const;
  int precision() const { return precision_; }
  int64_t n_dofs_coarse() const { return n_dofs_c_; }
  int64_t n_dofs_fine() const { return n_dofs_f_; }

 private:
  int precision_, nmc_, nmf_;
  int64_t n_cells_, n_dofs_c_, n_dofs_f_;
  Transfer1D<double> M64_{};
  Transfer1D<float> M32_{};
  DeviceArray<uint32_t> map_c_, map_f_;
  DeviceVector w_prolong_, w_restrict_;
};

class PreconditionChebyshev {
 public:
  PreconditionChebyshev(const LaplaceOperatorBK3 &op, int degree, double smoothing_range,
                        int eig_cg_n_iterations);
  void vmult(void *dst, const void *src, hipStream_t s) const;
  void step(void *dst, const void *src, hipStream_t s) const;
  double max_eigenvalue() const { return max_ev_; }
  double min_eigenvalue() const { return min_ev_; }
  int degree() const { return degree_; }
  const LaplaceOperatorBK3 &op() const { return op_; }

 private:
  void run(void *dst, const void *src, bool step, hipStream_t s) const;
  const LaplaceOperatorBK3 &op_;
  int degree_;
  double range_, max_ev_ = 1, min_ev_ = 1, theta_ = 1, delta_ = 0;
  mutable DeviceVector tmp_, old_;
};

// Scalars and work space of a CG run on the device.
struct CGWork {
  DeviceArray<double> sc, partials;
  DeviceVector r, z, p, q;
  void reinit(int precision, int64_t n);
};

// Preconditioned CG: precond 0 none, 1 Jacobi, 2 callback.
struct CGResult {
  int iterations = 0;
  double res0 = 0, res = 0;
  bool converged = false;
};

class VCycleMultigrid;

CGResult solve_cg(const LaplaceOperatorBK3 &op, int precond, const VCycleMultigrid *mg,
                  const void *b, void *x, double rel_tol, int max_it, CGWork &work,
                  hipStream_t s, int check_every = 1);

// Coarse levels with at most this many cells are solved by the one-workgroup PCG kernel.
constexpr int64_t kCoarseSingleWorkgroupCells = 4096;

class VCycleMultigrid {
 public:
  VCycleMultigrid(std::vector<const LaplaceOperatorBK3 *> ops,
                  std::vector<const MGTransfer *> transfers,
                  std::vector<const PreconditionChebyshev *> smoothers, int coarse_kind,
                  double coarse_tol, int outer_precision);
  void vmult(void *dst, const void *src, hipStream_t s) const;
  int outer_precision() const { return outer_precision_; }
  int levels() const { return static_cast<int>(ops_.size()); }

 private:
  void v_cycle(int level, hipStream_t s) const;
  std::vector<const LaplaceOperatorBK3 *> ops_;
  std::vector<const MGTransfer *> xfer_;
  std::vector<const PreconditionChebyshev *> sm_;
  int coarse_kind_, outer_precision_, precision_;
  double coarse_tol_;
  std::unique_ptr<PreconditionChebyshev> coarse_cheby_;
  mutable std::vector<DeviceVector> sol_, def_, t_;
  mutable CGWork coarse_work_;
};

}  // namespace pmg
