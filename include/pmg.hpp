/* pmg.hpp -- header-only C++17 mirror of the reference's operator, transfer,
 * smoother and V-cycle classes over the C ABI of pmg.h.
 *
 * The class and method names follow the reference so that a caller written
 * against it reads the same:
 *   LaplaceOperatorBK3 / LaplaceOperatorDG   Portable::LaplaceOperatorBase
 *       (include/base/portable_laplace_operator_base.h:16-61): vmult, Tvmult,
 *       compute_diagonal, get_matrix_diagonal_inverse, m
 *   GeometricTransfer / PolynomialTransfer / ContinuousTransfer
 *       Portable::MGTransferBase (include/base/portable_mg_transfer_base.h:15-36):
 *       prolongate_and_add, restrict_and_add
 *   PreconditionChebyshev  deal.II PreconditionChebyshev: vmult, step
 *   VCycleMultigrid        VCycleMultigrid::vmult (portable_v_cycle_multigrid.h:163-206)
 *   solve_cg               MultigridSolver::solve_cg (tests/common/portable_multigrid_solver.h:147-169)
 * Vectors are raw device pointers of the operator precision; every call is
 * ordered on the given HIP stream (0 = default stream).  Failures throw
 * pmg_hip::Error carrying the status and pmg_last_error().  Objects own their
 * handle and are move-only.
 */
#ifndef PMG_HPP
#define PMG_HPP

#include <cstdint>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "pmg.h"

namespace pmg_hip {

struct Error : std::runtime_error {
  int status;
  Error(int s, const std::string &what) : std::runtime_error(what), status(s) {}
};

inline void check(int status) {
  if (status != PMG_OK) throw Error(status, pmg_last_error());
}

template <typename H, int (*Destroy)(H)>
class Handle {
 public:
  Handle() = default;
  explicit Handle(H h) : h_(h) {}
  Handle(Handle &&o) noexcept : h_(std::exchange(o.h_, nullptr)) {}
  Handle &operator=(Handle &&o) noexcept {
    if (this != &o) {
      reset();
      h_ = std::exchange(o.h_, nullptr);
    }
    return *this;
  }
  Handle(const Handle &) = delete;
  Handle &operator=(const Handle &) = delete;
  ~Handle() { reset(); }
  H get() const { return h_; }

 private:
  void reset() {
    if (h_) Destroy(h_);
    h_ = nullptr;
  }
  H h_ = nullptr;
};

// ------------------------------------------------------------------ operators
class LaplaceOperatorBase {
 public:
  virtual ~LaplaceOperatorBase() = default;
  void vmult(void *dst, const void *src, pmg_stream s = nullptr) const {
    check(pmg_operator_vmult(h_.get(), dst, src, s));
  }
  void Tvmult(void *dst, const void *src, pmg_stream s = nullptr) const {
    check(pmg_operator_Tvmult(h_.get(), dst, src, s));
  }
  void apply_add(void *dst, const void *src, pmg_stream s = nullptr) const {
    check(pmg_operator_apply_add(h_.get(), dst, src, s));
  }
  // vmult_dummy (portable_laplace_operator_base.h:27-31): one GPU has no
  // ghost exchange, so ghost_exchange_on only adds the constrained-row copy
  void vmult_dummy(void *dst, const void *src, bool ghost_exchange_on, bool computation_on,
                   pmg_stream s = nullptr) const {
    check(pmg_operator_vmult_dummy(h_.get(), dst, src, ghost_exchange_on ? 1 : 0,
                                   computation_on ? 1 : 0, s));
  }
  void compute_diagonal() { check(pmg_operator_compute_diagonal(h_.get())); }
  // el(row, col) (portable_laplace_operator_bk3.h:586-595): diagonal entries only
  double el(int64_t row, int64_t col) const {
    if (row != col) throw Error(PMG_ERR_UNSUPPORTED, "el: only diagonal entries");
    double v = 0;
    check(pmg_operator_diagonal_entry(h_.get(), row, &v));
    return v;
  }
  // copies the inverse diagonal into out (device, m() entries)
  void get_matrix_diagonal_inverse(void *out, pmg_stream s = nullptr) const {
    check(pmg_operator_diagonal_inverse(h_.get(), out, s));
  }
  int64_t m() const {
    int64_t v = 0;
    check(pmg_operator_m(h_.get(), &v));
    return v;
  }
  int64_t n() const { return m(); }
  int precision() const {
    int p = 0;
    check(pmg_operator_info(h_.get(), nullptr, nullptr, &p, nullptr));
    return p;
  }
  void assemble_rhs(int f_kind, void *b, pmg_stream s = nullptr, int n_q_points_1d = 0) const {
    check(pmg_operator_assemble_rhs(h_.get(), f_kind, n_q_points_1d, b, s));
  }
  double l2_error(const void *u, int u_kind, pmg_stream s = nullptr, int n_q_points_1d = 0) const {
    double e = 0;
    check(pmg_operator_l2_error(h_.get(), u, u_kind, n_q_points_1d, &e, s));
    return e;
  }
  pmg_operator handle() const { return h_.get(); }

 protected:
  Handle<pmg_operator, pmg_operator_destroy> h_;
};

class LaplaceOperatorBK3 : public LaplaceOperatorBase {
 public:
  // dofmap [C][(p+1)^3] (0xFFFFFFFF = constrained), constrained DoF list,
  // cell_vertices [C][8][3]; all host arrays
  LaplaceOperatorBK3(int degree, int n_q_points_1d, int precision, int64_t n_cells,
                     const uint32_t *dofmap, int64_t n_dofs, const uint32_t *constrained,
                     int64_t n_constrained, const double *cell_vertices) {
    pmg_operator h = nullptr;
    check(pmg_operator_create(3, degree, n_q_points_1d, precision, n_cells, dofmap, n_dofs,
                              constrained, n_constrained, cell_vertices, &h));
    h_ = Handle<pmg_operator, pmg_operator_destroy>(h);
  }
  // structured box (dirichlet on the whole boundary), index data built on the device
  LaplaceOperatorBK3(const pmg_box &box, int degree, int n_q_points_1d, int precision) {
    pmg_operator h = nullptr;
    check(pmg_operator_create_box(&box, degree, n_q_points_1d, precision, &h));
    h_ = Handle<pmg_operator, pmg_operator_destroy>(h);
  }
  void set_coefficient(const double *kappa_device, pmg_stream s = nullptr) {
    check(pmg_operator_set_coefficient(h_.get(), kappa_device, s));
  }
  // affine cells: vmult reads 6 values per cell instead of G at every q-point
  void use_affine_geometry(bool enable = true) {
    check(pmg_operator_use_affine_geometry(h_.get(), enable ? 1 : 0));
  }
  void gradient_coefficient(const void *u, const void *lambda, void *out,
                            pmg_stream s = nullptr) const {
    check(pmg_operator_gradient_coefficient(h_.get(), u, lambda, out, s));
  }
  // scatter schedule 0 atomic, 1 colored, 2 patch; g_load -1 = the schedule's default
  void set_schedule(int schedule, int g_load = -1) {
    check(pmg_operator_set_schedule(h_.get(), schedule, g_load));
  }
  // SubdomainLaplaceOperator: this operator (every subdomain constraint) is
  // vmult_bk3; the plain index set and the interface cells (host arrays)
  // enable vmult_neumann and vmult_interface_cell_range
  void set_subdomain(const uint32_t *dofmap, const uint32_t *physical_dofs, int64_t n_physical,
                     const uint32_t *interface_cells, int64_t n_interface_cells) {
    check(pmg_operator_set_subdomain(h_.get(), dofmap, physical_dofs, n_physical,
                                     interface_cells, n_interface_cells));
  }
  void vmult_bk3(void *dst, const void *src, pmg_stream s = nullptr) const { vmult(dst, src, s); }
  void vmult_neumann(void *dst, const void *src, pmg_stream s = nullptr) const {
    check(pmg_operator_vmult_neumann(h_.get(), dst, src, s));
  }
  void vmult_interface_cell_range(void *dst, const void *src, pmg_stream s = nullptr) const {
    check(pmg_operator_vmult_interface_cells(h_.get(), dst, src, s));
  }
};

class LaplaceOperatorDG : public LaplaceOperatorBase {
 public:
  LaplaceOperatorDG(int degree, int n_q_points_1d, int precision, int64_t n_cells,
                    const double *cell_vertices, int64_t n_inner, const uint32_t *inner_faces,
                    int64_t n_boundary, const uint32_t *boundary_faces,
                    const double *penalty = nullptr, int penalty_kind = 0, int metric = 0) {
    pmg_operator h = nullptr;
    check(pmg_dg_operator_create(3, degree, n_q_points_1d, precision, n_cells, cell_vertices,
                                 n_inner, inner_faces, n_boundary, boundary_faces, penalty,
                                 penalty_kind, metric, &h));
    h_ = Handle<pmg_operator, pmg_operator_destroy>(h);
  }
  LaplaceOperatorDG(const pmg_box &box, int degree, int n_q_points_1d, int precision,
                    int penalty_kind = 0, int metric = 0) {
    const int64_t C = pmg_box_n_cells(&box);
    std::vector<double> cv(static_cast<size_t>(C * 24));
    check(pmg_box_cell_vertices(&box, cv.data()));
    int64_t ni = 0, nb = 0;
    check(pmg_box_dg_faces(&box, nullptr, &ni, nullptr, &nb));
    std::vector<uint32_t> inner(static_cast<size_t>(4 * ni)), bdry(static_cast<size_t>(3 * nb));
    check(pmg_box_dg_faces(&box, inner.data(), &ni, bdry.data(), &nb));
    pmg_operator h = nullptr;
    check(pmg_dg_operator_create(3, degree, n_q_points_1d, precision, C, cv.data(), ni,
                                 inner.data(), nb, bdry.data(), nullptr, penalty_kind, metric,
                                 &h));
    h_ = Handle<pmg_operator, pmg_operator_destroy>(h);
  }
};

// ------------------------------------------------------------------ transfers
class MGTransferBase {
 public:
  void prolongate_and_add(void *dst_fine, const void *src_coarse, pmg_stream s = nullptr) const {
    check(pmg_transfer_prolongate_and_add(h_.get(), dst_fine, src_coarse, s));
  }
  void restrict_and_add(void *dst_coarse, const void *src_fine, pmg_stream s = nullptr) const {
    check(pmg_transfer_restrict_and_add(h_.get(), dst_coarse, src_fine, s));
  }
  // MGTransferBase::reinit: new maps (coarse constraints as invalid entries)
  // and fine constraints; weights built on the device
  void reinit(int64_t n_cells, const uint32_t *map_coarse, const uint32_t *map_fine,
              const uint32_t *constrained_fine, int64_t n_constrained_fine, int64_t n_dofs_coarse,
              int64_t n_dofs_fine) {
    check(pmg_transfer_reinit(h_.get(), n_cells, map_coarse, map_fine, constrained_fine,
                              n_constrained_fine, n_dofs_coarse, n_dofs_fine));
  }
  pmg_transfer handle() const { return h_.get(); }

 protected:
  void create(int kind, int precision, int pc, int pf, int64_t n_cells, const uint32_t *mc,
              const uint32_t *mf, const double *w, int64_t nc, int64_t nf, double scale) {
    pmg_transfer h = nullptr;
    check(pmg_transfer_create(kind, precision, pc, pf, n_cells, mc, mf, w, nc, nf, scale, &h));
    h_ = Handle<pmg_transfer, pmg_transfer_destroy>(h);
  }
  void create_box(int kind, int precision, const pmg_box &box, int pc, int pf, double scale) {
    pmg_transfer h = nullptr;
    check(pmg_transfer_create_box(kind, precision, &box, pc, pf, scale, &h));
    h_ = Handle<pmg_transfer, pmg_transfer_destroy>(h);
  }
  Handle<pmg_transfer, pmg_transfer_destroy> h_;
};

// coarse box -> once-refined box at one degree (maps and weights built on the device)
class GeometricTransfer : public MGTransferBase {
 public:
  GeometricTransfer(const pmg_box &coarse, int degree, int precision, double prolongation_scale = 1.0) {
    create_box(PMG_TRANSFER_H, precision, coarse, degree, degree, prolongation_scale);
  }
};

// degree pc -> pf on one box
class PolynomialTransfer : public MGTransferBase {
 public:
  PolynomialTransfer(const pmg_box &box, int pc, int pf, int precision) {
    create_box(PMG_TRANSFER_P, precision, box, pc, pf, 1.0);
  }
};

// CG(p) -> DG(p) on one box (cell-major DG numbering)
class ContinuousTransfer : public MGTransferBase {
 public:
  ContinuousTransfer(const pmg_box &box, int degree, int precision) {
    create_box(PMG_TRANSFER_DG_CG, precision, box, degree, degree, 1.0);
  }
};

// ------------------------------------------------------------------ smoother, V-cycle, CG
class PreconditionChebyshev {
 public:
  PreconditionChebyshev(const LaplaceOperatorBase &op, int degree, double smoothing_range = 15.0,
                        int eig_cg_n_iterations = 10) {
    pmg_smoother h = nullptr;
    check(pmg_chebyshev_create(op.handle(), degree, smoothing_range, eig_cg_n_iterations, &h));
    h_ = Handle<pmg_smoother, pmg_chebyshev_destroy>(h);
  }
  void vmult(void *dst, const void *src, pmg_stream s = nullptr) const {
    check(pmg_chebyshev_vmult(h_.get(), dst, src, s));
  }
  void step(void *dst, const void *src, pmg_stream s = nullptr) const {
    check(pmg_chebyshev_step(h_.get(), dst, src, s));
  }
  pmg_smoother handle() const { return h_.get(); }

 private:
  Handle<pmg_smoother, pmg_chebyshev_destroy> h_;
};

class VCycleMultigrid {
 public:
  // levels from coarsest (0) to finest; transfers[l] joins l-1 -> l and
  // smoothers[l] smooths level l (entries 0 are ignored and may be NULL)
  VCycleMultigrid(const std::vector<const LaplaceOperatorBase *> &ops,
                  const std::vector<const MGTransferBase *> &transfers,
                  const std::vector<const PreconditionChebyshev *> &smoothers, int coarse_kind,
                  double coarse_tol, int outer_precision) {
    std::vector<pmg_operator> o;
    std::vector<pmg_transfer> t;
    std::vector<pmg_smoother> s;
    for (size_t l = 0; l < ops.size(); ++l) {
      o.push_back(ops[l]->handle());
      t.push_back(l > 0 && transfers[l] ? transfers[l]->handle() : nullptr);
      s.push_back(l > 0 && smoothers[l] ? smoothers[l]->handle() : nullptr);
    }
    pmg_multigrid h = nullptr;
    check(pmg_multigrid_create(static_cast<int>(o.size()), o.data(), t.data(), s.data(),
                               coarse_kind, coarse_tol, outer_precision, &h));
    h_ = Handle<pmg_multigrid, pmg_multigrid_destroy>(h);
  }
  void vmult(void *dst, const void *src, pmg_stream s = nullptr) const {
    check(pmg_multigrid_vmult(h_.get(), dst, src, s));
  }
  // true if vmult replays a captured HIP graph of the V-cycle body
  bool uses_graph() const {
    int g = 0;
    check(pmg_multigrid_uses_graph(h_.get(), &g));
    return g != 0;
  }
  pmg_multigrid handle() const { return h_.get(); }

 private:
  Handle<pmg_multigrid, pmg_multigrid_destroy> h_;
};

struct SolverControl {
  int iterations = 0;
  double initial_residual = 0, final_residual = 0;
};

// returns false if max_it was reached without convergence
inline bool solve_cg(const LaplaceOperatorBase &op, const void *b, void *x, int precond,
                     const VCycleMultigrid *mg, double rel_tol, int max_it,
                     SolverControl *control = nullptr, pmg_stream s = nullptr) {
  SolverControl c;
  const int st = pmg_cg_solve(op.handle(), precond, mg ? mg->handle() : nullptr, b, x, rel_tol,
                              max_it, &c.iterations, &c.initial_residual, &c.final_residual, s);
  if (control) *control = c;
  if (st == PMG_ERR_NOT_CONVERGED) return false;
  check(st);
  return true;
}

}  // namespace pmg_hip

#endif
