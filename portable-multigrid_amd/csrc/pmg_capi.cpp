// pmg_capi.cpp -- extern "C" boundary over the host classes.  Each entry point
// converts exceptions into a status code and records the message for
// pmg_last_error(); handles own their device metadata.
#include <algorithm>
#include <memory>
#include <string>

#include "pmg_library.hpp"

struct pmg_operator_s {
  std::unique_ptr<pmg::Operator> impl;
  pmg::CGWork cg_work;  // CG vectors of pmg_cg_solve on this operator, kept between solves
};
struct pmg_transfer_s {
  std::unique_ptr<pmg::MGTransfer> impl;
};
struct pmg_smoother_s {
  std::unique_ptr<pmg::PreconditionChebyshev> impl;
};
struct pmg_multigrid_s {
  std::unique_ptr<pmg::VCycleMultigrid> impl;
};

namespace pmg {
namespace {
thread_local std::string g_last_error;
}
void set_last_error(const std::string &msg) { g_last_error = msg; }
}  // namespace pmg

namespace {
template <typename F>
int guard(F &&f) {
  try {
    pmg::g_last_error.clear();
    f();
    return PMG_OK;
  } catch (const pmg::Error &e) {
    pmg::set_last_error(e.what());
    return e.status;
  } catch (const std::bad_alloc &) {
    pmg::set_last_error("host allocation failed");
    return PMG_ERR_ALLOC;
  } catch (const std::exception &e) {
    pmg::set_last_error(e.what());
    return PMG_ERR_INVALID_ARG;
  } catch (...) {
    pmg::set_last_error("unknown error");
    return PMG_ERR_INVALID_ARG;
  }
}
hipStream_t as_stream(pmg_stream s) { return static_cast<hipStream_t>(s); }
pmg::LaplaceOperatorBK3 &bk3_of(pmg_operator op) {
  auto *p = dynamic_cast<pmg::LaplaceOperatorBK3 *>(op->impl.get());
  PMG_REQUIRE(p != nullptr, PMG_ERR_UNSUPPORTED, "this entry point needs a continuous (BK3) operator");
  return *p;
}
#define PMG_NOT_NULL(p) PMG_REQUIRE((p) != nullptr, PMG_ERR_INVALID_ARG, #p " is NULL")
}  // namespace

extern "C" {

const char *pmg_last_error(void) { return pmg::g_last_error.c_str(); }

int pmg_version(void) { return 10000; }

int pmg_device_synchronize(void) {
  return guard([] { PMG_HIP_CHECK(hipDeviceSynchronize()); });
}

// ---------------------------------------------------------------- operator
int pmg_operator_create(int dim, int degree, int nq, int precision, int64_t n_cells,
                        const uint32_t *dofmap, int64_t n_dofs, const uint32_t *constrained,
                        int64_t n_constrained, const double *cell_vertices, pmg_operator *op) {
  return guard([&] {
    PMG_NOT_NULL(op);
    *op = nullptr;
    PMG_REQUIRE(dim == 3, PMG_ERR_UNSUPPORTED, "only dim = 3 is supported");
    PMG_NOT_NULL(cell_vertices);
    auto h = std::make_unique<pmg_operator_s>();
    h->impl = std::make_unique<pmg::LaplaceOperatorBK3>(degree, nq, precision, n_cells, dofmap, n_dofs,
                                                        constrained, n_constrained, cell_vertices,
                                                        nullptr);
    *op = h.release();
  });
}

int pmg_operator_create_box(const pmg_box *box, int degree, int nq, int precision,
                            pmg_operator *op) {
  return guard([&] {
    PMG_NOT_NULL(op);
    *op = nullptr;
    PMG_NOT_NULL(box);
    auto h = std::make_unique<pmg_operator_s>();
    h->impl = std::make_unique<pmg::LaplaceOperatorBK3>(*box, degree, nq, precision);
    *op = h.release();
  });
}

int pmg_operator_create_from_G(int dim, int degree, int nq, int precision, int64_t n_cells,
                               const uint32_t *dofmap, int64_t n_dofs, const uint32_t *constrained,
                               int64_t n_constrained, const double *G, pmg_operator *op) {
  return guard([&] {
    PMG_NOT_NULL(op);
    *op = nullptr;
    PMG_REQUIRE(dim == 3, PMG_ERR_UNSUPPORTED, "only dim = 3 is supported");
    PMG_NOT_NULL(G);
    auto h = std::make_unique<pmg_operator_s>();
    h->impl = std::make_unique<pmg::LaplaceOperatorBK3>(degree, nq, precision, n_cells, dofmap, n_dofs,
                                                        constrained, n_constrained, nullptr, G);
    *op = h.release();
  });
}

int pmg_dg_operator_create(int dim, int degree, int nq, int precision, int64_t n_cells,
                           const double *cell_vertices, int64_t n_inner_faces,
                           const uint32_t *inner_faces, int64_t n_boundary_faces,
                           const uint32_t *boundary_faces, const double *penalty, int penalty_kind,
                           int metric, pmg_operator *op) {
  return guard([&] {
    PMG_NOT_NULL(op);
    *op = nullptr;
    PMG_REQUIRE(dim == 3, PMG_ERR_UNSUPPORTED, "only dim = 3 is supported");
    auto h = std::make_unique<pmg_operator_s>();
    h->impl = std::make_unique<pmg::LaplaceOperatorDG>(degree, nq, precision, n_cells, cell_vertices,
                                                       n_inner_faces, inner_faces, n_boundary_faces,
                                                       boundary_faces, penalty, penalty_kind, metric);
    *op = h.release();
  });
}

int pmg_dg_operator_create_box(const pmg_box *box, int degree, int nq, int precision,
                               int penalty_kind, int metric, int neumann_sides, pmg_operator *op) {
  return guard([&] {
    PMG_NOT_NULL(op);
    *op = nullptr;
    PMG_NOT_NULL(box);
    auto h = std::make_unique<pmg_operator_s>();
    h->impl = std::make_unique<pmg::LaplaceOperatorDG>(*box, degree, nq, precision, penalty_kind,
                                                       metric, neumann_sides);
    *op = h.release();
  });
}

int pmg_dg_operator_penalty(pmg_operator op, double *penalty) {
  return guard([&] {
    PMG_NOT_NULL(op);
    PMG_NOT_NULL(penalty);
    auto *dg = dynamic_cast<pmg::LaplaceOperatorDG *>(op->impl.get());
    PMG_REQUIRE(dg != nullptr, PMG_ERR_UNSUPPORTED, "not a DG operator");
    std::copy(dg->penalty().begin(), dg->penalty().end(), penalty);
  });
}

int pmg_operator_destroy(pmg_operator op) {
  return guard([&] { delete op; });
}

int pmg_operator_m(pmg_operator op, int64_t *m) {
  return guard([&] {
    PMG_NOT_NULL(op);
    PMG_NOT_NULL(m);
    *m = op->impl->m();
  });
}

int pmg_operator_info(pmg_operator op, int *degree, int *nq, int *precision, int64_t *n_cells) {
  return guard([&] {
    PMG_NOT_NULL(op);
    if (degree) *degree = op->impl->degree();
    if (nq) *nq = op->impl->n_q_points_1d();
    if (precision) *precision = op->impl->precision();
    if (n_cells) *n_cells = op->impl->n_cells();
  });
}

int pmg_operator_vmult(pmg_operator op, void *dst, const void *src, pmg_stream s) {
  return guard([&] {
    PMG_NOT_NULL(op);
    op->impl->vmult(dst, src, as_stream(s));
  });
}

int pmg_operator_Tvmult(pmg_operator op, void *dst, const void *src, pmg_stream s) {
  return pmg_operator_vmult(op, dst, src, s);
}

int pmg_operator_apply_add(pmg_operator op, void *dst, const void *src, pmg_stream s) {
  return guard([&] {
    PMG_NOT_NULL(op);
    op->impl->apply_add(dst, src, as_stream(s));
  });
}

int pmg_operator_set_subdomain(pmg_operator op, const uint32_t *dofmap,
                               const uint32_t *physical_dofs, int64_t n_physical,
                               const uint32_t *interface_cells, int64_t n_interface_cells) {
  return guard([&] {
    PMG_NOT_NULL(op);
    bk3_of(op).set_subdomain(dofmap, physical_dofs, n_physical, interface_cells, n_interface_cells);
  });
}

int pmg_operator_vmult_neumann(pmg_operator op, void *dst, const void *src, pmg_stream s) {
  return guard([&] {
    PMG_NOT_NULL(op);
    bk3_of(op).vmult_subdomain(dst, src, false, as_stream(s));
  });
}

int pmg_operator_vmult_interface_cells(pmg_operator op, void *dst, const void *src,
                                       pmg_stream s) {
  return guard([&] {
    PMG_NOT_NULL(op);
    bk3_of(op).vmult_subdomain(dst, src, true, as_stream(s));
  });
}

int pmg_operator_compute_diagonal(pmg_operator op) {
  return guard([&] {
    PMG_NOT_NULL(op);
    op->impl->compute_diagonal();
  });
}

int pmg_operator_diagonal_inverse(pmg_operator op, void *out, pmg_stream s) {
  return guard([&] {
    PMG_NOT_NULL(op);
    PMG_NOT_NULL(out);
    const size_t bytes = op->impl->m() * pmg::element_size(op->impl->precision());
    PMG_HIP_CHECK(hipMemcpyAsync(out, op->impl->inverse_diagonal(), bytes, hipMemcpyDeviceToDevice,
                                 as_stream(s)));
  });
}

int pmg_operator_diagonal_entry(pmg_operator op, int64_t row, double *value) {
  return guard([&] {
    PMG_NOT_NULL(op);
    PMG_NOT_NULL(value);
    PMG_REQUIRE(row >= 0 && row < op->impl->m(), PMG_ERR_INVALID_ARG, "row out of range");
    const auto *inv = static_cast<const char *>(op->impl->inverse_diagonal());
    double d = 0.0;
    if (op->impl->precision() == PMG_FP64) {
      PMG_HIP_CHECK(hipMemcpy(&d, inv + row * sizeof(double), sizeof(double),
                              hipMemcpyDeviceToHost));
    } else {
      float f = 0.0f;
      PMG_HIP_CHECK(hipMemcpy(&f, inv + row * sizeof(float), sizeof(float), hipMemcpyDeviceToHost));
      d = f;
    }
    *value = 1.0 / d;
  });
}

int pmg_operator_vmult_dummy(pmg_operator op, void *dst, const void *src, int ghost_exchange_on,
                             int computation_on, pmg_stream s) {
  return guard([&] {
    PMG_NOT_NULL(op);
    PMG_NOT_NULL(dst);
    PMG_NOT_NULL(src);
    PMG_REQUIRE(dst != src, PMG_ERR_INVALID_ARG, "vmult_dummy: dst and src must not alias");
    const pmg::Operator &A = *op->impl;
    if (ghost_exchange_on && computation_on) {
      A.vmult(dst, src, as_stream(s));
      return;
    }
    if (A.precision() == PMG_FP64)
      pmg::launch_fill<double>(A.m(), static_cast<double *>(dst), 0.0, as_stream(s));
    else
      pmg::launch_fill<float>(A.m(), static_cast<float *>(dst), 0.0f, as_stream(s));
    if (computation_on) A.apply_add(dst, src, as_stream(s));
    if (ghost_exchange_on) A.copy_constrained(dst, src, as_stream(s));
  });
}

int pmg_operator_set_coefficient(pmg_operator op, const double *kappa, pmg_stream s) {
  return guard([&] {
    PMG_NOT_NULL(op);
    bk3_of(op).set_coefficient(kappa, as_stream(s));
  });
}

int pmg_operator_use_affine_geometry(pmg_operator op, int enable) {
  return guard([&] {
    PMG_NOT_NULL(op);
    bk3_of(op).use_affine_geometry(enable != 0);
  });
}

int pmg_operator_gradient_coefficient(pmg_operator op, const void *u, const void *lam, void *out,
                                      pmg_stream s) {
  return guard([&] {
    PMG_NOT_NULL(op);
    bk3_of(op).gradient_coefficient(u, lam, out, as_stream(s));
  });
}

int pmg_operator_assemble_rhs(pmg_operator op, int f_kind, int nq, void *b, pmg_stream s) {
  return guard([&] {
    PMG_NOT_NULL(op);
    op->impl->assemble_rhs(f_kind, nq, b, as_stream(s));
  });
}

int pmg_operator_l2_error(pmg_operator op, const void *u, int u_kind, int nq, double *err,
                          pmg_stream s) {
  return guard([&] {
    PMG_NOT_NULL(op);
    PMG_NOT_NULL(err);
    *err = op->impl->l2_error(u, u_kind, nq, as_stream(s));
  });
}

int pmg_operator_set_schedule(pmg_operator op, int schedule, int g_load) {
  return guard([&] {
    PMG_NOT_NULL(op);
    PMG_REQUIRE(g_load >= -1 && g_load <= 2, PMG_ERR_INVALID_ARG, "g_load must be -1, 0, 1 or 2");
    PMG_REQUIRE(schedule >= 0 && schedule <= 4, PMG_ERR_INVALID_ARG,
                "schedule must be 0 (atomic), 1 (colored), 2 (patch), 3 (gather) or 4 (grid)");
    bk3_of(op).set_schedule(schedule, g_load);
  });
}

int pmg_operator_set_box_structure(pmg_operator op, int nx, int ny, int nz) {
  return guard([&] {
    PMG_NOT_NULL(op);
    bk3_of(op).set_box_structure(nx, ny, nz);
  });
}

namespace {
void copy_index_array(const pmg::LaplaceOperatorBK3::IndexArray &a, void *host, int64_t *n) {
  PMG_NOT_NULL(n);
  if (host) {
    PMG_REQUIRE(*n == a.n, PMG_ERR_SIZE,
                "index array has " + std::to_string(a.n) + " elements, not " + std::to_string(*n));
    if (a.n > 0)
      PMG_HIP_CHECK(hipMemcpy(host, a.ptr, static_cast<size_t>(a.n) * a.elem,
                              hipMemcpyDeviceToHost));
  }
  *n = a.n;
}
}  // namespace

int pmg_operator_index_data(pmg_operator op, int which, void *host, int64_t *n) {
  return guard([&] {
    PMG_NOT_NULL(op);
    copy_index_array(bk3_of(op).index_array(which), host, n);
  });
}

int pmg_operator_algorithmic_bytes(pmg_operator op, double *bytes, double *flops) {
  return guard([&] {
    PMG_NOT_NULL(op);
    if (bytes) *bytes = op->impl->algorithmic_bytes();
    if (flops) *flops = op->impl->algorithmic_flops();
  });
}

// ---------------------------------------------------------------- transfers
int pmg_transfer_create_from_constraints(int kind, int precision, int pc, int pf,
                                         int64_t n_cells, const uint32_t *map_coarse,
                                         const uint32_t *map_fine,
                                         const uint32_t *constrained_fine,
                                         int64_t n_constrained_fine, int64_t n_dofs_coarse,
                                         int64_t n_dofs_fine, double scale, pmg_transfer *tr) {
  return guard([&] {
    PMG_NOT_NULL(tr);
    *tr = nullptr;
    PMG_REQUIRE(kind == PMG_TRANSFER_H || kind == PMG_TRANSFER_P, PMG_ERR_UNSUPPORTED,
                "weights from constraints exist for h- and p-transfers");
    PMG_REQUIRE(n_constrained_fine >= 0, PMG_ERR_INVALID_ARG, "negative constrained count");
    auto h = std::make_unique<pmg_transfer_s>();
    h->impl = std::make_unique<pmg::MGTransfer>(kind, precision, pc, pf, n_cells, map_coarse, map_fine,
                                                nullptr, n_dofs_coarse, n_dofs_fine, scale,
                                                constrained_fine, n_constrained_fine);
    *tr = h.release();
  });
}

int pmg_transfer_create(int kind, int precision, int pc, int pf, int64_t n_cells,
                        const uint32_t *map_coarse, const uint32_t *map_fine, const double *weights,
                        int64_t n_dofs_coarse, int64_t n_dofs_fine, double scale, pmg_transfer *tr) {
  return guard([&] {
    PMG_NOT_NULL(tr);
    *tr = nullptr;
    auto h = std::make_unique<pmg_transfer_s>();
    h->impl = std::make_unique<pmg::MGTransfer>(kind, precision, pc, pf, n_cells, map_coarse, map_fine,
                                                weights, n_dofs_coarse, n_dofs_fine, scale);
    *tr = h.release();
  });
}

int pmg_transfer_create_box(int kind, int precision, const pmg_box *box, int pc, int pf,
                            double scale, pmg_transfer *tr) {
  return guard([&] {
    PMG_NOT_NULL(tr);
    *tr = nullptr;
    PMG_NOT_NULL(box);
    auto h = std::make_unique<pmg_transfer_s>();
    h->impl = std::make_unique<pmg::MGTransfer>(kind, precision, *box, pc, pf, scale);
    *tr = h.release();
  });
}

int pmg_transfer_index_data(pmg_transfer tr, int which, void *host, int64_t *n) {
  return guard([&] {
    PMG_NOT_NULL(tr);
    copy_index_array(tr->impl->index_array(which), host, n);
  });
}

int pmg_transfer_destroy(pmg_transfer tr) {
  return guard([&] { delete tr; });
}

int pmg_transfer_info(pmg_transfer tr, int *kind, int *precision, int *degree_coarse,
                      int *degree_fine, int *layout) {
  return guard([&] {
    PMG_NOT_NULL(tr);
    const pmg::MGTransfer &t = *tr->impl;
    if (kind) *kind = t.kind();
    if (precision) *precision = t.precision();
    if (degree_coarse) *degree_coarse = t.degree_coarse();
    if (degree_fine) *degree_fine = t.degree_fine();
    if (layout) *layout = t.layout();
  });
}

int pmg_transfer_reinit(pmg_transfer tr, int64_t n_cells, const uint32_t *map_coarse,
                        const uint32_t *map_fine, const uint32_t *constrained_fine,
                        int64_t n_constrained_fine, int64_t n_dofs_coarse, int64_t n_dofs_fine) {
  return guard([&] {
    PMG_NOT_NULL(tr);
    PMG_REQUIRE(n_cells >= 0, PMG_ERR_INVALID_ARG, "negative cell count");
    tr->impl->reinit(n_cells, map_coarse, map_fine, constrained_fine, n_constrained_fine,
                     n_dofs_coarse, n_dofs_fine);
  });
}

int pmg_transfer_prolongate_and_add(pmg_transfer tr, void *dst, const void *src, pmg_stream s) {
  return guard([&] {
    PMG_NOT_NULL(tr);
    tr->impl->prolongate_and_add(dst, src, as_stream(s));
  });
}

int pmg_transfer_restrict_and_add(pmg_transfer tr, void *dst, const void *src, pmg_stream s) {
  return guard([&] {
    PMG_NOT_NULL(tr);
    tr->impl->restrict_and_add(dst, src, as_stream(s));
  });
}

// ---------------------------------------------------------------- smoother
int pmg_chebyshev_create(pmg_operator op, int degree, double range, int eig_its, pmg_smoother *sm) {
  return guard([&] {
    PMG_NOT_NULL(sm);
    *sm = nullptr;
    PMG_NOT_NULL(op);
    auto h = std::make_unique<pmg_smoother_s>();
    h->impl = std::make_unique<pmg::PreconditionChebyshev>(*op->impl, degree, range, eig_its);
    *sm = h.release();
  });
}

int pmg_chebyshev_destroy(pmg_smoother sm) {
  return guard([&] { delete sm; });
}

int pmg_chebyshev_vmult(pmg_smoother sm, void *dst, const void *src, pmg_stream s) {
  return guard([&] {
    PMG_NOT_NULL(sm);
    sm->impl->vmult(dst, src, as_stream(s));
  });
}

int pmg_chebyshev_step(pmg_smoother sm, void *dst, const void *src, pmg_stream s) {
  return guard([&] {
    PMG_NOT_NULL(sm);
    sm->impl->step(dst, src, as_stream(s));
  });
}

int pmg_chebyshev_info(pmg_smoother sm, double *mx, double *mn, int *degree) {
  return guard([&] {
    PMG_NOT_NULL(sm);
    if (mx) *mx = sm->impl->max_eigenvalue();
    if (mn) *mn = sm->impl->min_eigenvalue();
    if (degree) *degree = sm->impl->degree();
  });
}

// ---------------------------------------------------------------- multigrid
int pmg_multigrid_create(int n_levels, const pmg_operator *ops, const pmg_transfer *trs,
                         const pmg_smoother *sms, int coarse_kind, double coarse_tol,
                         int outer_precision, pmg_multigrid *mg) {
  return guard([&] {
    PMG_NOT_NULL(mg);
    *mg = nullptr;
    PMG_REQUIRE(n_levels >= 1 && ops && trs && sms, PMG_ERR_INVALID_ARG, "bad level arrays");
    std::vector<const pmg::Operator *> o(n_levels);
    std::vector<const pmg::MGTransfer *> t(n_levels, nullptr);
    std::vector<const pmg::PreconditionChebyshev *> m(n_levels, nullptr);
    for (int l = 0; l < n_levels; ++l) {
      PMG_REQUIRE(ops[l], PMG_ERR_INVALID_ARG, "NULL level operator");
      o[l] = ops[l]->impl.get();
      if (l > 0) {
        PMG_REQUIRE(trs[l] && sms[l], PMG_ERR_INVALID_ARG, "NULL transfer or smoother");
        t[l] = trs[l]->impl.get();
        m[l] = sms[l]->impl.get();
      }
    }
    auto h = std::make_unique<pmg_multigrid_s>();
    h->impl = std::make_unique<pmg::VCycleMultigrid>(o, t, m, coarse_kind, coarse_tol, outer_precision);
    *mg = h.release();
  });
}

int pmg_multigrid_destroy(pmg_multigrid mg) {
  return guard([&] { delete mg; });
}

int pmg_multigrid_uses_graph(pmg_multigrid mg, int *uses_graph) {
  return guard([&] {
    PMG_NOT_NULL(mg);
    PMG_NOT_NULL(uses_graph);
    *uses_graph = mg->impl->uses_graph() ? 1 : 0;
  });
}

int pmg_multigrid_level_times(pmg_multigrid mg, const void *src, int n_applications,
                              double *times, pmg_stream s) {
  return guard([&] {
    PMG_NOT_NULL(mg);
    mg->impl->level_times(src, n_applications, times, as_stream(s));
  });
}

int pmg_multigrid_vmult(pmg_multigrid mg, void *dst, const void *src, pmg_stream s) {
  return guard([&] {
    PMG_NOT_NULL(mg);
    mg->impl->vmult(dst, src, as_stream(s));
  });
}

// ---------------------------------------------------------------- CG
int pmg_cg_solve(pmg_operator op, int precond, pmg_multigrid mg, const void *b, void *x,
                 double rel_tol, int max_it, int *iterations, double *res0, double *res,
                 pmg_stream s) {
  int status = PMG_OK;
  const int g = guard([&] {
    PMG_NOT_NULL(op);
    const auto r = pmg::solve_cg(*op->impl, precond, mg ? mg->impl.get() : nullptr, b, x, rel_tol,
                                 max_it, op->cg_work, as_stream(s));
    if (iterations) *iterations = r.iterations;
    if (res0) *res0 = r.res0;
    if (res) *res = r.res;
    if (!r.converged) {
      status = PMG_ERR_NOT_CONVERGED;
      pmg::set_last_error("CG did not reach the tolerance in max_it iterations");
    }
  });
  return g != PMG_OK ? g : status;
}

int pmg_cg_solve_noalloc(pmg_operator op, int precond, pmg_multigrid mg, const void *b, void *x,
                         double rel_tol, int max_it, int *iterations, double *res0, double *res,
                         pmg_stream s) {
  int status = PMG_OK;
  const int g = guard([&] {
    PMG_NOT_NULL(op);
    const auto r = pmg::solve_cg_noalloc(*op->impl, precond, mg ? mg->impl.get() : nullptr, b, x,
                                         rel_tol, max_it, as_stream(s));
    if (iterations) *iterations = r.iterations;
    if (res0) *res0 = r.res0;
    if (res) *res = r.res;
    if (!r.converged) {
      status = PMG_ERR_NOT_CONVERGED;
      pmg::set_last_error("CG did not reach the tolerance in max_it iterations");
    }
  });
  return g != PMG_OK ? g : status;
}

// ---------------------------------------------------------------- vectors
int pmg_vector_dot(int precision, int64_t n, const void *a, const void *b, double *result,
                   pmg_stream s) {
  return guard([&] {
    pmg::check_precision(precision);
    PMG_NOT_NULL(result);
    PMG_REQUIRE(n >= 0 && (n == 0 || (a && b)), PMG_ERR_INVALID_ARG, "bad vectors");
    pmg::DeviceArray<double> partials(pmg::kDotBlocks), out(1);
    if (precision == PMG_FP64)
      pmg::launch_dot<double>(n, static_cast<const double *>(a), static_cast<const double *>(b),
                              partials.ptr, out.ptr, as_stream(s));
    else
      pmg::launch_dot<float>(n, static_cast<const float *>(a), static_cast<const float *>(b),
                             partials.ptr, out.ptr, as_stream(s));
    PMG_HIP_CHECK(hipMemcpyAsync(result, out.ptr, sizeof(double), hipMemcpyDeviceToHost, as_stream(s)));
    PMG_HIP_CHECK(hipStreamSynchronize(as_stream(s)));
  });
}

int pmg_vector_convert(int64_t n, int dst_precision, void *dst, int src_precision, const void *src,
                       pmg_stream s) {
  return guard([&] {
    pmg::check_precision(dst_precision);
    pmg::check_precision(src_precision);
    PMG_REQUIRE(n >= 0 && (n == 0 || (dst && src)), PMG_ERR_INVALID_ARG, "bad vectors");
    hipStream_t st = as_stream(s);
    if (dst_precision == PMG_FP64 && src_precision == PMG_FP64)
      pmg::launch_convert<double, double>(n, static_cast<double *>(dst), static_cast<const double *>(src), st);
    else if (dst_precision == PMG_FP64)
      pmg::launch_convert<double, float>(n, static_cast<double *>(dst), static_cast<const float *>(src), st);
    else if (src_precision == PMG_FP64)
      pmg::launch_convert<float, double>(n, static_cast<float *>(dst), static_cast<const double *>(src), st);
    else
      pmg::launch_convert<float, float>(n, static_cast<float *>(dst), static_cast<const float *>(src), st);
  });
}

int pmg_measure_fma_rate(int precision, double *flops_per_second, pmg_stream s) {
  return guard([&] {
    pmg::check_precision(precision);
    PMG_NOT_NULL(flops_per_second);
    *flops_per_second = pmg::measure_fma_rate(precision, as_stream(s));
  });
}

}  // extern "C"
