// poisson_mgcg.cpp -- C++ driver of the library through the header-only
// mirror include/pmg.hpp, shaped like the reference's LaplaceProblem drivers
// (tests/poisson_cube_bk3/program.cc, tests/poisson_cube_bk3_fp64_dg/program.cc):
// set up the level operators, transfers and Chebyshev smoothers of an h (and
// p) hierarchy on the unit cube, assemble the right-hand side of the
// manufactured solution sin(pi x) sin(pi y) sin(pi z), and solve with
// MG-preconditioned CG.  Prints one line "key=value ..." for the caller.
//
// usage: poisson_mgcg <cells per dir> <degree> <h levels> [cg|dg] [fp64|mixed]
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "pmg.hpp"

namespace {

void hip_check(hipError_t e) {
  if (e != hipSuccess) {
    std::fprintf(stderr, "HIP error: %s\n", hipGetErrorString(e));
    std::exit(2);
  }
}

struct DeviceBuffer {
  void *p = nullptr;
  explicit DeviceBuffer(size_t bytes) { hip_check(hipMalloc(&p, bytes)); }
  ~DeviceBuffer() { (void)hipFree(p); }
  DeviceBuffer(const DeviceBuffer &) = delete;
  DeviceBuffer &operator=(const DeviceBuffer &) = delete;
};

pmg_box make_box(int n) {
  pmg_box b{};
  b.nx = b.ny = b.nz = n;
  b.hi[0] = b.hi[1] = b.hi[2] = 1.0;
  return b;
}

}  // namespace

int main(int argc, char **argv) {
  if (argc < 4) {
    std::fprintf(stderr, "usage: %s <cells/dir> <degree> <h levels> [cg|dg] [fp64|mixed]\n", argv[0]);
    return 1;
  }
  const int n = std::atoi(argv[1]), p = std::atoi(argv[2]), n_h = std::atoi(argv[3]);
  const bool dg = argc > 4 && std::string(argv[4]) == "dg";
  const bool mixed = argc > 5 && std::string(argv[5]) == "mixed";
  const int lp = mixed ? PMG_FP32 : PMG_FP64;
  using namespace pmg_hip;
  try {
    // level list: Q1 h-levels up to the fine box, then p-levels p -> max(p-2,1)
    // reversed, then DG(p) (tests/poisson_cube_bk3_fp64_dg/program.cc:294-300)
    struct Level {
      int cells, degree;
      bool dg;
    };
    std::vector<Level> levels;
    for (int l = n_h - 1; l >= 0; --l) levels.push_back({n >> l, 1, false});
    std::vector<int> ps{p};
    while (ps.back() > 1) ps.push_back(std::max(ps.back() - 2, 1));
    for (int i = static_cast<int>(ps.size()) - 2; i >= 0; --i) levels.push_back({n, ps[i], false});
    if (dg) levels.push_back({n, p, true});

    std::vector<std::unique_ptr<LaplaceOperatorBase>> ops;
    std::vector<std::unique_ptr<MGTransferBase>> trs;
    std::vector<std::unique_ptr<PreconditionChebyshev>> sms;
    trs.emplace_back();
    sms.emplace_back();
    for (size_t l = 0; l < levels.size(); ++l) {
      const Level &L = levels[l];
      const pmg_box b = make_box(L.cells);
      if (L.dg) ops.emplace_back(new LaplaceOperatorDG(b, L.degree, L.degree + 1, lp));
      // the coarse-grid solve stays FP64 under FP32 levels
      else ops.emplace_back(new LaplaceOperatorBK3(b, L.degree, L.degree + 1, l == 0 ? PMG_FP64 : lp));
      ops.back()->compute_diagonal();
      if (l == 0) continue;
      const Level &C = levels[l - 1];
      if (L.dg) trs.emplace_back(new ContinuousTransfer(b, L.degree, lp));
      else if (C.cells == L.cells) trs.emplace_back(new PolynomialTransfer(b, C.degree, L.degree, lp));
      else trs.emplace_back(new GeometricTransfer(make_box(C.cells), L.degree, lp));
      sms.emplace_back(new PreconditionChebyshev(*ops.back(), 3, 15.0, 10));
    }
    std::vector<const LaplaceOperatorBase *> o;
    std::vector<const MGTransferBase *> t;
    std::vector<const PreconditionChebyshev *> s;
    for (size_t l = 0; l < levels.size(); ++l) {
      o.push_back(ops[l].get());
      t.push_back(trs[l].get());
      s.push_back(sms[l].get());
    }
    VCycleMultigrid mg(o, t, s, PMG_COARSE_PCG, 1e-12, PMG_FP64);

    // the outer operator is FP64 (the fine level itself in the all-FP64 run)
    std::unique_ptr<LaplaceOperatorBase> fine64;
    const LaplaceOperatorBase *fine = ops.back().get();
    if (mixed) {
      const pmg_box b = make_box(n);
      if (dg) fine64.reset(new LaplaceOperatorDG(b, p, p + 1, PMG_FP64));
      else fine64.reset(new LaplaceOperatorBK3(b, p, p + 1, PMG_FP64));
      fine = fine64.get();
    }
    const int64_t N = fine->m();
    DeviceBuffer b(N * sizeof(double)), x(N * sizeof(double));
    fine->assemble_rhs(1, b.p);
    hip_check(hipMemset(x.p, 0, N * sizeof(double)));
    hipEvent_t e0, e1;
    hip_check(hipEventCreate(&e0));
    hip_check(hipEventCreate(&e1));
    hip_check(hipEventRecord(e0, nullptr));
    SolverControl c;
    const bool ok = solve_cg(*fine, b.p, x.p, PMG_PRECOND_MG, &mg, 1e-10, 200, &c);
    hip_check(hipEventRecord(e1, nullptr));
    hip_check(hipEventSynchronize(e1));
    float ms = 0;
    hip_check(hipEventElapsedTime(&ms, e0, e1));
    const double err = fine->l2_error(x.p, 1);
    std::printf("dofs=%lld levels=%zu converged=%d iterations=%d rel_residual=%.3e l2_error=%.6e "
                "time_ms=%.3f\n",
                static_cast<long long>(N), levels.size(), ok ? 1 : 0, c.iterations,
                c.final_residual / c.initial_residual, err, ms);
    return ok ? 0 : 3;
  } catch (const Error &e) {
    std::fprintf(stderr, "pmg error %d: %s\n", e.status, e.what());
    return 4;
  }
}
