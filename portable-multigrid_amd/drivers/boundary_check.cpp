// boundary_check.cpp -- exercises, through the header-only C++ mirror
// include/pmg.hpp, the members of the reference's operator and transfer
// interfaces that a deal.II-side adapter forwards to the C ABI beyond
// vmult/prolongate/restrict:
//   LaplaceOperatorBase::el and ::vmult_dummy
//     (include/base/portable_laplace_operator_base.h:27-31,54),
//   MGTransferBase::reinit (include/base/portable_mg_transfer_base.h:32-35).
// Prints one line "key=value ..." of maximum deviations for the caller:
//   el            |el(i,i) - 1/(inverse diagonal)_i|, relative, over all rows
//   dummy_11..00  vmult_dummy(ghost, compute) against vmult / apply_add / the
//                 constrained-row copy composed by hand
//   reinit_p, reinit_r  a transfer reinit'ed to another box against a fresh
//                 transfer of that box (prolongation, restriction)
//   affine        vmult with use_affine_geometry on an undeformed box against
//                 the same operator with G at every q-point
//   affine_refused 1 if the deformed operator refused the affine geometry
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
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
  double *p = nullptr;
  size_t n = 0;
  explicit DeviceBuffer(size_t count) : n(count) {
    hip_check(hipMalloc(reinterpret_cast<void **>(&p), std::max<size_t>(count, 1) * sizeof(double)));
  }
  ~DeviceBuffer() { (void)hipFree(p); }
  DeviceBuffer(const DeviceBuffer &) = delete;
  DeviceBuffer &operator=(const DeviceBuffer &) = delete;
  void set(const std::vector<double> &v) {
    hip_check(hipMemcpy(p, v.data(), v.size() * sizeof(double), hipMemcpyHostToDevice));
  }
  std::vector<double> get() const {
    std::vector<double> v(n);
    hip_check(hipDeviceSynchronize());
    hip_check(hipMemcpy(v.data(), p, n * sizeof(double), hipMemcpyDeviceToHost));
    return v;
  }
};

pmg_box make_box(int n, int deform) {
  pmg_box b{};
  b.nx = b.ny = b.nz = n;
  b.hi[0] = b.hi[1] = b.hi[2] = 1.0;
  b.deform = deform;
  b.amplitude = 0.05;
  return b;
}

double max_abs_diff(const std::vector<double> &a, const std::vector<double> &b) {
  double m = 0;
  for (size_t i = 0; i < a.size(); ++i) m = std::max(m, std::fabs(a[i] - b[i]));
  return m;
}

double max_abs(const std::vector<double> &a) {
  double m = 0;
  for (double v : a) m = std::max(m, std::fabs(v));
  return m;
}

std::vector<double> pseudo_random(size_t n, unsigned seed) {
  std::vector<double> v(n);
  unsigned s = seed;
  for (auto &x : v) {
    s = s * 1664525u + 1013904223u;
    x = 2.0 * (s >> 8) / double(1u << 24) - 1.0;
  }
  return v;
}

}  // namespace

int main() {
  using namespace pmg_hip;
  try {
    // ---- operator: el and vmult_dummy
    const pmg_box b = make_box(4, 1);
    const int p = 2;
    LaplaceOperatorBK3 op(b, p, p + 1, PMG_FP64);
    op.compute_diagonal();
    const int64_t N = op.m();
    DeviceBuffer inv(N), x(N), y(N), z(N);
    op.get_matrix_diagonal_inverse(inv.p);
    const std::vector<double> d = inv.get();
    double el_err = 0;
    for (int64_t i = 0; i < N; ++i)
      el_err = std::max(el_err, std::fabs(op.el(i, i) * d[i] - 1.0));

    const std::vector<double> xs = pseudo_random(N, 7);
    x.set(xs);
    std::vector<uint32_t> cons(static_cast<size_t>(N));
    int64_t nc = N;
    check(pmg_box_boundary_dofs(&b, p, cons.data(), &nc));
    cons.resize(static_cast<size_t>(nc));
    op.vmult(y.p, x.p);
    const std::vector<double> Ax = y.get();
    std::vector<double> cell = Ax, copy(static_cast<size_t>(N), 0.0);
    for (uint32_t c : cons) {
      cell[c] = 0.0;   // the cell loop leaves constrained rows untouched
      copy[c] = xs[c]; // copy_constrained_values
    }
    const double scale = max_abs(Ax);
    double dummy[2][2];
    for (int g = 0; g < 2; ++g)
      for (int c = 0; c < 2; ++c) {
        z.set(pseudo_random(N, 11));  // stale contents: vmult_dummy zeroes dst first
        op.vmult_dummy(z.p, x.p, g == 1, c == 1);
        const std::vector<double> got = z.get();
        const std::vector<double> &exp =
            g && c ? Ax : (c ? cell : (g ? copy : std::vector<double>(static_cast<size_t>(N), 0.0)));
        dummy[g][c] = max_abs_diff(got, exp) / scale;
      }

    // ---- transfer reinit: a box(2) transfer rebuilt for box(3)
    const int pt = 2;
    const pmg_box c2 = make_box(2, 1), c3 = make_box(3, 1);
    GeometricTransfer tr(c2, pt, PMG_FP64), fresh(c3, pt, PMG_FP64);
    pmg_box f3 = c3;
    f3.nx *= 2;
    f3.ny *= 2;
    f3.nz *= 2;
    const int64_t C3 = pmg_box_n_cells(&c3);
    const int64_t nmc3 = int64_t(pt + 1) * (pt + 1) * (pt + 1);
    const int64_t nmf3 = int64_t(2 * pt + 1) * (2 * pt + 1) * (2 * pt + 1);
    std::vector<uint32_t> mc(static_cast<size_t>(C3 * nmc3)), mf(static_cast<size_t>(C3 * nmf3));
    check(pmg_box_h_transfer_maps(&c3, pt, mc.data(), mf.data(), nullptr));
    int64_t ncf = 0;
    check(pmg_box_boundary_dofs(&f3, pt, nullptr, &ncf));
    std::vector<uint32_t> consf(static_cast<size_t>(ncf));
    check(pmg_box_boundary_dofs(&f3, pt, consf.data(), &ncf));
    const int64_t Nc = pmg_box_n_dofs(&c3, pt), Nf = pmg_box_n_dofs(&f3, pt);
    tr.reinit(C3, mc.data(), mf.data(), consf.data(), ncf, Nc, Nf);
    DeviceBuffer xc(Nc), xf(Nf), y1(Nf), y2(Nf), r1(Nc), r2(Nc);
    xc.set(pseudo_random(Nc, 21));
    xf.set(pseudo_random(Nf, 22));
    const std::vector<double> zf(static_cast<size_t>(Nf), 0.0), zc(static_cast<size_t>(Nc), 0.0);
    y1.set(zf);
    y2.set(zf);
    r1.set(zc);
    r2.set(zc);
    tr.prolongate_and_add(y1.p, xc.p);
    fresh.prolongate_and_add(y2.p, xc.p);
    tr.restrict_and_add(r1.p, xf.p);
    fresh.restrict_and_add(r2.p, xf.p);
    const std::vector<double> a = y1.get(), bb = y2.get(), ra = r1.get(), rb = r2.get();
    const double reinit_p = max_abs_diff(a, bb) / max_abs(bb);
    const double reinit_r = max_abs_diff(ra, rb) / max_abs(rb);

    // ---- affine geometry (a library option with no reference counterpart)
    int refused = 0;
    try {
      op.use_affine_geometry();
    } catch (const Error &e) {
      refused = e.status == PMG_ERR_INVALID_ARG;
    }
    const pmg_box ba = make_box(4, 0);
    LaplaceOperatorBK3 opa(ba, 4, 5, PMG_FP64);
    const int64_t Na = opa.m();
    DeviceBuffer xa(Na), ya(Na), yb(Na);
    xa.set(pseudo_random(Na, 31));
    opa.vmult(ya.p, xa.p);
    opa.use_affine_geometry();
    opa.vmult(yb.p, xa.p);
    const std::vector<double> ga = ya.get(), gb = yb.get();
    const double affine = max_abs_diff(gb, ga) / max_abs(ga);

    std::printf("el=%.3e dummy_11=%.3e dummy_01=%.3e dummy_10=%.3e dummy_00=%.3e "
                "reinit_p=%.3e reinit_r=%.3e affine=%.3e affine_refused=%d\n",
                el_err, dummy[1][1], dummy[0][1], dummy[1][0], dummy[0][0], reinit_p, reinit_r,
                affine, refused);
    return 0;
  } catch (const Error &e) {
    std::fprintf(stderr, "pmg error %d: %s\n", e.status, e.what());
    return 4;
  }
}
