// pmg_operator_dg.cpp -- host side of the SIP-DG operator: face connectivity,
// penalty parameters and face geometry set-up (the reference builds the same
// from deal.II MatrixFree face data, portable_laplace_operator_dg.h:559-854),
// and the stream-ordered apply.
#include <algorithm>
#include <cmath>
#include <string>

#include "pmg_library.hpp"

namespace pmg {

namespace {
template <typename T>
Shape1D<T> make_shape(const std::vector<double> &S, const std::vector<double> &D) {
  Shape1D<T> sh{};
  for (size_t i = 0; i < S.size(); ++i) sh.S[i] = static_cast<T>(S[i]);
  for (size_t i = 0; i < D.size(); ++i) sh.D[i] = static_cast<T>(D[i]);
  return sh;
}

void trilinear_jacobian(const double *v, const double xi[3], double J[3][3]) {
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b) J[a][b] = 0;
  for (int vi = 0; vi < 8; ++vi) {
    const int bb[3] = {vi & 1, (vi >> 1) & 1, (vi >> 2) & 1};
    double f[3], df[3];
    for (int e = 0; e < 3; ++e) {
      f[e] = bb[e] ? xi[e] : 1.0 - xi[e];
      df[e] = bb[e] ? 1.0 : -1.0;
    }
    const double dN[3] = {df[0] * f[1] * f[2], f[0] * df[1] * f[2], f[0] * f[1] * df[2]};
    for (int d = 0; d < 3; ++d)
      for (int k = 0; k < 3; ++k) J[d][k] += dN[k] * v[3 * vi + d];
  }
}

// |K| and |F_f| of a trilinear cell by 4-point Gauss rules (the cell measure
// is exact; the face measure integrates |J e_t1 x J e_t2|)
void cell_measures(const double *v, double &vol, double area[6]) {
  const Quadrature1D g = gauss_legendre(4);
  vol = 0;
  for (int f = 0; f < 6; ++f) area[f] = 0;
  for (int a = 0; a < 4; ++a)
    for (int b = 0; b < 4; ++b)
      for (int c = 0; c < 4; ++c) {
        const double xi[3] = {g.points[a], g.points[b], g.points[c]};
        double J[3][3];
        trilinear_jacobian(v, xi, J);
        const double det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) -
                           J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
                           J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
        vol += std::fabs(det) * g.weights[a] * g.weights[b] * g.weights[c];
      }
  for (int f = 0; f < 6; ++f) {
    const int d = f / 2, t1 = (d + 1) % 3, t2 = (d + 2) % 3;
    for (int a = 0; a < 4; ++a)
      for (int b = 0; b < 4; ++b) {
        double xi[3], J[3][3];
        xi[d] = f % 2;
        xi[t1] = g.points[a];
        xi[t2] = g.points[b];
        trilinear_jacobian(v, xi, J);
        const double n0 = J[1][t1] * J[2][t2] - J[2][t1] * J[1][t2];
        const double n1 = J[2][t1] * J[0][t2] - J[0][t1] * J[2][t2];
        const double n2 = J[0][t1] * J[1][t2] - J[1][t1] * J[0][t2];
        area[f] += std::sqrt(n0 * n0 + n1 * n1 + n2 * n2) * g.weights[a] * g.weights[b];
      }
  }
}

template <typename T>
void upload_as(DeviceVector &dv, int precision, const std::vector<double> &h) {
  std::vector<T> t(h.begin(), h.end());
  dv.reinit(precision, static_cast<int64_t>(h.size()));
  if (!t.empty())
    PMG_HIP_CHECK(hipMemcpy(dv.data(), t.data(), t.size() * sizeof(T), hipMemcpyHostToDevice));
}
}  // namespace

LaplaceOperatorDG::LaplaceOperatorDG(int degree, int nq, int precision, int64_t n_cells,
                                     const double *cell_vertices, int64_t n_inner,
                                     const uint32_t *inner, int64_t n_boundary,
                                     const uint32_t *boundary, const double *penalty,
                                     int penalty_kind, int metric)
    : degree_(degree), nq_(nq), precision_(precision), n_cells_(n_cells) {
  check_precision(precision);
  PMG_REQUIRE(degree >= 1 && degree <= 6, PMG_ERR_UNSUPPORTED, "degree must be in [1,6]");
  PMG_REQUIRE(nq == degree + 1 || nq == degree + 2, PMG_ERR_UNSUPPORTED,
              "n_q_points_1d must be degree+1 or degree+2");
  PMG_REQUIRE(n_cells >= 0 && n_inner >= 0 && n_boundary >= 0, PMG_ERR_INVALID_ARG,
              "negative size");
  PMG_REQUIRE(cell_vertices || n_cells == 0, PMG_ERR_INVALID_ARG, "cell_vertices is NULL");
  PMG_REQUIRE(inner || n_inner == 0, PMG_ERR_INVALID_ARG, "inner face list is NULL");
  PMG_REQUIRE(boundary || n_boundary == 0, PMG_ERR_INVALID_ARG, "boundary face list is NULL");
  PMG_REQUIRE(metric == 0 || metric == 1, PMG_ERR_INVALID_ARG, "metric must be 0 or 1");
  PMG_REQUIRE(penalty || penalty_kind == 0 || penalty_kind == 1, PMG_ERR_INVALID_ARG,
              "penalty_kind must be 0 or 1");
  nm3_ = int64_t(degree + 1) * (degree + 1) * (degree + 1);
  PMG_REQUIRE(n_cells * nm3_ < int64_t(0xFFFFFFFFu), PMG_ERR_SIZE,
              "DG DoF count exceeds 32-bit indexing");
  n_faces_ = n_inner + n_boundary;
  PMG_REQUIRE(n_faces_ < (int64_t(1) << 30), PMG_ERR_SIZE, "too many faces");

  // cell -> face connectivity; every cell face must be covered exactly once
  std::vector<uint32_t> cface(static_cast<size_t>(6 * n_cells), kInvalid);
  std::vector<uint32_t> cnbr(static_cast<size_t>(6 * n_cells), 0u);
  std::vector<uint32_t> finfo(static_cast<size_t>(3 * n_faces_));
  auto claim = [&](int64_t c, uint32_t f, uint32_t code, uint32_t nb) {
    PMG_REQUIRE(c >= 0 && c < n_cells && f < 6, PMG_ERR_INVALID_ARG,
                "face entry names a cell or face number out of range");
    PMG_REQUIRE(cface[6 * c + f] == kInvalid, PMG_ERR_INVALID_ARG,
                "cell " + std::to_string(c) + " face " + std::to_string(f) + " listed twice");
    cface[6 * c + f] = code;
    cnbr[6 * c + f] = nb;
  };
  for (int64_t F = 0; F < n_inner; ++F) {
    const uint32_t cm = inner[4 * F], cp = inner[4 * F + 1], fm = inner[4 * F + 2],
                   fp = inner[4 * F + 3];
    PMG_REQUIRE(cm != cp, PMG_ERR_INVALID_ARG, "inner face joins a cell with itself");
    PMG_REQUIRE(fm < 6 && fp == (fm ^ 1u), PMG_ERR_UNSUPPORTED,
                "inner faces need face+ = face- ^ 1 (standard orientation)");
    claim(cm, fm, static_cast<uint32_t>(F), cp);
    claim(cp, fp, static_cast<uint32_t>(F) | (1u << 30), cm);
    finfo[3 * F] = cm;
    finfo[3 * F + 1] = cp;
    finfo[3 * F + 2] = fm;
  }
  for (int64_t B = 0; B < n_boundary; ++B) {
    const int64_t F = n_inner + B;
    const uint32_t c = boundary[3 * B], f = boundary[3 * B + 1], id = boundary[3 * B + 2];
    claim(c, f, static_cast<uint32_t>(F) | ((id == 0 ? 2u : 3u) << 30), 0u);
    finfo[3 * F] = c;
    finfo[3 * F + 1] = kInvalid;
    finfo[3 * F + 2] = f;
  }
  for (int64_t i = 0; i < 6 * n_cells; ++i)
    PMG_REQUIRE(cface[i] != kInvalid, PMG_ERR_INVALID_ARG,
                "cell " + std::to_string(i / 6) + " face " + std::to_string(i % 6) +
                    " has no face entry");

  // penalty parameters (get_penalty_factor, portable_laplace_operator_dg.h:147-151)
  sigma_.resize(static_cast<size_t>(n_faces_));
  if (penalty) {
    std::copy(penalty, penalty + n_faces_, sigma_.begin());
  } else {
    std::vector<double> h(static_cast<size_t>(6 * n_cells));
    for (int64_t c = 0; c < n_cells; ++c) {
      double vol, area[6];
      cell_measures(cell_vertices + 24 * c, vol, area);
      for (int f = 0; f < 6; ++f) h[6 * c + f] = vol / area[f];
    }
    const double fac = penalty_kind == 0 ? 0.5 * degree * (degree + 1.0)
                                         : (degree + 1.0) * (degree + 1.0);
    for (int64_t F = 0; F < n_faces_; ++F) {
      const uint32_t cm = finfo[3 * F], cp = finfo[3 * F + 1], fm = finfo[3 * F + 2];
      const double hm = h[6 * cm + fm], hp = cp == kInvalid ? hm : h[6 * cp + (fm ^ 1u)];
      sigma_[F] = fac * (1.0 / hm + 1.0 / hp);
    }
  }

  setup_1d();
  cface_.upload(cface);
  cnbr_.upload(cnbr);
  vertices_.upload(cell_vertices, static_cast<size_t>(24 * n_cells));
  DeviceArray<uint32_t> finfo_d;
  finfo_d.upload(finfo);
  setup_geometry(finfo_d.ptr, metric);
}

// A box on the device: vertices, connectivity (the numbering of
// pmg_box_dg_faces), the penalty from the cell and face measures, G and the
// face data, without host loops or host arrays.
LaplaceOperatorDG::LaplaceOperatorDG(const pmg_box &box, int degree, int nq, int precision,
                                     int penalty_kind, int metric, int neumann_sides)
    : degree_(degree), nq_(nq), precision_(precision),
      n_cells_(int64_t(box.nx) * box.ny * box.nz) {
  check_precision(precision);
  PMG_REQUIRE(degree >= 1 && degree <= 6, PMG_ERR_UNSUPPORTED, "degree must be in [1,6]");
  PMG_REQUIRE(nq == degree + 1 || nq == degree + 2, PMG_ERR_UNSUPPORTED,
              "n_q_points_1d must be degree+1 or degree+2");
  PMG_REQUIRE(box.nx > 0 && box.ny > 0 && box.nz > 0, PMG_ERR_INVALID_ARG,
              "box needs at least one cell per direction");
  PMG_REQUIRE(box.deform >= 0 && box.deform <= 2, PMG_ERR_INVALID_ARG, "unknown deformation");
  PMG_REQUIRE(metric == 0 || metric == 1, PMG_ERR_INVALID_ARG, "metric must be 0 or 1");
  PMG_REQUIRE(penalty_kind == 0 || penalty_kind == 1, PMG_ERR_INVALID_ARG,
              "penalty_kind must be 0 or 1");
  nm3_ = int64_t(degree + 1) * (degree + 1) * (degree + 1);
  PMG_REQUIRE(n_cells_ * nm3_ < int64_t(0xFFFFFFFFu), PMG_ERR_SIZE,
              "DG DoF count exceeds 32-bit indexing");
  const int64_t nx = box.nx, ny = box.ny, nz = box.nz;
  const int64_t n_inner = (nx - 1) * ny * nz + nx * (ny - 1) * nz + nx * ny * (nz - 1);
  n_faces_ = n_inner + 2 * (ny * nz + nx * nz + nx * ny);
  PMG_REQUIRE(n_faces_ < (int64_t(1) << 30), PMG_ERR_SIZE, "too many faces");
  vertices_.resize(static_cast<size_t>(24 * n_cells_));
  box_vertices_device(box, vertices_.ptr, nullptr);
  cface_.resize(static_cast<size_t>(6 * n_cells_));
  cnbr_.resize(static_cast<size_t>(6 * n_cells_));
  DeviceArray<uint32_t> finfo_d(static_cast<size_t>(3 * n_faces_));
  DeviceArray<double> h(static_cast<size_t>(6 * n_cells_)), sig(static_cast<size_t>(n_faces_));
  const Quadrature1D g4 = gauss_legendre(4);
  DeviceArray<double> g4p, g4w;
  g4p.upload(g4.points);
  g4w.upload(g4.weights);
  const double fac = penalty_kind == 0 ? 0.5 * degree * (degree + 1.0)
                                       : (degree + 1.0) * (degree + 1.0);
  launch_dg_box_setup(box, neumann_sides, vertices_.ptr, g4p.ptr, g4w.ptr, fac, cface_.ptr,
                      cnbr_.ptr, finfo_d.ptr, h.ptr, sig.ptr, n_faces_, nullptr);
  sigma_ = sig.download();
  setup_1d();
  setup_geometry(finfo_d.ptr, metric);
}

// 1D data: shape matrices, their even-odd tables, and the face tables
void LaplaceOperatorDG::setup_1d() {
  const int degree = degree_, nq = nq_;
  shape_ = make_shape_data(degree, nq);
  sh64_ = make_shape<double>(shape_.S, shape_.D);
  sh32_ = make_shape<float>(shape_.S, shape_.D);
  fill_even_odd(sh64_, shape_.S.data(), shape_.D.data(), degree + 1, shape_.n_q_points_1d);
  fill_even_odd(sh32_, shape_.S.data(), shape_.D.data(), degree + 1, shape_.n_q_points_1d);
  shd64_ = make_shape<double>(shape_.S, shape_.Sd);
  const int nm = degree + 1;
  const std::vector<double> gll = gauss_lobatto_points(nm);
  std::vector<double> val(nm), d0(nm), d1(nm);
  lagrange_eval(gll, 0.0, val.data(), d0.data());
  lagrange_eval(gll, 1.0, val.data(), d1.data());
  for (size_t i = 0; i < shape_.Sd.size(); ++i) {
    fs64_.Sd[i] = shape_.Sd[i];
    fs32_.Sd[i] = static_cast<float>(shape_.Sd[i]);
  }
  for (int i = 0; i < nm; ++i) {
    fs64_.dl[i] = d0[i];
    fs64_.dl[nm + i] = d1[i];
    fs32_.dl[i] = static_cast<float>(d0[i]);
    fs32_.dl[nm + i] = static_cast<float>(d1[i]);
  }
  // Lagrange basis on the Gauss points and its derivative at the cell ends
  {
    std::vector<double> lv(nq), ld(nq);
    for (int end = 0; end < 2; ++end) {
      lagrange_eval(shape_.quad.points, double(end), lv.data(), ld.data());
      for (int r = 0; r < nq; ++r) {
        fs64_.L[end * nq + r] = lv[r];
        fs64_.Ld[end * nq + r] = ld[r];
        fs32_.L[end * nq + r] = static_cast<float>(lv[r]);
        fs32_.Ld[end * nq + r] = static_cast<float>(ld[r]);
      }
    }
  }

}

// G and the face data (finfo: device [F][3]), the penalty in the operator
// precision
void LaplaceOperatorDG::setup_geometry(const uint32_t *finfo_d, int metric) {
  const int nq = nq_, precision = precision_;
  const int64_t n_cells = n_cells_;
  DeviceArray<double> qp, qw;
  qp.upload(shape_.quad.points);
  qw.upload(shape_.quad.weights);
  G_.reinit(precision, g_size(nq, n_cells));
  fdata_.reinit(precision, n_faces_ * 7 * nq * nq);
  if (precision == PMG_FP64) {
    launch_geometry<double>(nq, n_cells, vertices_.ptr, qp.ptr, qw.ptr, nullptr, nullptr,
                            static_cast<double *>(G_.data()), nullptr);
    launch_dg_face_geometry<double>(nq, n_faces_, finfo_d, vertices_.ptr, qp.ptr, qw.ptr,
                                    metric, static_cast<double *>(fdata_.data()), nullptr);
    upload_as<double>(fsig_, precision, sigma_);
  } else {
    launch_geometry<float>(nq, n_cells, vertices_.ptr, qp.ptr, qw.ptr, nullptr, nullptr,
                           static_cast<float *>(G_.data()), nullptr);
    launch_dg_face_geometry<float>(nq, n_faces_, finfo_d, vertices_.ptr, qp.ptr, qw.ptr,
                                   metric, static_cast<float *>(fdata_.data()), nullptr);
    upload_as<float>(fsig_, precision, sigma_);
  }
  PMG_HIP_CHECK(hipDeviceSynchronize());  // qp, qw and the caller's finfo are freed after
}

template <typename T>
void LaplaceOperatorDG::apply_t(bool add, T *dst, const T *src, hipStream_t s) const {
  const bool f64 = std::is_same<T, double>::value;
  const Shape1D<T> &sh = *reinterpret_cast<const Shape1D<T> *>(
      f64 ? static_cast<const void *>(&sh64_) : static_cast<const void *>(&sh32_));
  const DgShape<T> &fs = *reinterpret_cast<const DgShape<T> *>(
      f64 ? static_cast<const void *>(&fs64_) : static_cast<const void *>(&fs32_));
  launch_dg_apply<T>(degree_, nq_, add, src, dst, static_cast<const T *>(G_.data()), cface_.ptr,
                     cnbr_.ptr, static_cast<const T *>(fdata_.data()),
                     static_cast<const T *>(fsig_.data()), n_cells_, sh, fs, s);
}

void LaplaceOperatorDG::vmult(void *dst, const void *src, hipStream_t s) const {
  PMG_REQUIRE(dst && src, PMG_ERR_INVALID_ARG, "NULL vector");
  PMG_REQUIRE(dst != src, PMG_ERR_INVALID_ARG, "vmult: dst and src must not alias");
  if (precision_ == PMG_FP64) apply_t(false, static_cast<double *>(dst), static_cast<const double *>(src), s);
  else apply_t(false, static_cast<float *>(dst), static_cast<const float *>(src), s);
}

void LaplaceOperatorDG::apply_add(void *dst, const void *src, hipStream_t s) const {
  PMG_REQUIRE(dst && src && dst != src, PMG_ERR_INVALID_ARG, "bad vectors");
  if (precision_ == PMG_FP64) apply_t(true, static_cast<double *>(dst), static_cast<const double *>(src), s);
  else apply_t(true, static_cast<float *>(dst), static_cast<const float *>(src), s);
}

const uint32_t *LaplaceOperatorDG::identity_map() const {
  if (ident_.ptr == nullptr && n_cells_ > 0) {
    std::vector<uint32_t> id(static_cast<size_t>(n_cells_ * nm3_));
    for (size_t i = 0; i < id.size(); ++i) id[i] = static_cast<uint32_t>(i);
    ident_.upload(id);
  }
  return ident_.ptr;
}

void LaplaceOperatorDG::compute_diagonal() {
  const int64_t n = m();
  inv_diag_.reinit(precision_, n);
  const uint32_t *map = identity_map();
  // assembled and inverted in FP64 at both level precisions, stored in the
  // level precision
  auto run = [&](auto zero) {
    using T = decltype(zero);
    const bool f64 = std::is_same<T, double>::value;
    DeviceArray<double> tmp;
    if (!f64) tmp.resize(static_cast<size_t>(n));
    double *d = f64 ? static_cast<double *>(inv_diag_.data()) : tmp.ptr;
    launch_fill<double>(n, d, 0.0, nullptr);
    launch_bk3_diagonal<T>(degree_, nq_, n_cells_, static_cast<const T *>(G_.data()), map, shd64_,
                           d, nullptr);
    const int64_t bad = first_nonpositive<double>(n, d, nullptr);
    PMG_REQUIRE(bad < 0, PMG_ERR_INVALID_ARG,
                "non-positive diagonal entry at DoF " + std::to_string(bad));
    launch_invert<double>(n, d, nullptr);
    if (!f64) launch_convert<float, double>(n, static_cast<float *>(inv_diag_.data()), d, nullptr);
    PMG_HIP_CHECK(hipDeviceSynchronize());  // before tmp is released
  };
  if (precision_ == PMG_FP64) run(0.0);
  else run(0.0f);
  PMG_HIP_CHECK(hipDeviceSynchronize());
  has_diag_ = true;
}

const void *LaplaceOperatorDG::inverse_diagonal() const {
  PMG_REQUIRE(has_diag_, PMG_ERR_INVALID_ARG, "compute_diagonal() has not been called");
  return inv_diag_.data();
}

void LaplaceOperatorDG::assemble_rhs(int f_kind, int nq, void *b, hipStream_t s) const {
  assemble_rhs_cells(precision_, degree_, n_cells_, m(), vertices_.ptr, identity_map(), f_kind,
                     nq, b, s);
}

double LaplaceOperatorDG::l2_error(const void *u, int u_kind, int nq, hipStream_t s) const {
  return l2_error_cells(precision_, degree_, n_cells_, vertices_.ptr, identity_map(), u, u_kind,
                        nq, s);
}

// HBM bytes one apply must move: source and result vectors, cell G, face data
// (metrics of both sides and JxW per face q-point, the penalty) and the
// cell->face connectivity (two uint32 per cell face).
double LaplaceOperatorDG::algorithmic_bytes() const {
  const double s = element_size(precision_);
  const double nq2 = double(nq_) * nq_, nq3 = nq2 * nq_;
  return s * (2.0 * m() + 6.0 * n_cells_ * nq3 + n_faces_ * (7.0 * nq2 + 1.0)) +
         48.0 * n_cells_;
}

// cell part as the continuous operator plus, per cell face: nodal traces
// (4 nm^2 + 2 nm^3... counted as 4 nm^2 (2 + nm)), interpolation to the face
// q-points (8 nm^2 nq + 8 nm nq^2), the flux (~30 nq^2), and the test
// (8 nq^2 nm + 6 nq nm^2 + 2 nm^3)
double LaplaceOperatorDG::algorithmic_flops() const {
  const double nm = degree_ + 1, nq = nq_;
  const double cell = 2.0 * (2.0 * (nm * nm * nm * nq + nm * nm * nq * nq + nm * nq * nq * nq) +
                             6.0 * nq * nq * nq * nq + 9.0 * nq * nq * nq);
  const double face = 4.0 * nm * nm * (2.0 + nm) + 2.0 * (8.0 * nm * nm * nq + 8.0 * nm * nq * nq) +
                      30.0 * nq * nq + 2.0 * (8.0 * nq * nq * nm + 6.0 * nq * nm * nm) +
                      2.0 * nm * nm * nm;
  return n_cells_ * (cell + 6.0 * face);
}

}  // namespace pmg
