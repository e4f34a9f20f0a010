// pmg_mesh.cpp -- structured box meshes: the host-side inputs that deal.II's
// Triangulation/DoFHandler/MatrixFree build for the reference drivers
// (tests/poisson_cube_bk3/program.cc:281-566, source/geometric_multigrid_bk3/program.cc).
//
// Numbering: cells lexicographic (x fastest); Q_p DoFs lexicographic on the
// global (p nx+1)(p ny+1)(p nz+1) grid; cell-local DoFs lexicographic, as the
// reference's dof_indices views (portable_laplace_operator_bk3.h:360-437).
// Deformation 1: x += a sin(2 pi y) sin(2 pi z), y += a sin(2 pi z) sin(2 pi x),
// z += a sin(2 pi x) sin(2 pi y), evaluated at the undeformed vertex positions;
// cells stay trilinear (straight edges between displaced vertices).
// Deformation 2: the affine shear x += a y, y += a z, z += a x (parallelepiped
// cells whose Jacobian is not symmetric).
#include <cmath>
#include <vector>

#include "pmg_common.hpp"

namespace pmg {
void set_last_error(const std::string &msg);
}

namespace {

using pmg::Error;

void check_box(const pmg_box *b) {
  PMG_REQUIRE(b != nullptr, PMG_ERR_INVALID_ARG, "box is NULL");
  PMG_REQUIRE(b->nx > 0 && b->ny > 0 && b->nz > 0, PMG_ERR_INVALID_ARG,
              "box needs at least one cell per direction");
  PMG_REQUIRE(b->deform >= 0 && b->deform <= 2, PMG_ERR_INVALID_ARG, "unknown deformation");
}

void vertex(const pmg_box *b, int a, int c1, int c2, double *X) {
  const double x = b->lo[0] + (b->hi[0] - b->lo[0]) * a / b->nx;
  const double y = b->lo[1] + (b->hi[1] - b->lo[1]) * c1 / b->ny;
  const double z = b->lo[2] + (b->hi[2] - b->lo[2]) * c2 / b->nz;
  X[0] = x;
  X[1] = y;
  X[2] = z;
  if (b->deform == 1) {
    const double w = 2.0 * M_PI, A = b->amplitude;
    X[0] += A * std::sin(w * y) * std::sin(w * z);
    X[1] += A * std::sin(w * z) * std::sin(w * x);
    X[2] += A * std::sin(w * x) * std::sin(w * y);
  } else if (b->deform == 2) {
    const double A = b->amplitude;
    X[0] += A * y;
    X[1] += A * z;
    X[2] += A * x;
  }
}

struct Grid {
  int64_t NX, NY, NZ;
  Grid(const pmg_box *b, int p) : NX(int64_t(p) * b->nx + 1), NY(int64_t(p) * b->ny + 1), NZ(int64_t(p) * b->nz + 1) {}
  bool boundary(int64_t I, int64_t J, int64_t K) const {
    return I == 0 || J == 0 || K == 0 || I == NX - 1 || J == NY - 1 || K == NZ - 1;
  }
  int64_t index(int64_t I, int64_t J, int64_t K) const { return I + NX * (J + NY * K); }
  int64_t size() const { return NX * NY * NZ; }
};

void patch_weights(int64_t n_cells, int64_t n_local, const uint32_t *map_f, const Grid &g,
                   double *w) {
  std::vector<uint32_t> count(g.size(), 0);
  for (int64_t e = 0; e < n_cells * n_local; ++e) ++count[map_f[e]];
  for (int64_t e = 0; e < n_cells * n_local; ++e) {
    const int64_t d = map_f[e];
    const int64_t I = d % g.NX, J = (d / g.NX) % g.NY, K = d / (g.NX * g.NY);
    w[e] = g.boundary(I, J, K) ? 0.0 : 1.0 / count[d];
  }
}

template <typename F>
int guard(F &&f) {
  try {
    f();
    return PMG_OK;
  } catch (const Error &e) {
    pmg::set_last_error(e.what());
    return e.status;
  } catch (const std::exception &e) {
    pmg::set_last_error(e.what());
    return PMG_ERR_INVALID_ARG;
  }
}

}  // namespace

extern "C" {

int64_t pmg_box_n_cells(const pmg_box *b) {
  if (!b) return -1;
  return int64_t(b->nx) * b->ny * b->nz;
}

int64_t pmg_box_n_dofs(const pmg_box *b, int p) {
  if (!b || p < 1) return -1;
  return Grid(b, p).size();
}

int pmg_box_cell_vertices(const pmg_box *b, double *cv) {
  return guard([&] {
    check_box(b);
    PMG_REQUIRE(cv, PMG_ERR_INVALID_ARG, "cell_vertices is NULL");
    int64_t c = 0;
    for (int cz = 0; cz < b->nz; ++cz)
      for (int cy = 0; cy < b->ny; ++cy)
        for (int cx = 0; cx < b->nx; ++cx, ++c)
          for (int v = 0; v < 8; ++v)
            vertex(b, cx + (v & 1), cy + ((v >> 1) & 1), cz + ((v >> 2) & 1), cv + c * 24 + 3 * v);
  });
}

int pmg_box_dofmap(const pmg_box *b, int p, int dirichlet, uint32_t *map) {
  return guard([&] {
    check_box(b);
    PMG_REQUIRE(p >= 1 && map, PMG_ERR_INVALID_ARG, "bad degree or NULL map");
    const Grid g(b, p);
    PMG_REQUIRE(g.size() < int64_t(0xFFFFFFFFu), PMG_ERR_SIZE, "too many DoFs for 32-bit maps");
    const int nm = p + 1;
    int64_t c = 0;
    for (int cz = 0; cz < b->nz; ++cz)
      for (int cy = 0; cy < b->ny; ++cy)
        for (int cx = 0; cx < b->nx; ++cx, ++c) {
          uint32_t *m = map + c * nm * nm * nm;
          for (int k = 0; k < nm; ++k)
            for (int j = 0; j < nm; ++j)
              for (int i = 0; i < nm; ++i) {
                const int64_t I = int64_t(p) * cx + i, J = int64_t(p) * cy + j, K = int64_t(p) * cz + k;
                *m++ = (dirichlet && g.boundary(I, J, K)) ? pmg::kInvalid
                                                          : static_cast<uint32_t>(g.index(I, J, K));
              }
        }
  });
}

int pmg_box_boundary_dofs(const pmg_box *b, int p, uint32_t *dofs, int64_t *n) {
  return guard([&] {
    check_box(b);
    PMG_REQUIRE(p >= 1 && n, PMG_ERR_INVALID_ARG, "bad degree or NULL count");
    const Grid g(b, p);
    int64_t cnt = 0;
    for (int64_t K = 0; K < g.NZ; ++K)
      for (int64_t J = 0; J < g.NY; ++J)
        for (int64_t I = 0; I < g.NX; ++I)
          if (g.boundary(I, J, K)) {
            if (dofs) dofs[cnt] = static_cast<uint32_t>(g.index(I, J, K));
            ++cnt;
          }
    *n = cnt;
  });
}

int pmg_box_h_transfer_maps(const pmg_box *coarse, int p, uint32_t *map_c, uint32_t *map_f,
                            double *w) {
  return guard([&] {
    check_box(coarse);
    PMG_REQUIRE(map_c && map_f, PMG_ERR_INVALID_ARG, "NULL output");
    PMG_REQUIRE(pmg_box_dofmap(coarse, p, 1, map_c) == PMG_OK, PMG_ERR_INVALID_ARG,
                "coarse map failed");
    pmg_box fine = *coarse;
    fine.nx *= 2;
    fine.ny *= 2;
    fine.nz *= 2;
    const Grid g(&fine, p);
    const int nf = 2 * p + 1;
    const int64_t nf3 = int64_t(nf) * nf * nf;
    int64_t c = 0;
    for (int cz = 0; cz < coarse->nz; ++cz)
      for (int cy = 0; cy < coarse->ny; ++cy)
        for (int cx = 0; cx < coarse->nx; ++cx, ++c) {
          uint32_t *m = map_f + c * nf3;
          for (int k = 0; k < nf; ++k)
            for (int j = 0; j < nf; ++j)
              for (int i = 0; i < nf; ++i)
                *m++ = static_cast<uint32_t>(
                    g.index(2 * int64_t(p) * cx + i, 2 * int64_t(p) * cy + j, 2 * int64_t(p) * cz + k));
        }
    if (w) patch_weights(pmg_box_n_cells(coarse), nf3, map_f, g, w);
  });
}

int pmg_box_p_transfer_maps(const pmg_box *b, int pc, int pf, uint32_t *map_c, uint32_t *map_f,
                            double *w) {
  return guard([&] {
    check_box(b);
    PMG_REQUIRE(pc >= 1 && pf > pc, PMG_ERR_INVALID_ARG, "need 1 <= degree_coarse < degree_fine");
    PMG_REQUIRE(map_c && map_f, PMG_ERR_INVALID_ARG, "NULL output");
    PMG_REQUIRE(pmg_box_dofmap(b, pc, 1, map_c) == PMG_OK, PMG_ERR_INVALID_ARG, "coarse map failed");
    PMG_REQUIRE(pmg_box_dofmap(b, pf, 0, map_f) == PMG_OK, PMG_ERR_INVALID_ARG, "fine map failed");
    const int64_t nf3 = int64_t(pf + 1) * (pf + 1) * (pf + 1);
    if (w) patch_weights(pmg_box_n_cells(b), nf3, map_f, Grid(b, pf), w);
  });
}

int pmg_box_dg_faces(const pmg_box *b, uint32_t *inner, int64_t *n_inner, uint32_t *bdry,
                     int64_t *n_bdry) {
  return guard([&] {
    check_box(b);
    PMG_REQUIRE(n_inner && n_bdry, PMG_ERR_INVALID_ARG, "NULL count");
    const int64_t n[3] = {b->nx, b->ny, b->nz};
    PMG_REQUIRE(n[0] * n[1] * n[2] < int64_t(0xFFFFFFFFu), PMG_ERR_SIZE, "too many cells");
    auto cell = [&](int64_t i, int64_t j, int64_t k) { return uint32_t(i + n[0] * (j + n[1] * k)); };
    int64_t fi = 0, fb = 0;
    for (int d = 0; d < 3; ++d)
      for (int64_t k = 0; k < n[2]; ++k)
        for (int64_t j = 0; j < n[1]; ++j)
          for (int64_t i = 0; i < n[0]; ++i) {
            const int64_t c[3] = {i, j, k};
            if (c[d] == 0) {  // lower boundary face of the first cell layer
              if (bdry) {
                bdry[3 * fb] = cell(i, j, k);
                bdry[3 * fb + 1] = uint32_t(2 * d);
                bdry[3 * fb + 2] = 0;
              }
              ++fb;
            }
            if (c[d] + 1 < n[d]) {
              if (inner) {
                int64_t e[3] = {i, j, k};
                ++e[d];
                inner[4 * fi] = cell(i, j, k);
                inner[4 * fi + 1] = cell(e[0], e[1], e[2]);
                inner[4 * fi + 2] = uint32_t(2 * d + 1);
                inner[4 * fi + 3] = uint32_t(2 * d);
              }
              ++fi;
            } else {
              if (bdry) {
                bdry[3 * fb] = cell(i, j, k);
                bdry[3 * fb + 1] = uint32_t(2 * d + 1);
                bdry[3 * fb + 2] = 0;
              }
              ++fb;
            }
          }
    *n_inner = fi;
    *n_bdry = fb;
  });
}

}  // extern "C"
