// pmg_dg.hip -- symmetric interior penalty DG Laplacian kernels for gfx950.
//
// Operator (reference: LaplaceOperatorDG::vmult, portable_laplace_operator_dg.h:252-351,
// kernels BK3::DG::compute_cell / compute_inner_faces / compute_boundary_faces /
// distribute_face_to_global, laplace_dg_kokkos_kernels.h:23-2027):
//   (A u)_K = A_K u_K + sum_{faces F of K} int_F (sigma [u] - {d_n u}) phi - 1/2 [u] d_n phi
// The reference runs four kernels (cell part, inner faces, boundary faces,
// face-to-cell distribution) that exchange data through face buffers in HBM.
// Here ONE kernel does everything, cell-centric: each cell computes its cell
// integral and, for each of its six faces, the flux terms tested with its own
// basis functions, reading the neighbour's nodal values (from L2/MALL) to get
// u+ and d_n u+ on the face.  Every DG DoF belongs to one cell, so the result
// is written with plain coalesced stores: no atomics, no face buffers, no
// second pass.  Each inner face's flux is evaluated from both sides, which
// costs ALU time, not HBM bytes.  Per face, on the cell's Gauss-point grid:
//  - own value and normal derivative at the face point: 1D contraction of the
//    Gauss-point values along the normal with the Gauss-Lagrange basis L_r(s),
//    L_r'(s) (exact: u is a polynomial of degree < nq along every line);
//  - neighbour value and normal derivative: its nodal line along the normal,
//    then the 2D interpolation to the face points;
//  - tangential derivatives: collocation derivative on the face Gauss grid;
//  - test: tangential-gradient terms become value terms through the
//    transposed collocation derivative, and value and normal-derivative terms
//    are spread back along the normal with L_r(s), L_r'(s) into a value
//    coefficient V at the Gauss points, which the cell's final
//    interpolation sweeps integrate together with the cell flux.
//
// Layout: cell-major DoFs (cell c owns x[c*NM^3 .. (c+1)*NM^3), lexicographic
// inside); G batch-interleaved as for the continuous operator; per face F
//   fdata[F][7][NQ^2] = metric m- (3 comps), metric m+ (3 comps), JxW, with the
//   metric components in face-local order (normal axis, then the two tangential
//   axes in increasing order) and the face q-point index a + NQ b along those
//   tangential axes; fsig[F] the penalty.  m- = J-^{-1} n-, m+ = J+^{-1} n-
//   (or the J^{-T} n variant), n- the outward normal of the "-" cell.
// Per cell and face f (0: x=0, 1: x=1, 2: y=0, 3: y=1, 4: z=0, 5: z=1):
//   cface[c][f] = F | role << 30, role 0: "-" side of inner face F, 1: "+"
//   side, 2: boundary face with weak homogeneous Dirichlet (u+ = -u-,
//   d_n u+ = d_n u-), 3: boundary face with homogeneous Neumann (u+ = u-,
//   d_n u+ = -d_n u-) (laplace_dg_kokkos_kernels.h:1563-1576);
//   cnbr[c][f] = neighbour cell across f (inner faces).
#include "pmg_bk3_cell.hpp"
#include "pmg_common.hpp"
#include "pmg_kernels.hpp"

namespace pmg {

// same cell batching as the continuous operator, so G loads stay coalesced
template <int NQ>
using DgTile = Bk3Tile<NQ>;

template <int NM, int NQ, typename T, bool ADD>
__global__ void __launch_bounds__(256)
dg_apply_kernel(const T *__restrict__ x, T *__restrict__ y, const T *__restrict__ G,
                const uint32_t *__restrict__ cface, const uint32_t *__restrict__ cnbr,
                const T *__restrict__ fdata, const T *__restrict__ fsig, const int64_t n_cells,
                const Shape1D<T> sh, const DgShape<T> fs)
{
  constexpr int NQ2 = NQ * NQ, NQ3 = NQ2 * NQ, NM2 = NM * NM, NM3 = NM2 * NM;
  constexpr int E = DgTile<NQ>::cells;
  constexpr int CELL_LDS = lds_cell_stride<T>(4 * NQ3 + 4 * NQ2);
  __shared__ T lds[E * CELL_LDS];

  const unsigned lb = xcd_logical_block(blockIdx.x, gridDim.x);
  const int tid = threadIdx.x;
  const int e = tid / NQ2, t = tid - e * NQ2, ta = t % NQ, tb = t / NQ;
  const int64_t cell = static_cast<int64_t>(lb) * E + e;
  const bool active = (e < E) && (cell < n_cells);
  const bool col = active && ta < NM && tb < NM;
  T *P0 = lds + (active ? e : 0) * CELL_LDS;
  T *P1 = P0 + NQ3, *P2 = P1 + NQ3;
  T *V = P2 + NQ3;      // value coefficient at the Gauss points [r][q][p]
  T *A = V + NQ3;       // face scratch, 2 NQ^2
  T *B = A + 2 * NQ2;   // face scratch, 2 NQ^2
  const T *Gc = G + g_index<T>(NQ, active ? cell : 0, 0, t);

  // ---- own column (i,j) = (ta,tb) over k; V = 0
  T u[NM], out[NM];
  if (col) {
    const T *xc = x + cell * NM3 + tb * NM + ta;
#pragma unroll
    for (int k = 0; k < NM; ++k) u[k] = xc[k * NM2];
  }
  if (active)
    for (int i = t; i < NQ3; i += NQ2) V[i] = T(0);

  // face connectivity of all six faces, and the neighbour's nodal line for
  // the first face, are requested before the cell sweeps so their latency
  // overlaps them; inside the face loop the next face's neighbour line is
  // requested while the current face is computed
  uint32_t cfs[6], nbs[6];
#pragma unroll
  for (int f = 0; f < 6; ++f) {
    cfs[f] = active ? cface[cell * 6 + f] : 0u;
    nbs[f] = col ? cnbr[cell * 6 + f] : 0u;
  }
  // neighbour nodal line for face f (normal d, side s): entries k along the
  // normal at this lane's tangential position; zero where there is no neighbour
  auto load_line = [&](int f, T (&w)[NM]) {
    const int d = f >> 1;
    const int snn = d == 0 ? 1 : (d == 1 ? NM : NM2);
    const int t1n = d == 0 ? NM : 1, t2n = d == 2 ? NM : NM2;
    const bool has = col && (cfs[f] >> 30) < 2u;
    const T *xn = x + (has ? static_cast<int64_t>(nbs[f]) * NM3 + ta * t1n + tb * t2n : 0);
#pragma unroll
    for (int k = 0; k < NM; ++k) w[k] = has ? xn[k * snn] : T(0);
  };
  T wcur[NM], wnext[NM];
  load_line(0, wcur);
  bk3_forward<NM, NQ, T>(u, sh, P0, P1, P2, active, ta, tb);  // values at Gauss points in P0

#pragma unroll 1
  for (int f = 0; f < 6; ++f) {
    const int d = f >> 1, s = f & 1;
    // strides of the tangential axes (ascending) on the Gauss-point grid
    // [r][q][p] and of the normal
    const int snq = d == 0 ? 1 : (d == 1 ? NQ : NQ2);
    const int t1q = d == 0 ? NQ : 1, t2q = d == 2 ? NQ : NQ2;
    const uint32_t cf = cfs[f];
    const int role = static_cast<int>(cf >> 30);
    const int64_t F = cf & 0x3FFFFFFFu;
    if (f < 5) load_line(f + 1, wnext);

    // face geometry of this face point, requested now and used after two
    // barriers, so its latency overlaps steps A and B
    T fq[7], sig = 0;
    if (active) {
      const T *fd = fdata + F * 7 * NQ2 + t;
#pragma unroll
      for (int c = 0; c < 7; ++c) fq[c] = fd[c * NQ2];
      sig = fsig[F];
    }

    // A: neighbour trace and normal derivative on its nodal face grid, lane
    // (alpha,beta); own value and normal derivative at face point (a,b) by a
    // 1D contraction of the Gauss-point values along the normal
    T um = 0, dnu = 0;
    if (col) {
      T dp = 0;
#pragma unroll
      for (int k = 0; k < NM; ++k) dp += fs.dl[(1 - s) * NM + k] * wcur[k];
      A[tb * NM + ta] = s ? wcur[0] : wcur[NM - 1];
      A[NQ2 + tb * NM + ta] = dp;
    }
    if (active) {
      const T *Ul = P0 + ta * t1q + tb * t2q;
#pragma unroll
      for (int r = 0; r < NQ; ++r) {
        const T w = Ul[r * snq];
        um += fs.L[s * NQ + r] * w;
        dnu += fs.Ld[s * NQ + r] * w;
      }
      B[t] = um;
    }
    cell_sync<NQ>();  // face scratch and V are per cell

    // B: neighbour value and normal derivative at face point (a,b)
    T vpq = 0, dpq = 0;
    if (active) {
#pragma unroll
      for (int be = 0; be < NM; ++be) {
        T a0 = 0, a1 = 0;
#pragma unroll
        for (int al = 0; al < NM; ++al) {
          const T sa = sh.S[al * NQ + ta];
          a0 += sa * A[be * NM + al];
          a1 += sa * A[NQ2 + be * NM + al];
        }
        const T sb = sh.S[be * NQ + tb];
        vpq += sb * a0;
        dpq += sb * a1;
      }
      B[NQ2 + t] = vpq;
    }
    cell_sync<NQ>();  // face scratch and V are per cell

    // C: tangential derivatives by collocation on the face Gauss grid, flux
    T cval = 0, cn = 0;
    if (active) {
      T d1m = 0, d2m = 0, d1p = 0, d2p = 0;
#pragma unroll
      for (int n = 0; n < NQ; ++n) {
        const T da = sh.D[n * NQ + ta], db = sh.D[n * NQ + tb];
        d1m += da * B[tb * NQ + n];
        d2m += db * B[n * NQ + ta];
        d1p += da * B[NQ2 + tb * NQ + n];
        d2p += db * B[NQ2 + n * NQ + ta];
      }
      const T jxw = fq[6];
      T mo0, mo1, mo2, dnm, dnp, up = vpq;
      if (role == 1) {
        mo0 = -fq[3];
        mo1 = -fq[4];
        mo2 = -fq[5];
        dnm = mo0 * dnu + mo1 * d1m + mo2 * d2m;
        dnp = -(fq[0] * dpq + fq[1] * d1p + fq[2] * d2p);
      } else {
        mo0 = fq[0];
        mo1 = fq[1];
        mo2 = fq[2];
        dnm = mo0 * dnu + mo1 * d1m + mo2 * d2m;
        if (role == 0) {
          dnp = fq[3] * dpq + fq[4] * d1p + fq[5] * d2p;
        } else if (role == 2) {
          up = -um;
          dnp = dnm;
        } else {
          up = um;
          dnp = -dnm;
        }
      }
      const T jump = um - up, avg = T(0.5) * (dnm + dnp);
      const T td = T(-0.5) * jump * jxw;
      cval = (sig * jump - avg) * jxw;
      cn = td * mo0;
      A[t] = td * mo1;        // tangential test coefficients
      A[NQ2 + t] = td * mo2;
    }
    cell_sync<NQ>();  // face scratch and V are per cell

    // D: tangential test terms become value terms by the transposed
    // collocation derivative; inject value and normal-derivative tests into V
    if (active) {
      T c = cval;
#pragma unroll
      for (int n = 0; n < NQ; ++n)
        c += sh.D[ta * NQ + n] * A[tb * NQ + n] + sh.D[tb * NQ + n] * A[NQ2 + n * NQ + ta];
      T *Vl = V + ta * t1q + tb * t2q;
#pragma unroll
      for (int r = 0; r < NQ; ++r) Vl[r * snq] += c * fs.L[s * NQ + r] + cn * fs.Ld[s * NQ + r];
    }
    cell_sync<NQ>();  // face scratch and V are per cell
#pragma unroll
    for (int k = 0; k < NM; ++k) wcur[k] = wnext[k];
  }

  T g[6][NQ];
  bk3_backward<NM, NQ, T, kGLateNT, true>(out, g, Gc, sh, P0, P1, P2, V, active, ta, tb);
  if (col) {
    T *yc = y + cell * NM3 + tb * NM + ta;
#pragma unroll
    for (int k = 0; k < NM; ++k) {
      if (ADD) yc[k * NM2] += out[k];
      else yc[k * NM2] = out[k];
    }
  }
}

template <int NM, int NQ, typename T>
static void dg_launch(bool add, const T *x, T *y, const T *G, const uint32_t *cface,
                      const uint32_t *cnbr, const T *fdata, const T *fsig, int64_t n_cells,
                      const Shape1D<T> &sh, const DgShape<T> &fs, hipStream_t s) {
  if (n_cells <= 0) return;
  using Tile = DgTile<NQ>;
  const int64_t nb = (n_cells + Tile::cells - 1) / Tile::cells;
  PMG_REQUIRE(nb < (int64_t(1) << 31), PMG_ERR_SIZE, "too many cells for one launch");
  const unsigned grid = pad_to_xcd(static_cast<unsigned>(nb));
  if (add)
    hipLaunchKernelGGL((dg_apply_kernel<NM, NQ, T, true>), dim3(grid), dim3(Tile::lanes), 0, s, x,
                       y, G, cface, cnbr, fdata, fsig, n_cells, sh, fs);
  else
    hipLaunchKernelGGL((dg_apply_kernel<NM, NQ, T, false>), dim3(grid), dim3(Tile::lanes), 0, s,
                       x, y, G, cface, cnbr, fdata, fsig, n_cells, sh, fs);
  PMG_HIP_CHECK(hipGetLastError());
}

template <typename T>
void launch_dg_apply(int degree, int nq, bool add, const T *x, T *y, const T *G,
                     const uint32_t *cface, const uint32_t *cnbr, const T *fdata, const T *fsig,
                     int64_t n_cells, const Shape1D<T> &sh, const DgShape<T> &fs, hipStream_t s) {
#define PMG_DG_CASE(P, Q)                                                                    \
  if (degree == P && nq == Q) {                                                              \
    dg_launch<P + 1, Q, T>(add, x, y, G, cface, cnbr, fdata, fsig, n_cells, sh, fs, s);      \
    return;                                                                                  \
  }
  PMG_DG_CASE(1, 2) PMG_DG_CASE(1, 3) PMG_DG_CASE(2, 3) PMG_DG_CASE(2, 4)
  PMG_DG_CASE(3, 4) PMG_DG_CASE(3, 5) PMG_DG_CASE(4, 5) PMG_DG_CASE(4, 6)
  PMG_DG_CASE(5, 6) PMG_DG_CASE(5, 7) PMG_DG_CASE(6, 7) PMG_DG_CASE(6, 8)
#undef PMG_DG_CASE
  throw Error(PMG_ERR_UNSUPPORTED, "DG kernel: unsupported (degree, n_q_points_1d) = (" +
                                       std::to_string(degree) + ", " + std::to_string(nq) + ")");
}
template void launch_dg_apply<double>(int, int, bool, const double *, double *, const double *,
                                      const uint32_t *, const uint32_t *, const double *,
                                      const double *, int64_t, const Shape1D<double> &,
                                      const DgShape<double> &, hipStream_t);
template void launch_dg_apply<float>(int, int, bool, const float *, float *, const float *,
                                     const uint32_t *, const uint32_t *, const float *,
                                     const float *, int64_t, const Shape1D<float> &,
                                     const DgShape<float> &, hipStream_t);

// ---------------------------------------------------------------- face geometry
// One lane per (face, q): trilinear Jacobians of the two cells at the shared
// face point, outward unit normal n- of the "-" cell, area element, and the
// metric vectors m = J^{-1} n- (metric 0) or J^{-T} n- (metric 1) of both
// sides, permuted to face-local order (normal axis, tangential axes).  Setup
// only (the reference computes the same with FEFaceValues on the host,
// portable_laplace_operator_dg.h:746-846).
__device__ inline void trilinear_jacobian(const double *v, const double xi[3], double J[3][3]) {
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b) J[a][b] = 0;
  for (int vi = 0; vi < 8; ++vi) {
    const int b0 = vi & 1, b1 = (vi >> 1) & 1, b2 = (vi >> 2) & 1;
    const double f0 = b0 ? xi[0] : 1.0 - xi[0], f1 = b1 ? xi[1] : 1.0 - xi[1],
                 f2 = b2 ? xi[2] : 1.0 - xi[2];
    const double d0 = b0 ? 1.0 : -1.0, d1 = b1 ? 1.0 : -1.0, d2 = b2 ? 1.0 : -1.0;
    const double dN[3] = {d0 * f1 * f2, f0 * d1 * f2, f0 * f1 * d2};
    for (int d = 0; d < 3; ++d)
      for (int k = 0; k < 3; ++k) J[d][k] += dN[k] * v[3 * vi + d];
  }
}

// K = J^{-1}
__device__ inline void inverse3(const double J[3][3], double K[3][3]) {
  const double c00 = J[1][1] * J[2][2] - J[1][2] * J[2][1];
  const double c01 = J[1][2] * J[2][0] - J[1][0] * J[2][2];
  const double c02 = J[1][0] * J[2][1] - J[1][1] * J[2][0];
  const double c10 = J[0][2] * J[2][1] - J[0][1] * J[2][2];
  const double c11 = J[0][0] * J[2][2] - J[0][2] * J[2][0];
  const double c12 = J[0][1] * J[2][0] - J[0][0] * J[2][1];
  const double c20 = J[0][1] * J[1][2] - J[0][2] * J[1][1];
  const double c21 = J[0][2] * J[1][0] - J[0][0] * J[1][2];
  const double c22 = J[0][0] * J[1][1] - J[0][1] * J[1][0];
  const double id = 1.0 / (J[0][0] * c00 + J[0][1] * c01 + J[0][2] * c02);
  K[0][0] = c00 * id; K[0][1] = c10 * id; K[0][2] = c20 * id;
  K[1][0] = c01 * id; K[1][1] = c11 * id; K[1][2] = c21 * id;
  K[2][0] = c02 * id; K[2][1] = c12 * id; K[2][2] = c22 * id;
}

template <typename T>
__global__ void __launch_bounds__(256)
dg_face_geometry_kernel(const int nq, const int64_t n_faces, const uint32_t *__restrict__ finfo,
                        const double *__restrict__ cv, const double *__restrict__ qp,
                        const double *__restrict__ qw, const int metric, T *__restrict__ fdata)
{
  const int nq2 = nq * nq;
  const int64_t gid = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (gid >= n_faces * nq2) return;
  const int64_t F = gid / nq2;
  const int q = static_cast<int>(gid - F * nq2);
  const int a = q % nq, b = q / nq;
  // finfo[F] = (cell-, cell+ or 0xFFFFFFFF, face of the "-" cell)
  const uint32_t cm = finfo[3 * F], cp = finfo[3 * F + 1];
  const int f = static_cast<int>(finfo[3 * F + 2]);
  const int d = f >> 1, s = f & 1;
  const int t1 = d == 0 ? 1 : 0, t2 = d == 2 ? 1 : 2;
  double xi[3];
  xi[d] = s;
  xi[t1] = qp[a];
  xi[t2] = qp[b];
  double J[3][3], K[3][3];
  trilinear_jacobian(cv + 24 * static_cast<int64_t>(cm), xi, J);
  inverse3(J, K);
  // outward normal: orientation of (J e_c1) x (J e_c2), (c1,c2) cyclic after d
  const int c1 = (d + 1) % 3, c2 = (d + 2) % 3;
  double N[3] = {J[1][c1] * J[2][c2] - J[2][c1] * J[1][c2],
                 J[2][c1] * J[0][c2] - J[0][c1] * J[2][c2],
                 J[0][c1] * J[1][c2] - J[1][c1] * J[0][c2]};
  const double area = sqrt(N[0] * N[0] + N[1] * N[1] + N[2] * N[2]);
  const double sg = (s ? 1.0 : -1.0) / area;
  const double n[3] = {N[0] * sg, N[1] * sg, N[2] * sg};
  const int ax[3] = {d, t1, t2};
  T *out = fdata + F * 7 * nq2 + q;
  for (int c = 0; c < 3; ++c) {
    const int r = ax[c];
    const double m = metric == 0 ? K[r][0] * n[0] + K[r][1] * n[1] + K[r][2] * n[2]
                                 : K[0][r] * n[0] + K[1][r] * n[1] + K[2][r] * n[2];
    out[c * nq2] = static_cast<T>(m);
  }
  if (cp != 0xFFFFFFFFu) {
    xi[d] = 1 - s;
    trilinear_jacobian(cv + 24 * static_cast<int64_t>(cp), xi, J);
    inverse3(J, K);
    for (int c = 0; c < 3; ++c) {
      const int r = ax[c];
      const double m = metric == 0 ? K[r][0] * n[0] + K[r][1] * n[1] + K[r][2] * n[2]
                                   : K[0][r] * n[0] + K[1][r] * n[1] + K[2][r] * n[2];
      out[(3 + c) * nq2] = static_cast<T>(m);
    }
  } else {
    for (int c = 0; c < 3; ++c) out[(3 + c) * nq2] = T(0);
  }
  out[6 * nq2] = static_cast<T>(area * qw[a] * qw[b]);
}

template <typename T>
void launch_dg_face_geometry(int nq, int64_t n_faces, const uint32_t *finfo, const double *cv,
                             const double *qp, const double *qw, int metric, T *fdata,
                             hipStream_t s) {
  const int64_t n = n_faces * nq * nq;
  if (n == 0) return;
  hipLaunchKernelGGL(dg_face_geometry_kernel<T>, dim3(static_cast<unsigned>((n + 255) / 256)),
                     dim3(256), 0, s, nq, n_faces, finfo, cv, qp, qw, metric, fdata);
  PMG_HIP_CHECK(hipGetLastError());
}
template void launch_dg_face_geometry<double>(int, int64_t, const uint32_t *, const double *,
                                              const double *, const double *, int, double *,
                                              hipStream_t);
template void launch_dg_face_geometry<float>(int, int64_t, const uint32_t *, const double *,
                                             const double *, const double *, int, float *,
                                             hipStream_t);

// ---------------------------------------------------------------- box set-up
// SIP-DG topology of an nx x ny x nz box on the device, in the numbering of
// pmg_box_dg_faces: inner faces direction by direction (x, y, z), each the
// face (L, L + e_d) of the cells L with L_d < n_d - 1 in lexicographic order;
// then the boundary faces direction by direction in lexicographic cell
// order, a cell's lower face before its upper one.  One lane per cell:
// cface/cnbr of its six faces, and finfo of the faces it is the "-" side of
// (inner faces: the lower cell; boundary faces: the cell).  Boundary id of
// box side f: bit f of neumann_sides (0 Dirichlet, 1 Neumann).
__global__ void __launch_bounds__(256)
dg_box_topology_kernel(const int nx, const int ny, const int nz, const int neumann_sides,
                       uint32_t *__restrict__ cface, uint32_t *__restrict__ cnbr,
                       uint32_t *__restrict__ finfo)
{
  const int64_t nc = int64_t(nx) * ny * nz;
  const int64_t c = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (c >= nc) return;
  const int n[3] = {nx, ny, nz};
  const int64_t i = c % nx, j = (c / nx) % ny, k = c / (int64_t(nx) * ny);
  const int64_t ijk[3] = {i, j, k};
  const int64_t stride[3] = {1, nx, int64_t(nx) * ny};
  const int64_t off_in[3] = {0, int64_t(nx - 1) * ny * nz,
                             int64_t(nx - 1) * ny * nz + int64_t(nx) * (ny - 1) * nz};
  const int64_t n_inner = off_in[2] + int64_t(nx) * ny * (nz - 1);
  const int64_t off_bd[3] = {0, 2 * int64_t(ny) * nz, 2 * int64_t(ny) * nz + 2 * int64_t(nx) * nz};
  // rank of cell (a, b, e) among the cells with x_d < n_d - 1, lexicographic
  auto inner_index = [&](int d, int64_t a, int64_t b, int64_t e) -> int64_t {
    if (d == 0) return off_in[0] + a + int64_t(nx - 1) * (b + int64_t(ny) * e);
    if (d == 1) return off_in[1] + a + int64_t(nx) * (b + int64_t(ny - 1) * e);
    return off_in[2] + a + int64_t(nx) * (b + int64_t(ny) * e);
  };
  auto bdry_index = [&](int d, int s) -> int64_t {
    if (d == 0) return off_bd[0] + 2 * (j + int64_t(ny) * k) + s;
    if (d == 1)
      return off_bd[1] + 2 * int64_t(nx) * k + (ny > 1 ? (s ? nx + i : i) : 2 * i + s);
    return off_bd[2] + (nz > 1 ? (s ? int64_t(nx) * ny + i + int64_t(nx) * j : i + int64_t(nx) * j)
                               : 2 * (i + int64_t(nx) * j) + s);
  };
  for (int f = 0; f < 6; ++f) {
    const int d = f >> 1, s = f & 1;
    const bool nb = s ? ijk[d] + 1 < n[d] : ijk[d] > 0;
    if (nb) {
      const int64_t other = c + (s ? stride[d] : -stride[d]);
      // the lower cell of the pair names the face
      int64_t low[3] = {i, j, k};
      if (!s) --low[d];
      const int64_t F = inner_index(d, low[0], low[1], low[2]);
      cface[6 * c + f] = static_cast<uint32_t>(F) | (s ? 0u : (1u << 30));
      cnbr[6 * c + f] = static_cast<uint32_t>(other);
      if (s) {
        finfo[3 * F] = static_cast<uint32_t>(c);
        finfo[3 * F + 1] = static_cast<uint32_t>(other);
        finfo[3 * F + 2] = static_cast<uint32_t>(f);
      }
    } else {
      const int64_t F = n_inner + bdry_index(d, s);
      const uint32_t role = ((neumann_sides >> f) & 1) ? 3u : 2u;
      cface[6 * c + f] = static_cast<uint32_t>(F) | (role << 30);
      cnbr[6 * c + f] = 0u;
      finfo[3 * F] = static_cast<uint32_t>(c);
      finfo[3 * F + 1] = kInvalid;
      finfo[3 * F + 2] = static_cast<uint32_t>(f);
    }
  }
}

// h[c][f] = |K| / |F_f| of every trilinear cell: |K| and the face areas by
// the 4-point Gauss rule per direction (g4p, g4w) of the host set-up
__global__ void __launch_bounds__(256)
dg_cell_h_kernel(const int64_t n_cells, const double *__restrict__ cv,
                 const double *__restrict__ g4p, const double *__restrict__ g4w,
                 double *__restrict__ h)
{
  const int64_t c = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (c >= n_cells) return;
  const double *v = cv + 24 * c;
  double vol = 0;
  for (int a = 0; a < 4; ++a)
    for (int b = 0; b < 4; ++b)
      for (int e = 0; e < 4; ++e) {
        const double xi[3] = {g4p[a], g4p[b], g4p[e]};
        double J[3][3];
        trilinear_jacobian(v, xi, J);
        const double det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) -
                           J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
                           J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
        vol += fabs(det) * g4w[a] * g4w[b] * g4w[e];
      }
  for (int f = 0; f < 6; ++f) {
    const int d = f / 2, t1 = (d + 1) % 3, t2 = (d + 2) % 3;
    double area = 0;
    for (int a = 0; a < 4; ++a)
      for (int b = 0; b < 4; ++b) {
        double xi[3], J[3][3];
        xi[d] = f % 2;
        xi[t1] = g4p[a];
        xi[t2] = g4p[b];
        trilinear_jacobian(v, xi, J);
        const double n0 = J[1][t1] * J[2][t2] - J[2][t1] * J[1][t2];
        const double n1 = J[2][t1] * J[0][t2] - J[0][t1] * J[2][t2];
        const double n2 = J[0][t1] * J[1][t2] - J[1][t1] * J[0][t2];
        area += sqrt(n0 * n0 + n1 * n1 + n2 * n2) * g4w[a] * g4w[b];
      }
    h[6 * c + f] = vol / area;
  }
}

// sigma[F] = fac (1/h- + 1/h+), h+ = h- on boundary faces
__global__ void __launch_bounds__(256)
dg_penalty_kernel(const int64_t n_faces, const uint32_t *__restrict__ finfo,
                  const double *__restrict__ h, const double fac, double *__restrict__ sigma)
{
  const int64_t F = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (F >= n_faces) return;
  const uint32_t cm = finfo[3 * F], cp = finfo[3 * F + 1], fm = finfo[3 * F + 2];
  const double hm = h[6 * int64_t(cm) + fm];
  const double hp = cp == kInvalid ? hm : h[6 * int64_t(cp) + (fm ^ 1u)];
  sigma[F] = fac * (1.0 / hm + 1.0 / hp);
}

void launch_dg_box_setup(const pmg_box &box, int neumann_sides, const double *cv,
                         const double *g4p, const double *g4w, double penalty_fac,
                         uint32_t *cface, uint32_t *cnbr, uint32_t *finfo, double *h,
                         double *sigma, int64_t n_faces, hipStream_t s) {
  const int64_t nc = int64_t(box.nx) * box.ny * box.nz;
  if (nc <= 0) return;
  const dim3 gc(static_cast<unsigned>((nc + 255) / 256)), b(256);
  hipLaunchKernelGGL(dg_box_topology_kernel, gc, b, 0, s, box.nx, box.ny, box.nz, neumann_sides,
                     cface, cnbr, finfo);
  PMG_HIP_CHECK(hipGetLastError());
  hipLaunchKernelGGL(dg_cell_h_kernel, gc, b, 0, s, nc, cv, g4p, g4w, h);
  PMG_HIP_CHECK(hipGetLastError());
  hipLaunchKernelGGL(dg_penalty_kernel, dim3(static_cast<unsigned>((n_faces + 255) / 256)), b, 0,
                     s, n_faces, finfo, h, penalty_fac, sigma);
  PMG_HIP_CHECK(hipGetLastError());
}

// ---------------------------------------------------------------- DG <-> CG
// ContinuousTransfer (portable_continuous_transfer.h:112-160, 219-266): DG DoF
// e = c*NM^3 + i (or dg_map[e]) and CG DoF cg_map[e] share a support point.
template <typename T>
__global__ void __launch_bounds__(256)
dg_cg_prolongate_kernel(const int64_t n, const uint32_t *__restrict__ cg_map,
                        const uint32_t *__restrict__ dg_map, const T *__restrict__ src,
                        T *__restrict__ dst)
{
  const int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const uint32_t c = cg_map[e];
  if (c == kInvalid) return;
  const int64_t fdof = dg_map ? static_cast<int64_t>(dg_map[e]) : e;
  dst[fdof] += src[c];
}

template <typename T>
__global__ void __launch_bounds__(256)
dg_cg_restrict_kernel(const int64_t n, const uint32_t *__restrict__ cg_map,
                      const uint32_t *__restrict__ dg_map, const T *__restrict__ src,
                      const T *__restrict__ sub, T *__restrict__ dst)
{
  const int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const uint32_t c = cg_map[e];
  if (c == kInvalid) return;
  const int64_t fdof = dg_map ? static_cast<int64_t>(dg_map[e]) : e;
  atomic_add(dst + c, sub ? src[fdof] - sub[fdof] : src[fdof]);
}

// Restriction as a gather over the inverse map: CG DoF i sums the DG values
// of its copies list[off[i] .. off[i+1]) (CSR built at set-up), no atomics.
// DoFs with no copy (constrained) are left untouched, as by the scatter form.
// The copies are read in groups (entry 0; 1-3 where there are more; 4-7):
// all list entries of the groups first, then all values, so a DoF with 8
// copies waits two memory latencies instead of two per copy, and only lanes
// whose DoF has that many copies issue a group.
template <typename T, bool SUB>
__global__ void __launch_bounds__(256)
dg_cg_restrict_gather_kernel(const int64_t n_cg, const uint32_t *__restrict__ off,
                             const uint32_t *__restrict__ list, const T *__restrict__ src,
                             const T *__restrict__ sub, T *__restrict__ dst)
{
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n_cg) return;
  const uint32_t k0 = off[i], k1 = off[i + 1];
  if (k0 == k1) return;
  const uint32_t ne = k1 - k0, kl = k1 - 1;
  auto val = [&](uint32_t e) { return SUB ? src[e] - sub[e] : src[e]; };
  uint32_t e[8] = {};
  e[0] = list[k0];
  if (ne > 1) {
#pragma unroll
    for (int j = 1; j < 4; ++j) e[j] = list[k0 + j <= kl ? k0 + j : kl];
  }
  if (ne > 4) {
#pragma unroll
    for (int j = 4; j < 8; ++j) e[j] = list[k0 + j <= kl ? k0 + j : kl];
  }
  T a[8] = {};
  a[0] = val(e[0]);
  if (ne > 1) {
#pragma unroll
    for (int j = 1; j < 4; ++j) a[j] = val(e[j]);
  }
  if (ne > 4) {
#pragma unroll
    for (int j = 4; j < 8; ++j) a[j] = val(e[j]);
  }
  T acc = a[0];
#pragma unroll
  for (int j = 1; j < 8; ++j)
    if (static_cast<uint32_t>(j) < ne) acc += a[j];
  for (uint32_t k = k0 + 8; k < k1; ++k) acc += val(list[k]);
  dst[i] += acc;
}

template <typename T>
void launch_dg_cg_restrict_gather(int64_t n_cg, const uint32_t *off, const uint32_t *list,
                                  const T *src, T *dst, hipStream_t s, const T *sub) {
  if (n_cg == 0) return;
  const dim3 grid(static_cast<unsigned>((n_cg + 255) / 256));
  if (sub)
    hipLaunchKernelGGL((dg_cg_restrict_gather_kernel<T, true>), grid, dim3(256), 0, s, n_cg, off,
                       list, src, sub, dst);
  else
    hipLaunchKernelGGL((dg_cg_restrict_gather_kernel<T, false>), grid, dim3(256), 0, s, n_cg, off,
                       list, src, sub, dst);
  PMG_HIP_CHECK(hipGetLastError());
}
template void launch_dg_cg_restrict_gather<double>(int64_t, const uint32_t *, const uint32_t *,
                                                   const double *, double *, hipStream_t,
                                                   const double *);
template void launch_dg_cg_restrict_gather<float>(int64_t, const uint32_t *, const uint32_t *,
                                                  const float *, float *, hipStream_t,
                                                  const float *);

template <typename T>
void launch_dg_cg_transfer(bool prolongate, int64_t n_entries, const uint32_t *cg_map,
                           const uint32_t *dg_map, const T *src, T *dst, hipStream_t s,
                           const T *sub) {
  if (n_entries == 0) return;
  const dim3 grid(static_cast<unsigned>((n_entries + 255) / 256));
  if (prolongate)
    hipLaunchKernelGGL(dg_cg_prolongate_kernel<T>, grid, dim3(256), 0, s, n_entries, cg_map,
                       dg_map, src, dst);
  else
    hipLaunchKernelGGL(dg_cg_restrict_kernel<T>, grid, dim3(256), 0, s, n_entries, cg_map, dg_map,
                       src, sub, dst);
  PMG_HIP_CHECK(hipGetLastError());
}
template void launch_dg_cg_transfer<double>(bool, int64_t, const uint32_t *, const uint32_t *,
                                            const double *, double *, hipStream_t, const double *);
template void launch_dg_cg_transfer<float>(bool, int64_t, const uint32_t *, const uint32_t *,
                                           const float *, float *, hipStream_t, const float *);

}  // namespace pmg
