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
// second pass.  The price is that each inner face's flux is evaluated twice
// (once from each side), which costs FP64 ALU time, not HBM bytes.
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
  constexpr int CELL_LDS = 3 * NQ3 + NM3;
  __shared__ T lds[E * CELL_LDS];

  const unsigned lb = xcd_logical_block(blockIdx.x, gridDim.x);
  const int tid = threadIdx.x;
  const int e = tid / NQ2, t = tid - e * NQ2, ta = t % NQ, tb = t / NQ;
  const int64_t cell = static_cast<int64_t>(lb) * E + e;
  const bool active = (e < E) && (cell < n_cells);
  const bool col = active && ta < NM && tb < NM;
  T *P0 = lds + (active ? e : 0) * CELL_LDS;
  T *P1 = P0 + NQ3, *P2 = P1 + NQ3, *U = P2 + NQ3;
  const T *Gc = G + g_index(NQ, active ? cell : 0, 0, t);

  // ---- own column (i,j) = (ta,tb) over k: contiguous cell-major loads
  T u[NM], out[NM];
  if (col) {
    const T *xc = x + cell * NM3 + tb * NM + ta;
#pragma unroll
    for (int k = 0; k < NM; ++k) {
      u[k] = xc[k * NM2];
      U[k * NM2 + tb * NM + ta] = u[k];
    }
  }
  T g[6][NQ];
  bk3_cell_apply<NM, NQ, T, kGLateNT>(u, out, g, Gc, sh, P0, P1, P2, active, ta, tb);
  __syncthreads();

  // ---- face terms; Y (= P0) accumulates them on the nodal grid
  T *Y = P0;
  if (active)
    for (int i = t; i < NM3; i += NQ2) Y[i] = T(0);
  T *NV = P1;  // 4 nodal face arrays [NM][NM]: v-, d_n v-, v+, d_n v+
  T *TA = P1;  // 4 face q-point arrays [NQ][NQ] (aliases NV after a barrier)

#pragma unroll 1
  for (int f = 0; f < 6; ++f) {
    const int d = f >> 1, s = f & 1;
    const int sn = d == 0 ? 1 : (d == 1 ? NM : NM2);
    const int st1 = d == 0 ? NM : 1;
    const int st2 = d == 2 ? NM : NM2;
    const uint32_t cf = active ? cface[cell * 6 + f] : 0u;
    const int role = static_cast<int>(cf >> 30);
    const int64_t F = cf & 0x3FFFFFFFu;

    // F1: traces and normal derivatives on the nodal face grid, lane (alpha,beta)
    if (col) {
      const int base = ta * st1 + tb * st2;
      T vm = 0, dm = 0, vp = 0, dp = 0;
#pragma unroll
      for (int k = 0; k < NM; ++k) {
        const T w = U[base + k * sn];
        dm += fs.dl[s * NM + k] * w;
        if (k == (s ? NM - 1 : 0)) vm = w;
      }
      if (role < 2) {
        const int64_t nb = cnbr[cell * 6 + f];
        const T *xn = x + nb * NM3 + base;
#pragma unroll
        for (int k = 0; k < NM; ++k) {
          const T w = xn[k * sn];
          dp += fs.dl[(1 - s) * NM + k] * w;
          if (k == (s ? 0 : NM - 1)) vp = w;
        }
      }
      const int id = tb * NM + ta;
      NV[id] = vm;
      NV[NM2 + id] = dm;
      NV[2 * NM2 + id] = vp;
      NV[3 * NM2 + id] = dp;
    }
    __syncthreads();

    // F2: values and reference gradients at face q-point (a,b) = (ta,tb), flux
    T cval = 0, cn = 0, c1 = 0, c2 = 0;
    if (active) {
      T um = 0, um1 = 0, um2 = 0, umn = 0, up = 0, up1 = 0, up2 = 0, upn = 0;
#pragma unroll
      for (int be = 0; be < NM; ++be) {
        T a0 = 0, a1 = 0, a2 = 0, b0 = 0, b1 = 0, b2 = 0;
#pragma unroll
        for (int al = 0; al < NM; ++al) {
          const T sa = sh.S[al * NQ + ta], da = fs.Sd[al * NQ + ta];
          const T w0 = NV[be * NM + al], w1 = NV[NM2 + be * NM + al];
          const T w2 = NV[2 * NM2 + be * NM + al], w3 = NV[3 * NM2 + be * NM + al];
          a0 += sa * w0;
          a1 += da * w0;
          a2 += sa * w1;
          b0 += sa * w2;
          b1 += da * w2;
          b2 += sa * w3;
        }
        const T sb = sh.S[be * NQ + tb], db = fs.Sd[be * NQ + tb];
        um += sb * a0;
        um1 += sb * a1;
        um2 += db * a0;
        umn += sb * a2;
        up += sb * b0;
        up1 += sb * b1;
        up2 += db * b0;
        upn += sb * b2;
      }
      const T *fd = fdata + F * 7 * NQ2 + t;
      const T jxw = fd[6 * NQ2];
      const T sig = fsig[F];
      T mo0, mo1, mo2;
      T dnm, dnp;
      if (role == 1) {
        mo0 = -fd[3 * NQ2];
        mo1 = -fd[4 * NQ2];
        mo2 = -fd[5 * NQ2];
        dnm = mo0 * umn + mo1 * um1 + mo2 * um2;
        dnp = -(fd[0] * upn + fd[NQ2] * up1 + fd[2 * NQ2] * up2);
      } else {
        mo0 = fd[0];
        mo1 = fd[NQ2];
        mo2 = fd[2 * NQ2];
        dnm = mo0 * umn + mo1 * um1 + mo2 * um2;
        if (role == 0) {
          dnp = fd[3 * NQ2] * upn + fd[4 * NQ2] * up1 + fd[5 * NQ2] * up2;
        } else if (role == 2) {
          up = -um;
          dnp = dnm;
        } else {
          up = um;
          dnp = -dnm;
        }
      }
      const T jump = um - up, avg = T(0.5) * (dnm + dnp);
      const T tv = (sig * jump - avg) * jxw, td = T(-0.5) * jump * jxw;
      cval = tv;
      cn = td * mo0;
      c1 = td * mo1;
      c2 = td * mo2;
    }
    __syncthreads();
    if (active) {
      TA[t] = cval;
      TA[NQ2 + t] = cn;
      TA[2 * NQ2 + t] = c1;
      TA[3 * NQ2 + t] = c2;
    }
    __syncthreads();

    // F3: test with the own basis, lane (alpha,beta); add to the normal line
    if (col) {
      T W = 0, Z = 0;
#pragma unroll
      for (int b = 0; b < NQ; ++b) {
        T r1 = 0, r2 = 0, r3 = 0;
#pragma unroll
        for (int a = 0; a < NQ; ++a) {
          const T sa = sh.S[ta * NQ + a], da = fs.Sd[ta * NQ + a];
          r1 += TA[b * NQ + a] * sa + TA[2 * NQ2 + b * NQ + a] * da;
          r2 += TA[3 * NQ2 + b * NQ + a] * sa;
          r3 += TA[NQ2 + b * NQ + a] * sa;
        }
        const T sb = sh.S[tb * NQ + b], db = fs.Sd[tb * NQ + b];
        W += r1 * sb + r2 * db;
        Z += r3 * sb;
      }
      const int base = ta * st1 + tb * st2;
#pragma unroll
      for (int k = 0; k < NM; ++k) {
        T v = fs.dl[s * NM + k] * Z;
        if (k == (s ? NM - 1 : 0)) v += W;
        Y[base + k * sn] += v;
      }
    }
    __syncthreads();
  }

  // ---- write the cell's result: cell part (registers) + face part (Y)
  if (col) {
    T *yc = y + cell * NM3 + tb * NM + ta;
#pragma unroll
    for (int k = 0; k < NM; ++k) {
      const T v = out[k] + Y[k * NM2 + tb * NM + ta];
      if (ADD) yc[k * NM2] += v;
      else yc[k * NM2] = v;
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
                      T *__restrict__ dst)
{
  const int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const uint32_t c = cg_map[e];
  if (c == kInvalid) return;
  const int64_t fdof = dg_map ? static_cast<int64_t>(dg_map[e]) : e;
  atomic_add(dst + c, src[fdof]);
}

template <typename T>
void launch_dg_cg_transfer(bool prolongate, int64_t n_entries, const uint32_t *cg_map,
                           const uint32_t *dg_map, const T *src, T *dst, hipStream_t s) {
  if (n_entries == 0) return;
  const dim3 grid(static_cast<unsigned>((n_entries + 255) / 256));
  if (prolongate)
    hipLaunchKernelGGL(dg_cg_prolongate_kernel<T>, grid, dim3(256), 0, s, n_entries, cg_map,
                       dg_map, src, dst);
  else
    hipLaunchKernelGGL(dg_cg_restrict_kernel<T>, grid, dim3(256), 0, s, n_entries, cg_map, dg_map,
                       src, dst);
  PMG_HIP_CHECK(hipGetLastError());
}
template void launch_dg_cg_transfer<double>(bool, int64_t, const uint32_t *, const uint32_t *,
                                            const double *, double *, hipStream_t);
template void launch_dg_cg_transfer<float>(bool, int64_t, const uint32_t *, const uint32_t *,
                                           const float *, float *, hipStream_t);

}  // namespace pmg
