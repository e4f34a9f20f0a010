/*
 * pmg_oracle.h -- CPU oracle for the matrix-free high-order multigrid hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * algorithms (gshubhamk/portable-multigrid, deal.II Portable::MatrixFree +
 * Kokkos) used as the checker for the HIP library.  Only tests/, the smoke()
 * entry and the CPU-baseline leg of bench.py may load it; the product never
 * links against it.
 *
 * Every routine cites the reference file:line it restates.  Pinning (what
 * each routine is checked against, since the reference ships no golden
 * vectors and cannot be built here: it needs deal.II 9.8, Kokkos, p4est):
 *   1D data           closed forms (quadrature exactness, partition of unity,
 *                     D*1 = 0, GLL endpoint/symmetry)          tests/test_oracle_basis.py
 *   G tensors         closed forms on affine boxes, volume = sum JxW       test_oracle_bk3.py
 *   BK3 cell apply    dense element matrices with exact Lagrange
 *                     derivatives, Q1 27-point stencil, x^T A x = int |grad u|^2
 *                                                             test_oracle_bk3.py
 *   diagonal          dense element matrix diagonal            test_oracle_bk3.py
 *   subdomain vmult   dense assembly over the listed cells with the physical
 *                     rows/columns replaced by identity; A 1 = 0 without
 *                     physical DoFs; all-boundary case = the BK3 vmult
 *                                                             test_oracle_bk3.py
 *   h/p transfers     exact prolongation of polynomials, <Px,y> = <x,Ry>
 *                                                             test_oracle_transfer.py
 *   Chebyshev / CG    textbook Chebyshev error bound on a 1D Laplacian,
 *                     CG on a closed-form SPD system           test_oracle_solver.py
 *   SIP-DG            reduces to the continuous operator for smooth u,
 *                     symmetry, J^{-1}n vs J^{-T}n sheared cell test_oracle_dg.py
 *   adjoint gradient  difference of the bilinear form under a kappa change,
 *                     central finite differences of J          test_oracle_bk3.py,
 *                                                             test_oracle_adjoint.py
 *   eigenvalue estimate  dense Rayleigh-Ritz on the same Krylov space,
 *   and auto degree      closed-form Chebyshev bound             test_oracle_pinning.py
 *   SIP face measure     20x20 Gauss area of non-planar faces   test_oracle_pinning.py
 *   RHS and L2 error     closed-form integrals                  test_oracle_pinning.py
 * Every function is pinned.  The eigenvalue estimate restates deal.II
 * PreconditionChebyshev's published algorithm (start vector, stopping rule,
 * safety factor); its Ritz values are fixed by the Rayleigh-Ritz check.
 */
#ifndef PMG_ORACLE_H
#define PMG_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PMO_INVALID 0xFFFFFFFFu

/* ---------------- 1D data (pmo_basis.c) ---------------- */
void pmo_gauss_legendre(int n, double *x, double *w);   /* on [0,1] */
void pmo_gauss_lobatto(int n, double *x);               /* on [0,1] */
/* Lagrange basis on nodes z[0..m-1]: values and first derivatives at x */
void pmo_lagrange(int m, const double *z, double x, double *val, double *der);
/* S[i*nq+q] = phi_i(xi_q), phi on the p+1 GLL points, xi Gauss points;
 * D[n*nq+q] = l_n'(xi_q), l Lagrange on the nq Gauss points (collocation). */
void pmo_shape_matrices(int p, int nq, double *S, double *D);
/* Sd[i*nq+q] = phi_i'(xi_q) exactly (for dense checks) */
void pmo_shape_gradients(int p, int nq, double *Sd);

/* ---------------- mesh, DoFs, geometry (pmo_mesh.c) ---------------- */
typedef struct {
  int nx, ny, nz;          /* cells per direction */
  double lo[3], hi[3];     /* box before deformation */
  int deform;              /* 0: none, 1: sinusoidal cyclic displacement, 2: affine shear */
  double amplitude;        /* displacement amplitude / shear factor */
} pmo_mesh;

long pmo_n_cells(const pmo_mesh *m);
long pmo_n_vertices(const pmo_mesh *m);
void pmo_vertices(const pmo_mesh *m, double *xyz);           /* [V][3] */
void pmo_cell_vertices(const pmo_mesh *m, double *cv);       /* [C][8][3] */
long pmo_n_dofs_cg(const pmo_mesh *m, int p);
/* lexicographic cell map [C][(p+1)^3]; if dirichlet, boundary DoFs -> INVALID */
void pmo_dofmap_cg(const pmo_mesh *m, int p, int dirichlet, uint32_t *map);
/* list of boundary DoFs of the global lexicographic grid; returns count */
long pmo_boundary_dofs_cg(const pmo_mesh *m, int p, uint32_t *out);
/* G[c][6][nq^3] = JxW * J^{-1} J^{-T} (xx,xy,xz,yy,yz,zz) of the trilinear
 * map of each cell; optionally JxW[c][nq^3] and physical points X[c][nq^3][3] */
void pmo_geometry(const pmo_mesh *m, int nq, double *G, double *JxW, double *X);

/* ---------------- BK3 operator (pmo_bk3.c) ---------------- */
/* y += sum_cells A_e x  (reference KokkosKernel order) */
void pmo_bk3_cells(int p, int nq, long n_cells, const double *S, const double *D,
                   const double *G, const uint32_t *map, const double *x, double *y,
                   int n_threads, const int *colors /* NULL or [C] in 0..7 */);
/* dst = A src with constrained rows copied (LaplaceOperatorBK3::vmult) */
void pmo_bk3_vmult(int p, int nq, long n_cells, const double *S, const double *D,
                   const double *G, const uint32_t *map, long n_dofs,
                   const uint32_t *constrained, long n_constrained,
                   const double *src, double *dst, int n_threads, const int *colors);
/* subdomain vmult (Neumann / interface cell range): plain map with the
 * physical boundary DoFs eliminated, over every cell (cells NULL) or the
 * n_list listed cells, dst_b = src_b on the physical boundary DoFs */
void pmo_bk3_vmult_subdomain(int p, int nq, long n_cells, const double *S, const double *D,
                             const double *G, const uint32_t *plain_map, long n_dofs,
                             const uint32_t *physical, long n_physical, const uint32_t *cells,
                             long n_list, const double *src, double *dst);
/* dense element matrices K[c][nm^3][nm^3] with exact derivatives */
void pmo_bk3_element_matrices(int p, int nq, long n_cells, const double *G, double *K);
/* diag(A) (1 on constrained rows) with exact derivatives */
void pmo_bk3_diagonal(int p, int nq, long n_cells, const double *G, const uint32_t *map,
                      long n_dofs, const uint32_t *constrained, long n_constrained,
                      double *diag);
/* colors for the structured mesh: (cx&1) + 2(cy&1) + 4(cz&1) */
void pmo_cell_colors(const pmo_mesh *m, int *colors);
/* right-hand side b_i = int phi_i f with Gauss(nq) (f kinds: 0 const 1, 1 3pi^2 sin sin sin) */
void pmo_rhs(const pmo_mesh *m, int p, int nq, int f_kind, const uint32_t *map, long n_dofs,
             double *b);
/* L2 error of the Q_p field u against sin(pi x)sin(pi y)sin(pi z) (kind 1) or 0 (kind 0) */
double pmo_l2_error(const pmo_mesh *m, int p, int nq, int u_kind, const uint32_t *plain_map,
                    const double *u);
/* variable coefficient: G' = kappa * G (kappa[c][nq^3]) */
void pmo_scale_G(long n_cells, int nq, const double *kappa, const double *G, double *Gk);
/* adjoint gradient: out[c][q] = -kappa_q (grad_xi u)^T G_q (grad_xi lambda) */
void pmo_grad_kappa(int p, int nq, long n_cells, const double *S, const double *D,
                    const double *G, const double *kappa, const uint32_t *map,
                    const double *u, const double *lam, double *out);

/* ---------------- transfers (pmo_transfer.c) ---------------- */
/* h: coarse mesh m (nx,ny,nz), fine = 2x refined; patch (2p+1)^3 per coarse cell */
void pmo_h_matrix(int p, double *M /*(p+1)*(2p+1)*/);
void pmo_p_matrix(int pc, int pf, double *M /*(pc+1)*(pf+1)*/);
void pmo_h_maps(const pmo_mesh *coarse, int p, uint32_t *map_c /*[C][(p+1)^3]*/,
                uint32_t *map_f /*[C][(2p+1)^3]*/, double *w /*[C][(2p+1)^3]*/);
void pmo_p_maps(const pmo_mesh *m, int pc, int pf, uint32_t *map_c, uint32_t *map_f,
                double *w);
/* y_f += scale * w .* (M x M x M) x_c  (KokkosProlongationBatchedKernel) */
void pmo_prolongate(int nmc, int nmf, long n_cells, const double *M, const uint32_t *map_c,
                    const uint32_t *map_f, const double *w, double scale, const double *xc,
                    double *yf);
/* y_c += (M x M x M)^T (w .* x_f)  (KokkosRestrictionBatchedKernel) */
void pmo_restrict(int nmc, int nmf, long n_cells, const double *M, const uint32_t *map_c,
                  const uint32_t *map_f, const double *w, const double *xf, double *yc);

/* ---------------- SIP-DG (pmo_dg.c) ---------------- */
/* cell-major DG(p) numbering: DoF c*(p+1)^3 + lexicographic local index */
long pmo_dg_n_dofs(const pmo_mesh *m, int p);
/* sigma[c][f] of the face f (0: x=0, 1: x=1, 2: y=0, ...) of cell c;
 * kind 0: p(p+1)/2 (1/h- + 1/h+), kind 1: (p+1)^2 (1/h- + 1/h+), h = |K|/|F| */
void pmo_dg_penalty(const pmo_mesh *m, int p, int kind, double *sigma /*[C][6]*/);
/* dst = A_DG src; metric 0: d_n = (J^{-1} n).grad_xi, 1: (J^{-T} n).grad_xi;
 * bit f of neumann_mask marks box side f as Neumann (else weak Dirichlet) */
void pmo_dg_vmult(const pmo_mesh *m, int p, int nq, const double *sigma, int metric,
                  int neumann_mask, const double *src, double *dst);
/* DG(p) <-> CG(p) on one mesh; cg_map[C][(p+1)^3] with 0xFFFFFFFF = constrained */
void pmo_dg_cg_prolongate(int p, long n_cells, const uint32_t *cg_map, const double *src_cg,
                          double *dst_dg);
void pmo_dg_cg_restrict(int p, long n_cells, const uint32_t *cg_map, const double *src_dg,
                        double *dst_cg);

/* ---------------- solvers (pmo_solver.c) ---------------- */
typedef struct pmo_level pmo_level;
/* A level of a geometric/polynomial hierarchy on the unit-cube family */
typedef struct {
  int n_levels;
  int n_threads;
  int cheby_degree;          /* pre = post degree */
  double smoothing_range;    /* 15 */
  int eig_cg_n_iterations;   /* 10 */
  int coarse_kind;           /* 0: Chebyshev auto-degree to 1e-3, 1: PCG to coarse_tol */
  double coarse_tol;         /* 1e-12 */
  double h_prolong_scale;    /* 1 (mathematical P) or 2 (reference HEAD's doubled h-prolongation) */
  int dg_top;                /* 1: the finest level is SIP-DG on the mesh/degree of the level below */
  int dg_penalty_kind;       /* see pmo_dg_penalty */
  int dg_metric;             /* see pmo_dg_vmult */
} pmo_mg_params;

typedef struct pmo_mg pmo_mg;
/* hierarchy description: level l has mesh ms[l], degree ps[l]; level l-1 -> l is
 * an h-transfer if ms differ (2x), else a p-transfer.  nq = p+1 per level
 * unless nqs != NULL. */
pmo_mg *pmo_mg_create(int n_levels, const pmo_mesh *ms, const int *ps, const int *nqs,
                      const pmo_mg_params *prm);
void pmo_mg_destroy(pmo_mg *mg);
long pmo_mg_n_dofs(const pmo_mg *mg, int level);
void pmo_mg_vmult(pmo_mg *mg, double *dst, const double *src);
/* Level-wise access for tests */
void pmo_mg_level_vmult(pmo_mg *mg, int level, double *dst, const double *src);
void pmo_mg_smoother_info(pmo_mg *mg, int level, double *max_ev, double *min_ev,
                          int *degree);
void pmo_mg_smooth(pmo_mg *mg, int level, double *dst, const double *src, int step);
/* PCG: precond 0 none, 1 Jacobi (inverse diagonal of level), 2 MG V-cycle */
int pmo_pcg(pmo_mg *mg, int level, int precond, const double *b, double *x, double rel_tol,
            int max_it, double *res0, double *res);

/* Chebyshev / Lanczos helpers (exposed for tests) */
void pmo_tridiag_eigs(int n, const double *diag, const double *off, double *eig);

#ifdef __cplusplus
}
#endif
#endif
