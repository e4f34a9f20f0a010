/*
 * pmg.h -- C ABI of the MI355X-native matrix-free multigrid library (libpmg.so).
 *
 * Drop-in boundary for the hot path of gshubhamk/portable-multigrid: the
 * BK3-style continuous Q_p Laplace operator, the SIP-DG operator, h-/p- and
 * DG->CG transfers, the Chebyshev/Jacobi smoother, the V-cycle and the
 * (mixed-precision) preconditioned CG solve.  Every entry point names the
 * reference interface it replaces (paths relative to the reference root).
 *
 * Conventions
 *  - Return value: PMG_OK (0) or a PMG_ERR_* status; pmg_last_error() gives a
 *    message (thread-local).  No C++ exception crosses this boundary.
 *  - Vectors are caller-owned DEVICE pointers (hipMalloc / torch.cuda) of the
 *    handle's precision (double for PMG_FP64, float for PMG_FP32), length m().
 *  - Setup arrays (DoF maps, vertices, weights, G tensors) are HOST pointers;
 *    the library copies them to device memory it owns.  Operator, transfer,
 *    smoother and V-cycle applications allocate nothing (the first V-cycle
 *    application instantiates its graph); the first pmg_cg_solve on an
 *    operator allocates the CG vectors, which the operator handle keeps for
 *    later solves of the same size; pmg_vector_dot and the set-up/
 *    post-processing calls (create, compute_diagonal, assemble_rhs,
 *    l2_error) allocate temporaries.
 *  - stream: a hipStream_t (NULL = the default stream).  All work is
 *    stream-ordered and asynchronous unless stated otherwise.  A handle must
 *    not be used from two host threads at once.
 *  - DoF maps are lexicographic per cell (x fastest), [n_cells][(p+1)^3],
 *    with 0xFFFFFFFF marking a constrained (Dirichlet) DoF, exactly as the
 *    reference's dof_indices views (portable_laplace_operator_bk3.h:360-437).
 */
#ifndef PMG_H
#define PMG_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PMG_INVALID_DOF 0xFFFFFFFFu

enum pmg_status {
  PMG_OK = 0,
  PMG_ERR_INVALID_ARG = 1,
  PMG_ERR_SIZE = 2,
  PMG_ERR_HIP = 3,
  PMG_ERR_NOT_CONVERGED = 4,
  PMG_ERR_UNSUPPORTED = 5,
  PMG_ERR_ALLOC = 6
};

enum pmg_precision { PMG_FP64 = 0, PMG_FP32 = 1 };

enum pmg_transfer_kind { PMG_TRANSFER_H = 0, PMG_TRANSFER_P = 1, PMG_TRANSFER_DG_CG = 2 };

enum pmg_precond_kind { PMG_PRECOND_NONE = 0, PMG_PRECOND_JACOBI = 1, PMG_PRECOND_MG = 2 };

enum pmg_coarse_kind { PMG_COARSE_CHEBYSHEV = 0, PMG_COARSE_PCG = 1 };

/* Index arrays of an operator (pmg_operator_index_data), element type in
 * brackets, cells in the operator's internal order (cell order array) */
enum pmg_index_array {
  PMG_INDEX_MAP = 0,               /* [uint32] [C][(p+1)^3], constrained entries invalid */
  PMG_INDEX_CONSTRAINED = 1,       /* [uint32] the constrained DoF list */
  PMG_INDEX_CELL_ORDER = 2,        /* [uint32] internal -> caller cell index */
  PMG_INDEX_LINES = 3,             /* [uint32] [C][(p+1)^2] x-line starts (0 where not compressed) */
  PMG_INDEX_LINE_FLAGS = 4,        /* [uint8]  [C] 1 where every x-line is compressed */
  PMG_INDEX_PATCH_DOFS = 5,        /* [uint32] [batches][E (p+1)^3] patch DoFs, bit 31 shared */
  PMG_INDEX_PATCH_COUNTS = 6,      /* [uint32] [batches] */
  PMG_INDEX_PATCH_SLOTS = 7,       /* [uint16] [C][(p+1)^3] patch slot, 0xFFFF constrained */
  PMG_INDEX_PATCH_LINES = 8,       /* [uint16] [C][(p+1)^2] first slot of each x-line */
  PMG_INDEX_PATCH_LINE_FLAGS = 9,  /* [uint8]  [C] */
  PMG_INDEX_VERTICES = 10,         /* [double] [C][8][3] */
  PMG_INDEX_PATCH_COLORS = 11,     /* [uint8]  [C] colour of the cell within its batch */
  PMG_INDEX_PATCH_NCOLORS = 12,    /* [uint8]  [batches] colours (0: the batch adds with atomics) */
  PMG_INDEX_PATCH_CHUNKS = 13      /* [uint32] [batches][ceil(E (p+1)^3 / 64)][8] the patch DoFs of
                                      every 64 slots as <= 4 runs: first DoFs of runs 0..3,
                                      s1 | s2 << 8 | s3 << 16 | n_runs << 24 (run k starts at
                                      slot s_k, 64 if absent; n_runs 0: more than 4 runs), 0, 0,
                                      0; empty when fewer than half of the chunks are runs */
};
/* Index arrays of a transfer (pmg_transfer_index_data) */
enum pmg_transfer_index_array {
  PMG_TRANSFER_INDEX_COARSE_MAP = 0,        /* [uint32] [C][nc^3], constrained invalid */
  PMG_TRANSFER_INDEX_FINE_MAP = 1,          /* [uint32] [C][nf^3], bit 31 owner mark */
  PMG_TRANSFER_INDEX_FINE_LINES = 2,        /* [uint32] [C][nf^2][2] {first DoF, owner bits} */
  PMG_TRANSFER_INDEX_RESTRICT_OFFSETS = 3,  /* [uint32] coarse DoF i -> its restriction buffer
                                               range [off[i], off[i+1]) */
  PMG_TRANSFER_INDEX_RESTRICT_POSITIONS = 4, /* [uint32] [C][nc^3] buffer position of each coarse
                                               map entry (entries ordered by coarse DoF, then
                                               by entry index), constrained entries invalid */
  PMG_TRANSFER_INDEX_DG_OFFSETS = 5,        /* [uint32] CG DoF -> DG copies (CSR) */
  PMG_TRANSFER_INDEX_DG_LIST = 6            /* [uint32] */
};

typedef struct pmg_operator_s *pmg_operator;
typedef struct pmg_transfer_s *pmg_transfer;
typedef struct pmg_smoother_s *pmg_smoother;
typedef struct pmg_multigrid_s *pmg_multigrid;
typedef void *pmg_stream;

/* ------------------------------------------------------------ library */
/* Message of the last failing call on this thread ("" if none). */
const char *pmg_last_error(void);
/* ABI version (major*10000 + minor*100 + patch). */
int pmg_version(void);
/* hipDeviceSynchronize(); reports asynchronous kernel errors. */
int pmg_device_synchronize(void);

/* ------------------------------------------------------------ structured meshes
 * Host-side helpers that build the inputs deal.II's DoFHandler/MatrixFree
 * produce in the reference (tests/poisson_cube_bk3/program.cc:281-566): a
 * box [lo,hi] of nx*ny*nz hexahedra, optionally displaced by
 * x += a sin(2 pi y) sin(2 pi z) and cyclic permutations (deform = 1) or
 * sheared by x += a y, y += a z, z += a x (deform = 2), with a
 * lexicographic global numbering on the (p nx+1)(p ny+1)(p nz+1) grid.
 * All outputs are caller-allocated host arrays. */
typedef struct {
  int nx, ny, nz;
  double lo[3], hi[3];
  int deform;
  double amplitude;
} pmg_box;

int64_t pmg_box_n_cells(const pmg_box *box);
int64_t pmg_box_n_dofs(const pmg_box *box, int degree);
/* [C][8][3], vertices of each cell in lexicographic order */
int pmg_box_cell_vertices(const pmg_box *box, double *cell_vertices);
/* [C][(p+1)^3]; dirichlet != 0 marks boundary DoFs 0xFFFFFFFF */
int pmg_box_dofmap(const pmg_box *box, int degree, int dirichlet, uint32_t *dofmap);
/* boundary DoF list; pass NULL to query the count in *n */
int pmg_box_boundary_dofs(const pmg_box *box, int degree, uint32_t *dofs, int64_t *n);
/* h-transfer patch maps of GeometricTransfer<dim,p>::setup_dof_indices /
 * setup_weights (include/multigrid/portable_geometric_transfer.h:1344-1486):
 * coarse map [C_c][(p+1)^3] (constrained -> invalid), fine patch map
 * [C_c][(2p+1)^3] (plain), weights [C_c][(2p+1)^3] (1/multiplicity, 0 on
 * constrained fine DoFs; weights = NULL skips them, for
 * pmg_transfer_create_from_constraints).  The fine box is the coarse box
 * refined once. */
int pmg_box_h_transfer_maps(const pmg_box *coarse, int degree, uint32_t *map_coarse,
                            uint32_t *map_fine, double *weights);
/* p-transfer maps of PolynomialTransfer<dim,pc,pf>::setup_weights_and_dof_indices
 * (include/multigrid/portable_polynomial_transfer.h:1155-1360); weights may be NULL */
int pmg_box_p_transfer_maps(const pmg_box *box, int degree_coarse, int degree_fine,
                            uint32_t *map_coarse, uint32_t *map_fine, double *weights);
/* DG face lists of the box (the face_info of LaplaceOperatorDG::compute_face_info,
 * include/operators/portable_laplace_operator_dg.h:623-675): inner faces
 * [F][4] = (cell-, cell+, face-, face+) with face- = 2d+1, face+ = 2d between
 * neighbours along axis d; boundary faces [Fb][3] = (cell, face, boundary id 0).
 * Pass NULL arrays to query the counts. */
int pmg_box_dg_faces(const pmg_box *box, uint32_t *inner_faces, int64_t *n_inner,
                     uint32_t *boundary_faces, int64_t *n_boundary);

/* ------------------------------------------------------------ operator
 * Replaces Portable::LaplaceOperatorBK3<dim,p,number>
 * (include/operators/portable_laplace_operator_bk3.h:19-96, ctor :104-126) and
 * the LaplaceOperatorBase<dim,Number> interface
 * (include/base/portable_laplace_operator_base.h:16-61).  dim = 3,
 * 1 <= degree <= 6 (OperatorDispatchFactory, max 6), n_q_points_1d = degree+1
 * (the reference, portable_laplace_operator_bk3.h:120) or degree+2.
 * Geometry: the trilinear map of each cell's 8 vertices (what MappingQ(p)
 * reproduces on straight-sided hexahedra), G computed on the device
 * (compute_G_tensors, portable_laplace_operator_bk3.h:456-515).
 * Constrained DoFs are eliminated: a map entry naming a DoF of the
 * constrained list acts as 0xFFFFFFFF, and vmult returns dst = src on those
 * rows (setup_dof_indices_per_color, portable_laplace_operator_bk3.h:360-437). */
int pmg_operator_create(int dim, int degree, int n_q_points_1d, int precision,
                        int64_t n_cells, const uint32_t *dofmap, int64_t n_dofs,
                        const uint32_t *constrained_dofs, int64_t n_constrained,
                        const double *cell_vertices, pmg_operator *op);
/* Same, with G[C][6][nq^3] (xx,xy,xz,yy,yz,zz, already times JxW) given. */
/* Same for the Q_p operator of a structured box (pmg_box_dofmap's
 * numbering, every boundary DoF constrained), its index data generated on the
 * device -- what pmg_box_dofmap, pmg_box_boundary_dofs and
 * pmg_box_cell_vertices followed by pmg_operator_create give, without a
 * host pass over the cells: the DoFHandler/MatrixFree set-up of the
 * reference's drivers (tests/poisson_cube_bk3/program.cc:281-566,
 * setup_dof_indices_per_color, portable_laplace_operator_bk3.h:360-437).
 * Declares the box structure (pmg_operator_set_box_structure) where the grid
 * schedule's index limits allow.  Cells stay in box order (PMG_CELL_ORDER is
 * not consulted). */
int pmg_operator_create_box(const pmg_box *box, int degree, int n_q_points_1d, int precision,
                            pmg_operator *op);
int pmg_operator_create_from_G(int dim, int degree, int n_q_points_1d, int precision,
                               int64_t n_cells, const uint32_t *dofmap, int64_t n_dofs,
                               const uint32_t *constrained_dofs, int64_t n_constrained,
                               const double *G, pmg_operator *op);
int pmg_operator_destroy(pmg_operator op);
/* m() / n() (portable_laplace_operator_bk3.h:565-577) */
int pmg_operator_m(pmg_operator op, int64_t *m);
int pmg_operator_info(pmg_operator op, int *degree, int *n_q_points_1d, int *precision,
                      int64_t *n_cells);
/* vmult: dst = A src with dst_c = src_c on constrained rows
 * (LaplaceOperatorBK3::vmult, portable_laplace_operator_bk3.h:128-231).
 * src and dst must not alias. */
int pmg_operator_vmult(pmg_operator op, void *dst, const void *src, pmg_stream stream);
/* Tvmult == vmult (portable_laplace_operator_bk3.h:440-453) */
int pmg_operator_Tvmult(pmg_operator op, void *dst, const void *src, pmg_stream stream);
/* dst += sum_cells A_e src, no framing (the raw cell loop a subdomain
 * operator such as SubdomainLaplaceOperator::vmult_bk3 calls) */
int pmg_operator_apply_add(pmg_operator op, void *dst, const void *src, pmg_stream stream);
/* Domain-decomposition subdomain operator: replaces the index sets and cell
 * loops of Portable::SubdomainLaplaceOperator<dim,p,number>
 * (include/operators/portable_subdomain_laplace_operator.h:23-175).  The
 * subdomain's operator is created with every subdomain constraint (physical
 * boundary and interface DoFs): its vmult is vmult_bk3 (:311-382) and its
 * apply_add the bare cell loop.  set_subdomain adds the plain index set of the
 * set-up at :1056-1061 -- dofmap (host, caller cell order, as given to create)
 * with only the physical_dofs eliminated -- and the interface cells (host,
 * caller indices of the cells with a face on the interface, :1181-1203).
 * Copies what it needs; may be called again to replace the sets.  The Neumann
 * vmult uses the operator's scatter schedule as it is at this call (patch or
 * atomic; colored falls back to atomic). */
int pmg_operator_set_subdomain(pmg_operator op, const uint32_t *dofmap,
                               const uint32_t *physical_dofs, int64_t n_physical,
                               const uint32_t *interface_cells, int64_t n_interface_cells);
/* vmult_neumann (:384-454): dst = sum over every cell of A_e src with the
 * plain indices, then dst_b = src_b on the physical boundary DoFs.  Interface
 * DoFs are free rows.  src and dst must not alias. */
int pmg_operator_vmult_neumann(pmg_operator op, void *dst, const void *src, pmg_stream stream);
/* vmult_interface_cell_range (:456-533): the same over the interface cells
 * only (rows no interface cell touches are 0, physical rows = src). */
int pmg_operator_vmult_interface_cells(pmg_operator op, void *dst, const void *src,
                                       pmg_stream stream);
/* compute_diagonal(): diag(A), 1 on constrained rows, then inverted
 * (portable_laplace_operator_bk3.h:533-561). Synchronous. */
int pmg_operator_compute_diagonal(pmg_operator op);
/* get_matrix_diagonal_inverse(): copies the inverse diagonal into out (device, m()) */
int pmg_operator_diagonal_inverse(pmg_operator op, void *out, pmg_stream stream);
/* el(i, i) (portable_laplace_operator_bk3.h:586-595): 1 / (inverse diagonal)_ii
 * of the last compute_diagonal() (constrained rows: 1).  Synchronous. */
int pmg_operator_diagonal_entry(pmg_operator op, int64_t row, double *value);
/* vmult_dummy(dst, src, ghost_exchange_on, computation_on)
 * (include/base/portable_laplace_operator_base.h:27-31, implementation
 * portable_laplace_operator_bk3.h:233-357), the reference's instrument for
 * timing the halo exchange and the cell loop apart.  On one GPU there is no
 * ghost exchange, so it is: dst = 0; computation_on: dst += sum_cells A_e src;
 * ghost_exchange_on: dst_c = src_c on the constrained rows (both on: vmult). */
int pmg_operator_vmult_dummy(pmg_operator op, void *dst, const void *src, int ghost_exchange_on,
                             int computation_on, pmg_stream stream);
/* Variable coefficient a(u,v) = sum_q kappa_q grad u . grad v JxW: kappa is a
 * DEVICE array [C][nq^3] of doubles; G <- kappa * G_geometry.  Requires an
 * operator created from cell vertices.  kappa = NULL restores kappa = 1. */
int pmg_operator_set_coefficient(pmg_operator op, const double *kappa, pmg_stream stream);
/* Affine geometry for vmult (enable 1) or G at every q-point (enable 0, the
 * default).  On affine cells G_q = A_c w_p w_q w_r with A_c = |det J| J^-1 J^-T
 * constant per cell, so the atomic and patch schedules read 6 values per
 * cell instead of 6 nq^3; the result equals the G form up to rounding.
 * Requires an operator created from cell vertices whose cells are all affine
 * (PMG_ERR_INVALID_ARG otherwise) and kappa = 1; setting a coefficient turns
 * it off.  No reference counterpart: the reference stores G at every
 * quadrature point (portable_laplace_operator_bk3.h:456-515); this serves
 * callers with affine meshes (boxes, sheared boxes). */
int pmg_operator_use_affine_geometry(pmg_operator op, int enable);
/* Adjoint gradient: out[c][q] = -(grad_xi u)^T G_q (grad_xi lambda) with the
 * operator's current G (= kappa G_geometry), i.e. dJ/dg_q for kappa = exp(g)
 * when lambda solves A lambda = dJ/du.  u, lambda: device vectors of the
 * operator precision; out: device [C][nq^3] of the operator precision.
 * No reference counterpart (new in this library). */
int pmg_operator_gradient_coefficient(pmg_operator op, const void *u, const void *lambda,
                                      void *out, pmg_stream stream);
/* Right-hand side b_i = int phi_i f dx with Gauss(n_q_points_1d) (<= 0: degree+1)
 * on the operator's cells, constrained entries 0 (LaplaceProblem::compute_rhs,
 * tests/poisson_cube_bk3/program.cc:517-566).  f_kind 0: f = 1 (the reference);
 * 1: f = 3 pi^2 sin(pi x) sin(pi y) sin(pi z).  b: device, op precision.
 * Needs an operator created from cell vertices.  Synchronous. */
int pmg_operator_assemble_rhs(pmg_operator op, int f_kind, int n_q_points_1d, void *b,
                              pmg_stream stream);
/* ||u_h - u||_L2 with Gauss(n_q_points_1d) (<= 0: degree+2), u_kind 0: u = 0,
 * 1: u = sin(pi x) sin(pi y) sin(pi z) (VectorTools::integrate_difference,
 * source/geometric_multigrid_bk3/program.cc:391-400).  Synchronous. */
int pmg_operator_l2_error(pmg_operator op, const void *u, int u_kind, int n_q_points_1d,
                          double *error, pmg_stream stream);
/* Cell-loop schedule (BK3 operators):
 *  0 atomic: one launch, every cell scatter-adds its (p+1)^3 values with
 *    device FP atomics (as the reference's Kokkos::atomic_add,
 *    bk3_kokkos_kernel.h:732);
 *  1 colored: one launch per colour of a greedy cell colouring with a plain
 *    read-modify-write scatter (atomics-free);
 *  2 patch: one launch; each workgroup's batch of cells accumulates the union
 *    of its DoFs in LDS and adds it once per DoF, with device atomics in
 *    sorted DoF order (fewer, longer-run atomics);
 *  3 gather: two launches and no atomics on the output: the cells store
 *    their values to a per-cell buffer, then every DoF sums its entries
 *    through the inverse map (built on the device at the first selection;
 *    extra device memory n_cells (p+1)^3 (4 + s) + 4 n_dofs bytes); the
 *    summation order is fixed, so results repeat bit for bit;
 *  4 grid: the same two passes without the inverse map, for an operator
 *    declared a lexicographic box (pmg_operator_set_box_structure): the cells
 *    store to the box's exploded grid and every DoF sums its <= 8 adjacent
 *    copies (coalesced, no atomics, bit-reproducible; n_cells (p+1)^3 s bytes).
 * Default: patch for n_q_points_1d <= 5 (FP64) or <= 7 (FP32) on operators
 * of more than 4096 cells, atomic otherwise (the faster of the two on MI355X
 * for each level type and size of the MG hierarchies);
 * PMG_BK3_SCHEDULE=atomic|colored|patch|gather|grid overrides it.
 * g_load: 0 non-temporal G loads at kernel entry, 1 ordinary loads at entry,
 * 2 non-temporal loads at first use, -1 the schedule's default (0 for patch,
 * 2 otherwise; PMG_BK3_GLOAD overrides the constructor's choice).  Variant 1,
 * and variant 0 of the atomic schedule, exist for degree 4 only; elsewhere
 * they fall back to 2. */
int pmg_operator_set_schedule(pmg_operator op, int schedule, int g_load);
/* Declares the operator's cells the lexicographic box of nx x ny x nz cells
 * (x fastest) with the lexicographic DoF numbering on the (p nx+1)(p ny+1)
 * (p nz+1) grid -- what pmg_box_dofmap produces; constrained DoFs may appear
 * as 0xFFFFFFFF.  Verified on the device against the operator's map
 * (PMG_ERR_INVALID_ARG if it differs); enables schedule 4 (grid).  The
 * reference's MatrixFree knows no such structure: no counterpart. */
int pmg_operator_set_box_structure(pmg_operator op, int nx, int ny, int nz);
/* Bytes one vmult must move (BK3: s(2N + 6 C nq^3); DG: adds the face data
 * and connectivity) and its flop count. */
int pmg_operator_algorithmic_bytes(pmg_operator op, double *bytes, double *flops);
/* Copies the operator's device index array `which` (pmg_index_array) to the
 * host array `host` of *n elements; host NULL only sets *n (0 when the array
 * is not built: patch arrays exist once the patch schedule was selected).
 * The views of the index data get_matrix_free() exposes
 * (include/base/portable_laplace_operator_base.h:56-59), for inspection and
 * tests. */
int pmg_operator_index_data(pmg_operator op, int which, void *host, int64_t *n);

/* ------------------------------------------------------------ SIP-DG operator
 * Replaces Portable::LaplaceOperatorDG<dim,p,n_q_points_1d,number>
 * (include/operators/portable_laplace_operator_dg.h:20-151, ctor :154-200,
 * vmult :252-351, compute_diagonal :887-915) on trilinear cells.  The handle
 * is a pmg_operator: vmult, apply_add (= vmult added to dst), m, info,
 * compute_diagonal / diagonal_inverse (cell integrals only, as the reference),
 * assemble_rhs, l2_error, Chebyshev, multigrid and CG accept it.
 * DoFs are cell-major: cell c owns [c (p+1)^3, (c+1) (p+1)^3), lexicographic.
 * inner_faces [F][4] = (cell-, cell+, face-, face+), face+ = face- ^ 1;
 * boundary_faces [Fb][3] = (cell, face, boundary_id): id 0 weak homogeneous
 * Dirichlet (u+ = -u-), other ids homogeneous Neumann
 * (laplace_dg_kokkos_kernels.h:1563-1576).  Faces numbered 0: x=0, 1: x=1,
 * 2: y=0, 3: y=1, 4: z=0, 5: z=1 in the reference cell.
 * penalty: NULL, or [F + Fb] values (inner faces first); NULL computes
 * penalty_kind 0: p(p+1)/2 (1/h- + 1/h+) (get_penalty_factor, :147-151) or
 * penalty_kind 1: (p+1)^2 (1/h- + 1/h+), h = |K| / |F|, h+ = h- on the boundary.
 * metric 0: d_n u = (J^{-1} n) . grad_xi u (exact normal derivative);
 * metric 1: (J^{-T} n) . grad_xi u as the reference GPU face data stores it
 * (portable_laplace_operator_dg.h:781-789); equal on axis-aligned cells. */
int pmg_dg_operator_create(int dim, int degree, int n_q_points_1d, int precision,
                           int64_t n_cells, const double *cell_vertices, int64_t n_inner_faces,
                           const uint32_t *inner_faces, int64_t n_boundary_faces,
                           const uint32_t *boundary_faces, const double *penalty,
                           int penalty_kind, int metric, pmg_operator *op);
/* The same operator for a box (pmg_box_dg_faces' faces; boundary id of box
 * side f: bit f of neumann_sides, 0 Dirichlet, 1 Neumann) with vertices,
 * connectivity, penalty (penalty_kind 0 or 1), G and face data built on the
 * device: no host arrays and no host loops over cells or faces. */
int pmg_dg_operator_create_box(const pmg_box *box, int degree, int n_q_points_1d, int precision,
                               int penalty_kind, int metric, int neumann_sides, pmg_operator *op);
/* penalty values in use, [F + Fb] (host array) */
int pmg_dg_operator_penalty(pmg_operator op, double *penalty);

/* ------------------------------------------------------------ transfers
 * Replaces Portable::MGTransferBase<dim,number> (include/base/portable_mg_transfer_base.h:15-36)
 * as implemented by GeometricTransfer<dim,p,number> (h), PolynomialTransfer<dim,pc,pf,number> (p)
 * and ContinuousTransfer<dim,p,number> (DG->CG).
 * kind H: degree_fine must equal degree_coarse; per coarse cell a (2p+1)^3 patch.
 * kind P: same mesh, degree_coarse < degree_fine.
 * kind DG_CG: same mesh and degree; map_coarse is the CG map (invalid on
 *   constrained), map_fine the DG DoF of each cell-local index, or NULL for the
 *   cell-major numbering of pmg_dg_operator_create; weights are ignored
 *   (ContinuousTransfer, include/multigrid/portable_continuous_transfer.h:112-266).
 * prolongation_scale multiplies the weights of the prolongation only: 1.0 is
 * the mathematical P; 2.0 reproduces the reference HEAD, whose h-prolongation
 * launches two kernels into the same vector (portable_geometric_transfer.h:805-855). */
int pmg_transfer_create(int kind, int precision, int degree_coarse, int degree_fine,
                        int64_t n_cells, const uint32_t *map_coarse, const uint32_t *map_fine,
                        const double *weights, int64_t n_dofs_coarse, int64_t n_dofs_fine,
                        double prolongation_scale, pmg_transfer *tr);
/* Kinds H and P with the weights built on the device from the fine level's
 * constraints, as GeometricTransfer::setup_weights and its PolynomialTransfer
 * counterpart compute them (include/multigrid/portable_geometric_transfer.h:1344-1380,
 * portable_polynomial_transfer.h:1155-1360): 1/multiplicity of each fine DoF
 * over map_fine, 0 on the constrained_fine DoFs (host list).  The library
 * keeps them as owner marks, so the caller builds and passes no weights. */
int pmg_transfer_create_from_constraints(int kind, int precision, int degree_coarse,
                                         int degree_fine, int64_t n_cells,
                                         const uint32_t *map_coarse, const uint32_t *map_fine,
                                         const uint32_t *constrained_fine,
                                         int64_t n_constrained_fine, int64_t n_dofs_coarse,
                                         int64_t n_dofs_fine, double prolongation_scale,
                                         pmg_transfer *tr);
/* The transfer between levels of structured boxes (every boundary DoF
 * constrained), maps and weights generated on the device: kind H from `box`
 * to the box refined once (degree_coarse == degree_fine), P on `box` from
 * degree_coarse to degree_fine, DG_CG on `box` (CG(p) -> cell-major DG(p)).
 * Equals pmg_transfer_create_from_constraints (H, P) or pmg_transfer_create
 * (DG_CG) over pmg_box_h/p_transfer_maps / pmg_box_dofmap and the fine
 * boundary DoFs (GeometricTransfer::setup_dof_indices / setup_weights,
 * portable_geometric_transfer.h:1344-1486; PolynomialTransfer,
 * portable_polynomial_transfer.h:1247-1360). */
int pmg_transfer_create_box(int kind, int precision, const pmg_box *box, int degree_coarse,
                            int degree_fine, double prolongation_scale, pmg_transfer *tr);
/* As pmg_operator_index_data, for pmg_transfer_index_array. */
int pmg_transfer_index_data(pmg_transfer tr, int which, void *host, int64_t *n);
int pmg_transfer_destroy(pmg_transfer tr);
/* kind, precision, degrees, and how the fine side is stored (layout 0:
 * explicit weights, 1: owner marks in the full fine map, 2: owner marks as
 * one word per fine x-line, used when every fine x-line is a run of
 * consecutive DoFs).  Any output may be NULL. */
int pmg_transfer_info(pmg_transfer tr, int *kind, int *precision, int *degree_coarse,
                      int *degree_fine, int *layout);
/* MGTransferBase::reinit(mf_coarse, mf_fine, constraints_coarse,
 * constraints_fine) (include/base/portable_mg_transfer_base.h:32-35): rebuild
 * the transfer in place for new maps (coarse constraints as invalid entries
 * of map_coarse) and new fine constraints (host list; weights built on the
 * device as pmg_transfer_create_from_constraints does; ignored for DG_CG),
 * keeping kind, precision, degrees and prolongation scale.  A V-cycle built
 * over the handle notices the rebuild at its next application and captures
 * its graph again; its level sizes must still match. */
int pmg_transfer_reinit(pmg_transfer tr, int64_t n_cells, const uint32_t *map_coarse,
                        const uint32_t *map_fine, const uint32_t *constrained_fine,
                        int64_t n_constrained_fine, int64_t n_dofs_coarse, int64_t n_dofs_fine);
/* dst_fine += P src_coarse (prolongate_and_add) */
int pmg_transfer_prolongate_and_add(pmg_transfer tr, void *dst_fine, const void *src_coarse,
                                    pmg_stream stream);
/* dst_coarse += P^T src_fine (restrict_and_add) */
int pmg_transfer_restrict_and_add(pmg_transfer tr, void *dst_coarse, const void *src_fine,
                                  pmg_stream stream);

/* ------------------------------------------------------------ smoother
 * Replaces deal.II PreconditionChebyshev<LaplaceOperatorBase, Vector> with the
 * operator's inverse diagonal as preconditioner, configured as at
 * tests/poisson_cube_bk3/program.cc:474-515.  degree <= 0 selects the automatic
 * degree for the relative accuracy smoothing_range < 1 (coarse level).
 * Eigenvalues are estimated at creation (Lanczos through eig_cg_n_iterations
 * CG steps on P^{-1}A); compute_diagonal() must have been called. Synchronous. */
int pmg_chebyshev_create(pmg_operator op, int degree, double smoothing_range,
                         int eig_cg_n_iterations, pmg_smoother *sm);
int pmg_chebyshev_destroy(pmg_smoother sm);
/* vmult: zero initial guess (dst overwritten) */
int pmg_chebyshev_vmult(pmg_smoother sm, void *dst, const void *src, pmg_stream stream);
/* step: dst is the initial guess */
int pmg_chebyshev_step(pmg_smoother sm, void *dst, const void *src, pmg_stream stream);
int pmg_chebyshev_info(pmg_smoother sm, double *max_eigenvalue, double *min_eigenvalue,
                       int *degree);

/* ------------------------------------------------------------ multigrid
 * Replaces VCycleMultigrid<dim,number,Transfer>::vmult
 * (include/multigrid/portable_v_cycle_multigrid.h:163-206) and
 * MultigridSolver::vmult (tests/common/portable_multigrid_solver.h:171-271).
 * Level l = 0 is the coarsest.  transfers[l] maps level l-1 -> l
 * (transfers[0] ignored), smoothers[l] for l >= 1.  All levels share one
 * precision (the V-cycle precision), except that the coarsest operator may be
 * FP64 under FP32 levels: the coarse-grid solve then runs in FP64 on its own
 * vectors, converted from and to FP32 around it (transfers[1] stays FP32).
 * outer_precision is the precision of the vectors passed to
 * pmg_multigrid_vmult (FP64 outer with FP32 levels is the mixed-precision
 * preconditioner; conversion happens at the V-cycle boundary).
 * coarse: PMG_COARSE_PCG (Jacobi-PCG to coarse_tol relative) or
 * PMG_COARSE_CHEBYSHEV (Chebyshev auto-degree to 1e-3, the reference). */
int pmg_multigrid_create(int n_levels, const pmg_operator *level_operators,
                         const pmg_transfer *transfers, const pmg_smoother *smoothers,
                         int coarse_kind, double coarse_tol, int outer_precision,
                         pmg_multigrid *mg);
int pmg_multigrid_destroy(pmg_multigrid mg);
/* The V-cycle body (everything between the conversion of src and of the
 * result) is captured into a HIP graph on the first call and replayed on the
 * caller's stream, unless the coarse solve needs host round trips (Jacobi-PCG
 * on a coarse level with more than 4096 cells) or PMG_VCYCLE_GRAPH=0. */
int pmg_multigrid_vmult(pmg_multigrid mg, void *dst, const void *src, pmg_stream stream);
/* *uses_graph = 1 if vmult replays a captured HIP graph (see above) */
int pmg_multigrid_uses_graph(pmg_multigrid mg, int *uses_graph);
/* Per-level wall times of the V-cycle, as MultigridSolver::print_wall_times
 * tabulates them (tests/common/portable_multigrid_solver.h:108-132,215-268):
 * n_applications eager V-cycles on src (outer precision; the defect is
 * staged once) with device events around every stage.
 * times[l * 6 + stage] = seconds per application at level l (0 = coarsest),
 * stage 0 pre-smoothing, 1 residual matvec, 2 restriction (the defect is
 * formed inside it), 3 coarse solve (level 0 only), 4 prolongation,
 * 5 post-smoothing.  A measurement call: it creates events.  Synchronous. */
int pmg_multigrid_level_times(pmg_multigrid mg, const void *src, int n_applications,
                              double *times, pmg_stream stream);

/* ------------------------------------------------------------ CG
 * Preconditioned CG of deal.II SolverCG as driven by MultigridSolver::solve_cg
 * (tests/common/portable_multigrid_solver.h:147-169): x is the initial guess
 * (device, op precision), stop when ||r|| <= rel_tol ||r_0|| or max_it.
 * precond: NONE, JACOBI (op's inverse diagonal) or MG (mg, outer precision
 * = op precision).  Returns PMG_ERR_NOT_CONVERGED if max_it was reached.
 * The CG work vectors (4 of length m()) are allocated on the first call and
 * kept with op until pmg_operator_destroy. */
int pmg_cg_solve(pmg_operator op, int precond, pmg_multigrid mg, const void *b, void *x,
                 double rel_tol, int max_it, int *iterations, double *res0, double *res,
                 pmg_stream stream);

/* The same CG without any allocation (no hipMalloc, hipFree, host pinning
 * or allocating library call inside the call).  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (tight
 * serving loops, harnesses that forbid hipMalloc).  r, z, p, q, the scalars
 * and the dot-product partial sums live in a fixed file-scope device array
 * of 64 MiB that every GPU the library runs on holds until the process
 * ends: m() may be at most 2 096 096 (FP64 operator) or 4 192 192 (FP32),
 * larger sizes return PMG_ERR_SIZE.  Every call on one GPU shares that
 * array: calls that overlap on one GPU (two host threads, or unordered
 * streams) race.  Preconditioners allocate nothing either, once set up: a
 * V-cycle captures its HIP graph at its first application, so apply it
 * once (pmg_multigrid_vmult or pmg_cg_solve) before calling this from code
 * that must not allocate.  Results equal pmg_cg_solve's. */
int pmg_cg_solve_noalloc(pmg_operator op, int precond, pmg_multigrid mg, const void *b, void *x,
                         double rel_tol, int max_it, int *iterations, double *res0, double *res,
                         pmg_stream stream);

/* ------------------------------------------------------------ vectors (device)
 * Small helpers for callers without a device-vector library. */
int pmg_vector_dot(int precision, int64_t n, const void *a, const void *b, double *result,
                   pmg_stream stream);
int pmg_vector_convert(int64_t n, int dst_precision, void *dst, int src_precision,
                       const void *src, pmg_stream stream);
/* Measurement only (no reference counterpart): the device's vector FMA rate
 * in flop/s for the given precision, from a kernel of independent FMA chains
 * in every lane (the FLOP peak the roofline fractions are quoted against is
 * checked with it).  Allocates a scratch array.  Synchronous. */
int pmg_measure_fma_rate(int precision, double *flops_per_second, pmg_stream stream);

#ifdef __cplusplus
}
#endif
#endif /* PMG_H */
