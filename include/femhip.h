/*
 * femhip.h -- C-ABI of the MI355X (gfx950) finite-element hot path.
 *
 * One GPU, one subdomain: structured hex mesh -> first-touch DOF numbering and
 * CSR pattern (on the device) -> cell-wise Galerkin integration -> deterministic
 * owner-computes assembly into CSR (each row summed over its cells in ascending
 * order and written once) -> (Jacobi-)preconditioned CG, in plain FP64 or
 * FP32-refined form.
 *
 * Conventions follow FEMPAR's Fortran side so the ISO_C_BINDING shim passes its
 * arrays unchanged: DOF ids are 1-based, strong-Dirichlet DOFs carry negative
 * ids, CSR irp/ja are 1-based, element matrices are column-major.  Host pointers
 * are borrowed for the duration of a call and never freed by the library;
 * device state is owned by the handles and released by the *_destroy calls.
 * Every entry point returns FEMHIP_OK or an error code; femhip_last_error()
 * gives the message of the last failure on the calling thread.  Handles are
 * reference counted along ctx <- space <- operator <- solver: destroying a
 * parent first only drops its reference, and it is released with the last
 * dependent.  One thread drives a context at a time.
 *
 * Reference interfaces each entry point replaces are cited as file:line
 * relative to FEMPAR's Sources/Lib.
 */
#ifndef FEMHIP_H
#define FEMHIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes (the reference aborts through check/mcheck -> runend) ---- */
#define FEMHIP_OK                 0
#define FEMHIP_ERR_INVALID        1   /* bad argument / unsupported configuration   */
#define FEMHIP_ERR_HIP            2   /* HIP runtime failure                         */
#define FEMHIP_ERR_OOM            3   /* device or host allocation failed            */
#define FEMHIP_ERR_NOT_CONVERGED  4   /* max iterations reached (cg.f90:214)          */
#define FEMHIP_ERR_ZERO_DIAGONAL  5   /* Jacobi on a zero pivot (serial_scalar_array.f90:585) */
#define FEMHIP_ERR_STATE          6   /* call out of order (e.g. solve before compute) */
#define FEMHIP_ERR_ADMISSION      7   /* request exceeds the HBM budget (serving)     */
#define FEMHIP_ERR_NO_DEVICE      8   /* no HIP device: the library has no CPU path   */

/* ---- enums ---- */
#define FEMHIP_BC_ALL_BOUNDARY    0   /* set_id = 1 on every boundary vef (test_poisson_driver.f90:193-204) */
#define FEMHIP_BC_X0_FACE         1   /* set_id = 1 on vefs in the plane x = x_min (clamped face) */
#define FEMHIP_BC_NONE            2

#define FEMHIP_PHYS_POISSON       0   /* poisson_discrete_integration.f90:62-141 (+ coefficient k) */
#define FEMHIP_PHYS_ELASTICITY    1   /* sbm_irreducible_discrete_integration.i90:46-134 */

#define FEMHIP_K_CONST            0   /* k = k0 */
#define FEMHIP_K_SIN3             1   /* k = 1 + 0.5 sin(2 pi x) sin(2 pi y) sin(2 pi z) */
#define FEMHIP_K_CELL             2   /* k = k_cell[cell], piecewise constant */
#define FEMHIP_K_QPOINT           3   /* k = k_qp[cell * nq + q] at each quadrature point */

#define FEMHIP_F_CONST            0   /* f = f0 */
#define FEMHIP_F_MANUFACTURED     1   /* f = -div(k grad u), u = prod_d sin(pi x_d) */
#define FEMHIP_F_CELL_SIN         2   /* f = f0 k_cell d pi^2 prod_d sin(pi x_d) */
#define FEMHIP_F_QPOINT           3   /* f = f_qp[cell * nq + q]: any source_term%get_value
                                         evaluated by the caller at the points of
                                         femhip_space_quadrature_points
                                         (poisson_discrete_integration.f90:109-128) */

#define FEMHIP_G_ZERO             0   /* homogeneous Dirichlet */
#define FEMHIP_G_LINEAR           1   /* g = x + y (+ z), poisson_analytical_functions.f90:77-104 */

/* stopping criteria, same meaning and order as iterative_solver_parameters.f90:37-46 */
#define FEMHIP_RES_NRMGIVEN_RHS_NRMGIVEN 0
#define FEMHIP_RES_NRMGIVEN_RES_NRMGIVEN 1
#define FEMHIP_DELTA_RHS                 2
#define FEMHIP_DELTA_DELTA               3
#define FEMHIP_RES_RES                   4
#define FEMHIP_RES_RHS                   5
#define FEMHIP_DELTA_RHS_AND_RES_RES     6
#define FEMHIP_DELTA_RHS_AND_RES_RHS     7
#define FEMHIP_DELTA_DELTA_AND_RES_RES   8
#define FEMHIP_DELTA_DELTA_AND_RES_RHS   9

#define FEMHIP_PREC_IDENTITY      0   /* .identity. operator (operator.f90:39-180) */
#define FEMHIP_PREC_JACOBI        1   /* jacobi_preconditioner.f90:136-450 */

#define FEMHIP_PRECISION_FP64        0   /* plain drop-in path */
#define FEMHIP_PRECISION_FP32_REFINED 1  /* FP32 inner PCG inside FP64 iterative refinement */
#define FEMHIP_PRECISION_FP32_STRICT  2  /* FP32_REFINED with the FP32 matrix integrated in FP32 on the
                                            matrix cores (scalar Q1, symmetric-diagonal layout; rounded
                                            from FP64 elsewhere) and FP32 row sums in the inner products */

/* ---- opaque handles ---- */
typedef struct femhip_ctx femhip_ctx_t;
typedef struct femhip_space femhip_space_t;
typedef struct femhip_operator femhip_operator_t;
typedef struct femhip_solver femhip_solver_t;

/* ---- plain-data descriptors ---- */
typedef struct {
  int32_t dim;              /* 2 or 3 (STRUCT_HEX_MESH_GENERATOR_NUM_DIMS) */
  int32_t n[3];             /* cells per direction (..._NUM_CELLS_X_DIM) */
  double limits[6];         /* x0 x1 y0 y1 z0 z1 (..._DOMAIN_LIMITS) */
  int32_t order;            /* Lagrangian order 1 or 2 (FES_REF_FE_ORDERS) */
  int32_t num_components;   /* 1 scalar, dim for vector (elasticity) */
  int32_t dirichlet;        /* FEMHIP_BC_* */
} femhip_mesh_desc_t;

typedef struct {
  int32_t kind;             /* FEMHIP_PHYS_* */
  int32_t k_kind;           /* FEMHIP_K_* */
  double k0;
  const double *k_cell;     /* host, num_cells values, read during set_physics */
  int32_t f_kind;           /* FEMHIP_F_* */
  double f0;
  int32_t g_kind;           /* FEMHIP_G_* (used by femhip_space_interpolate_dirichlet) */
  double E, nu;             /* elasticity: lambda = nu E/((1+nu)(1-2nu)), mu = E/(2(1+nu)) */
  double body[3];           /* elasticity body force */
  const double *k_qp;       /* host, num_cells x nq values for FEMHIP_K_QPOINT, read during set_physics */
  const double *f_qp;       /* host, num_cells x nq values for FEMHIP_F_QPOINT, read during set_physics */
} femhip_physics_t;

typedef struct {
  double rtol;              /* ILS_RTOL (default 1e-6, iterative_solver_parameters.f90:167) */
  double atol;              /* ILS_ATOL (default 0) */
  int32_t max_num_iterations; /* ILS_MAX_NUM_ITERATIONS (default 1000) */
  int32_t stopping_criterion; /* FEMHIP_RES_RES etc. (CG default RES_RES) */
  int32_t preconditioner;   /* FEMHIP_PREC_* */
  int32_t precision;        /* FEMHIP_PRECISION_* */
  double inner_rtol;        /* FP32_REFINED: relative reduction per inner solve (0 = default) */
  int32_t max_outer;        /* FP32_REFINED: max refinement steps (0 = default) */
} femhip_cg_params_t;

typedef struct {
  int32_t num_iterations;   /* CG iterations (inner iterations summed for FP32_REFINED) */
  int32_t did_converge;
  int32_t outer_iterations; /* refinement steps (0 for FP64) */
  int32_t status;           /* FEMHIP_OK or FEMHIP_ERR_NOT_CONVERGED */
  double error_estimate;    /* last error_estimate_convergence_test */
  double rhs_convergence_test;
  double true_residual;     /* ||b - A x||_2 in FP64 after the solve */
  double rhs_norm;          /* ||b||_2 */
  double solve_ms;          /* device time of the solve (hipEvents) */
} femhip_cg_result_t;

/* ---- context ---- */
/* Opens HIP device `device`, creates the context's stream. */
int femhip_ctx_create(int32_t device, femhip_ctx_t **ctx);
int femhip_ctx_destroy(femhip_ctx_t *ctx);
/* Message of the last failing call on this thread ("" if none). */
const char *femhip_last_error(void);

/* ---- FE space: serial_fe_space_t on a struct-hex triangulation ----
 * Replaces serial_triangulation_t%create with the structured generator
 * (sbm_serial_triangulation.i90:61-118, uniform_hex_mesh_generator.f90:145-647)
 * + serial_fe_space_t%create: strong Dirichlet numbering
 * (sbm_serial_fe_space.i90:525-715) and first-touch free numbering
 * (sbm_serial_fe_space.i90:2623-2922), computed on the device (the cell DOF
 * lists stay in HBM; femhip_space_get_dofs copies them out). */
int femhip_space_create_structured(femhip_ctx_t *ctx, const femhip_mesh_desc_t *mesh,
                                   femhip_space_t **space);
/* Drop-in form: the caller (FEMPAR's fe_space) already numbered the DOFs and
 * passes lst_dofs_gids (num_cells x nd, cell-major, reference local order,
 * fe_space.f90:636-637).  Every free id must sit on one vef of the mesh at one
 * position of its own-DOF list (FEMHIP_ERR_INVALID otherwise). */
int femhip_space_create_from_dofs(femhip_ctx_t *ctx, const femhip_mesh_desc_t *mesh,
                                  const int32_t *lst_dofs_gids, int32_t n_free, int32_t n_fixed,
                                  femhip_space_t **space);
int femhip_space_info(const femhip_space_t *space, int64_t *num_cells, int32_t *nd,
                      int32_t *nq, int32_t *n_free, int32_t *n_fixed);
int femhip_space_get_dofs(const femhip_space_t *space, int32_t *lst_dofs_gids);
/* Vertex coordinates of the triangulation (serial_triangulation_t's vertex
 * coordinates, read by fe_cell_iterator_t%get_nodes_coordinates,
 * sbm_fe_cell_iterator.i90:966-967): num_vertices x dim doubles, vertex
 * (i, j, k) at index i + (n0 + 1) (j + (n1 + 1) k), x fastest.  The topology
 * (cells, vefs, DOF numbering, CSR pattern) stays that of the structured mesh;
 * every cell then has its own trilinear map, recomputed per cell as the
 * reference does when a cell is not a translate of the one before
 * (cell_ressemblance = no_ressemblance, sbm_fe_cell_iterator.i90:970-987 with
 * sbm_cell_map.i90:245-369).  Set before femhip_operator_compute; operators
 * computed on the space afterwards integrate every cell with its own Jacobian.
 * Scalar Poisson (Q1/Q2, 2-D/3-D) and vector Q1 elasticity. */
int femhip_space_set_vertex_coords(femhip_space_t *space, const double *coords);
/* fe_cell_iterator_t%get_quadrature_points_coordinates (sbm_cell_map.i90:673-687)
 * for every cell: num_cells x nq x dim doubles (cell-major, then quadrature
 * point in the reference order, first index fastest), so a caller can evaluate
 * its own coefficient or source function there and pass FEMHIP_K_QPOINT /
 * FEMHIP_F_QPOINT values. */
int femhip_space_quadrature_points(const femhip_space_t *space, double *xq);
/* serial_fe_space_t%interpolate_dirichlet_values (sbm_serial_fe_space.i90:997-1176):
 * nodal interpolation of g (phys->g_kind) into fixed_values (n_fixed). */
int femhip_space_interpolate_dirichlet(const femhip_space_t *space, const femhip_physics_t *phys,
                                       double *fixed_values);
int femhip_space_destroy(femhip_space_t *space);

/* ---- operator: fe_affine_operator_t + sparse_assembler_t + CSR sparse_matrix_t ----
 * femhip_operator_create replaces fe_affine_operator_t%create with
 * sparse_matrix_storage 'CSR' (sbm_fe_operator.i90:28-71, 162-308): it builds,
 * on the device, the CSR pattern that sort_and_compress would produce
 * (base_sparse_matrix.f90:4121-4829, csr_sparse_matrix.f90:416-479).
 * symmetric_storage = diagonal_blocks_symmetric_storage (upper triangle kept). */
int femhip_operator_create(femhip_ctx_t *ctx, femhip_space_t *space, int32_t symmetric_storage,
                           femhip_operator_t **op);
/* n = num_free_dofs; nnz = stored entries (upper if symmetric_storage); nnz_full = solver copy */
int femhip_operator_info(const femhip_operator_t *op, int32_t *n, int64_t *nnz, int64_t *nnz_full);
/* csr_sparse_matrix_t%get_irp/get_ja (csr_sparse_matrix.f90:192-225), 1-based */
int femhip_operator_get_pattern(const femhip_operator_t *op, int32_t *irp, int32_t *ja);
/* The discrete_integration_t plugin (discrete_integration.f90:44-73): which
 * integrate_galerkin body runs and with which coefficients. */
int femhip_operator_set_physics(femhip_operator_t *op, const femhip_physics_t *phys);
/* fe_affine_operator_t%compute (sbm_fe_affine_operator.i90:15-34): integrate
 * all cells, eliminate strong Dirichlet columns into the RHS with the given
 * fixed values (n_fixed, may be NULL for homogeneous), scatter-add into the
 * device CSR values and RHS.  Device-resident; elapsed device ms in *ms (may be NULL). */
int femhip_operator_compute(femhip_operator_t *op, const double *fixed_values, double *ms);
/* get_matrix()%val and get_translation() (sbm_fe_operator.i90:499-513) copied to host */
int femhip_operator_get_values(const femhip_operator_t *op, double *val, double *rhs);
/* sparse_matrix_t%apply (sparse_matrix.f90:1264-1276): y = A x, host vectors */
int femhip_operator_apply(femhip_operator_t *op, const double *x, double *y);
/* No counterpart in the reference: femhip_operator_apply for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  Same result bit for bit; x and y
 * are staged in a fixed 32 MiB device array per GPU instead of per-call
 * allocations, so the operator may have at most 2,097,152 rows
 * (FEMHIP_ERR_INVALID beyond).  Calls on one GPU share that array: they are
 * serialised within the process, and calls from two processes on one GPU
 * never share it (each process has its own copy). */
int femhip_operator_apply_noalloc(femhip_operator_t *op, const double *x, double *y);
int femhip_operator_destroy(femhip_operator_t *op);

/* ---- iterative_linear_solver_t of type 'CG' ----
 * Replaces iterative_linear_solver_t%create/set_type_from_string('CG')/
 * set_parameters_from_pl/set_operators (iterative_linear_solver.f90:62-241,
 * base_iterative_linear_solver.f90:263-296, 714-774). */
int femhip_solver_create(femhip_ctx_t *ctx, femhip_operator_t *op, const femhip_cg_params_t *params,
                         femhip_solver_t **solver);
/* iterative_linear_solver_t%solve(b, x) -> cg_solve_body (cg.f90:101-238).
 * b == NULL uses the operator's assembled RHS (device-resident); x is the host
 * initial guess on entry and the solution on exit (x == NULL: start from 0 and
 * keep the solution on the device). */
int femhip_solver_solve(femhip_solver_t *solver, const double *b, double *x, femhip_cg_result_t *res);
/* Measurement aid (the reference's timer_t, Sources/Lib/timer.f90:42-207, has
 * no kernel-level counterpart): launches the solver's SpMV kernel `reps` times
 * on the context stream and reports the average launch duration from HIP
 * events and the kernel's algorithmic HBM bytes per launch (matrix stream +
 * row pointers + one read of p and one write of q). */
int femhip_solver_profile_spmv(femhip_solver_t *solver, int32_t reps, double *ms_per_launch,
                               double *bytes_per_launch);
/* Measurement aid: one device-resident solve (operator RHS, x0 = 0) launched
 * kernel by kernel, with HIP events around every SpMV on the solver's stream;
 * mean SpMV duration over the iterations that ran (inside the CG sequence,
 * with the vector kernels in between), the iteration count, and the SpMV's
 * algorithmic bytes per launch. */
int femhip_solver_profile_cg(femhip_solver_t *solver, double *spmv_ms, int32_t *iterations,
                             double *bytes_per_launch);
/* copy the last solution to the host (n values) */
int femhip_solver_get_solution(const femhip_solver_t *solver, double *x);

/* error_estimate_history_convergence_test of the last FP64 solve, the
 * quantity the stopping criterion compared per iteration (cg.f90:388-389,
 * base_iterative_linear_solver.f90:520-525): the first min(count, iterations)
 * entries into hist (NULL allowed with count 0), the iteration count in
 * *num_iterations.  FEMHIP_ERR_STATE for an FP32_REFINED solver. */
int femhip_solver_get_convergence_history(const femhip_solver_t *solver, int32_t count, double *hist,
                                          int32_t *num_iterations);

/* Error of the last solution (fixed DOFs: the values given to
 * femhip_operator_compute, 0 if none) against an exact solution, scalar
 * spaces; replaces error_norms_scalar_t%compute
 * (Integration/ErrorNorms/sbm_error_norms_scalar.i90:67-241) for the norm
 * types l2_norm, h1_seminorm, h1_norm and linfty_norm
 * (Integration/ErrorNorms/error_norms.f90:43-52). */
enum { FEMHIP_EXACT_SIN = 0,      /* u = prod_d sin(pi x_d), the manufactured solution */
       FEMHIP_EXACT_LINEAR = 1 }; /* u = x + y (+ z) (poisson_analytical_functions.f90:77-104) */
typedef struct {
  double l2_norm, h1_seminorm, h1_norm, linfty_norm;
} femhip_error_norms_t;
int femhip_solver_error_norms(femhip_solver_t *solver, int32_t exact, femhip_error_norms_t *norms);
int femhip_solver_destroy(femhip_solver_t *solver);

/* ---- multi-field spaces: block layout, block assembler, block CG ----
 * A serial_fe_space_t with several fields (sbm_serial_fe_space.i90:40-140)
 * given as one femhip space per field on the same mesh (each numbered as if
 * alone), its block_layout_t (Integration/FEOperator/block_layout.f90:66-117:
 * field -> block, field coupling) and the block_sparse_assembler_t
 * (Integration/FEOperator/block_sparse_assembler.f90:47-224) filling a
 * block_sparse_matrix_t and a serial_block_array_t.  Free ids of a field are
 * its own ids shifted by the free DOFs of the earlier fields of its block
 * (sbm_serial_fe_space.i90:2676-2682); fixed ids follow those of all earlier
 * fields (:584-640).  Scalar fields, homogeneous Dirichlet per field. */
#define FEMHIP_MAX_FIELDS 4
typedef struct femhip_block_operator femhip_block_operator_t;
typedef struct {
  int32_t num_fields;
  femhip_space_t *fields[FEMHIP_MAX_FIELDS];   /* scalar spaces on one mesh (retained) */
  int32_t field_blocks[FEMHIP_MAX_FIELDS];     /* 1-based block of each field (FES_FIELD_BLOCKS) */
  int32_t field_coupling[FEMHIP_MAX_FIELDS * FEMHIP_MAX_FIELDS]; /* [f * num_fields + g] != 0: rows of
                                                   field f take columns of field g; [f][f] must be set */
} femhip_block_layout_t;
/* The coupled diffusion-reaction plugin: sum_f kappa_f (grad u_f, grad v_f) +
 * sigma sum_{f<g} (u_f - u_g, v_f - v_g), source f_f per field; every field
 * integrated with the quadrature of the highest order (sbm_serial_fe_space.i90:1891-2006). */
typedef struct {
  double kappa[FEMHIP_MAX_FIELDS];
  double sigma;
  double f[FEMHIP_MAX_FIELDS];
} femhip_multifield_physics_t;
/* fe_affine_operator_t%create with a block layout: the pattern of every
 * coupled block, built on the device; diagonal_blocks_symmetric_storage is the
 * view the host gets of the diagonal blocks (upper triangle). */
int femhip_block_operator_create(femhip_ctx_t *ctx, const femhip_block_layout_t *layout,
                                 int32_t diagonal_blocks_symmetric_storage, femhip_block_operator_t **bop);
/* number of blocks, rows per block (num_blocks values), stored entries per
 * block pair (num_blocks^2, row-major; 0 for an uncoupled pair) */
int femhip_block_operator_info(const femhip_block_operator_t *bop, int32_t *num_blocks, int32_t *block_sizes,
                               int64_t *nnz);
/* lst_dofs_gids of field `field` (1-based) in the block numbering, cells x nd */
int femhip_block_operator_get_dofs(const femhip_block_operator_t *bop, int32_t field, int32_t *lst_dofs_gids);
/* block_sparse_matrix_t%get_block(ib, jb) -> get_irp/get_ja (1-based) */
int femhip_block_operator_get_block_pattern(const femhip_block_operator_t *bop, int32_t ib, int32_t jb,
                                            int32_t *irp, int32_t *ja);
int femhip_block_operator_set_physics(femhip_block_operator_t *bop, const femhip_multifield_physics_t *phys);
/* fe_affine_operator_t%compute: integrate every cell, assemble every coupled block and the block RHS */
int femhip_block_operator_compute(femhip_block_operator_t *bop, double *ms);
int femhip_block_operator_get_block_values(const femhip_block_operator_t *bop, int32_t ib, int32_t jb, double *val);
/* serial_block_array_t%get_block(ib) of the translation */
int femhip_block_operator_get_rhs(const femhip_block_operator_t *bop, int32_t ib, double *rhs);
/* block_sparse_matrix_t%apply: y = A x, vectors of all blocks' rows in block order (host) */
int femhip_block_operator_apply(femhip_block_operator_t *bop, const double *x, double *y);
/* cg_solve_body (cg.f90:101-238) on the block operator (FP64, RES_RES or
 * RES_RHS, identity or Jacobi on the diagonal blocks); b == NULL: the
 * assembled RHS; x: x0 in, solution out (host, block order) */
int femhip_block_operator_solve(femhip_block_operator_t *bop, const femhip_cg_params_t *params, const double *b,
                                double *x, femhip_cg_result_t *res);
int femhip_block_operator_destroy(femhip_block_operator_t *bop);

/* ---- batched form: many independent small problems in one persistent launch ----
 * Each problem is FEMPAR's test_poisson setting on its own mesh: Q1 on the unit
 * square/cube with n cells per direction, strong Dirichlet (g = 0) on the whole
 * boundary, constant coefficient k, source f = f0 k d pi^2 prod_i sin(pi x_i)
 * (so u = prod_i sin(pi x_i)).  The integration plugin, numbering, assembly and
 * CG semantics are those of the single-problem path (cited above); solves use
 * the FP32-refined CG with reliable updates, RES_RES criterion. */
typedef struct {
  int32_t dim;              /* 2 or 3 */
  int32_t n;                /* cells per direction (>= 2) */
  double k;                 /* constant coefficient */
  double f0;                /* source scale */
} femhip_batch_problem_t;

typedef struct {
  int32_t num_iterations;   /* FP32 CG iterations */
  int32_t updates;          /* reliable updates (FP64 residual replacements) */
  int32_t did_converge;
  int32_t num_dofs;         /* free DOFs of the problem */
  double true_residual;     /* ||b - A x||_2 in FP64 at exit */
  double rhs_norm;          /* ||b||_2 */
  double latency_us;        /* completion time after the solve launch started (device clock) */
} femhip_batch_result_t;

typedef struct femhip_batch femhip_batch_t;

/* Sizes the device pools for `count` problems and builds the work order
 * (largest first). params: rtol, max_num_iterations, inner_rtol = reliable-update
 * factor (0: 0.1); precision must be FEMHIP_PRECISION_FP32_REFINED. */
int femhip_batch_create(femhip_ctx_t *ctx, int32_t count, const femhip_batch_problem_t *problems,
                        const femhip_cg_params_t *params, femhip_batch_t **batch);
/* Integrate + assemble every problem (one launch; FP64 and FP32 values, RHS, Jacobi). */
int femhip_batch_assemble(femhip_batch_t *batch, double *ms);
/* Solve every problem from x0 = 0 in one persistent, hipGraph-captured launch;
 * results (count entries, may be NULL) and elapsed device ms. */
int femhip_batch_solve(femhip_batch_t *batch, femhip_batch_result_t *results, double *ms);
/* Solution of problem i (num_dofs values, lexicographic free-node order = the
 * first-touch numbering of this mesh) and its assembled matrix as 1-based CSR
 * (full storage) with the RHS, for checking. */
int femhip_batch_get_solution(const femhip_batch_t *batch, int32_t i, double *x);
int femhip_batch_get_problem_matrix(const femhip_batch_t *batch, int32_t i, int32_t *irp, int32_t *ja,
                                    double *val, double *rhs, int64_t *nnz);
/* Autotuning: time the persistent solve on a sample of the batch for each
 * block shape (64, 128, 256, 512 or 1024 threads: 1 to 16 wavefronts per
 * problem, 16 wavefronts per CU) and keep the fastest (also used by
 * femhip_batch_solve when FEMHIP_AUTOTUNE is set).  Chosen threads-per-block
 * in *threads. */
int femhip_batch_autotune(femhip_batch_t *batch, int32_t *threads);
int femhip_batch_destroy(femhip_batch_t *batch);

/* ---- serving form: a request loop with continuous batching ----
 * A request is one linear elasticity solve (the irreducible integration of
 * Tests/Par/par_test_pb_bddc_linear_elasticity/modules/
 * sbm_irreducible_discrete_integration.i90:64-130 with lambda, mu from E, nu as
 * in par_test_elasticity_constitutive_models.f90:34-37) on the unit cube with n
 * hexahedral Q1 cells per direction, vector Q1 (3 DOFs per node), all
 * components clamped (g = 0) on the x = 0 face, constant body force, solved
 * by the FP32-refined Jacobi-PCG of femhip_solver_solve.  Since every cell
 * has the same element matrix, the operator is built on the device as a
 * 27-point block stencil in the solver's sliced format (no DOF lists, no CSR
 * pattern, no host pass over the matrix).
 *
 * The server owns `max_active` slots, each with its own HIP stream.  A request
 * is admitted when its arrival time has come, a slot is free and its device
 * memory estimate fits in what is left of `hbm_budget_bytes`, in the order of
 * the admission policy (first come, first served, or the smallest estimate
 * first with an aging bound).  Admitted requests are set up by worker threads
 * (operator build and assembly on the slot's stream) and then iterate on the
 * device in chunks of CG iterations; a scheduler thread launches and
 * interleaves the chunks of every running request, so requests that arrive
 * while others are on the GPU join as slots and memory free up, and leave as
 * they converge.
 * Mesh, matrix and Krylov state stay on the device until the request ends;
 * its device memory is then released and the solution kept on the host until
 * femhip_serve_release.  A request's arithmetic does not depend on the other
 * requests in flight: its result is the same alone and in a mix. */
#define FEMHIP_ADMIT_FCFS         0   /* arrived requests in arrival order */
#define FEMHIP_ADMIT_SMALLEST     1   /* smallest estimate first; a request waiting aging_s goes first */

typedef struct {
  double hbm_budget_bytes;  /* device memory admitted requests may hold together */
  int32_t max_active;       /* slots (requests set up or iterating at once), >= 1 */
  int32_t setup_threads;    /* host threads that set up and retire requests, >= 1 */
  femhip_cg_params_t cg;    /* rtol, max_num_iterations, inner_rtol; FP32_REFINED, RES_RES/RES_RHS */
  int32_t admission;        /* FEMHIP_ADMIT_* */
  double aging_s;           /* FEMHIP_ADMIT_SMALLEST: wait after which arrival order wins (0: never) */
} femhip_serve_params_t;

typedef struct {
  int32_t n;                /* cells per direction of the unit cube, 1..206 */
  double E;                 /* Young's modulus (> 0) */
  double nu;                /* Poisson ratio (-1, 0.5) */
  double body[3];           /* body force */
  double arrival_s;         /* arrival time in s since femhip_serve_create; < 0: now */
} femhip_serve_request_t;

typedef struct {
  int32_t status;           /* FEMHIP_OK, FEMHIP_ERR_NOT_CONVERGED or the setup error */
  int32_t num_iterations;   /* FP32 CG iterations */
  int32_t outer_iterations; /* reliable updates */
  int32_t did_converge;
  int32_t num_dofs;         /* free DOFs */
  int32_t pad;
  double true_residual;     /* ||b - A x||_2 in FP64 at exit */
  double rhs_norm;          /* ||b||_2 */
  double queue_ms;          /* arrival -> admitted */
  double setup_ms;          /* admitted -> first CG chunk queued (numbering, pattern, assembly) */
  double solve_ms;          /* first chunk -> solution on the host */
  double latency_ms;        /* arrival -> solution on the host */
  double hbm_bytes;         /* device memory the request held */
  double hbm_estimate;      /* the admission estimate */
} femhip_serve_result_t;

typedef struct {
  int64_t submitted, completed, failed;
  int32_t queued, active, peak_active, running_peak;  /* running_peak: most requests iterating at once */
  double hbm_in_use, hbm_peak;   /* admitted estimates (the budget's currency) */
  double arena_bytes;            /* device memory reserved at creation for requests (0: none) */
  int64_t arena_fallbacks;       /* request buffers that did not fit their block (allocated apart) */
} femhip_serve_stats_t;

typedef struct femhip_server femhip_server_t;

int femhip_serve_create(femhip_ctx_t *ctx, const femhip_serve_params_t *params, femhip_server_t **server);
/* Device memory the server would reserve for a request of n cells per
 * direction (what admission compares with the budget). */
double femhip_serve_estimate_bytes(int32_t n, int32_t max_num_iterations);
/* Queue a request; FEMHIP_ERR_ADMISSION if its estimate exceeds the whole
 * budget (it could never be admitted), FEMHIP_ERR_INVALID for bad fields. */
int femhip_serve_submit(femhip_server_t *server, const femhip_serve_request_t *request, int64_t *ticket);
/* Wait up to timeout_s (< 0: forever) for a request; FEMHIP_OK with the
 * result, or FEMHIP_ERR_STATE if it has not finished in time. */
int femhip_serve_wait(femhip_server_t *server, int64_t ticket, double timeout_s, femhip_serve_result_t *result);
/* Solution of a finished request: num_dofs values, entry 3 v + c is
 * component c of free vertex v, the free vertices (x index 1..n) numbered
 * x-fastest, then y, then z. */
int femhip_serve_get_solution(femhip_server_t *server, int64_t ticket, double *x);
/* Forget a finished request (its host solution). */
int femhip_serve_release(femhip_server_t *server, int64_t ticket);
int femhip_serve_stats(femhip_server_t *server, femhip_serve_stats_t *stats);
/* Seconds since femhip_serve_create on the server's clock (arrival times). */
double femhip_serve_now(femhip_server_t *server);
/* Stops admitting, lets admitted requests finish, drops queued ones. */
int femhip_serve_destroy(femhip_server_t *server);

#ifdef __cplusplus
}
#endif
#endif
