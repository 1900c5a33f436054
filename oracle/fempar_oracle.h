/*
 * fempar_oracle.h -- CPU restatement of FEMPAR's one-subdomain finite-element
 * hot path (structured hex mesh -> first-touch DOF numbering -> Galerkin
 * integration -> COO append -> sort/compress to CSR -> (Jacobi-)PCG).
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP product
 * path: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * may load it, and never as the thing measured or shipped.
 *
 * Every routine follows the reference algorithm in the reference's loop order
 * and cites the reference file:line it restates (paths relative to
 * Sources/Lib unless noted).  Arrays crossing this interface use the
 * reference's conventions: 1-based DOF ids, fixed (strong Dirichlet) DOFs as
 * negative ids, CSR irp/ja 1-based.
 */
#ifndef FEMPAR_ORACLE_H
#define FEMPAR_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- problem description (the structured Poisson and elasticity drivers) ---- */

enum { ORC_BC_ALL_BOUNDARY = 0,  /* set_id=1 on every boundary vef (test_poisson_driver.f90:193-204) */
       ORC_BC_X0_FACE      = 1,  /* set_id=1 on vefs lying in the plane x = x_min (clamped face)      */
       ORC_BC_NONE         = 2 };

enum { ORC_PHYS_POISSON = 0, ORC_PHYS_ELASTICITY = 1 };

enum { ORC_K_CONST = 0,        /* k = k0                                                      */
       ORC_K_SIN3 = 1,         /* k = 1 + 0.5 sin(2pi x) sin(2pi y) sin(2pi z)                */
       ORC_K_CELL = 2,         /* k = k_cell[cell] (piecewise constant field)                 */
       ORC_K_QPOINT = 3 };     /* k = k_qp[cell * nq + q] (values at the quadrature points)   */

enum { ORC_F_CONST = 0,        /* f = f0                                                      */
       ORC_F_MANUFACTURED = 1, /* f = -div(k grad u), u = prod sin(pi x_d), k of k_kind         */
       ORC_F_CELL_SIN = 2,     /* f = f0 * k_cell * d pi^2 prod sin(pi x_d) (cell-wise source)  */
       ORC_F_QPOINT = 3 };     /* f = f_qp[cell * nq + q]: source_term%get_value at each point
                                  (poisson_discrete_integration.f90:109-128)                    */

enum { ORC_G_ZERO = 0,         /* homogeneous Dirichlet                                        */
       ORC_G_LINEAR = 1 };     /* g = x + y (+ z)  (poisson_analytical_functions.f90:77-104)   */

typedef struct {
  int dim;           /* 2 or 3 */
  int n[3];          /* cells per direction */
  double lim[6];     /* x0 x1 y0 y1 z0 z1 */
  int order;         /* Lagrangian order: 1 or 2 */
  int ncomp;         /* 1 (scalar) or dim (vector, elasticity) */
  int bc_kind;       /* ORC_BC_* */
  const double *vcoords;  /* NULL: the generator's uniform vertices; else num_vertices x dim
                             coordinates (vertex (i,j,k) at i + (n0+1)(j + (n1+1) k)), and every
                             cell's map is computed from its own vertices
                             (sbm_fe_cell_iterator.i90:970-987, no_ressemblance) */
} orc_problem_t;

typedef struct {
  int kind;          /* ORC_PHYS_* */
  int k_kind;        /* ORC_K_* (Poisson) */
  double k0;
  const double *k_cell;   /* num_cells values when k_kind == ORC_K_CELL */
  int f_kind;        /* ORC_F_* (Poisson) */
  double f0;
  int g_kind;        /* ORC_G_* */
  double E, nu;      /* elasticity (lambda, mu as par_test_elasticity_constitutive_models.f90:34-37) */
  double body[3];    /* elasticity body force */
  const double *k_qp;     /* num_cells x nq values when k_kind == ORC_K_QPOINT */
  const double *f_qp;     /* num_cells x nq values when f_kind == ORC_F_QPOINT */
} orc_physics_t;

typedef struct orc_space orc_space_t;

/* mesh + reference FE + strong Dirichlet + free DOF numbering */
orc_space_t *orc_space_create(const orc_problem_t *prob);
void orc_space_free(orc_space_t *s);
void orc_space_sizes(const orc_space_t *s, int64_t *num_cells, int *nd, int *nq,
                     int *n_free, int *n_fixed);
/* lst_dofs_gids, cell-major (cells x nd), reference local DOF order, signed 1-based */
void orc_space_get_dofs(const orc_space_t *s, int32_t *out);
/* nodal interpolation of g at fixed DOFs (sbm_serial_fe_space.i90:997-1176) */
void orc_space_fixed_values(const orc_space_t *s, const orc_physics_t *phys, double *out);

/* reference tables: quadrature points (dim x nq, first index fastest), weights,
   shape values (nd_scalar x nq) and reference gradients (dim x nd_scalar x nq) */
void orc_reference_tables(const orc_space_t *s, double *qpts, double *qw, double *shape,
                          double *dshape);
/* 1-D Gauss-Legendre rule, ascending abscissae (sbm_quadrature.i90:767-896) */
void orc_gauss_legendre(int n, double *x, double *w);

/* One element of cell `cell` (0-based): elmat (nd x nd, column-major) and
   elvec (nd), before strong-Dirichlet elimination. */
/* physical coordinates of every quadrature point, num_cells x nq x dim
   (cell_map_compute_quadrature_points_coordinates, sbm_cell_map.i90:673-687) */
void orc_quadrature_points(const orc_space_t *s, double *xq);
void orc_element(const orc_space_t *s, const orc_physics_t *phys, int64_t cell,
                 double *elmat, double *elvec);

/* Integrate + assemble into COO and compress to CSR (fe_affine_operator%compute).
   Returns nnz; irp (n+1), ja (nnz), val (nnz), rhs (n) are malloc'ed and owned by
   the caller (free with orc_free).  Indices are 1-based like csr_sparse_matrix_t. */
int64_t orc_assemble(const orc_space_t *s, const orc_physics_t *phys, int symmetric_storage,
                     const double *fixed_values, int32_t **irp, int32_t **ja, double **val,
                     double **rhs);
/* Same, restricted to cells [cell_begin, cell_end) (partial assembly of a slab). */
int64_t orc_assemble_cells(const orc_space_t *s, const orc_physics_t *phys, int symmetric_storage,
                           const double *fixed_values, int64_t cell_begin, int64_t cell_end, int32_t **irp,
                           int32_t **ja, double **val, double **rhs);
void orc_free(void *p);

/* y = A x with the reference's matvec / matvec_symmetric_storage (csr_sparse_matrix.f90:545-623) */
void orc_spmv(int n, const int32_t *irp, const int32_t *ja, const double *val, int symmetric_storage,
              const double *x, double *y);
/* diagonal of a CSR matrix (csr_sparse_matrix.f90:4231-4256) */
void orc_extract_diagonal(int n, const int32_t *irp, const int32_t *ja, const double *val, double *d);

/* stopping criteria of iterative_solver_parameters.f90:37-46 */
enum { ORC_RES_NRMGIVEN_RHS_NRMGIVEN = 0, ORC_RES_NRMGIVEN_RES_NRMGIVEN = 1, ORC_DELTA_RHS = 2,
       ORC_DELTA_DELTA = 3, ORC_RES_RES = 4, ORC_RES_RHS = 5, ORC_DELTA_RHS_AND_RES_RES = 6,
       ORC_DELTA_RHS_AND_RES_RHS = 7, ORC_DELTA_DELTA_AND_RES_RES = 8, ORC_DELTA_DELTA_AND_RES_RHS = 9 };

typedef struct {
  int iterations;
  int converged;
  double error_estimate;   /* last error_estimate_convergence_test */
  double rhs_test;         /* rhs_convergence_test */
} orc_cg_result_t;

/* cg_solve_body (IterativeLinearSolver/cg.f90:101-238); precond 0 = identity,
   1 = Jacobi (jacobi_preconditioner.f90:356-412).  x holds x0 on entry. */
void orc_cg(int n, const int32_t *irp, const int32_t *ja, const double *val, int symmetric_storage,
            int precond, int criterion, double rtol, double atol, int maxit, const double *b,
            double *x, orc_cg_result_t *res, double *history /* maxit entries or NULL */);

/* exact solutions for the error norms */
enum { ORC_EXACT_SIN = 0,      /* u = prod_d sin(pi x_d)   (the manufactured solution)     */
       ORC_EXACT_LINEAR = 1 }; /* u = x + y (+ z)          (poisson_analytical_functions.f90:77-104) */

/* error_norms_scalar_t%compute (Integration/ErrorNorms/sbm_error_norms_scalar.i90:67-241)
   of the FE function with free values x_free (n_free) and fixed values
   fixed_values (n_fixed, NULL: 0) against `exact`, scalar spaces:
   out[0] = l2_norm, out[1] = h1_seminorm, out[2] = h1_norm, out[3] = linfty_norm
   (norm types of Integration/ErrorNorms/error_norms.f90:43-52). */
void orc_error_norms(const orc_space_t *s, const double *x_free, const double *fixed_values, int exact,
                     double out[4]);

/* ---- multi-field spaces, block layout and block assembly ----
 * serial_fe_space_t with several scalar Lagrangian fields on one structured
 * mesh, their block_layout_t (field -> block, 1-based) and the
 * block_sparse_assembler_t.  The integration plugin is the coupled
 * diffusion-reaction system sum_f kappa_f (grad u_f, grad v_f) +
 * sigma sum_{f<g} (u_f - u_g, v_f - v_g) + source f_f per field (entries
 * in orc_mf_element); homogeneous Dirichlet per field (bc_kind of each
 * field's problem). */
#define ORC_MF_MAXF 4
typedef struct orc_mf orc_mf_t;
typedef struct {
  double kappa[ORC_MF_MAXF];
  double sigma;
  double f[ORC_MF_MAXF];
} orc_mf_physics_t;
orc_mf_t *orc_mf_create(int nfields, const orc_problem_t *fields, const int *field_blocks);
void orc_mf_free(orc_mf_t *m);
int orc_mf_num_blocks(const orc_mf_t *m);
void orc_mf_field_sizes(const orc_mf_t *m, int field, int *nd, int *n_free, int *n_fixed);
int orc_mf_block_size(const orc_mf_t *m, int block);
/* field's lst_dofs_gids (cells x nd): free ids 1-based inside the field's block, fixed ids negative */
void orc_mf_get_dofs(const orc_mf_t *m, int field, int32_t *out);
void orc_mf_element(const orc_mf_t *m, const orc_mf_physics_t *ph, int64_t cell, double *elmat,
                    double *elvec);
int64_t orc_mf_assemble_block(const orc_mf_t *m, const orc_mf_physics_t *ph, const int *field_coupling, int ib,
                              int jb, int symmetric_storage, int32_t **irp, int32_t **ja, double **val);
void orc_mf_assemble_rhs(const orc_mf_t *m, const orc_mf_physics_t *ph, int ib, double *rhs);

/* Jacobi-PCG, RES_RES, full-storage CSR, loops over the host threads (OpenMP):
   the all-cores timing baseline only (not a restatement of the serial reference). */
void orc_cg_threads(int n, const int32_t *irp, const int32_t *ja, const double *val, double rtol, int maxit,
                    const double *b, double *x, orc_cg_result_t *res);

#ifdef __cplusplus
}
#endif
#endif
