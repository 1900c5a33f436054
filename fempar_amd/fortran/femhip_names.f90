! femhip_names -- ISO_C_BINDING interface to libfemhip (include/femhip.h).
!
! The binding a FEMPAR maintainer adds so that the unchanged Fortran API
! (discrete_integration_t%integrate_galerkin, fe_affine_operator_t,
! sparse_assembler_t, iterative_linear_solver_t) reaches the HIP kernels:
! derived types mirror the C structs field by field (bind(c)), every entry
! point is declared with the C signature, and 1-based signed DOF ids and
! 1-based CSR arrays cross unchanged.  Status codes follow femhip.h; the
! wrapper femhip_check turns a non-zero code into the reference's abort
! behaviour (check/mcheck -> runend, Sources/Include/debug.i90).
module femhip_names
  use iso_c_binding
  implicit none
  private

  integer(c_int32_t), parameter, public :: FEMHIP_OK = 0
  integer(c_int32_t), parameter, public :: FEMHIP_ERR_NOT_CONVERGED = 4
  integer(c_int32_t), parameter, public :: FEMHIP_BC_ALL_BOUNDARY = 0, FEMHIP_BC_X0_FACE = 1, FEMHIP_BC_NONE = 2
  integer(c_int32_t), parameter, public :: FEMHIP_PHYS_POISSON = 0, FEMHIP_PHYS_ELASTICITY = 1
  integer(c_int32_t), parameter, public :: FEMHIP_K_CONST = 0, FEMHIP_K_SIN3 = 1, FEMHIP_K_CELL = 2
  integer(c_int32_t), parameter, public :: FEMHIP_K_QPOINT = 3
  integer(c_int32_t), parameter, public :: FEMHIP_F_CONST = 0, FEMHIP_F_MANUFACTURED = 1, FEMHIP_F_CELL_SIN = 2
  integer(c_int32_t), parameter, public :: FEMHIP_F_QPOINT = 3
  integer(c_int32_t), parameter, public :: FEMHIP_G_ZERO = 0, FEMHIP_G_LINEAR = 1
  integer(c_int32_t), parameter, public :: FEMHIP_RES_RES = 4, FEMHIP_RES_RHS = 5
  integer(c_int32_t), parameter, public :: FEMHIP_PREC_IDENTITY = 0, FEMHIP_PREC_JACOBI = 1
  integer(c_int32_t), parameter, public :: FEMHIP_PRECISION_FP64 = 0, FEMHIP_PRECISION_FP32_REFINED = 1, &
                                            FEMHIP_PRECISION_FP32_STRICT = 2

  type, bind(c), public :: femhip_mesh_desc_t
    integer(c_int32_t) :: dim = 3
    integer(c_int32_t) :: n(3) = 10
    real(c_double)     :: limits(6) = [0.0_c_double, 1.0_c_double, 0.0_c_double, 1.0_c_double, &
                                       0.0_c_double, 1.0_c_double]
    integer(c_int32_t) :: order = 1
    integer(c_int32_t) :: num_components = 1
    integer(c_int32_t) :: dirichlet = 0
  end type

  type, bind(c), public :: femhip_physics_t
    integer(c_int32_t) :: kind = 0
    integer(c_int32_t) :: k_kind = 0
    real(c_double)     :: k0 = 1.0_c_double
    type(c_ptr)        :: k_cell = c_null_ptr
    integer(c_int32_t) :: f_kind = 0
    real(c_double)     :: f0 = 0.0_c_double
    integer(c_int32_t) :: g_kind = 0
    real(c_double)     :: E = 1.0_c_double, nu = 0.3_c_double
    real(c_double)     :: body(3) = [0.0_c_double, 0.0_c_double, -1.0_c_double]
    type(c_ptr)        :: k_qp = c_null_ptr     ! num_cells x nq values (FEMHIP_K_QPOINT)
    type(c_ptr)        :: f_qp = c_null_ptr     ! num_cells x nq values (FEMHIP_F_QPOINT)
  end type

  type, bind(c), public :: femhip_cg_params_t
    real(c_double)     :: rtol = 1.0e-6_c_double          ! default_rtol (iterative_solver_parameters.f90:167)
    real(c_double)     :: atol = 0.0_c_double
    integer(c_int32_t) :: max_num_iterations = 1000
    integer(c_int32_t) :: stopping_criterion = 4         ! RES_RES, default of CG
    integer(c_int32_t) :: preconditioner = 0
    integer(c_int32_t) :: precision = 0
    real(c_double)     :: inner_rtol = 0.0_c_double
    integer(c_int32_t) :: max_outer = 0
  end type

  type, bind(c), public :: femhip_cg_result_t
    integer(c_int32_t) :: num_iterations, did_converge, outer_iterations, status
    real(c_double)     :: error_estimate, rhs_convergence_test, true_residual, rhs_norm, solve_ms
  end type

  integer(c_int32_t), parameter, public :: FEMHIP_EXACT_SIN = 0, FEMHIP_EXACT_LINEAR = 1

  type, bind(c), public :: femhip_error_norms_t
    real(c_double) :: l2_norm, h1_seminorm, h1_norm, linfty_norm
  end type

  public :: femhip_ctx_create, femhip_ctx_destroy, femhip_last_error, femhip_check
  public :: femhip_space_create_structured, femhip_space_create_from_dofs, femhip_space_info
  public :: femhip_space_get_dofs, femhip_space_interpolate_dirichlet, femhip_space_destroy
  public :: femhip_space_set_vertex_coords, femhip_space_quadrature_points
  public :: femhip_operator_create, femhip_operator_info, femhip_operator_get_pattern
  public :: femhip_operator_set_physics, femhip_operator_compute, femhip_operator_get_values
  public :: femhip_operator_apply, femhip_operator_destroy
  public :: femhip_solver_create, femhip_solver_solve, femhip_solver_get_solution, femhip_solver_destroy
  public :: femhip_solver_error_norms

  interface
    integer(c_int) function femhip_ctx_create(device, ctx) bind(c, name="femhip_ctx_create")
      import :: c_int, c_int32_t, c_ptr
      integer(c_int32_t), value :: device
      type(c_ptr), intent(out)  :: ctx
    end function
    integer(c_int) function femhip_ctx_destroy(ctx) bind(c, name="femhip_ctx_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    type(c_ptr) function femhip_last_error_c() bind(c, name="femhip_last_error")
      import :: c_ptr
    end function
    integer(c_int) function femhip_space_create_structured(ctx, mesh, space) &
        bind(c, name="femhip_space_create_structured")
      import :: c_int, c_ptr, femhip_mesh_desc_t
      type(c_ptr), value :: ctx
      type(femhip_mesh_desc_t), intent(in) :: mesh
      type(c_ptr), intent(out) :: space
    end function
    integer(c_int) function femhip_space_create_from_dofs(ctx, mesh, lst_dofs_gids, n_free, n_fixed, space) &
        bind(c, name="femhip_space_create_from_dofs")
      import :: c_int, c_int32_t, c_ptr, femhip_mesh_desc_t
      type(c_ptr), value :: ctx
      type(femhip_mesh_desc_t), intent(in) :: mesh
      integer(c_int32_t), intent(in) :: lst_dofs_gids(*)
      integer(c_int32_t), value :: n_free, n_fixed
      type(c_ptr), intent(out) :: space
    end function
    integer(c_int) function femhip_space_info(space, num_cells, nd, nq, n_free, n_fixed) &
        bind(c, name="femhip_space_info")
      import :: c_int, c_int32_t, c_int64_t, c_ptr
      type(c_ptr), value :: space
      integer(c_int64_t), intent(out) :: num_cells
      integer(c_int32_t), intent(out) :: nd, nq, n_free, n_fixed
    end function
    integer(c_int) function femhip_space_get_dofs(space, lst_dofs_gids) bind(c, name="femhip_space_get_dofs")
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: space
      integer(c_int32_t), intent(out) :: lst_dofs_gids(*)
    end function
    integer(c_int) function femhip_space_interpolate_dirichlet(space, phys, fixed_values) &
        bind(c, name="femhip_space_interpolate_dirichlet")
      import :: c_int, c_ptr, c_double, femhip_physics_t
      type(c_ptr), value :: space
      type(femhip_physics_t), intent(in) :: phys
      real(c_double), intent(out) :: fixed_values(*)
    end function
    integer(c_int) function femhip_space_set_vertex_coords(space, coords) bind(c, name="femhip_space_set_vertex_coords")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: space
      real(c_double), intent(in) :: coords(*)
    end function
    integer(c_int) function femhip_space_quadrature_points(space, xq) bind(c, name="femhip_space_quadrature_points")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: space
      real(c_double), intent(out) :: xq(*)
    end function
    integer(c_int) function femhip_space_destroy(space) bind(c, name="femhip_space_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: space
    end function
    integer(c_int) function femhip_operator_create(ctx, space, symmetric_storage, op) &
        bind(c, name="femhip_operator_create")
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: ctx, space
      integer(c_int32_t), value :: symmetric_storage
      type(c_ptr), intent(out) :: op
    end function
    integer(c_int) function femhip_operator_info(op, n, nnz, nnz_full) bind(c, name="femhip_operator_info")
      import :: c_int, c_int32_t, c_int64_t, c_ptr
      type(c_ptr), value :: op
      integer(c_int32_t), intent(out) :: n
      integer(c_int64_t), intent(out) :: nnz, nnz_full
    end function
    integer(c_int) function femhip_operator_get_pattern(op, irp, ja) bind(c, name="femhip_operator_get_pattern")
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: op
      integer(c_int32_t), intent(out) :: irp(*), ja(*)
    end function
    integer(c_int) function femhip_operator_set_physics(op, phys) bind(c, name="femhip_operator_set_physics")
      import :: c_int, c_ptr, femhip_physics_t
      type(c_ptr), value :: op
      type(femhip_physics_t), intent(in) :: phys
    end function
    integer(c_int) function femhip_operator_compute(op, fixed_values, ms) bind(c, name="femhip_operator_compute")
      import :: c_int, c_ptr
      type(c_ptr), value :: op, fixed_values, ms
    end function
    integer(c_int) function femhip_operator_get_values(op, val, rhs) bind(c, name="femhip_operator_get_values")
      import :: c_int, c_ptr
      type(c_ptr), value :: op, val, rhs
    end function
    integer(c_int) function femhip_operator_apply(op, x, y) bind(c, name="femhip_operator_apply")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: op
      real(c_double), intent(in) :: x(*)
      real(c_double), intent(out) :: y(*)
    end function
    integer(c_int) function femhip_operator_destroy(op) bind(c, name="femhip_operator_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: op
    end function
    integer(c_int) function femhip_solver_create(ctx, op, params, solver) bind(c, name="femhip_solver_create")
      import :: c_int, c_ptr, femhip_cg_params_t
      type(c_ptr), value :: ctx, op
      type(femhip_cg_params_t), intent(in) :: params
      type(c_ptr), intent(out) :: solver
    end function
    integer(c_int) function femhip_solver_solve(solver, b, x, res) bind(c, name="femhip_solver_solve")
      import :: c_int, c_ptr, femhip_cg_result_t
      type(c_ptr), value :: solver, b, x
      type(femhip_cg_result_t), intent(out) :: res
    end function
    integer(c_int) function femhip_solver_get_solution(solver, x) bind(c, name="femhip_solver_get_solution")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: solver
      real(c_double), intent(out) :: x(*)
    end function
    integer(c_int) function femhip_solver_destroy(solver) bind(c, name="femhip_solver_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: solver
    end function
    integer(c_int) function femhip_solver_error_norms(solver, exact, norms) &
        bind(c, name="femhip_solver_error_norms")
      import :: c_int, c_ptr, c_int32_t, femhip_error_norms_t
      type(c_ptr), value :: solver
      integer(c_int32_t), value :: exact
      type(femhip_error_norms_t), intent(out) :: norms
    end function
  end interface

contains

  function femhip_last_error() result(msg)
    character(len=:), allocatable :: msg
    character(kind=c_char), pointer :: cs(:)
    type(c_ptr) :: p
    integer :: n
    p = femhip_last_error_c()
    n = 0
    call c_f_pointer(p, cs, [4096])
    do while (n < 4096)
      if (cs(n + 1) == c_null_char) exit
      n = n + 1
    end do
    allocate(character(len=n) :: msg)
    msg = transfer(cs(1:n), msg)
  end function

  ! mcheck-like wrapper: abort with the library's message on a non-zero status
  subroutine femhip_check(rc, where)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: where
    if (rc /= FEMHIP_OK) then
      write(0, '(a,a,a,i0,a,a)') 'femhip: ', where, ' failed (status ', rc, '): ', femhip_last_error()
      error stop 1
    end if
  end subroutine

end module femhip_names
