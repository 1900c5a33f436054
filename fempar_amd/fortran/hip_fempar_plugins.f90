! hip_fempar_plugins -- the HIP back end behind FEMPAR's plug-in seams: the
! subtypes a FEMPAR build adds so that unchanged drivers reach libfemhip.
!   * hip_sparse_assembler_t extends assembler_t: allocate builds the device
!     CSR pattern (femhip_operator_create, the pattern sort_and_compress would
!     give); compress_storage_matrix has nothing left to do; the element-wise
!     assembly_array / assembly_matrix are not used because the integration
!     plug-in assembles every cell on the device at once.
!   * hip_poisson_discrete_integration_t extends discrete_integration_t:
!     integrate_galerkin is the Poisson plug-in of test_poisson
!     (Sources/Tests/Serial/test_poisson/modules/poisson_discrete_integration.f90:62-141)
!     run for all cells by femhip_operator_compute, with the strong Dirichlet
!     values of the FE space lifted into the RHS.
!   * hip_cg_t extends base_iterative_linear_solver_t: solve_body is the CG
!     of cg.f90:101-238 on the device (femhip_solver_solve); create_hip_cg is
!     registered under 'CG' in the creational-methods dictionary.
! hip_fempar_plugins_init installs both (the dictionary entry and the serial
! assembler creator), as FEMPAR's own initialisation registers its solvers.
module hip_fempar_plugins
  use iso_c_binding
  use femhip_names
  use fempar_seams
  implicit none
  private

  type, extends(assembler_t), public :: hip_sparse_assembler_t
    type(c_ptr) :: op = c_null_ptr
    integer(ip) :: n = 0
    integer(c_int64_t) :: nnz = 0
  contains
    procedure :: assembly_array => hip_assembly_array
    procedure :: assembly_matrix => hip_assembly_matrix
    procedure :: compress_storage_matrix => hip_compress_storage_matrix
    procedure :: allocate => hip_assembler_allocate
    procedure :: free => hip_assembler_free
    procedure :: get_translation => hip_assembler_get_translation
    procedure :: get_csr => hip_assembler_get_csr
  end type

  type, extends(discrete_integration_t), public :: hip_poisson_discrete_integration_t
    type(femhip_physics_t) :: phys
    ! when associated: the plug-in's source term, evaluated at every cell's
    ! quadrature points as poisson_discrete_integration.f90:109-128 does
    class(scalar_function_t), pointer :: source_term => null()
  contains
    procedure :: integrate_galerkin => hip_poisson_integrate_galerkin
  end type

  type, extends(base_iterative_linear_solver_t), public :: hip_cg_t
    type(c_ptr) :: solver = c_null_ptr
    type(femhip_cg_params_t) :: prm
    type(c_ptr) :: op_of_solver = c_null_ptr
  contains
    procedure :: set_parameters_from_pl => hip_cg_set_parameters_from_pl
    procedure :: solve_body => hip_cg_solve_body
    procedure :: free => hip_cg_free
  end type

  public :: hip_fempar_plugins_init, create_hip_cg, create_hip_serial_assembler

contains

  subroutine hip_fempar_plugins_init()
    call the_iterative_linear_solver_creational_methods_dictionary%set('CG', create_hip_cg)
    call set_serial_assembler_creator(create_hip_serial_assembler)
  end subroutine

  subroutine create_hip_serial_assembler(sparse_matrix_storage_format, assembler)
    character(*), intent(in) :: sparse_matrix_storage_format
    class(assembler_t), allocatable, intent(inout) :: assembler
    if (sparse_matrix_storage_format /= 'CSR') then
      write(*, '(3a)') 'create_hip_serial_assembler: storage format ''', sparse_matrix_storage_format, &
          ''' is not available (CSR only)'
      error stop 1
    end if
    allocate(hip_sparse_assembler_t :: assembler)
  end subroutine

  subroutine hip_assembly_array(this, num_fields, field_blocks, num_dofs, fe_dofs, elvec)
    class(hip_sparse_assembler_t), intent(inout) :: this
    integer(ip), intent(in) :: num_fields
    integer(ip), intent(in) :: field_blocks(num_fields)
    integer(ip), intent(in) :: num_dofs(num_fields)
    type(i1p_t), intent(in) :: fe_dofs(num_fields)
    real(rp), intent(in) :: elvec(:)
    write(*, '(a)') 'hip_sparse_assembler_t: element-wise assembly_array is not used (the HIP plug-in ' // &
        'integrates and assembles every cell on the device)'
    error stop 1
  end subroutine

  subroutine hip_assembly_matrix(this, num_fields, field_blocks, field_coupling, num_row_dofs, num_col_dofs, &
                                 fe_dofs_row, fe_dofs_col, elmat)
    class(hip_sparse_assembler_t), intent(inout) :: this
    integer(ip), intent(in) :: num_fields
    integer(ip), intent(in) :: field_blocks(num_fields)
    logical, intent(in) :: field_coupling(num_fields, num_fields)
    integer(ip), intent(in) :: num_row_dofs(num_fields), num_col_dofs(num_fields)
    type(i1p_t), intent(in) :: fe_dofs_row(num_fields), fe_dofs_col(num_fields)
    real(rp), intent(in) :: elmat(:, :)
    write(*, '(a)') 'hip_sparse_assembler_t: element-wise assembly_matrix is not used (the HIP plug-in ' // &
        'integrates and assembles every cell on the device)'
    error stop 1
  end subroutine

  ! the device pattern already is the compressed CSR of sort_and_compress
  subroutine hip_compress_storage_matrix(this, sparse_matrix_storage_format)
    class(hip_sparse_assembler_t), intent(inout) :: this
    character(*), intent(in) :: sparse_matrix_storage_format
  end subroutine

  subroutine hip_assembler_allocate(this, fe_space, diagonal_blocks_symmetric_storage)
    class(hip_sparse_assembler_t), intent(inout) :: this
    class(serial_fe_space_t), intent(in) :: fe_space
    logical, intent(in) :: diagonal_blocks_symmetric_storage
    integer(c_int32_t) :: sym
    integer(c_int64_t) :: nnz_full
    sym = 0
    if (diagonal_blocks_symmetric_storage) sym = 1
    call femhip_check(femhip_operator_create(fe_space%triangulation%environment%ctx, fe_space%h, sym, this%op), &
                      'hip_sparse_assembler_t%allocate')
    call femhip_check(femhip_operator_info(this%op, this%n, this%nnz, nnz_full), 'hip_sparse_assembler_t%allocate')
  end subroutine

  subroutine hip_assembler_free(this)
    class(hip_sparse_assembler_t), intent(inout) :: this
    if (c_associated(this%op)) call femhip_check(femhip_operator_destroy(this%op), 'hip_sparse_assembler_t%free')
    this%op = c_null_ptr
  end subroutine

  ! get_translation(): the assembled RHS copied to the host
  subroutine hip_assembler_get_translation(this, b)
    class(hip_sparse_assembler_t), intent(in) :: this
    real(rp), intent(out), target :: b(:)
    call femhip_check(femhip_operator_get_values(this%op, c_null_ptr, c_loc(b)), 'get_translation')
  end subroutine

  ! get_matrix()%get_irp/get_ja/get_val (csr_sparse_matrix.f90:192-225)
  subroutine hip_assembler_get_csr(this, irp, ja, val)
    class(hip_sparse_assembler_t), intent(in) :: this
    integer(ip), intent(out), target :: irp(:), ja(:)
    real(rp), intent(out), target :: val(:)
    call femhip_check(femhip_operator_get_pattern(this%op, irp, ja), 'get_matrix')
    call femhip_check(femhip_operator_get_values(this%op, c_loc(val), c_null_ptr), 'get_matrix')
  end subroutine

  subroutine hip_poisson_integrate_galerkin(this, fe_space, assembler)
    class(hip_poisson_discrete_integration_t), intent(in) :: this
    class(serial_fe_space_t), intent(inout) :: fe_space
    class(assembler_t), intent(inout) :: assembler
    type(femhip_physics_t) :: phys
    real(rp), allocatable, target :: quad_coords(:), source_values(:)
    real(rp) :: point(3)
    integer(c_int64_t) :: i, npts
    integer :: d, dim
    phys = this%phys
    if (associated(this%source_term)) then
      ! quad_coords => fe%get_quadrature_points_coordinates() for every cell,
      ! then source_term%get_value(quad_coords(qpoint), source_term_value)
      dim = fe_space%triangulation%num_dims
      npts = fe_space%num_cells * fe_space%nq
      allocate(quad_coords(npts * dim), source_values(npts))
      call femhip_check(femhip_space_quadrature_points(fe_space%h, quad_coords), 'integrate_galerkin')
      do i = 0, npts - 1
        point = 0.0_rp
        do d = 1, dim
          point(d) = quad_coords(i * dim + d)
        end do
        call this%source_term%get_value(point, source_values(i + 1))
      end do
      phys%f_kind = FEMHIP_F_QPOINT
      phys%f_qp = c_loc(source_values)
    end if
    select type (assembler)
    type is (hip_sparse_assembler_t)
      call femhip_check(femhip_operator_set_physics(assembler%op, phys), 'integrate_galerkin')
      call femhip_check(femhip_operator_compute(assembler%op, c_loc(fe_space%fixed_dof_values), c_null_ptr), &
                        'integrate_galerkin')
    class default
      write(*, '(a)') 'hip_poisson_discrete_integration_t needs the HIP sparse assembler'
      error stop 1
    end select
  end subroutine

  subroutine create_hip_cg(environment, base_iterative_linear_solver)
    type(environment_t), intent(in), target :: environment
    class(base_iterative_linear_solver_t), pointer, intent(inout) :: base_iterative_linear_solver
    allocate(hip_cg_t :: base_iterative_linear_solver)
    base_iterative_linear_solver%environment => environment
  end subroutine

  subroutine hip_cg_set_parameters_from_pl(this, parameter_list)
    class(hip_cg_t), intent(inout) :: this
    type(solver_parameters_t), intent(in) :: parameter_list
    this%prm%rtol = parameter_list%ils_rtol
    this%prm%atol = parameter_list%ils_atol
    this%prm%max_num_iterations = parameter_list%ils_max_num_iterations
    this%prm%stopping_criterion = parameter_list%ils_stopping_criterium
    this%prm%precision = FEMHIP_PRECISION_FP64
    call hip_cg_release(this)
  end subroutine

  subroutine hip_cg_release(this)
    class(hip_cg_t), intent(inout) :: this
    if (c_associated(this%solver)) call femhip_check(femhip_solver_destroy(this%solver), 'hip_cg_t%free')
    this%solver = c_null_ptr
    this%op_of_solver = c_null_ptr
  end subroutine

  ! cg_solve_body (cg.f90:101-238) on the device: the solver is (re)made when
  ! the operator or the preconditioner changed since the last solve
  subroutine hip_cg_solve_body(this, b, x)
    class(hip_cg_t), intent(inout) :: this
    real(rp), intent(in), target :: b(:)
    real(rp), intent(inout), target :: x(:)
    type(femhip_cg_result_t) :: res
    type(c_ptr) :: op
    if (.not. associated(this%A)) error stop 'hip_cg_t: set_operators first'
    select type (asm => this%A%assembler)
    type is (hip_sparse_assembler_t)
      op = asm%op
    class default
      write(*, '(a)') 'hip_cg_t needs an operator assembled by the HIP sparse assembler'
      error stop 1
    end select
    if (.not. c_associated(this%op_of_solver, op) .or. this%prm%preconditioner /= this%M) then
      call hip_cg_release(this)
      this%prm%preconditioner = this%M
      call femhip_check(femhip_solver_create(this%environment%ctx, op, this%prm, this%solver), 'hip_cg_t%solve')
      this%op_of_solver = op
    end if
    call femhip_check(femhip_solver_solve(this%solver, c_loc(b), c_loc(x), res), 'hip_cg_t%solve')
    this%num_iterations = res%num_iterations
    this%did_converge = res%did_converge == 1
  end subroutine

  subroutine hip_cg_free(this)
    class(hip_cg_t), intent(inout) :: this
    call hip_cg_release(this)
  end subroutine

end module hip_fempar_plugins
