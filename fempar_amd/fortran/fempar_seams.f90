! fempar_seams -- the FEMPAR types a HIP plug-in attaches to, restated as a
! standalone Fortran module (FEMPAR itself needs the FPL/FLAP submodules and
! MPI to build; their absence is why this mirror exists).  Names, argument
! lists and call order follow the reference:
!   * discrete_integration_t with the deferred integrate_galerkin(fe_space,
!     assembler) (Sources/Lib/Integration/FEOperator/discrete_integration.f90:44-73);
!   * assembler_t with the deferred assembly_array, assembly_matrix and
!     compress_storage_matrix (Integration/FEOperator/assembler.f90:37-113);
!   * fe_affine_operator_t, whose create picks the serial assembler in
!     fe_operator_create_serial_assembler (Integration/FEOperator/
!     sbm_fe_operator.i90:162-201) and whose compute runs the plug-in and then
!     compresses (sbm_fe_affine_operator.i90:15-34);
!   * base_iterative_linear_solver_t with the deferred set_parameters_from_pl
!     and solve_body (LinearAlgebra/IterativeLinearSolver/
!     base_iterative_linear_solver.f90:64-175), the creational-methods
!     dictionary Set/Get by key (iterative_linear_solver_creational_methods_
!     dictionary.f90:79-105) and the iterative_linear_solver_t wrapper
!     (iterative_linear_solver.f90:62-241).
! The mesh, FE space and parameter list are reduced to what the structured
! test_poisson driver uses; the concrete types that do the work are the HIP
! plug-ins of hip_fempar_plugins.
module fempar_seams
  use iso_c_binding
  use femhip_names
  implicit none
  private

  integer, parameter, public :: ip = c_int32_t, rp = c_double

  ! environment_t: one GPU context per process
  type, public :: environment_t
    type(c_ptr) :: ctx = c_null_ptr
  contains
    procedure :: create => environment_create
    procedure :: free => environment_free
  end type

  ! serial_triangulation_t from the struct-hex generator parameters
  ! (STRUCT_HEX_MESH_GENERATOR_NUM_DIMS / _NUM_CELLS_X_DIM / _DOMAIN_LIMITS),
  ! optionally with moved vertex coordinates (num_dims x num_vertices, vertex
  ! (i, j, k) at column i + (n0+1)(j + (n1+1) k) + 1)
  type, public :: serial_triangulation_t
    type(environment_t), pointer :: environment => null()
    integer(ip) :: num_dims = 3
    integer(ip) :: num_cells_x_dim(3) = 1
    real(rp)    :: domain_limits(6) = [0.0_rp, 1.0_rp, 0.0_rp, 1.0_rp, 0.0_rp, 1.0_rp]
    real(rp), allocatable :: vertex_coordinates(:, :)
  contains
    procedure :: create => triangulation_create
  end type

  ! scalar_function_t: an analytical function evaluated at a point
  ! (Integration/Function/function.f90:59-82, get_value_space)
  type, public :: scalar_function_t
  contains
    procedure :: get_value_space => scalar_function_get_value_space
    generic :: get_value => get_value_space
  end type

  ! serial_fe_space_t: one scalar Lagrangian field with strong Dirichlet data
  ! on the boundary (set id 1), numbered as the reference numbers it
  type, public :: serial_fe_space_t
    type(serial_triangulation_t), pointer :: triangulation => null()
    type(c_ptr)        :: h = c_null_ptr
    integer(c_int64_t) :: num_cells = 0
    integer(ip)        :: order = 1, nd = 0, nq = 0, num_free_dofs = 0, num_fixed_dofs = 0
    real(rp), pointer  :: fixed_dof_values(:) => null()   ! the fe_function's strong Dirichlet values
  contains
    procedure :: create => fe_space_create
    procedure :: get_num_dofs => fe_space_get_num_dofs
    procedure :: get_lst_dofs_gids => fe_space_get_lst_dofs_gids
    procedure :: interpolate_dirichlet_values => fe_space_interpolate_dirichlet_values
    procedure :: free => fe_space_free
  end type

  ! i1p_t of the reference (types_names): a pointer to an integer array
  type, public :: i1p_t
    integer(ip), pointer :: p(:) => null()
  end type

  type, abstract, public :: assembler_t
  contains
    procedure(assembly_array_interface), deferred :: assembly_array
    procedure(assembly_matrix_interface), deferred :: assembly_matrix
    procedure(compress_storage_matrix_interface), deferred :: compress_storage_matrix
    procedure(assembler_allocate_interface), deferred :: allocate
    procedure(assembler_free_interface), deferred :: free
  end type

  abstract interface
    subroutine assembly_array_interface(this, num_fields, field_blocks, num_dofs, fe_dofs, elvec)
      import :: assembler_t, ip, rp, i1p_t
      class(assembler_t), intent(inout) :: this
      integer(ip), intent(in) :: num_fields
      integer(ip), intent(in) :: field_blocks(num_fields)
      integer(ip), intent(in) :: num_dofs(num_fields)
      type(i1p_t), intent(in) :: fe_dofs(num_fields)
      real(rp), intent(in) :: elvec(:)
    end subroutine
    subroutine assembly_matrix_interface(this, num_fields, field_blocks, field_coupling, num_row_dofs, &
                                         num_col_dofs, fe_dofs_row, fe_dofs_col, elmat)
      import :: assembler_t, ip, rp, i1p_t
      class(assembler_t), intent(inout) :: this
      integer(ip), intent(in) :: num_fields
      integer(ip), intent(in) :: field_blocks(num_fields)
      logical, intent(in) :: field_coupling(num_fields, num_fields)
      integer(ip), intent(in) :: num_row_dofs(num_fields), num_col_dofs(num_fields)
      type(i1p_t), intent(in) :: fe_dofs_row(num_fields), fe_dofs_col(num_fields)
      real(rp), intent(in) :: elmat(:, :)
    end subroutine
    subroutine compress_storage_matrix_interface(this, sparse_matrix_storage_format)
      import :: assembler_t
      class(assembler_t), intent(inout) :: this
      character(*), intent(in) :: sparse_matrix_storage_format
    end subroutine
    subroutine assembler_allocate_interface(this, fe_space, diagonal_blocks_symmetric_storage)
      import :: assembler_t, serial_fe_space_t
      class(assembler_t), intent(inout) :: this
      class(serial_fe_space_t), intent(in) :: fe_space
      logical, intent(in) :: diagonal_blocks_symmetric_storage
    end subroutine
    subroutine assembler_free_interface(this)
      import :: assembler_t
      class(assembler_t), intent(inout) :: this
    end subroutine
  end interface

  type, abstract, public :: discrete_integration_t
  contains
    procedure(integrate_galerkin_interface), deferred :: integrate_galerkin
  end type

  abstract interface
    subroutine integrate_galerkin_interface(this, fe_space, assembler)
      import :: discrete_integration_t, serial_fe_space_t, assembler_t
      class(discrete_integration_t), intent(in) :: this
      class(serial_fe_space_t), intent(inout) :: fe_space
      class(assembler_t), intent(inout) :: assembler
    end subroutine
  end interface

  ! the serial assembler factory of fe_operator_create_serial_assembler: a
  ! back end installs its creator (the reference selects by storage format)
  abstract interface
    subroutine create_serial_assembler_interface(sparse_matrix_storage_format, assembler)
      import :: assembler_t
      character(*), intent(in) :: sparse_matrix_storage_format
      class(assembler_t), allocatable, intent(inout) :: assembler
    end subroutine
  end interface
  procedure(create_serial_assembler_interface), pointer, save :: the_serial_assembler_creator => null()
  public :: set_serial_assembler_creator

  integer, parameter :: created = 1, assembler_computed = 2   ! fe_operator states (fe_operator.f90)

  type, public :: fe_affine_operator_t
    class(assembler_t), allocatable :: assembler
    class(discrete_integration_t), pointer :: discrete_integration => null()
    class(serial_fe_space_t), pointer :: fe_space => null()
    character(len=16) :: sparse_matrix_storage_format = 'CSR'
    integer :: state = 0
  contains
    procedure :: create => fe_affine_operator_create
    procedure :: compute => fe_affine_operator_compute
    procedure :: free => fe_affine_operator_free
  end type

  ! the parameter list entries the CG of test_poisson reads (ILS_*)
  type, public :: solver_parameters_t
    real(rp)    :: ils_rtol = 1.0e-6_rp                  ! default_rtol (iterative_solver_parameters.f90:167)
    real(rp)    :: ils_atol = 0.0_rp
    integer(ip) :: ils_max_num_iterations = 1000
    integer(ip) :: ils_stopping_criterium = 4            ! res_res, the CG default
  end type

  integer(ip), parameter, public :: identity_preconditioner = 0, jacobi_preconditioner = 1

  type, abstract, public :: base_iterative_linear_solver_t
    type(environment_t), pointer :: environment => null()
    type(fe_affine_operator_t), pointer :: A => null()
    integer(ip) :: M = identity_preconditioner
    integer(ip) :: num_iterations = 0
    logical     :: did_converge = .false.
  contains
    procedure(set_parameters_from_pl_interface), deferred :: set_parameters_from_pl
    procedure(solve_body_interface), deferred :: solve_body
    procedure(solver_free_interface), deferred :: free
    procedure :: set_operators => base_solver_set_operators
  end type

  abstract interface
    subroutine set_parameters_from_pl_interface(this, parameter_list)
      import :: base_iterative_linear_solver_t, solver_parameters_t
      class(base_iterative_linear_solver_t), intent(inout) :: this
      type(solver_parameters_t), intent(in) :: parameter_list
    end subroutine
    subroutine solve_body_interface(this, b, x)
      import :: base_iterative_linear_solver_t, rp
      class(base_iterative_linear_solver_t), intent(inout) :: this
      real(rp), intent(in), target :: b(:)
      real(rp), intent(inout), target :: x(:)
    end subroutine
    subroutine solver_free_interface(this)
      import :: base_iterative_linear_solver_t
      class(base_iterative_linear_solver_t), intent(inout) :: this
    end subroutine
    subroutine create_iterative_linear_solver_interface(environment, base_iterative_linear_solver)
      import :: environment_t, base_iterative_linear_solver_t
      type(environment_t), intent(in), target :: environment
      class(base_iterative_linear_solver_t), pointer, intent(inout) :: base_iterative_linear_solver
    end subroutine
  end interface
  public :: create_iterative_linear_solver_interface

  ! the creational methods dictionary: key -> creator
  integer, parameter :: max_solver_keys = 16
  type :: solver_creator_t
    character(len=32) :: key = ''
    procedure(create_iterative_linear_solver_interface), pointer, nopass :: proc => null()
  end type
  type, public :: iterative_linear_solver_creational_methods_dictionary_t
    type(solver_creator_t) :: entries(max_solver_keys)
    integer :: n = 0
  contains
    procedure :: set => dictionary_set
    procedure :: get => dictionary_get
    procedure :: is_present => dictionary_is_present
  end type
  type(iterative_linear_solver_creational_methods_dictionary_t), save, public :: &
      the_iterative_linear_solver_creational_methods_dictionary

  type, public :: iterative_linear_solver_t
    type(environment_t), pointer :: environment => null()
    class(base_iterative_linear_solver_t), pointer :: solver => null()
  contains
    procedure :: create => ils_create
    procedure :: set_type_from_string => ils_set_type_from_string
    procedure :: set_parameters_from_pl => ils_set_parameters_from_pl
    procedure :: set_operators => ils_set_operators
    procedure :: solve => ils_solve
    procedure :: get_num_iterations => ils_get_num_iterations
    procedure :: did_converge => ils_did_converge
    procedure :: free => ils_free
  end type

contains

  subroutine environment_create(this, device)
    class(environment_t), intent(inout) :: this
    integer(ip), intent(in) :: device
    call femhip_check(femhip_ctx_create(device, this%ctx), 'environment_t%create')
  end subroutine

  subroutine environment_free(this)
    class(environment_t), intent(inout) :: this
    if (c_associated(this%ctx)) call femhip_check(femhip_ctx_destroy(this%ctx), 'environment_t%free')
    this%ctx = c_null_ptr
  end subroutine

  subroutine triangulation_create(this, environment, num_dims, num_cells_x_dim, domain_limits, vertex_coordinates)
    class(serial_triangulation_t), intent(inout) :: this
    type(environment_t), intent(in), target :: environment
    integer(ip), intent(in) :: num_dims, num_cells_x_dim(:)
    real(rp), intent(in) :: domain_limits(:)
    real(rp), intent(in), optional :: vertex_coordinates(:, :)
    this%environment => environment
    this%num_dims = num_dims
    this%num_cells_x_dim = 1
    this%num_cells_x_dim(1:num_dims) = num_cells_x_dim(1:num_dims)
    this%domain_limits = 0.0_rp
    this%domain_limits(1:2 * num_dims) = domain_limits(1:2 * num_dims)
    if (allocated(this%vertex_coordinates)) deallocate(this%vertex_coordinates)
    if (present(vertex_coordinates)) this%vertex_coordinates = vertex_coordinates
  end subroutine

  ! the base type's value: 0 (concrete functions override get_value_space)
  subroutine scalar_function_get_value_space(this, point, result)
    class(scalar_function_t), intent(in) :: this
    real(rp), intent(in) :: point(3)
    real(rp), intent(inout) :: result
    result = 0.0_rp
  end subroutine

  subroutine fe_space_create(this, triangulation, order)
    class(serial_fe_space_t), intent(inout) :: this
    type(serial_triangulation_t), intent(in), target :: triangulation
    integer(ip), intent(in) :: order
    type(femhip_mesh_desc_t) :: mesh
    integer(ip) :: nf, nx
    this%triangulation => triangulation
    this%order = order
    mesh%dim = triangulation%num_dims
    mesh%n = triangulation%num_cells_x_dim
    mesh%limits = triangulation%domain_limits
    mesh%order = order
    mesh%num_components = 1
    mesh%dirichlet = FEMHIP_BC_ALL_BOUNDARY
    call femhip_check(femhip_space_create_structured(triangulation%environment%ctx, mesh, this%h), &
                      'serial_fe_space_t%create')
    if (allocated(triangulation%vertex_coordinates)) &
      call femhip_check(femhip_space_set_vertex_coords(this%h, triangulation%vertex_coordinates), &
                        'serial_fe_space_t%create')
    call femhip_check(femhip_space_info(this%h, this%num_cells, this%nd, this%nq, nf, nx), 'serial_fe_space_t%create')
    this%num_free_dofs = nf
    this%num_fixed_dofs = nx
    allocate(this%fixed_dof_values(max(1, nx)))
    this%fixed_dof_values = 0.0_rp
  end subroutine

  function fe_space_get_num_dofs(this) result(n)
    class(serial_fe_space_t), intent(in) :: this
    integer(ip) :: n
    n = this%num_free_dofs
  end function

  subroutine fe_space_get_lst_dofs_gids(this, lst_dofs_gids)
    class(serial_fe_space_t), intent(in) :: this
    integer(ip), intent(out), target :: lst_dofs_gids(:)
    call femhip_check(femhip_space_get_dofs(this%h, lst_dofs_gids), 'serial_fe_space_t%get_lst_dofs_gids')
  end subroutine

  ! interpolate_dirichlet_values (sbm_serial_fe_space.i90:997-1176) of the
  ! boundary function phys%g_kind into the space's fixed DOF values
  subroutine fe_space_interpolate_dirichlet_values(this, phys)
    class(serial_fe_space_t), intent(inout) :: this
    type(femhip_physics_t), intent(in) :: phys
    call femhip_check(femhip_space_interpolate_dirichlet(this%h, phys, this%fixed_dof_values), &
                      'serial_fe_space_t%interpolate_dirichlet_values')
  end subroutine

  subroutine fe_space_free(this)
    class(serial_fe_space_t), intent(inout) :: this
    if (c_associated(this%h)) call femhip_check(femhip_space_destroy(this%h), 'serial_fe_space_t%free')
    this%h = c_null_ptr
    if (associated(this%fixed_dof_values)) deallocate(this%fixed_dof_values)
    this%fixed_dof_values => null()
  end subroutine

  subroutine set_serial_assembler_creator(proc)
    procedure(create_serial_assembler_interface) :: proc
    the_serial_assembler_creator => proc
  end subroutine

  ! fe_operator_create_serial_assembler (sbm_fe_operator.i90:162-201)
  subroutine fe_operator_create_serial_assembler(sparse_matrix_storage_format, assembler)
    character(*), intent(in) :: sparse_matrix_storage_format
    class(assembler_t), allocatable, intent(inout) :: assembler
    if (.not. associated(the_serial_assembler_creator)) then
      write(*, '(a)') 'fe_operator_create_serial_assembler: no serial assembler installed'
      error stop 1
    end if
    call the_serial_assembler_creator(sparse_matrix_storage_format, assembler)
  end subroutine

  ! fe_affine_operator_t%create (sbm_fe_operator.i90:28-71)
  subroutine fe_affine_operator_create(this, sparse_matrix_storage_format, diagonal_blocks_symmetric_storage, &
                                       fe_space, discrete_integration)
    class(fe_affine_operator_t), intent(inout) :: this
    character(*), intent(in) :: sparse_matrix_storage_format
    logical, intent(in) :: diagonal_blocks_symmetric_storage(:)
    class(serial_fe_space_t), target, intent(inout) :: fe_space
    class(discrete_integration_t), target, intent(in) :: discrete_integration
    this%sparse_matrix_storage_format = sparse_matrix_storage_format
    this%fe_space => fe_space
    this%discrete_integration => discrete_integration
    call fe_operator_create_serial_assembler(sparse_matrix_storage_format, this%assembler)
    call this%assembler%allocate(fe_space, diagonal_blocks_symmetric_storage(1))
    this%state = created
  end subroutine

  ! fe_affine_operator_t%compute (sbm_fe_affine_operator.i90:15-34)
  subroutine fe_affine_operator_compute(this)
    class(fe_affine_operator_t), intent(inout) :: this
    if (this%state < created) then
      write(*, '(a)') 'fe_affine_operator_t%compute: operator not created'
      error stop 1
    end if
    call this%discrete_integration%integrate_galerkin(this%fe_space, this%assembler)
    call this%assembler%compress_storage_matrix(trim(this%sparse_matrix_storage_format))
    this%state = assembler_computed
  end subroutine

  subroutine fe_affine_operator_free(this)
    class(fe_affine_operator_t), intent(inout) :: this
    if (allocated(this%assembler)) then
      call this%assembler%free()
      deallocate(this%assembler)
    end if
    this%state = 0
  end subroutine

  subroutine base_solver_set_operators(this, A, M)
    class(base_iterative_linear_solver_t), intent(inout) :: this
    type(fe_affine_operator_t), target, intent(in) :: A
    integer(ip), intent(in) :: M
    this%A => A
    this%M = M
  end subroutine

  subroutine dictionary_set(this, key, proc)
    class(iterative_linear_solver_creational_methods_dictionary_t), intent(inout) :: this
    character(*), intent(in) :: key
    procedure(create_iterative_linear_solver_interface) :: proc
    integer :: i
    do i = 1, this%n
      if (trim(this%entries(i)%key) == key) then
        this%entries(i)%proc => proc
        return
      end if
    end do
    if (this%n == max_solver_keys) error stop 'creational methods dictionary full'
    this%n = this%n + 1
    this%entries(this%n)%key = key
    this%entries(this%n)%proc => proc
  end subroutine

  subroutine dictionary_get(this, key, proc)
    class(iterative_linear_solver_creational_methods_dictionary_t), intent(in) :: this
    character(*), intent(in) :: key
    procedure(create_iterative_linear_solver_interface), pointer, intent(out) :: proc
    integer :: i
    proc => null()
    do i = 1, this%n
      if (trim(this%entries(i)%key) == key) proc => this%entries(i)%proc
    end do
  end subroutine

  function dictionary_is_present(this, key) result(present_)
    class(iterative_linear_solver_creational_methods_dictionary_t), intent(in) :: this
    character(*), intent(in) :: key
    logical :: present_
    integer :: i
    present_ = .false.
    do i = 1, this%n
      present_ = present_ .or. trim(this%entries(i)%key) == key
    end do
  end function

  subroutine ils_create(this, environment)
    class(iterative_linear_solver_t), intent(inout) :: this
    type(environment_t), target, intent(in) :: environment
    this%environment => environment
  end subroutine

  ! iterative_linear_solver_set_type_from_string: the creator registered
  ! under the key allocates the concrete solver (iterative_linear_solver.f90)
  subroutine ils_set_type_from_string(this, key)
    class(iterative_linear_solver_t), intent(inout) :: this
    character(*), intent(in) :: key
    procedure(create_iterative_linear_solver_interface), pointer :: proc
    call the_iterative_linear_solver_creational_methods_dictionary%get(key, proc)
    if (.not. associated(proc)) then
      write(*, '(3a)') 'iterative_linear_solver_t%set_type_from_string: no solver registered as ''', key, ''''
      error stop 1
    end if
    call this%free()
    call proc(this%environment, this%solver)
  end subroutine

  subroutine ils_set_parameters_from_pl(this, parameter_list)
    class(iterative_linear_solver_t), intent(inout) :: this
    type(solver_parameters_t), intent(in) :: parameter_list
    call this%solver%set_parameters_from_pl(parameter_list)
  end subroutine

  subroutine ils_set_operators(this, A, M)
    class(iterative_linear_solver_t), intent(inout) :: this
    type(fe_affine_operator_t), target, intent(in) :: A
    integer(ip), intent(in) :: M
    call this%solver%set_operators(A, M)
  end subroutine

  subroutine ils_solve(this, b, x)
    class(iterative_linear_solver_t), intent(inout) :: this
    real(rp), intent(in), target :: b(:)
    real(rp), intent(inout), target :: x(:)
    call this%solver%solve_body(b, x)
  end subroutine

  function ils_get_num_iterations(this) result(n)
    class(iterative_linear_solver_t), intent(in) :: this
    integer(ip) :: n
    n = this%solver%num_iterations
  end function

  function ils_did_converge(this) result(c)
    class(iterative_linear_solver_t), intent(in) :: this
    logical :: c
    c = this%solver%did_converge
  end function

  subroutine ils_free(this)
    class(iterative_linear_solver_t), intent(inout) :: this
    if (associated(this%solver)) then
      call this%solver%free()
      deallocate(this%solver)
    end if
    this%solver => null()
  end subroutine

end module fempar_seams
