! test_poisson_seams -- FEMPAR's test_poisson driver sequence written only
! against the reference's plug-in types (module fempar_seams), with the HIP
! back end installed at start-up (module hip_fempar_plugins):
!   triangulation%create -> fe_space%create -> interpolate_dirichlet_values
!   -> fe_affine_operator%create('CSR', [.true.], fe_space, discrete_integration)
!   -> fe_affine_operator%compute() -> iterative_linear_solver%create /
!   set_type_from_string('CG') / set_parameters_from_pl / set_operators /
!   solve(translation, x),
! as Sources/Tests/Serial/test_poisson/modules/test_poisson_driver.f90 does
! (:131-560).  The check is the driver's check_solution (:562-600): with
! f = 0 and g = x + y (+ z) the solution reproduces u = x + y (+ z) at every
! free node.  No femhip_* call appears here; the driver would read the same
! against FEMPAR's own serial_fe_space_t and sparse assembler.
! With moved = 1 the interior vertices are moved (every cell then has its own
! map) and the source term is a scalar_function_t the plug-in evaluates at
! every quadrature point (zero here, as the linear solution needs); the
! nodes of the check are then the mapped reference nodes.
! Usage: test_poisson_seams [dim] [n] [order] [precond 0|1] [moved 0|1]
module test_poisson_seams_functions
  use fempar_seams
  implicit none
  integer, save :: source_calls = 0
  type, extends(scalar_function_t) :: zero_source_t
  contains
    procedure :: get_value_space => zero_source_get_value_space
  end type
contains
  ! f = 0 (the source of the linear solution), counting the evaluations
  subroutine zero_source_get_value_space(this, point, result)
    class(zero_source_t), intent(in) :: this
    real(rp), intent(in) :: point(3)
    real(rp), intent(inout) :: result
    result = 0.0_rp
    source_calls = source_calls + 1
  end subroutine
end module test_poisson_seams_functions

program test_poisson_seams
  use iso_c_binding
  use femhip_names, only: femhip_physics_t, FEMHIP_PHYS_POISSON, FEMHIP_K_CONST, FEMHIP_F_CONST, FEMHIP_G_LINEAR
  use fempar_seams
  use hip_fempar_plugins
  use test_poisson_seams_functions
  implicit none
  type(environment_t), target :: environment
  type(serial_triangulation_t), target :: triangulation
  type(serial_fe_space_t), target :: fe_space
  type(hip_poisson_discrete_integration_t), target :: poisson_integration
  type(fe_affine_operator_t), target :: fe_affine_operator
  type(iterative_linear_solver_t) :: iterative_linear_solver
  type(solver_parameters_t) :: parameter_list
  type(zero_source_t), target :: source_term
  integer(ip), allocatable, target :: gids(:)
  real(rp), allocatable, target :: x(:), b(:), vc(:, :)
  integer :: dim, ncell, order, prec, moved, c, l, k, m, rem, ijk(3), np1, nd, nv, v, a, vid, mul
  real(rp) :: u, err, xi(3), xn(3), shp
  logical :: converged, interior
  character(len=32) :: arg

  dim = 3; ncell = 8; order = 1; prec = 0; moved = 0
  if (command_argument_count() >= 1) then; call get_command_argument(1, arg); read(arg, *) dim; end if
  if (command_argument_count() >= 2) then; call get_command_argument(2, arg); read(arg, *) ncell; end if
  if (command_argument_count() >= 3) then; call get_command_argument(3, arg); read(arg, *) order; end if
  if (command_argument_count() >= 4) then; call get_command_argument(4, arg); read(arg, *) prec; end if
  if (command_argument_count() >= 5) then; call get_command_argument(5, arg); read(arg, *) moved; end if

  ! vertex coordinates: the generator's, interior ones moved by up to a fifth
  ! of a cell in an irregular (deterministic) pattern when moved = 1
  nv = (ncell + 1) ** dim
  allocate(vc(dim, nv))
  do v = 0, nv - 1
    rem = v
    interior = .true.
    do k = 1, dim
      ijk(k) = mod(rem, ncell + 1)
      rem = rem / (ncell + 1)
      interior = interior .and. ijk(k) > 0 .and. ijk(k) < ncell
    end do
    do k = 1, dim
      vc(k, v + 1) = real(ijk(k), rp) / real(ncell, rp)
      if (moved == 1 .and. interior) vc(k, v + 1) = vc(k, v + 1) + &
          0.2_rp / real(ncell, rp) * sin(12.9898_rp * ijk(1) + 78.233_rp * ijk(2) + 37.719_rp * ijk(min(3, dim)) + k)
    end do
  end do

  call hip_fempar_plugins_init()
  call environment%create(0_ip)
  if (moved == 1) then
    call triangulation%create(environment, int(dim, ip), [ncell, ncell, ncell], &
                              [0.0_rp, 1.0_rp, 0.0_rp, 1.0_rp, 0.0_rp, 1.0_rp], vc)
  else
    call triangulation%create(environment, int(dim, ip), [ncell, ncell, ncell], &
                              [0.0_rp, 1.0_rp, 0.0_rp, 1.0_rp, 0.0_rp, 1.0_rp])
  end if
  call fe_space%create(triangulation, int(order, ip))

  ! the Poisson plug-in: k = 1, f = 0, g = x + y (+ z) on the boundary
  poisson_integration%phys%kind = FEMHIP_PHYS_POISSON
  poisson_integration%phys%k_kind = FEMHIP_K_CONST
  poisson_integration%phys%k0 = 1.0_rp
  poisson_integration%phys%f_kind = FEMHIP_F_CONST
  poisson_integration%phys%f0 = 0.0_rp
  poisson_integration%phys%g_kind = FEMHIP_G_LINEAR
  call fe_space%interpolate_dirichlet_values(poisson_integration%phys)
  if (moved == 1) poisson_integration%source_term => source_term

  call fe_affine_operator%create('CSR', [.true.], fe_space, poisson_integration)
  call fe_affine_operator%compute()

  parameter_list%ils_rtol = 1.0e-12_rp
  parameter_list%ils_max_num_iterations = 5000
  call iterative_linear_solver%create(environment)
  call iterative_linear_solver%set_type_from_string('CG')
  call iterative_linear_solver%set_parameters_from_pl(parameter_list)
  if (prec == 1) then
    call iterative_linear_solver%set_operators(fe_affine_operator, jacobi_preconditioner)
  else
    call iterative_linear_solver%set_operators(fe_affine_operator, identity_preconditioner)
  end if

  allocate(x(fe_space%get_num_dofs()), b(fe_space%get_num_dofs()))
  select type (asm => fe_affine_operator%assembler)
  type is (hip_sparse_assembler_t)
    call asm%get_translation(b)          ! fe_affine_operator%get_translation()
  end select
  x = 0.0_rp
  call iterative_linear_solver%solve(b, x)

  ! check_solution: nodal values of u = x + y (+ z)
  nd = fe_space%nd
  allocate(gids(fe_space%num_cells * nd))
  call fe_space%get_lst_dofs_gids(gids)
  np1 = order + 1
  err = 0.0_rp
  do c = 0, int(fe_space%num_cells) - 1
    rem = c
    do k = 1, dim
      ijk(k) = mod(rem, ncell)
      rem = rem / ncell
    end do
    do l = 0, nd - 1
      if (gids(c * nd + l + 1) <= 0) cycle
      ! the node: trilinear map of its reference position xi = (2/order) m - 1
      rem = l
      xi = 0.0_rp
      do k = 1, dim
        m = mod(rem, np1)
        rem = rem / np1
        xi(k) = (2.0_rp / real(order, rp)) * real(m, rp) - 1.0_rp
      end do
      xn = 0.0_rp
      do a = 0, 2 ** dim - 1
        shp = 1.0_rp
        vid = 0
        mul = 1
        do k = 1, dim
          if (btest(a, k - 1)) then
            shp = shp * 0.5_rp * (1.0_rp + xi(k))
            vid = vid + mul * (ijk(k) + 1)
          else
            shp = shp * 0.5_rp * (1.0_rp - xi(k))
            vid = vid + mul * ijk(k)
          end if
          mul = mul * (ncell + 1)
        end do
        xn(1:dim) = xn(1:dim) + shp * vc(:, vid + 1)
      end do
      u = sum(xn(1:dim))
      err = max(err, abs(x(gids(c * nd + l + 1)) - u))
    end do
  end do
  write(*, '(a,i0,a,i0,a,i0,a,i0,a,i0,a,i0,a,es10.3)') 'test_poisson_seams dim=', dim, ' n=', ncell, &
      ' order=', order, ' moved=', moved, ' free=', fe_space%get_num_dofs(), &
      ' cg_iterations=', iterative_linear_solver%get_num_iterations(), ' max_nodal_error=', err
  converged = iterative_linear_solver%did_converge()
  if (moved == 1 .and. source_calls /= int(fe_space%num_cells) * fe_space%nq) then
    write(*, '(a,i0)') 'source term evaluations: ', source_calls
    converged = .false.
  end if
  call iterative_linear_solver%free()
  call fe_affine_operator%free()
  call fe_space%free()
  call environment%free()
  if (.not. converged .or. err > 1.0e-6_rp) then
    write(*, '(a)') 'FAILED'
    error stop 1
  end if
  write(*, '(a)') 'PASSED'
end program test_poisson_seams
