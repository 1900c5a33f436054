! test_poisson_hip -- FEMPAR's test_poisson run sequence through femhip_names.
!
! Mirrors Sources/Tests/Serial/test_poisson/modules/test_poisson_driver.f90
! (setup_triangulation :131-269, setup_fe_space :324-386, setup_system
! :414-458, assemble_system :516-541, setup_solver/solve_system :460-560,
! check_solution :562-600): structured hex mesh, Q1/Q2 Lagrangian space with
! strong Dirichlet data g = x + y (+ z) on the boundary, f = 0, CG with
! rtol 1e-12; the discrete solution must reproduce u = x + y (+ z) at the nodes.
! Usage: test_poisson_hip [dim] [n] [order]
program test_poisson_hip
  use iso_c_binding
  use femhip_names
  implicit none
  type(c_ptr) :: ctx, space, op, solver
  type(femhip_mesh_desc_t) :: mesh
  type(femhip_physics_t) :: phys
  type(femhip_cg_params_t) :: prm
  type(femhip_cg_result_t) :: res
  type(femhip_error_norms_t) :: en
  integer(c_int64_t) :: num_cells, nnz, nnz_full
  integer(c_int32_t) :: nd, nq, n_free, n_fixed, n
  integer(c_int32_t), allocatable, target :: gids(:)
  real(c_double), allocatable, target :: fixed(:), x(:), b(:)
  integer :: dim, ncell, order, c, l, k, m, rem, ijk(3), np1
  real(c_double) :: u, err, t
  character(len=32) :: arg

  dim = 3; ncell = 8; order = 1
  if (command_argument_count() >= 1) then; call get_command_argument(1, arg); read(arg, *) dim; end if
  if (command_argument_count() >= 2) then; call get_command_argument(2, arg); read(arg, *) ncell; end if
  if (command_argument_count() >= 3) then; call get_command_argument(3, arg); read(arg, *) order; end if

  call femhip_check(femhip_ctx_create(0_c_int32_t, ctx), 'ctx_create')

  ! triangulation + fe_space (struct-hex generator, boundary vefs -> Dirichlet)
  mesh%dim = dim
  mesh%n = 1
  mesh%n(1:dim) = ncell
  mesh%limits = 0.0_c_double
  do k = 1, dim
    mesh%limits(2 * k) = 1.0_c_double
  end do
  mesh%order = order
  mesh%num_components = 1
  mesh%dirichlet = FEMHIP_BC_ALL_BOUNDARY
  call femhip_check(femhip_space_create_structured(ctx, mesh, space), 'space_create_structured')
  call femhip_check(femhip_space_info(space, num_cells, nd, nq, n_free, n_fixed), 'space_info')
  allocate(gids(num_cells * nd), fixed(max(1, n_fixed)))
  call femhip_check(femhip_space_get_dofs(space, gids), 'space_get_dofs')

  ! discrete integration: Poisson, f = 0, g = x + y (+ z)
  phys%kind = FEMHIP_PHYS_POISSON
  phys%k_kind = FEMHIP_K_CONST
  phys%k0 = 1.0_c_double
  phys%f_kind = FEMHIP_F_CONST
  phys%f0 = 0.0_c_double
  phys%g_kind = FEMHIP_G_LINEAR
  call femhip_check(femhip_space_interpolate_dirichlet(space, phys, fixed), 'interpolate_dirichlet')

  ! fe_affine_operator: CSR, symmetric storage, compute
  call femhip_check(femhip_operator_create(ctx, space, 1_c_int32_t, op), 'operator_create')
  call femhip_check(femhip_operator_info(op, n, nnz, nnz_full), 'operator_info')
  call femhip_check(femhip_operator_set_physics(op, phys), 'operator_set_physics')
  call femhip_check(femhip_operator_compute(op, c_loc(fixed), c_null_ptr), 'operator_compute')

  ! iterative_linear_solver: CG, identity preconditioner, rtol 1e-12, maxit 5000
  prm%rtol = 1.0e-12_c_double
  prm%max_num_iterations = 5000
  prm%stopping_criterion = FEMHIP_RES_RES
  prm%preconditioner = FEMHIP_PREC_IDENTITY
  prm%precision = FEMHIP_PRECISION_FP64
  call femhip_check(femhip_solver_create(ctx, op, prm, solver), 'solver_create')
  allocate(x(n), b(n))
  call femhip_check(femhip_operator_get_values(op, c_null_ptr, c_loc(b)), 'get_translation')
  x = 0.0_c_double
  call femhip_check(femhip_solver_solve(solver, c_loc(b), c_loc(x), res), 'solve')

  ! check_solution: nodal values of u = x + y (+ z)
  np1 = order + 1
  err = 0.0_c_double
  do c = 0, int(num_cells) - 1
    rem = c
    do k = 1, dim
      ijk(k) = mod(rem, ncell)
      rem = rem / ncell
    end do
    do l = 0, nd - 1
      if (gids(c * nd + l + 1) <= 0) cycle
      rem = l
      u = 0.0_c_double
      do k = 1, dim
        m = mod(rem, np1)
        rem = rem / np1
        t = real(m, c_double) / real(order, c_double)
        u = u + (real(ijk(k), c_double) + t) / real(ncell, c_double)
      end do
      err = max(err, abs(x(gids(c * nd + l + 1)) - u))
    end do
  end do
  write(*, '(a,i0,a,i0,a,i0,a,i0,a,i0,a,es10.3)') 'test_poisson_hip dim=', dim, ' n=', ncell, ' order=', order, &
      ' free=', n_free, ' cg_iterations=', res%num_iterations, ' max_nodal_error=', err
  ! error norms of u - u_h, as test_poisson_driver's check_solution computes them
  call femhip_check(femhip_solver_error_norms(solver, FEMHIP_EXACT_LINEAR, en), 'error_norms')
  write(*, '(a,es10.3,a,es10.3,a,es10.3)') 'l2_norm=', en%l2_norm, ' h1_seminorm=', en%h1_seminorm, &
      ' linfty_norm=', en%linfty_norm
  call femhip_check(femhip_solver_destroy(solver), 'solver_destroy')
  call femhip_check(femhip_operator_destroy(op), 'operator_destroy')
  call femhip_check(femhip_space_destroy(space), 'space_destroy')
  call femhip_check(femhip_ctx_destroy(ctx), 'ctx_destroy')
  if (res%did_converge /= 1 .or. err > 1.0e-6_c_double .or. en%h1_norm > 1.0e-6_c_double &
      .or. en%linfty_norm > 1.0e-6_c_double) then
    write(*, '(a)') 'FAILED'
    error stop 1
  end if
  write(*, '(a)') 'PASSED'
end program test_poisson_hip
