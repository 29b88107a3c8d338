!=======================================================================
! swm_mi355x.F90 -- Fortran side of the MI355X dynamical core.
!
!   module swm_mi355x_c     ISO_C_BINDING interfaces of include/swm_mi355x.h,
!                           one for one
!   module swm_mi355x       the device context of a run and the drop-in
!                           replacements of the reference's hot path:
!                             step_dt, wrap_prognostics (swm/source/step.F:
!                             114-212, 438-493), elliptic (utilities/layer_3/
!                             elliptic_solver.F:40-195), multigrid
!                             (utilities/layer_3/multigrid_solver.F:476-591),
!                             wrap (utilities/layer_2/wrap_data.F:1790-1974)
!   module operators_mi355x the operators of utilities/layer_2/operators.F
!                           under their own names and argument lists
!   module mass_fct_mi355x, tracer_fct_mi355x, global_reductions_mi355x
!                           set_mass_fct (utilities/layer_2/mass_fct.F:23),
!                           set_tracer_fct (utilities/layer_2/tracer_fct.F:
!                           22), global_real_sum (utilities/layer_2/
!                           global_reductions.F:77-98)
!
! A non-zero status from the library stops the program, as the reference's
! abort_code does (utilities/layer_1/abort.F:26-49).
!=======================================================================
module swm_mi355x_c
  use iso_c_binding
  implicit none
  public

  integer(c_int), parameter :: SWM_OK = 0, SWM_FP64 = 0, SWM_MIXED = 1
  integer(c_int), parameter :: SWM_ADV_FCT = 0, SWM_ADV_CENTERED = 1, SWM_ADV_UPWIND = 2, SWM_ADV_REMAP = 3

  type, bind(C) :: swm_config
    integer(c_int) :: ntracer
    real(c_double) :: dt
    real(c_double) :: grav
    integer(c_int) :: precision
    integer(c_int) :: device
    integer(c_int) :: advection
  end type swm_config

  interface
    function swm_last_error() bind(C, name="swm_last_error") result(p)
      import :: c_ptr
      type(c_ptr) :: p
    end function
    function swm_grid_points(g, p_list) bind(C, name="swm_grid_points") result(st)
      import :: c_int, c_double
      integer(c_int), value :: g
      real(c_double) :: p_list(*)
      integer(c_int) :: st
    end function
    function swm_grid_create(g, p_list, radius, omega, grid) bind(C, name="swm_grid_create") result(st)
      import :: c_int, c_double, c_ptr
      integer(c_int), value :: g
      type(c_ptr), value :: p_list
      real(c_double), value :: radius, omega
      type(c_ptr) :: grid
      integer(c_int) :: st
    end function
    subroutine swm_grid_destroy(grid) bind(C, name="swm_grid_destroy")
      import :: c_ptr
      type(c_ptr), value :: grid
    end subroutine
    function swm_grid_dims(grid, g, iim, jjm, max_ig) bind(C, name="swm_grid_dims") result(st)
      import :: c_int, c_ptr
      type(c_ptr), value :: grid
      integer(c_int) :: g, iim, jjm, max_ig
      integer(c_int) :: st
    end function
    function swm_grid_array(grid, name, out) bind(C, name="swm_grid_array") result(st)
      import :: c_int, c_ptr, c_char
      type(c_ptr), value :: grid
      character(kind=c_char) :: name(*)
      type(c_ptr), value :: out
      integer(c_int) :: st
    end function
    function swm_create(grid, cfg, ctx) bind(C, name="swm_create") result(st)
      import :: c_int, c_ptr, swm_config
      type(c_ptr), value :: grid
      type(swm_config) :: cfg
      type(c_ptr) :: ctx
      integer(c_int) :: st
    end function
    subroutine swm_destroy(ctx) bind(C, name="swm_destroy")
      import :: c_ptr
      type(c_ptr), value :: ctx
    end subroutine
    function swm_init_state(ctx, h, hs, u, tracer) bind(C, name="swm_init_state") result(st)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double) :: h(*), u(*), tracer(*)
      type(c_ptr), value :: hs
      integer(c_int) :: st
    end function
    function swm_init_synthetic(ctx, seed_index, lmax, umax, h0, h_rms, nbumps, bump_radius) &
        bind(C, name="swm_init_synthetic") result(st)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: seed_index, lmax, nbumps
      real(c_double), value :: umax, h0, h_rms, bump_radius
      integer(c_int) :: st
    end function
    function swm_step(ctx, nsteps) bind(C, name="swm_step") result(st)
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: nsteps
      integer(c_int) :: st
    end function
    function swm_get_field(ctx, name, out) bind(C, name="swm_get_field") result(st)
      import :: c_int, c_ptr, c_char, c_double
      type(c_ptr), value :: ctx
      character(kind=c_char) :: name(*)
      real(c_double) :: out(*)
      integer(c_int) :: st
    end function
    function swm_synchronize(ctx) bind(C, name="swm_synchronize") result(st)
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int) :: st
    end function
    function swm_step_count(ctx) bind(C, name="swm_step_count") result(n)
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int) :: n
    end function
    function swm_last_vcycles(ctx) bind(C, name="swm_last_vcycles") result(n)
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int) :: n
    end function
    function swm_total_mass(ctx, mass) bind(C, name="swm_total_mass") result(st)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double) :: mass
      integer(c_int) :: st
    end function
    function swm_time_seconds(ctx) bind(C, name="swm_time_seconds") result(t)
      import :: c_ptr, c_long_long
      type(c_ptr), value :: ctx
      integer(c_long_long) :: t
    end function
    function swm_write_restart(ctx, path) bind(C, name="swm_write_restart") result(st)
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, path
      integer(c_int) :: st
    end function
    function swm_read_restart(ctx, path) bind(C, name="swm_read_restart") result(st)
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, path
      integer(c_int) :: st
    end function
    function swm_init_diss_op(ctx, grid) bind(C, name="swm_init_diss_op") result(st)
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, grid
      integer(c_int) :: st
    end function
    function swm_op_dissipation(ctx, norder, visc, xin, xout) bind(C, name="swm_op_dissipation") result(st)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: norder
      real(c_double), value :: visc
      real(c_double) :: xin(*), xout(*)
      integer(c_int) :: st
    end function
    function swm_set_remap_limiter(ctx, limit_min, limit_max) bind(C, name="swm_set_remap_limiter") result(st)
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: limit_min, limit_max
      integer(c_int) :: st
    end function
    function swm_op_remap_mass_and_tracers(ctx, iim, jjm, nmass, nsdm, ntrace, mass, trace, dt, uc, edgeflux) &
        bind(C, name="swm_op_remap_mass_and_tracers") result(st)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: iim, jjm, nmass, nsdm, ntrace
      real(c_double) :: mass(*), trace(*), uc(*)
      real(c_double), value :: dt
      type(c_ptr), value :: edgeflux
      integer(c_int) :: st
    end function
    function swm_qp_create_coords(grid, tag, nsecond, path, clat, clon, crlat, crlon, qp) &
        bind(C, name="swm_qp_create_coords") result(st)
      import :: c_int, c_ptr, c_char, c_long_long, c_double
      type(c_ptr), value :: grid, path
      character(kind=c_char) :: tag(*)
      integer(c_long_long), value :: nsecond
      real(c_double) :: clat(*), clon(*), crlat(*), crlon(*)
      type(c_ptr) :: qp
      integer(c_int) :: st
    end function
    function swm_qp_pack(qp, ctx) bind(C, name="swm_qp_pack") result(st)
      import :: c_int, c_ptr
      type(c_ptr), value :: qp, ctx
      integer(c_int) :: st
    end function
    function swm_qp_write(qp, nsecond) bind(C, name="swm_qp_write") result(st)
      import :: c_int, c_ptr, c_long_long
      type(c_ptr), value :: qp
      integer(c_long_long), value :: nsecond
      integer(c_int) :: st
    end function
    subroutine swm_qp_destroy(qp) bind(C, name="swm_qp_destroy")
      import :: c_ptr
      type(c_ptr), value :: qp
    end subroutine
    function swm_wrap(ctx, kind, nspace, nlayer, x) bind(C, name="swm_wrap") result(st)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: kind, nspace, nlayer
      real(c_double) :: x(*)
      integer(c_int) :: st
    end function
    function swm_elliptic(ctx, nvar, rhs, xout) bind(C, name="swm_elliptic") result(st)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: nvar
      real(c_double) :: rhs(*), xout(*)
      integer(c_int) :: st
    end function
    function swm_multigrid(ctx, nvar, beta, alpha) bind(C, name="swm_multigrid") result(st)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: nvar
      real(c_double) :: beta(*), alpha(*)
      integer(c_int) :: st
    end function
    function swm_global_real_sum(ctx, x, mask, n, out) bind(C, name="swm_global_real_sum") result(st)
      import :: c_int, c_ptr, c_double, c_int64_t
      type(c_ptr), value :: ctx
      real(c_double) :: x(*), mask(*)
      integer(c_int64_t), value :: n
      real(c_double) :: out
      integer(c_int) :: st
    end function
    function swm_op_set_mass_fct(ctx, iim, jjm, nlm, nsdm, ntprog, nttend, ncorners, ndata, nedges, nspace, &
        npad, r3, n3, n4, tend_index, tend_weights, dt, area_inv, area, dc, normal, mss, mss_c, u, mss_low, &
        mss_flux_horz, msh_f, mss_flux_forced, mask) bind(C, name="swm_op_set_mass_fct") result(st)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: iim, jjm, nlm, nsdm, ntprog, nttend, ncorners, ndata, nedges, nspace, npad, r3, n3, n4
      integer(c_int) :: tend_index(*)
      real(c_double) :: tend_weights(*)
      real(c_double), value :: dt
      real(c_double) :: area_inv(*), area(*), dc(*), normal(*), mss(*), mss_c(*), u(*), mss_low(*), &
                        mss_flux_horz(*), msh_f(*)
      type(c_ptr), value :: mss_flux_forced, mask
      integer(c_int) :: st
    end function
    function swm_op_set_tracer_fct(ctx, iim, jjm, nlm, nsdm, nedges, dt, area_inv, area, trc, trc_high, mss, &
        mss_low, mss_horz, mss_flux_horz, trc_fct, trc_flux_horz, msstrc_low_order, mask) &
        bind(C, name="swm_op_set_tracer_fct") result(st)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: iim, jjm, nlm, nsdm, nedges
      real(c_double), value :: dt
      real(c_double) :: area_inv(*), area(*), trc(*), trc_high(*), mss(*), mss_low(*), mss_horz(*), &
                        mss_flux_horz(*), trc_fct(*)
      type(c_ptr), value :: trc_flux_horz, msstrc_low_order, mask
      integer(c_int) :: st
    end function
    ! ---- operators.F, argument lists of the reference plus nsdm ----
    function swm_op_curl_of_a_vector(ctx, npad, nx, ny, nsdm, ndata, nspace, ncorners, area_inv, &
        v_weights, xin, xout) bind(C, name="swm_op_curl_of_a_vector") result(st)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: npad, nx, ny, nsdm, ndata, nspace, ncorners
      real(c_double) :: area_inv(*), v_weights(*), xin(*), xout(*)
      integer(c_int) :: st
    end function
    function swm_op_div_of_a_vector(ctx, npad, nx, ny, nsdm, ndata, nspace, ncorners, area_inv, &
        v_weights, xin, xout) bind(C, name="swm_op_div_of_a_vector") result(st)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: npad, nx, ny, nsdm, ndata, nspace, ncorners
      real(c_double) :: area_inv(*), v_weights(*), xin(*), xout(*)
      integer(c_int) :: st
    end function
    function swm_op_div_of_an_edge_flux(ctx, npad, nx, ny, nsdm, nedges, area_inv, xin, xout) &
        bind(C, name="swm_op_div_of_an_edge_flux") result(st)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: npad, nx, ny, nsdm, nedges
      real(c_double) :: area_inv(*), xin(*), xout(*)
      integer(c_int) :: st
    end function
    function swm_op_laplacian_of_a_scalar(ctx, npad, nx, ny, nsdm, nweights, l_weights, xin, xout) &
        bind(C, name="swm_op_laplacian_of_a_scalar") result(st)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: npad, nx, ny, nsdm, nweights
      real(c_double) :: l_weights(*), xin(*), xout(*)
      integer(c_int) :: st
    end function
    function swm_op_k_cross_v(ctx, nx, ny, nsdm, nspace, ncorners, xin, xout) &
        bind(C, name="swm_op_k_cross_v") result(st)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: nx, ny, nsdm, nspace, ncorners
      real(c_double) :: xin(*), xout(*)
      integer(c_int) :: st
    end function
    function swm_op_gradient(ctx, npad, nx, ny, nsdm, ndata, nspace, ncorners, corner_area_inv, &
        v_weights, xin, xout) bind(C, name="swm_op_gradient") result(st)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: npad, nx, ny, nsdm, ndata, nspace, ncorners
      real(c_double) :: corner_area_inv(*), v_weights(*), xin(*), xout(*)
      integer(c_int) :: st
    end function
    function swm_op_average_to_cell_corners(ctx, npad, nx, ny, nsdm, ndata, ncorners, s_weights, &
        xin, xout, flag, min_value) bind(C, name="swm_op_average_to_cell_corners") result(st)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: npad, nx, ny, nsdm, ndata, ncorners, flag, min_value
      real(c_double) :: s_weights(*), xin(*), xout(*)
      integer(c_int) :: st
    end function
    function swm_op_calculate_center_kinetic_energy(ctx, npad, nx, ny, nsdm, ndata, nspace, &
        ncorners, sub_area, area_inv, ke_c, ke) bind(C, name="swm_op_calculate_center_kinetic_energy") result(st)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: npad, nx, ny, nsdm, ndata, nspace, ncorners
      real(c_double) :: sub_area(*), area_inv(*), ke_c(*), ke(*)
      integer(c_int) :: st
    end function
    function swm_op_calculate_velocity(ctx, npad, nx, ny, nsdm, ndata, nspace, ncorners, &
        corner_area_inv, v_weights, psi, chi, xout) bind(C, name="swm_op_calculate_velocity") result(st)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: npad, nx, ny, nsdm, ndata, nspace, ncorners
      real(c_double) :: corner_area_inv(*), v_weights(*), psi(*), chi(*), xout(*)
      integer(c_int) :: st
    end function
    function swm_op_centered_mass_flux(ctx, npad, nx, ny, nsdm, ndata, ncorners, nedges, nspace, r3, &
        dc, normal, mass_c, u, mass, flux) bind(C, name="swm_op_centered_mass_flux") result(st)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: npad, nx, ny, nsdm, ndata, ncorners, nedges, nspace, r3
      real(c_double) :: dc(*), normal(*), mass_c(*), u(*), mass(*), flux(*)
      integer(c_int) :: st
    end function
    function swm_op_upstream_mass_flux(ctx, npad, nx, ny, nsdm, ndata, ncorners, nedges, nspace, r3, &
        dc, normal, u, mass, flux) bind(C, name="swm_op_upstream_mass_flux") result(st)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: npad, nx, ny, nsdm, ndata, ncorners, nedges, nspace, r3
      real(c_double) :: dc(*), normal(*), u(*), mass(*), flux(*)
      integer(c_int) :: st
    end function
    function swm_op_centered_vorticity_flux(ctx, npad, nx, ny, nsdm, nedges, mass_flux, mass, xin, &
        xout) bind(C, name="swm_op_centered_vorticity_flux") result(st)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: npad, nx, ny, nsdm, nedges
      real(c_double) :: mass_flux(*), mass(*), xin(*), xout(*)
      integer(c_int) :: st
    end function
    function swm_op_centered_divergence_curl(ctx, npad, nx, ny, nsdm, ndata, ncorners, nedges, &
        nspace, r3, dc, tangent, mass_c, eta, u, mass, flux) &
        bind(C, name="swm_op_centered_divergence_curl") result(st)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: npad, nx, ny, nsdm, ndata, ncorners, nedges, nspace, r3
      real(c_double) :: dc(*), tangent(*), mass_c(*), eta(*), u(*), mass(*), flux(*)
      integer(c_int) :: st
    end function
  end interface
end module swm_mi355x_c

!=======================================================================
module swm_mi355x
  use iso_c_binding
  use swm_mi355x_c
  implicit none
  private
  public :: swm_gpu_initialize, swm_gpu_finalize, swm_gpu_check, swm_gpu_context, swm_gpu_grid, swm_gpu_grid_array, &
            step_dt, wrap_prognostics, sync_prognostics_to_host, elliptic, multigrid, wrap, wrap_face, &
            wrap_vrtx, wrap_edge, write_swm_restart, read_swm_restart, swm_gpu_time, initialize_wrap, &
            initialize_multigrid

  ! elliptic_solver.F:30 and multigrid_solver.F:54: the Helmholtz constant of
  ! del^2 x - helm x = rhs; the device solver is the Poisson solver the
  ! shallow-water step uses (0), so elliptic and multigrid stop on any other
  real(c_double), public, save :: helm_const = 0.0_c_double
  real(c_double), public, save :: helm_const_mg = 0.0_c_double

  type(c_ptr), save :: ctx = c_null_ptr
  type(c_ptr), save :: grid = c_null_ptr
  integer, save :: iim_ = 0, jjm_ = 0, ntracer_ = 0

contains

  ! status check: a failure stops the program, as abort_code does
  subroutine swm_gpu_check(st, where)
    integer(c_int), intent(in) :: st
    character(len=*), intent(in) :: where
    character(kind=c_char), pointer :: msg(:)
    character(len=512) :: text
    integer :: k
    if (st == SWM_OK) return
    call c_f_pointer(swm_last_error(), msg, [512])
    text = ''
    do k = 1, 512
      if (msg(k) == c_null_char) exit
      text(k:k) = msg(k)
    end do
    write(*, '(a,a,a,i0,a,a)') 'swm_mi355x: ', where, ' failed, status ', st, ': ', trim(text)
    stop 1
  end subroutine swm_gpu_check

  function swm_gpu_context() result(p)
    type(c_ptr) :: p
    p = ctx
  end function

  function swm_gpu_grid() result(p)
    type(c_ptr) :: p
    p = grid
  end function

  ! Replaces the device-relevant part of initialize_model (swm/source/
  ! initialize.F:82-140): builds the level-g grid (the reference's own cell
  ! centres may be passed in p_list(3, 10*4**g+2)), the device state and the
  ! initial fields h, hs (iim,jjm,10), u (2,2,iim,jjm,10), tracer
  ! (iim,jjm,ntracer,10).
  subroutine swm_gpu_initialize(g, ntracer, dt, grav, radius, omega, h, hs, u, tracer, p_list, device, &
                                l_fct, l_centered, l_upwind, l_remap)
    integer, intent(in) :: g, ntracer
    real(c_double), intent(in) :: dt, grav, radius, omega
    real(c_double), intent(in), target :: h(:,:,:), hs(:,:,:), u(:,:,:,:,:), tracer(:,:,:,:)
    real(c_double), intent(in), optional, target :: p_list(:,:)
    integer, intent(in), optional :: device
    logical, intent(in), optional :: l_fct, l_centered, l_upwind, l_remap   ! the namelist flags (default l_fct)
    type(swm_config) :: cfg
    type(c_ptr) :: pl
    integer(c_int) :: gg, mig
    pl = c_null_ptr
    if (present(p_list)) pl = c_loc(p_list)
    call swm_gpu_check(swm_grid_create(int(g, c_int), pl, radius, omega, grid), 'swm_grid_create')
    call swm_gpu_check(swm_grid_dims(grid, gg, iim_, jjm_, mig), 'swm_grid_dims')
    cfg%ntracer = ntracer
    cfg%dt = dt
    cfg%grav = grav
    cfg%precision = SWM_FP64
    cfg%device = 0
    if (present(device)) cfg%device = device
    cfg%advection = SWM_ADV_FCT
    if (present(l_centered)) then
      if (l_centered) cfg%advection = SWM_ADV_CENTERED
    end if
    if (present(l_upwind)) then
      if (l_upwind) cfg%advection = SWM_ADV_UPWIND
    end if
    if (present(l_remap)) then
      if (l_remap) cfg%advection = SWM_ADV_REMAP
    end if
    if (present(l_fct)) then
      if (l_fct) cfg%advection = SWM_ADV_FCT
    end if
    ntracer_ = ntracer
    call swm_gpu_check(swm_create(grid, cfg, ctx), 'swm_create')
    call swm_gpu_check(swm_init_state(ctx, h, c_loc(hs), u, tracer), 'swm_init_state')
  end subroutine swm_gpu_initialize

  ! a named set-up array of the run's grid (swm_grid_array), e.g. the
  ! metrics the operator subroutines take as arguments
  subroutine swm_gpu_grid_array(name, out)
    character(len=*), intent(in) :: name
    real(c_double), intent(out), target, contiguous :: out(..)
    call swm_gpu_check(swm_grid_array(grid, trim(name)//c_null_char, c_loc(out)), 'swm_grid_array '//name)
  end subroutine

  subroutine swm_gpu_finalize()
    if (c_associated(ctx)) call swm_destroy(ctx)
    if (c_associated(grid)) call swm_grid_destroy(grid)
    ctx = c_null_ptr
    grid = c_null_ptr
  end subroutine

  ! step::step_dt with get_pointers/get_weights of the driver loop
  ! (swm/source/driver.F:103-121): one time step on the device
  subroutine step_dt()
    call swm_gpu_check(swm_step(ctx, 1_c_int), 'step_dt')
  end subroutine step_dt

  ! step::wrap_prognostics: the device step leaves every halo filled
  subroutine wrap_prognostics()
  end subroutine wrap_prognostics

  ! current time level of the prognostic fields to host arrays (before
  ! output or restart writes of the driver)
  subroutine sync_prognostics_to_host(h, u, tracer, eta, div, psi, chi, ke)
    real(c_double), intent(out) :: h(:,:,:), u(:,:,:,:,:), tracer(:,:,:,:)
    real(c_double), intent(out), optional :: eta(:,:,:), div(:,:,:), psi(:,:,:), chi(:,:,:), ke(:,:,:)
    call swm_gpu_check(swm_get_field(ctx, 'h'//c_null_char, h), 'get h')
    call swm_gpu_check(swm_get_field(ctx, 'u'//c_null_char, u), 'get u')
    call swm_gpu_check(swm_get_field(ctx, 'tracer'//c_null_char, tracer), 'get tracer')
    if (present(eta)) call swm_gpu_check(swm_get_field(ctx, 'eta'//c_null_char, eta), 'get eta')
    if (present(div)) call swm_gpu_check(swm_get_field(ctx, 'div'//c_null_char, div), 'get div')
    if (present(psi)) call swm_gpu_check(swm_get_field(ctx, 'psi'//c_null_char, psi), 'get psi')
    if (present(chi)) call swm_gpu_check(swm_get_field(ctx, 'chi'//c_null_char, chi), 'get chi')
    if (present(ke)) call swm_gpu_check(swm_get_field(ctx, 'ke'//c_null_char, ke), 'get ke')
  end subroutine sync_prognostics_to_host

  ! swm_restart::write_swm_restart (swm/source/swm_restart.F:113-175): the
  ! device state to restarts/swm_restart.dDDDDDD.hh:mm:ss of the model time
  subroutine write_swm_restart()
    call swm_gpu_check(swm_write_restart(ctx, c_null_ptr), 'write_swm_restart')
  end subroutine write_swm_restart

  ! swm_restart::read_swm_restart (swm/source/swm_restart.F:25-81): the
  ! device state and model time from restarts/swm_restart; the surface
  ! height stays that of swm_gpu_initialize
  subroutine read_swm_restart()
    call swm_gpu_check(swm_read_restart(ctx, c_null_ptr), 'read_swm_restart')
  end subroutine read_swm_restart

  ! the model time in seconds (time_manager_module.F nsecond_dyn)
  function swm_gpu_time() result(t)
    integer(c_long_long) :: t
    t = swm_time_seconds(ctx)
  end function swm_gpu_time

  ! wrap_data::initialize_wrap (wrap_data.F:34-157) and multigrid_solver_xxx::
  ! initialize_multigrid (multigrid_solver.F:61-270): the halo tables and the
  ! multigrid hierarchy are built by swm_gpu_initialize, with the grid; these
  ! only check that it ran
  subroutine initialize_wrap(component_name)
    character(len=*), intent(in) :: component_name
    if (.not. c_associated(ctx)) call swm_gpu_check(7_c_int, 'initialize_wrap '//component_name)
  end subroutine initialize_wrap
  subroutine initialize_multigrid(component_name)
    character(len=*), intent(in) :: component_name
    if (.not. c_associated(ctx)) call swm_gpu_check(7_c_int, 'initialize_multigrid '//component_name)
  end subroutine initialize_multigrid

  subroutine check_poisson(where)
    character(len=*), intent(in) :: where
    if (helm_const /= 0.0_c_double .or. helm_const_mg /= 0.0_c_double) then
      write(*, '(a,a)') 'swm_mi355x: ', where//': helm_const must be 0 (the device solver is the Poisson solver)'
      stop 1
    end if
  end subroutine check_poisson

  ! elliptic_solver_xxx::elliptic (component_name is not needed)
  subroutine elliptic(component_name, nvar, rhs, xout)
    character(len=*), intent(in) :: component_name
    integer, intent(in) :: nvar
    real(c_double), intent(inout) :: rhs(:,:,:,:), xout(:,:,:,:)
    call check_poisson('elliptic')
    call swm_gpu_check(swm_elliptic(ctx, int(nvar, c_int), rhs, xout), 'elliptic '//component_name)
  end subroutine elliptic

  ! multigrid_solver_xxx::multigrid: one V-cycle
  subroutine multigrid(nvar, beta, alpha)
    integer, intent(in) :: nvar
    real(c_double), intent(in) :: beta(:,:,:,:)
    real(c_double), intent(inout) :: alpha(:,:,:,:)
    call check_poisson('multigrid')
    call swm_gpu_check(swm_multigrid(ctx, int(nvar, c_int), beta, alpha), 'multigrid')
  end subroutine multigrid

  ! wrap_data::wrap (wrap_data.F:1790-1974) with the reference's optional
  ! keyword arguments: every array present is halo-filled in place; with
  ! l_HA_stencil, the two edges next to the pentagon of edge_scalar are
  ! zeroed afterwards (every panel is next to one with 10 panels)
  subroutine wrap(component_name, l_HA_stencil, face, face_1L, vrtx, vrtx_1L, vrtx_scalar, vrtx_scalar_1L, &
                  edge, edge_1L, edge_scalar, edge_scalar_1L)
    character(len=*), intent(in) :: component_name
    logical, intent(in), optional :: l_HA_stencil
    real(c_double), optional, contiguous :: face(:,:,:,:), face_1L(:,:,:), vrtx(:,:,:,:,:,:), vrtx_1L(:,:,:,:,:), &
                                            vrtx_scalar(:,:,:,:,:), vrtx_scalar_1L(:,:,:,:), edge(:,:,:,:,:,:), &
                                            edge_1L(:,:,:,:,:), edge_scalar(:,:,:,:,:), edge_scalar_1L(:,:,:,:)
    character(len=64) :: w
    w = 'wrap '//component_name
    if (present(face)) call swm_gpu_check(swm_wrap(ctx, 0_c_int, 1_c_int, int(size(face, 3), c_int), face), w)
    if (present(face_1L)) call swm_gpu_check(swm_wrap(ctx, 0_c_int, 1_c_int, 1_c_int, face_1L), w)
    if (present(vrtx)) call swm_gpu_check(swm_wrap(ctx, 1_c_int, int(size(vrtx, 1), c_int), &
                                                   int(size(vrtx, 5), c_int), vrtx), w)
    if (present(vrtx_1L)) call swm_gpu_check(swm_wrap(ctx, 1_c_int, int(size(vrtx_1L, 1), c_int), 1_c_int, &
                                                      vrtx_1L), w)
    if (present(vrtx_scalar)) call swm_gpu_check(swm_wrap(ctx, 1_c_int, 1_c_int, &
                                                          int(size(vrtx_scalar, 4), c_int), vrtx_scalar), w)
    if (present(vrtx_scalar_1L)) call swm_gpu_check(swm_wrap(ctx, 1_c_int, 1_c_int, 1_c_int, vrtx_scalar_1L), w)
    if (present(edge)) call swm_gpu_check(swm_wrap(ctx, 2_c_int, int(size(edge, 1), c_int), &
                                                   int(size(edge, 5), c_int), edge), w)
    if (present(edge_1L)) call swm_gpu_check(swm_wrap(ctx, 2_c_int, int(size(edge_1L, 1), c_int), 1_c_int, &
                                                      edge_1L), w)
    if (present(edge_scalar)) call swm_gpu_check(swm_wrap(ctx, 2_c_int, 1_c_int, &
                                                          int(size(edge_scalar, 4), c_int), edge_scalar), w)
    if (present(edge_scalar_1L)) call swm_gpu_check(swm_wrap(ctx, 2_c_int, 1_c_int, 1_c_int, edge_scalar_1L), w)
    if (present(edge_scalar) .and. present(l_HA_stencil)) then
      if (l_HA_stencil) then
        edge_scalar(1, iim_, jjm_ - 1, :, :) = 0.0_c_double
        edge_scalar(3, iim_ - 1, jjm_, :, :) = 0.0_c_double
      end if
    end if
  end subroutine wrap

  ! wrap_data::wrap, one keyword at a time: face (iim,jjm,nlayer,10)
  subroutine wrap_face(x, nlayer)
    real(c_double), intent(inout) :: x(*)
    integer, intent(in) :: nlayer
    call swm_gpu_check(swm_wrap(ctx, 0_c_int, 1_c_int, int(nlayer, c_int), x), 'wrap face')
  end subroutine
  ! vrtx (nspace,2,iim,jjm,nlayer,10)
  subroutine wrap_vrtx(x, nspace, nlayer)
    real(c_double), intent(inout) :: x(*)
    integer, intent(in) :: nspace, nlayer
    call swm_gpu_check(swm_wrap(ctx, 1_c_int, int(nspace, c_int), int(nlayer, c_int), x), 'wrap vrtx')
  end subroutine
  ! edge (3,iim,jjm,nlayer,10)
  subroutine wrap_edge(x, nlayer)
    real(c_double), intent(inout) :: x(*)
    integer, intent(in) :: nlayer
    call swm_gpu_check(swm_wrap(ctx, 2_c_int, 1_c_int, int(nlayer, c_int), x), 'wrap edge')
  end subroutine
end module swm_mi355x

!=======================================================================
! operators.F under the reference's names and argument lists (one panel per
! call, as the reference calls them); the device context is the run's.
module operators_mi355x
  use iso_c_binding
  use swm_mi355x_c
  use swm_mi355x, only: swm_gpu_check, swm_gpu_context
  implicit none
  private
  public :: curl_of_a_vector, div_of_a_vector, div_of_an_edge_flux, laplacian_of_a_scalar, k_cross_v, &
            gradient, average_to_cell_corners, calculate_center_kinetic_energy, calculate_velocity, &
            centered_mass_flux, upstream_mass_flux, centered_vorticity_flux, centered_divergence_curl

contains

  subroutine curl_of_a_vector(npad, nx, ny, ndata, nspace, ncorners, area_inv, v_weights, xin, xout)
    integer, intent(in) :: npad, nx, ny, ndata, nspace, ncorners
    real(c_double), intent(in) :: area_inv(nx,ny), v_weights(ndata,nspace,ncorners,nx,ny), &
                                  xin(nspace,ncorners,nx,ny)
    real(c_double), intent(out) :: xout(nx,ny)
    call swm_gpu_check(swm_op_curl_of_a_vector(swm_gpu_context(), npad, nx, ny, 1, ndata, nspace, ncorners, &
                       area_inv, v_weights, xin, xout), 'curl_of_a_vector')
  end subroutine

  subroutine div_of_a_vector(npad, nx, ny, ndata, nspace, ncorners, area_inv, v_weights, xin, xout)
    integer, intent(in) :: npad, nx, ny, ndata, nspace, ncorners
    real(c_double), intent(in) :: area_inv(nx,ny), v_weights(ndata,nspace,ncorners,nx,ny), &
                                  xin(nspace,ncorners,nx,ny)
    real(c_double), intent(out) :: xout(nx,ny)
    call swm_gpu_check(swm_op_div_of_a_vector(swm_gpu_context(), npad, nx, ny, 1, ndata, nspace, ncorners, &
                       area_inv, v_weights, xin, xout), 'div_of_a_vector')
  end subroutine

  subroutine div_of_an_edge_flux(npad, nx, ny, nedges, area_inv, xin, xout)
    integer, intent(in) :: npad, nx, ny, nedges
    real(c_double), intent(in) :: area_inv(nx,ny), xin(nedges,nx,ny)
    real(c_double), intent(out) :: xout(nx,ny)
    call swm_gpu_check(swm_op_div_of_an_edge_flux(swm_gpu_context(), npad, nx, ny, 1, nedges, area_inv, xin, &
                       xout), 'div_of_an_edge_flux')
  end subroutine

  subroutine laplacian_of_a_scalar(npad, nx, ny, nweights, l_weights, xin, xout)
    integer, intent(in) :: npad, nx, ny, nweights
    real(c_double), intent(in) :: l_weights(nweights,nx,ny), xin(nx,ny)
    real(c_double), intent(out) :: xout(nx,ny)
    call swm_gpu_check(swm_op_laplacian_of_a_scalar(swm_gpu_context(), npad, nx, ny, 1, nweights, l_weights, &
                       xin, xout), 'laplacian_of_a_scalar')
  end subroutine

  subroutine k_cross_v(nx, ny, nspace, ncorners, xin, xout)
    integer, intent(in) :: nx, ny, nspace, ncorners
    real(c_double), intent(in) :: xin(nspace,ncorners,nx,ny)
    real(c_double), intent(out) :: xout(nspace,ncorners,nx,ny)
    call swm_gpu_check(swm_op_k_cross_v(swm_gpu_context(), nx, ny, 1, nspace, ncorners, xin, xout), 'k_cross_v')
  end subroutine

  subroutine gradient(npad, nx, ny, ndata, nspace, ncorners, corner_area_inv, v_weights, xin, xout)
    integer, intent(in) :: npad, nx, ny, ndata, nspace, ncorners
    real(c_double), intent(in) :: corner_area_inv(ncorners,nx,ny), v_weights(ndata,nspace,ncorners,nx,ny), &
                                  xin(nx,ny)
    real(c_double), intent(out) :: xout(nspace,ncorners,nx,ny)
    call swm_gpu_check(swm_op_gradient(swm_gpu_context(), npad, nx, ny, 1, ndata, nspace, ncorners, &
                       corner_area_inv, v_weights, xin, xout), 'gradient')
  end subroutine

  subroutine average_to_cell_corners(npad, nx, ny, ndata, ncorners, s_weights, xin, xout, flag, min_value)
    integer, intent(in) :: npad, nx, ny, ndata, ncorners
    real(c_double), intent(in) :: s_weights(ndata,ncorners,nx,ny), xin(nx,ny)
    real(c_double), intent(out) :: xout(ncorners,nx,ny)
    logical :: flag
    logical, optional :: min_value
    integer(c_int) :: f, mv
    f = merge(1, 0, flag)
    mv = -1
    if (present(min_value)) mv = merge(1, 0, min_value)
    call swm_gpu_check(swm_op_average_to_cell_corners(swm_gpu_context(), npad, nx, ny, 1, ndata, ncorners, &
                       s_weights, xin, xout, f, mv), 'average_to_cell_corners')
  end subroutine

  subroutine calculate_center_kinetic_energy(npad, nx, ny, ndata, nspace, ncorners, sub_area, area_inv, ke_c, ke)
    integer, intent(in) :: npad, nx, ny, ndata, nspace, ncorners
    real(c_double), intent(in) :: sub_area(ndata,ncorners,nx,ny), area_inv(nx,ny), ke_c(ncorners,nx,ny)
    real(c_double), intent(out) :: ke(nx,ny)
    call swm_gpu_check(swm_op_calculate_center_kinetic_energy(swm_gpu_context(), npad, nx, ny, 1, ndata, nspace, &
                       ncorners, sub_area, area_inv, ke_c, ke), 'calculate_center_kinetic_energy')
  end subroutine

  subroutine calculate_velocity(npad, nx, ny, ndata, nspace, ncorners, corner_area_inv, v_weights, psi, chi, xout)
    integer, intent(in) :: npad, nx, ny, ndata, nspace, ncorners
    real(c_double), intent(in) :: corner_area_inv(ncorners,nx,ny), v_weights(ndata,nspace,ncorners,nx,ny), &
                                  psi(nx,ny), chi(nx,ny)
    real(c_double), intent(out) :: xout(nspace,ncorners,nx,ny)
    call swm_gpu_check(swm_op_calculate_velocity(swm_gpu_context(), npad, nx, ny, 1, ndata, nspace, ncorners, &
                       corner_area_inv, v_weights, psi, chi, xout), 'calculate_velocity')
  end subroutine

  subroutine centered_mass_flux(npad, nx, ny, ndata, ncorners, nedges, nspace, r3, dc, normal, mass_c, u, mass, flux)
    integer, intent(in) :: npad, nx, ny, ndata, ncorners, nedges, nspace, r3
    real(c_double), intent(in) :: dc(ndata,ncorners,nx,ny), normal(ndata,r3,ncorners,nx,ny), &
                                  mass_c(ncorners,nx,ny), u(nspace,ncorners,nx,ny), mass(nx,ny)
    real(c_double), intent(out) :: flux(nedges,nx,ny)
    call swm_gpu_check(swm_op_centered_mass_flux(swm_gpu_context(), npad, nx, ny, 1, ndata, ncorners, nedges, &
                       nspace, r3, dc, normal, mass_c, u, mass, flux), 'centered_mass_flux')
  end subroutine

  subroutine upstream_mass_flux(npad, nx, ny, ndata, ncorners, nedges, nspace, r3, dc, normal, u, mass, flux)
    integer, intent(in) :: npad, nx, ny, ndata, ncorners, nedges, nspace, r3
    real(c_double), intent(in) :: dc(ndata,ncorners,nx,ny), normal(ndata,r3,ncorners,nx,ny), &
                                  u(nspace,ncorners,nx,ny), mass(nx,ny)
    real(c_double), intent(out) :: flux(nedges,nx,ny)
    call swm_gpu_check(swm_op_upstream_mass_flux(swm_gpu_context(), npad, nx, ny, 1, ndata, ncorners, nedges, &
                       nspace, r3, dc, normal, u, mass, flux), 'upstream_mass_flux')
  end subroutine

  subroutine centered_vorticity_flux(npad, nx, ny, nedges, mass_flux, mass, xin, xout)
    integer, intent(in) :: npad, nx, ny, nedges
    real(c_double), intent(in) :: mass_flux(nedges,nx,ny), mass(nx,ny), xin(nx,ny)
    real(c_double), intent(out) :: xout(nedges,nx,ny)
    call swm_gpu_check(swm_op_centered_vorticity_flux(swm_gpu_context(), npad, nx, ny, 1, nedges, mass_flux, &
                       mass, xin, xout), 'centered_vorticity_flux')
  end subroutine

  subroutine centered_divergence_curl(npad, nx, ny, ndata, ncorners, nedges, nspace, r3, dc, tangent, mass_c, &
                                      eta, u, mass, flux)
    integer, intent(in) :: npad, nx, ny, ndata, ncorners, nedges, nspace, r3
    real(c_double), intent(in) :: dc(ndata,ncorners,nx,ny), tangent(ndata,r3,ncorners,nx,ny), &
                                  mass_c(ncorners,nx,ny), eta(nx,ny), u(nspace,ncorners,nx,ny), mass(nx,ny)
    real(c_double), intent(out) :: flux(nedges,nx,ny)
    call swm_gpu_check(swm_op_centered_divergence_curl(swm_gpu_context(), npad, nx, ny, 1, ndata, ncorners, &
                       nedges, nspace, r3, dc, tangent, mass_c, eta, u, mass, flux), 'centered_divergence_curl')
  end subroutine
end module operators_mi355x

!=======================================================================
! swm/source/swm_qp_module.F under its own names and arguments: the
! quick-plot netCDF file of the run, qp_output/hswm_dDDDDDDp000.g2.nc (the
! directory must exist), with the coordinates the caller passes, the fields
! packed from the device state's current time level
module swm_qp_module_mi355x
  use iso_c_binding
  use swm_mi355x_c
  use swm_mi355x, only: swm_gpu_check, swm_gpu_context, swm_gpu_grid
  implicit none
  private
  public :: initialize_swm_qp, write_swm_qp, pack_swm_qp
  type(c_ptr), save :: qp = c_null_ptr
contains
  ! swm_qp_module.F:70-214
  subroutine initialize_swm_qp(simulation_tag, nsecond, grid_center_lat, grid_center_lon, grid_corner_lat, &
                               grid_corner_lon)
    character(len=*), intent(in) :: simulation_tag
    integer(c_long_long), intent(in) :: nsecond
    real(c_double), intent(in) :: grid_center_lat(:,:,:), grid_center_lon(:,:,:)
    real(c_double), intent(in) :: grid_corner_lat(:,:,:,:), grid_corner_lon(:,:,:,:)
    if (c_associated(qp)) call swm_qp_destroy(qp)
    qp = c_null_ptr
    call swm_gpu_check(swm_qp_create_coords(swm_gpu_grid(), trim(simulation_tag)//c_null_char, nsecond, &
                                            c_null_ptr, grid_center_lat, grid_center_lon, grid_corner_lat, &
                                            grid_corner_lon, qp), 'initialize_swm_qp')
  end subroutine initialize_swm_qp

  ! swm_qp_module.F:309-370
  subroutine pack_swm_qp()
    call swm_gpu_check(swm_qp_pack(qp, swm_gpu_context()), 'pack_swm_qp')
  end subroutine pack_swm_qp

  ! swm_qp_module.F:218-300
  subroutine write_swm_qp(nsecond)
    integer(c_long_long), intent(in) :: nsecond
    call swm_gpu_check(swm_qp_write(qp, nsecond), 'write_swm_qp')
  end subroutine write_swm_qp
end module swm_qp_module_mi355x

!=======================================================================
! swm/source/swm_dissipation_operator.F under its own names and arguments
! (init_diss_op :50-203, dissipation :212-443) on the run's grid
module swm_dissipation_operator_mi355x
  use iso_c_binding
  use swm_mi355x_c
  use swm_mi355x, only: swm_gpu_check, swm_gpu_context, swm_gpu_grid
  implicit none
  private
  public :: init_diss_op, dissipation
contains
  subroutine init_diss_op(component_name)
    character(len=*), intent(in) :: component_name
    call swm_gpu_check(swm_init_diss_op(swm_gpu_context(), swm_gpu_grid()), 'init_diss_op '//component_name)
  end subroutine init_diss_op

  subroutine dissipation(component_name, norder, visc, xin, xout)
    character(len=*), intent(in) :: component_name
    integer, intent(in) :: norder
    real(c_double), intent(in) :: visc
    real(c_double), intent(in) :: xin(:,:,:,:,:)
    real(c_double), intent(out) :: xout(:,:,:,:,:)
    call swm_gpu_check(swm_op_dissipation(swm_gpu_context(), int(norder, c_int), visc, xin, xout), &
                       'dissipation '//component_name)
  end subroutine dissipation
end module swm_dissipation_operator_mi355x

!=======================================================================
! utilities/layer_2/remap.F under its own names and arguments: init_remap
! (:176-1063; the quantities are built with the device model when
! swm_gpu_initialize is given l_remap) and remap_mass_and_tracers
! (:1072-1604).  Of the optional arguments the tracer limiter's
! lin_limit_tracers, lin_limit_tracers_min and lin_limit_tracers_max are
! taken (both switches on when absent, as l_limit_tracers = .true. intends);
! the checks lin_check_* only print in the reference and are not made; the
! others stop the program unless they hold get_remap_flux's values (no mask,
! independent tracers, linear integrals, first-order departure points, mass
! gradients unlimited, no positive-definite-only limits).
module remap_mi355x
  use iso_c_binding
  use swm_mi355x_c
  use swm_mi355x, only: swm_gpu_check, swm_gpu_context
  implicit none
  private
  public :: init_remap, remap_mass_and_tracers
contains
  subroutine init_remap(component_name)
    character(len=*), intent(in) :: component_name
    if (.not. c_associated(swm_gpu_context())) call swm_gpu_check(7_c_int, 'init_remap '//component_name)
  end subroutine init_remap

  subroutine remap_mass_and_tracers(component_name, nmass, ntrace, mass, trace, dt, uc, pmask_in, depend_in, &
                                    integral_order_in, lin_dp2ord, lin_check_conservation, &
                                    lin_check_monotonicity, lin_check_divergence, lin_check_velocity, &
                                    lin_mask_gradient, lin_limit_mass, lin_limit_mass_min, lin_limit_mass_max, &
                                    lin_mass_pd, lin_limit_tracers, lin_limit_tracers_min, lin_limit_tracers_max, &
                                    lin_tracers_pd, edgeflux)
    character(len=*), intent(in) :: component_name
    integer, intent(in) :: nmass, ntrace
    real(c_double), intent(inout) :: mass(:,:,:,:), trace(:,:,:,:,:)
    real(c_double), intent(in) :: dt
    real(c_double), intent(in) :: uc(:,:,:,:,:)
    real(c_double), intent(in), optional :: pmask_in(:,:,:)
    integer, intent(in), optional :: depend_in(:), integral_order_in
    logical, intent(in), optional :: lin_dp2ord, lin_check_conservation, lin_check_monotonicity, &
        lin_check_divergence, lin_check_velocity, lin_mask_gradient, lin_limit_mass, lin_limit_mass_min, &
        lin_limit_mass_max, lin_mass_pd, lin_limit_tracers, lin_limit_tracers_min, lin_limit_tracers_max, &
        lin_tracers_pd
    real(c_double), intent(out), optional, target :: edgeflux(:,:,:,:,:)
    integer(c_int) :: lmin, lmax
    type(c_ptr) :: ef
    logical :: off
    off = .false.
    if (present(pmask_in)) off = off .or. any(pmask_in /= 1.0_c_double)
    if (present(depend_in)) off = off .or. any(depend_in /= 0)
    if (present(integral_order_in)) off = off .or. integral_order_in /= 1
    if (present(lin_dp2ord)) off = off .or. lin_dp2ord
    if (present(lin_mask_gradient)) off = off .or. lin_mask_gradient
    if (present(lin_limit_mass)) off = off .or. lin_limit_mass
    if (present(lin_mass_pd)) off = off .or. lin_mass_pd
    if (present(lin_tracers_pd)) off = off .or. lin_tracers_pd
    if (off) then
      write(*, '(a)') 'swm_mi355x: remap_mass_and_tracers '//component_name// &
          ': only the optional arguments get_remap_flux uses are supported'
      stop 1
    end if
    lmin = 1
    lmax = 1
    if (present(lin_limit_tracers_min)) lmin = merge(1_c_int, 0_c_int, lin_limit_tracers_min)
    if (present(lin_limit_tracers_max)) lmax = merge(1_c_int, 0_c_int, lin_limit_tracers_max)
    if (present(lin_limit_tracers)) then
      if (.not. lin_limit_tracers) then
        lmin = 0
        lmax = 0
      end if
    end if
    call swm_gpu_check(swm_set_remap_limiter(swm_gpu_context(), lmin, lmax), 'remap_mass_and_tracers '//component_name)
    ef = c_null_ptr
    if (present(edgeflux)) ef = c_loc(edgeflux)
    call swm_gpu_check(swm_op_remap_mass_and_tracers(swm_gpu_context(), int(size(mass, 1), c_int), &
                       int(size(mass, 2), c_int), int(nmass, c_int), int(size(mass, 4), c_int), int(ntrace, c_int), &
                       mass, trace, dt, uc, ef), 'remap_mass_and_tracers '//component_name)
  end subroutine remap_mass_and_tracers
end module remap_mi355x

!=======================================================================
! global_reductions.F:77-98 global_real_sum under its own name and
! arguments: sum(X * MASK), on the device with a compensated final sum
module global_reductions_mi355x
  use iso_c_binding
  use swm_mi355x_c
  use swm_mi355x, only: swm_gpu_check, swm_gpu_context
  implicit none
  private
  public :: global_real_sum
contains
  real(c_double) function global_real_sum(X, MASK, size, component_name)
    integer, intent(in) :: size
    character(len=*), intent(in) :: component_name
    real(c_double), intent(in) :: X(size), MASK(size)
    real(c_double) :: out
    call swm_gpu_check(swm_global_real_sum(swm_gpu_context(), X, MASK, int(size, c_int64_t), out), &
                       'global_real_sum '//component_name)
    global_real_sum = out
  end function global_real_sum
end module global_reductions_mi355x

!=======================================================================
! mass_fct.F:23-539 set_mass_fct under its own name and arguments (the whole
! sphere of the run's grid; the optional mss_flux_forced and mask, which
! step.F never passes, stop the program)
module mass_fct_mi355x
  use iso_c_binding
  use swm_mi355x_c
  use swm_mi355x, only: swm_gpu_check, swm_gpu_context
  implicit none
  private
  public :: set_mass_fct
contains
  subroutine set_mass_fct(component_name, iim, jjm, nlm, nsdm, ntprog, nttend, ncorners, ndata, nedges, nspace, &
                          npad, r3, n3, n4, tend_index, tend_weights, dt, area_inv, area, dc, normal, mss, mss_c, u, &
                          mss_low, mss_flux_horz, msh_f, mss_flux_forced, mask)
    character(len=*), intent(in) :: component_name
    integer, intent(in) :: iim, jjm, nlm, nsdm, ntprog, nttend, ncorners, ndata, nedges, nspace, npad, r3, n3, n4
    integer, intent(in) :: tend_index(nttend)
    real(c_double), intent(in) :: dt, tend_weights(nttend)
    real(c_double), intent(in) :: area_inv(iim,jjm,nsdm), area(iim,jjm,nsdm), &
                                  normal(ndata,r3,ncorners,iim,jjm,nsdm), dc(ndata,ncorners,iim,jjm,nsdm)
    real(c_double), intent(inout) :: mss(iim,jjm,nlm,ntprog,nsdm)
    real(c_double), intent(in) :: mss_c(ncorners,iim,jjm,nlm,nsdm), u(nspace,ncorners,iim,jjm,nlm,nsdm)
    real(c_double), intent(out) :: mss_low(iim,jjm,nlm,nsdm), mss_flux_horz(nedges,iim,jjm,nlm,nsdm)
    real(c_double), intent(inout) :: msh_f(nedges,iim,jjm,nlm,nttend,nsdm)
    real(c_double), intent(in), optional :: mss_flux_forced(nedges,iim,jjm,nlm,nsdm), mask(nedges,iim,jjm,nlm,nsdm)
    integer(c_int) :: ti(3)
    if (present(mss_flux_forced) .or. present(mask)) &
      call swm_gpu_check(1_c_int, 'set_mass_fct '//component_name//' (mss_flux_forced, mask)')
    ti = int(tend_index(1:3), c_int)
    call swm_gpu_check(swm_op_set_mass_fct(swm_gpu_context(), iim, jjm, nlm, nsdm, ntprog, nttend, ncorners, &
                       ndata, nedges, nspace, npad, r3, n3, n4, ti, tend_weights, dt, area_inv, area, dc, normal, &
                       mss, mss_c, u, mss_low, mss_flux_horz, msh_f, c_null_ptr, c_null_ptr), &
                       'set_mass_fct '//component_name)
  end subroutine set_mass_fct
end module mass_fct_mi355x

!=======================================================================
! tracer_fct.F:22-670 set_tracer_fct under its own name and arguments (the
! whole sphere; the optional mask, which step.F never passes, stops the
! program)
module tracer_fct_mi355x
  use iso_c_binding
  use swm_mi355x_c
  use swm_mi355x, only: swm_gpu_check, swm_gpu_context
  implicit none
  private
  public :: set_tracer_fct
contains
  subroutine set_tracer_fct(component_name, iim, jjm, nlm, nsdm, nedges, dt, area_inv, area, trc, trc_high, mss, &
                            mss_low, mss_horz, mss_flux_horz, trc_fct, trc_flux_horz, msstrc_low_order, mask)
    character(len=*), intent(in) :: component_name
    integer, intent(in) :: iim, jjm, nlm, nsdm, nedges
    real(c_double), intent(in) :: dt
    real(c_double), intent(in) :: area_inv(iim,jjm,nsdm), area(iim,jjm,nsdm)
    real(c_double), intent(in) :: trc(iim,jjm,nlm,nsdm), trc_high(iim,jjm,nlm,nsdm), mss(iim,jjm,nlm,nsdm), &
                                  mss_low(iim,jjm,nlm,nsdm), mss_horz(iim,jjm,nlm,nsdm), &
                                  mss_flux_horz(nedges,iim,jjm,nlm,nsdm)
    real(c_double), intent(out) :: trc_fct(iim,jjm,nlm,nsdm)
    real(c_double), intent(out), optional, target :: trc_flux_horz(nedges,iim,jjm,nlm,nsdm), &
                                                     msstrc_low_order(iim,jjm,nlm,nsdm)
    real(c_double), intent(in), optional :: mask(nedges,iim,jjm,nlm,nsdm)
    type(c_ptr) :: pf, pl
    if (present(mask)) call swm_gpu_check(1_c_int, 'set_tracer_fct '//component_name//' (mask)')
    pf = c_null_ptr
    pl = c_null_ptr
    if (present(trc_flux_horz)) pf = c_loc(trc_flux_horz)
    if (present(msstrc_low_order)) pl = c_loc(msstrc_low_order)
    call swm_gpu_check(swm_op_set_tracer_fct(swm_gpu_context(), iim, jjm, nlm, nsdm, nedges, dt, area_inv, area, &
                       trc, trc_high, mss, mss_low, mss_horz, mss_flux_horz, trc_fct, pf, pl, c_null_ptr), &
                       'set_tracer_fct '//component_name)
  end subroutine set_tracer_fct
end module tracer_fct_mi355x
