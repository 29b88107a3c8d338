! ref_harness.F90 -- TEST INFRASTRUCTURE ONLY (part of the oracle).
!
! A driver program of our own that links the reference shallow-water model's
! Fortran modules, compiled unmodified from where they lie (see
! oracle/Makefile.ref), and writes golden vectors for the parity tests.
! Nothing in the shipped library links, loads or calls this program.
!
! It repeats the setup order of the reference initialize_model
! (swm/source/initialize.F:67-278) without its I/O, the Coriolis
! set-up of set_coriolis_parameter (initialize.F:286-321), the IC flow of
! initial_conditions_swtc5 (initialize.F:330-513) with caller-supplied
! fields, and the time loop of program swm (swm/source/driver.F:90-166).
!
! usage: ref_harness <mode> <outdir> [args]
!   setup  : grid metrics, masks, index maps          -> outdir/setup.bin
!   wrap   : halo-exchange probes (unique-id fields)  -> outdir/wrap.bin
!   ops    : one application of each operator         -> outdir/ops.bin
!   mg     : one V-cycle and one elliptic solve       -> outdir/mg.bin
!   diss   : init_diss_op, then dissipation of a pseudo-random corner
!            vector field with norder 1 and 2         -> outdir/diss.bin
!   run    : <icfile> <dt> <nsteps> <ndump> [scheme] [budgets] [wrestart] [rrestart]
!            integrates nsteps, dumping the state every ndump steps
!            (outdir/run.bin); scheme fct | centered | upwind | remap
!            (remap: l_remap with init_remap first, initialize.F:216); with the
!            word budgets, the reference's compute_budgets runs at step 0
!            and at every dump (its records go to unit 31, fort.31); with
!            wrestart, write_swm_restart runs after the last step (file
!            restarts/swm_restart.dDDDDDD.hh:mm:ss); with rrestart,
!            read_swm_restart replaces h, u and tracers of icfile by those
!            of restarts/swm_restart after the initial state is set, as
!            initialize.F:222-231 does (swm/source/swm_restart.F)
!
! All files are raw little-endian stream records of the Fortran arrays in
! column-major order; the Python side reshapes them with order='F'.

program ref_harness
  use budgets, only: compute_budgets
  use remap_xxx, only: init_remap
  use swm_dissipation_operator, only: init_diss_op, dissipation
  use swm_restart, only: read_swm_restart, write_swm_restart
  use kinds
  use params_physical
  use params_grid_horizontal_xxx
  use params_parallel_xxx
  use component_intra
  use assign_owner
  use wrap_data
  use geodesic_grid_utilities
  use grid_metrics_variables_xxx
  use grid_metrics
  use grid_mask_variables_xxx
  use grid_mask
  use operators
  use operators_hl_xxx
  use multigrid_solver_xxx
  use elliptic_solver_xxx
  use swm_parameters
  use prognostics
  use time_manager_module
  use model_control
  use step
  implicit none

  character(len=512) :: mode, outdir, arg
  integer :: fu

  call get_command_argument(1, mode)
  call get_command_argument(2, outdir)

  call setup_wrap()
  if (trim(mode) /= 'wrap') call harness_setup()

  select case (trim(mode))
  case ('setup')
    call dump_setup()
  case ('wrap')
    call dump_wrap()
  case ('ops')
    call dump_ops()
  case ('diss')
    call dump_diss()
  case ('mg')
    call dump_mg()
  case ('run')
    call do_run()
  case default
    print *, 'unknown mode ', trim(mode)
    stop 2
  end select

contains

  subroutine open_out(name)
    character(len=*), intent(in) :: name
    open(newunit=fu, file=trim(outdir)//'/'//name, access='stream', &
         form='unformatted', status='replace')
  end subroutine open_out

  ! setup order of initialize_model (initialize.F:82-140), minus output;
  ! the halo tables need no grid file
  subroutine setup_wrap()
    call init_swm_parameters()
    call initialize_component(swm_name)
    call initialize_params_horizontal_xxx()
    call set_component_horizontal(swm_name, max_level, even_odd_factor, &
         max_ig, im, jm, iim, jjm, isdm, jsdm)
    call set_component_parallel(swm_name, npe, my_task, 0, 0, 0)
    call set_assign_owner(isdm, jsdm, max_level, max_level, &
         partition_factor, swm_name, './assign_file')
    call initialize_wrap(swm_name)
  end subroutine setup_wrap

  subroutine harness_setup()
    integer :: i, j, ksdm
    real(kind=dbl_kind) :: s(2)

    call set_grid_center_mask(swm_name, im, jm, iim, jjm, nsdm, &
         grid_center_mask, grid_center_index, polygon_type)
    call set_grid_corner_mask(swm_name, ncorners, iim, jjm, nsdm, &
         grid_center_index, grid_corner_mask, grid_corner_index)
    call generate_grid_metrics(swm_name, iim, jjm, nsdm, im, jm, max_ig, &
         max_level, ndata, ncorners, nspace, p, c, w, area, area_inv, &
         corner_area, corner_area_inv, sub_area, sub_area_inv, dc, normal, &
         tangent, v_weights, s_weights, l_weights, trans_to_X2, trans_to_X1, &
         grid_center_lon, grid_center_lat, grid_corner_lon, grid_corner_lat, &
         grid_center_centroid, grid_corner_centroid, tot_area, inv_tot_area)
    call initialize_multigrid(swm_name)
    call init_prognostics()
    ! Coriolis, as set_coriolis_parameter (initialize.F:286-321)
    do ksdm = 1, nsdm
      do j = 1, jjm
        do i = 1, iim
          if (grid_center_mask(i,j,ksdm) .eq. 1.0_dbl_kind) then
            call xyz_to_lonlat(p(:,i,j,ksdm), s(:))
            f(i,j,ksdm) = c2*sin(s(2))*omega
          endif
        enddo
      enddo
    enddo
    call wrap(swm_name, face_1L=f(:,:,:))
    call init_time_manager()
  end subroutine harness_setup

  subroutine dump_setup()
    integer :: hdr(8)
    hdr = (/ max_level, iim, jjm, nsdm, max_ig, im, jm, 0 /)
    call open_out('setup.bin')
    write(fu) hdr
    write(fu) tot_area
    write(fu) p, c, area, area_inv, corner_area, corner_area_inv
    write(fu) sub_area, dc, normal, tangent, v_weights, s_weights, l_weights
    write(fu) f, grid_center_mask, grid_corner_mask, polygon_type
    write(fu) grid_center_index, grid_corner_index
    write(fu) w, trans_to_X1, trans_to_X2
    close(fu)
  end subroutine dump_setup

  ! every slot and component holds a unique id; after the wrap each slot
  ! holds the id of the slot it was copied from (or 0 / its own id)
  subroutine dump_wrap()
    real(kind=dbl_kind), allocatable :: xf(:,:,:,:), xv(:,:,:,:,:,:), xe(:,:,:,:,:,:)
    integer :: n, i
    allocate(xf(iim,jjm,1,nsdm), xv(2,2,iim,jjm,1,nsdm), xe(1,3,iim,jjm,1,nsdm))
    n = size(xf)
    xf = reshape((/ (real(i,dbl_kind), i=1,n) /), shape(xf))
    n = size(xv)
    xv = reshape((/ (real(i,dbl_kind), i=1,n) /), shape(xv))
    n = size(xe)
    xe = reshape((/ (real(i,dbl_kind), i=1,n) /), shape(xe))
    call wrap(swm_name, face=xf)
    call wrap(swm_name, vrtx=xv)
    call wrap(swm_name, edge=xe)
    call open_out('wrap.bin')
    write(fu) xf, xv, xe
    close(fu)
  end subroutine dump_wrap

  ! deterministic pseudo-random numbers in [-1,1) (64-bit LCG)
  real(kind=dbl_kind) function rnd(state)
    integer(kind=8), intent(inout) :: state
    state = state*6364136223846793005_8 + 1442695040888963407_8
    rnd = real(ishft(state, -11), dbl_kind) / 4503599627370496.0_dbl_kind - 1.0_dbl_kind
  end function rnd

  subroutine fill(x, n, state)
    integer, intent(in) :: n
    real(kind=dbl_kind), intent(out) :: x(n)
    integer(kind=8), intent(inout) :: state
    integer :: i
    do i = 1, n
      x(i) = rnd(state)
    enddo
  end subroutine fill

  subroutine dump_ops()
    real(kind=dbl_kind), allocatable :: s1(:,:,:), s2(:,:,:), e1(:,:,:,:), vc(:,:,:,:,:), &
         cc(:,:,:,:), o1(:,:,:), o2(:,:,:), oe(:,:,:,:), ov(:,:,:,:,:), oc(:,:,:,:), &
         oe2(:,:,:,:), oe3(:,:,:,:), oe4(:,:,:,:), o3(:,:,:), o4(:,:,:), o5(:,:,:), oc2(:,:,:,:)
    integer(kind=8) :: st
    integer :: k
    allocate(s1(iim,jjm,nsdm), s2(iim,jjm,nsdm), e1(3,iim,jjm,nsdm), &
         vc(2,2,iim,jjm,nsdm), cc(2,iim,jjm,nsdm), o1(iim,jjm,nsdm), o2(iim,jjm,nsdm), &
         o3(iim,jjm,nsdm), o4(iim,jjm,nsdm), o5(iim,jjm,nsdm), &
         oe(3,iim,jjm,nsdm), oe2(3,iim,jjm,nsdm), oe3(3,iim,jjm,nsdm), oe4(3,iim,jjm,nsdm), &
         ov(2,2,iim,jjm,nsdm), oc(2,iim,jjm,nsdm), oc2(2,iim,jjm,nsdm))
    st = 20261203_8
    call fill(s1, size(s1), st)
    call fill(s2, size(s2), st)
    s2 = s2 + 2.0_dbl_kind          ! positive "mass"
    call fill(e1, size(e1), st)
    call fill(vc, size(vc), st)
    call fill(cc, size(cc), st)
    cc = cc + 2.0_dbl_kind
    do k = 1, nsdm
      call div_of_an_edge_flux(npad, iim, jjm, nedges, area_inv(:,:,k), e1(:,:,:,k), o1(:,:,k))
      call laplacian_of_a_scalar(npad, iim, jjm, nweights, l_weights(:,:,:,k), s1(:,:,k), o2(:,:,k))
      call curl_of_a_vector(npad, iim, jjm, ndata, nspace, ncorners, area_inv(:,:,k), &
           v_weights(:,:,:,:,:,k), vc(:,:,:,:,k), o3(:,:,k))
      call div_of_a_vector(npad, iim, jjm, ndata, nspace, ncorners, area_inv(:,:,k), &
           v_weights(:,:,:,:,:,k), vc(:,:,:,:,k), o4(:,:,k))
      call calculate_center_kinetic_energy(npad, iim, jjm, ndata, nspace, ncorners, &
           sub_area(:,:,:,:,k), area_inv(:,:,k), cc(:,:,:,k), o5(:,:,k))
      call average_to_cell_corners(npad, iim, jjm, ndata, ncorners, s_weights(:,:,:,:,k), &
           s1(:,:,k), oc(:,:,:,k), .true.)
      ov(:,:,:,:,k) = 0.0_dbl_kind
      call calculate_velocity(npad, iim, jjm, ndata, nspace, ncorners, corner_area_inv(:,:,:,k), &
           v_weights(:,:,:,:,:,k), s1(:,:,k), s2(:,:,k), ov(:,:,:,:,k))
      call centered_mass_flux(npad, iim, jjm, ndata, ncorners, nedges, nspace, r3, &
           dc(:,:,:,:,k), normal(:,:,:,:,:,k), cc(:,:,:,k), vc(:,:,:,:,k), s2(:,:,k), oe(:,:,:,k))
      call upstream_mass_flux(npad, iim, jjm, ndata, ncorners, nedges, nspace, r3, &
           dc(:,:,:,:,k), normal(:,:,:,:,:,k), vc(:,:,:,:,k), s2(:,:,k), oe2(:,:,:,k))
      call centered_vorticity_flux(npad, iim, jjm, nedges, e1(:,:,:,k), s2(:,:,k), s1(:,:,k), &
           oe3(:,:,:,k))
      call centered_divergence_curl(npad, iim, jjm, ndata, ncorners, nedges, nspace, r3, &
           dc(:,:,:,:,k), tangent(:,:,:,:,:,k), cc(:,:,:,k), s1(:,:,k), vc(:,:,:,:,k), &
           s2(:,:,k), oe4(:,:,:,k))
    enddo
    call open_out('ops.bin')
    write(fu) s1, s2, e1, vc, cc
    write(fu) o1, o2, o3, o4, o5, oc, ov, oe, oe2, oe3, oe4
    close(fu)
  end subroutine dump_ops

  subroutine dump_diss()
    real(kind=dbl_kind), allocatable :: vin(:,:,:,:,:), o1(:,:,:,:,:), o2(:,:,:,:,:)
    integer(kind=8) :: st
    allocate(vin(2,2,iim,jjm,nsdm), o1(2,2,iim,jjm,nsdm), o2(2,2,iim,jjm,nsdm))
    st = 20261204_8
    call fill(vin, size(vin), st)
    vin = 30.0_dbl_kind * vin
    call init_diss_op(swm_name)
    call dissipation(swm_name, 1, 2.5e12_dbl_kind, vin, o1)
    call dissipation(swm_name, 2, 2.5e12_dbl_kind, vin, o2)
    call open_out('diss.bin')
    write(fu) vin, o1, o2
    close(fu)
  end subroutine dump_diss

  subroutine dump_mg()
    real(kind=dbl_kind), allocatable :: beta(:,:,:,:), alpha(:,:,:,:), rhs(:,:,:,:), xo(:,:,:,:)
    integer :: i, j, k
    allocate(beta(iim,jjm,2,nsdm), alpha(iim,jjm,2,nsdm), rhs(iim,jjm,2,nsdm), xo(iim,jjm,2,nsdm))
    ! smooth right-hand sides built from the cell-centre positions
    do k = 1, nsdm
      do j = 1, jjm
        do i = 1, iim
          beta(i,j,1,k) = 1.0e-5_dbl_kind*(p(3,i,j,k) + 0.3_dbl_kind*p(1,i,j,k)*p(2,i,j,k))
          beta(i,j,2,k) = 2.0e-6_dbl_kind*(p(1,i,j,k)*p(3,i,j,k) - 0.5_dbl_kind*p(2,i,j,k))
        enddo
      enddo
    enddo
    call wrap(swm_name, face=beta)
    alpha = 0.0_dbl_kind
    call multigrid(2, beta, alpha)
    rhs = beta
    xo = 0.0_dbl_kind
    call elliptic('swm', 2, rhs, xo)
    call open_out('mg.bin')
    write(fu) beta, alpha, rhs, xo
    close(fu)
  end subroutine dump_mg

  ! ICs given as h, hs, u(corners), tracers; the derived fields follow
  ! initial_conditions_swtc5 (initialize.F:418-468); the loop follows
  ! driver.F:87-166
  subroutine do_run()
    integer :: nsteps, ndump, istep, ui, nt
    real(kind=dbl_kind) :: dtin
    real(kind=dbl_kind), allocatable :: h0(:,:,:), hs0(:,:,:), u0(:,:,:,:,:), t0(:,:,:,:)
    character(len=512) :: icfile
    logical :: do_budgets, do_wrestart
    integer :: iarg

    call get_command_argument(3, icfile)
    call get_command_argument(4, arg); read(arg,*) dtin
    call get_command_argument(5, arg); read(arg,*) nsteps
    call get_command_argument(6, arg); read(arg,*) ndump

    allocate(h0(iim,jjm,nsdm), hs0(iim,jjm,nsdm), u0(2,2,iim,jjm,nsdm), t0(iim,jjm,ntracer,nsdm))
    open(newunit=ui, file=trim(icfile), access='stream', form='unformatted', status='old')
    read(ui) h0, hs0, u0, t0
    close(ui)

    dt = dtin
    invdt = c1/dt
    ! advection scheme (argument 7): fct (default), centered or upwind
    arg = 'fct'
    if (command_argument_count() >= 7) call get_command_argument(7, arg)
    l_fct = trim(arg) == 'fct'
    l_centered = trim(arg) == 'centered'
    l_upwind = trim(arg) == 'upwind'
    l_remap = trim(arg) == 'remap'
    if (l_remap) call init_remap(swm_name)
    n_restart = 100000000
    n_output = 100000000
    n_diagnostics = 100000000
    tau_end = 1.0e30_dbl_kind
    call get_pointers()

    hs = hs0
    h(:,:,1,:) = h0
    h(:,:,2,:) = h0
    u(:,:,:,:,1,:) = u0
    u(:,:,:,:,2,:) = u0
    call wrap(swm_name, face=h(:,:,:,:))
    call wrap(swm_name, face_1L=hs(:,:,:))
    call wrap(swm_name, vrtx=u(:,:,:,:,:,:))
    call calculate_rel_div_and_ke('swm', u(:,:,:,:,1,:), relative(:,:,:), div(:,:,1,:), ke(:,:,:))
    div(:,:,2,:) = div(:,:,1,:)
    eta(:,:,1,:) = f(:,:,:) + relative(:,:,:)
    eta(:,:,2,:) = eta(:,:,1,:)
    do nt = 1, ntracer
      tracer(:,:,nt,1,:) = t0(:,:,nt,:)
    enddo
    tracer(:,:,:,2,:) = tracer(:,:,:,1,:)

    do_budgets = .false.
    do_wrestart = .false.
    do iarg = 8, command_argument_count()
      call get_command_argument(iarg, arg)
      if (trim(arg) == 'budgets') do_budgets = .true.
      if (trim(arg) == 'wrestart') do_wrestart = .true.
      if (trim(arg) == 'rrestart') call read_swm_restart()
    end do
    call wrap_prognostics()
    call open_out('run.bin')
    call dump_state(0)
    if (do_budgets) call compute_budgets()
    do istep = 1, nsteps
      call time_manager()
      call get_pointers()
      call get_weights()
      call step_dt()
      call wrap_prognostics()
      if (mod(istep, ndump) == 0) then
        call dump_state(istep)
        if (do_budgets) call compute_budgets()
      end if
    enddo
    close(fu)
    if (do_wrestart) call write_swm_restart()
  end subroutine do_run

  subroutine dump_state(istep)
    integer, intent(in) :: istep
    write(fu) istep, n4
    write(fu) h(:,:,n4,:), eta(:,:,n4,:), div(:,:,n4,:), tracer(:,:,:,n4,:)
    write(fu) u(:,:,:,:,n4,:), psi, chi, ke, relative
  end subroutine dump_state

end program ref_harness
