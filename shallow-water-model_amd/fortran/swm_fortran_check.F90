!=======================================================================
! swm_fortran_check -- drives the device model through the Fortran modules
! the way the reference driver does (swm/source/driver.F:78-121: wrap the
! initial state, then step_dt + wrap_prognostics per step), and applies
! laplacian_of_a_scalar and centered_vorticity_flux panel by panel as
! swm/source/step.F does.
!
!   swm_fortran_check <input> <output> [restart]
!
! input  (stream): int32 g, ntracer, nsteps; float64 dt, grav, radius,
!        omega; h, hs (iim,jjm,10); u (2,2,iim,jjm,10);
!        tracer (iim,jjm,ntracer,10)
! output (stream): after nsteps steps h, eta, div, psi, chi (iim,jjm,10),
!        u, tracer; then the Laplacian of h and the vorticity flux of
!        (h, eta) with a unit mass flux, panel by panel
! restart: write_swm_restart after the steps (into restarts/ of the
!        working directory), and the quick-plot file of the run
!        (qp_output/ of the working directory) with the initial state and
!        the state after the steps
!
!   swm_fortran_check remap <input> <output>
!
! remap_mass_and_tracers and dissipation through their modules, then steps
! of the l_remap model; input as the default mode's, output in the order
! written by run_remap.
!
!   swm_fortran_check ops <input> <output>
!
! calls every drop-in under the reference's names and argument lists: the 13
! operators of operators.F panel by panel, wrap with its keyword arguments,
! elliptic, multigrid, set_mass_fct, set_tracer_fct and global_real_sum.
! input (stream): int32 g; float64 dt, grav, radius, omega; h, hs
!        (iim,jjm,10); u (2,2,iim,jjm,10); tracer (iim,jjm,1,10); then
!        s1, s2 (iim,jjm,10), e1 (3,iim,jjm,10), vc (2,2,iim,jjm,10), cc
!        (2,iim,jjm,10), hist (3,iim,jjm,1,3,10)
! output (stream): in the order written by run_ops
!=======================================================================
program swm_fortran_check
  use iso_c_binding
  use swm_mi355x
  use swm_qp_module_mi355x
  use operators_mi355x
  implicit none
  character(len=512) :: fin, fout, opt
  integer(c_int32_t) :: g, ntracer, nsteps
  real(c_double) :: dt, grav, radius, omega
  real(c_double), allocatable, target :: h(:,:,:), hs(:,:,:), u(:,:,:,:,:), tracer(:,:,:,:)
  real(c_double), allocatable, target :: eta(:,:,:), div(:,:,:), psi(:,:,:), chi(:,:,:)
  real(c_double), allocatable, target :: lw(:,:,:,:), lap(:,:,:), mf(:,:,:,:), vf(:,:,:,:)
  integer :: n, k, fu

  call get_command_argument(1, fin)
  if (trim(fin) == 'ops') then
    call get_command_argument(2, fin)
    call get_command_argument(3, fout)
    call run_ops(fin, fout)
    stop
  end if
  if (trim(fin) == 'remap') then
    call get_command_argument(2, fin)
    call get_command_argument(3, fout)
    call run_remap(fin, fout)
    stop
  end if
  call get_command_argument(2, fout)
  opt = ''
  if (command_argument_count() >= 3) call get_command_argument(3, opt)
  open(newunit=fu, file=trim(fin), access='stream', form='unformatted', status='old')
  read(fu) g, ntracer, nsteps, dt, grav, radius, omega
  n = 2 + 2**g
  allocate(h(n,n,10), hs(n,n,10), u(2,2,n,n,10), tracer(n,n,ntracer,10))
  read(fu) h, hs, u, tracer
  close(fu)

  call swm_gpu_initialize(int(g), int(ntracer), dt, grav, radius, omega, h, hs, u, tracer)
  call wrap_prognostics()
  if (trim(opt) == 'restart') call qp_init()
  do k = 1, nsteps
    call step_dt()
    call wrap_prognostics()
  end do
  if (trim(opt) == 'restart') then
    call write_swm_restart()
    call pack_swm_qp()
    call write_swm_qp(swm_gpu_time())
  end if
  allocate(eta(n,n,10), div(n,n,10), psi(n,n,10), chi(n,n,10))
  call sync_prognostics_to_host(h, u, tracer, eta=eta, div=div, psi=psi, chi=chi)

  allocate(lw(7,n,n,10), lap(n,n,10), mf(3,n,n,10), vf(3,n,n,10))
  call swm_gpu_grid_array('l_weights', lw)
  mf = 1.0_c_double
  do k = 1, 10
    call laplacian_of_a_scalar(1, n, n, 7, lw(:,:,:,k), h(:,:,k), lap(:,:,k))
    call centered_vorticity_flux(1, n, n, 3, mf(:,:,:,k), h(:,:,k), eta(:,:,k), vf(:,:,:,k))
  end do

  open(newunit=fu, file=trim(fout), access='stream', form='unformatted', status='replace')
  write(fu) h, eta, div, psi, chi, u, tracer, lap, vf
  close(fu)
  call swm_gpu_finalize()

contains

  ! initialize.F:153-159 and :275-276: the qp file with the grid's
  ! coordinates (xyz_to_lonlat of the centres and corners, grid_metrics.F:
  ! 194-206) and the initial state as its first record
  subroutine qp_init()
    real(c_double), allocatable :: pxyz(:,:,:,:), cxyz(:,:,:,:,:), clat(:,:,:), clon(:,:,:), crlat(:,:,:,:), &
                                   crlon(:,:,:,:)
    integer :: i, j, kk, ic
    allocate(pxyz(3,n,n,10), cxyz(3,2,n,n,10), clat(n,n,10), clon(n,n,10), crlat(2,n,n,10), crlon(2,n,n,10))
    call swm_gpu_grid_array('p', pxyz)
    call swm_gpu_grid_array('c', cxyz)
    do kk = 1, 10
      do j = 1, n
        do i = 1, n
          call lonlat(pxyz(:,i,j,kk), clon(i,j,kk), clat(i,j,kk))
          do ic = 1, 2
            call lonlat(cxyz(:,ic,i,j,kk), crlon(ic,i,j,kk), crlat(ic,i,j,kk))
          end do
        end do
      end do
    end do
    call initialize_swm_qp('SWTC#5', 0_c_long_long, clat, clon, crlat, crlon)
    call pack_swm_qp()
    call write_swm_qp(swm_gpu_time())
  end subroutine qp_init

  subroutine lonlat(p, lon, lat)
    real(c_double), intent(in) :: p(3)
    real(c_double), intent(out) :: lon, lat
    lat = asin(p(3))
    if (p(1) /= 0.0_c_double .or. p(2) /= 0.0_c_double) then
      lon = atan2(p(2), p(1))
    else
      lon = 0.0_c_double
    end if
  end subroutine lonlat

  ! incremental remapping and the dissipation operator through their
  ! modules: remap_mass_and_tracers of the initial state (default limiter,
  ! then with lin_limit_tracers = .false.), dissipation of the initial
  ! velocity (norder 2), then nsteps steps of the l_remap model
  subroutine run_remap(fin, fout)
    use remap_mi355x
    use swm_dissipation_operator_mi355x
    character(len=*), intent(in) :: fin, fout
    integer(c_int32_t) :: g, nt, ns
    real(c_double) :: dt, grav, radius, omega
    real(c_double), allocatable, target :: h(:,:,:), hs(:,:,:), u(:,:,:,:,:), tracer(:,:,:,:)
    real(c_double), allocatable :: mass(:,:,:,:), trace(:,:,:,:,:), mass2(:,:,:,:), trace2(:,:,:,:,:), &
        ef(:,:,:,:,:), visc(:,:,:,:,:)
    integer :: n, k, fu
    open(newunit=fu, file=trim(fin), access='stream', form='unformatted', status='old')
    read(fu) g, nt, ns, dt, grav, radius, omega
    n = 2 + 2**g
    allocate(h(n,n,10), hs(n,n,10), u(2,2,n,n,10), tracer(n,n,nt,10))
    read(fu) h, hs, u, tracer
    close(fu)
    call swm_gpu_initialize(int(g), int(nt), dt, grav, radius, omega, h, hs, u, tracer, l_remap=.true.)
    call init_remap('swm')
    ! the initialised fields, halos filled
    call sync_prognostics_to_host(h, u, tracer)
    allocate(mass(n,n,1,10), trace(n,n,nt,1,10), mass2(n,n,1,10), trace2(n,n,nt,1,10), ef(1,3,n,n,10), &
             visc(2,2,n,n,10))
    mass(:,:,1,:) = h
    trace(:,:,:,1,:) = tracer
    mass2 = mass
    trace2 = trace
    call remap_mass_and_tracers('swm', 1, int(nt), mass, trace, dt, u, edgeflux=ef)
    call remap_mass_and_tracers('swm', 1, int(nt), mass2, trace2, dt, u, lin_limit_tracers=.false.)
    call init_diss_op('swm')
    call dissipation('swm', 2, 2.5e12_c_double, u, visc)
    ! remapped fields are valid at the real cells: fill the halos before the
    ! next remapping, as the reference's wraps do
    call wrap('swm', face=mass2)
    do k = 1, int(nt)
      call wrap('swm', face_1L=trace2(:,:,k,1,:))
    end do
    call remap_mass_and_tracers('swm', 1, int(nt), mass2, trace2, dt, u, lin_limit_tracers_min=.true., &
                                lin_limit_tracers_max=.true.)
    do k = 1, ns
      call step_dt()
      call wrap_prognostics()
    end do
    call sync_prognostics_to_host(h, u, tracer)
    open(newunit=fu, file=trim(fout), access='stream', form='unformatted', status='replace')
    write(fu) mass, trace, ef, mass2, trace2, visc, h, tracer
    close(fu)
    call swm_gpu_finalize()
  end subroutine run_remap

  subroutine run_ops(fin, fout)
    use mass_fct_mi355x
    use tracer_fct_mi355x
    use global_reductions_mi355x
    character(len=*), intent(in) :: fin, fout
    integer(c_int32_t) :: g
    real(c_double) :: dt, grav, radius, omega, total
    integer :: n, k, fu
    real(c_double), allocatable, target :: h(:,:,:), hs(:,:,:), u(:,:,:,:,:), tracer(:,:,:,:)
    real(c_double), allocatable :: s1(:,:,:), s2(:,:,:), e1(:,:,:,:), vc(:,:,:,:,:), cc(:,:,:,:), hist(:,:,:,:,:,:)
    real(c_double), allocatable, target :: area_inv(:,:,:), area(:,:,:), vw(:,:,:,:,:,:), lw(:,:,:,:), &
        cai(:,:,:,:), sw(:,:,:,:,:), sa(:,:,:,:,:), dc(:,:,:,:,:), nrm(:,:,:,:,:,:), tng(:,:,:,:,:,:), cmask(:,:,:)
    real(c_double), allocatable :: o_curl(:,:,:), o_divv(:,:,:), o_dive(:,:,:), o_lap(:,:,:), o_kxv(:,:,:,:,:), &
        o_grad(:,:,:,:,:), o_avg(:,:,:,:), o_ke(:,:,:), o_vel(:,:,:,:,:), o_cmf(:,:,:,:), o_umf(:,:,:,:), &
        o_cvf(:,:,:,:), o_cdc(:,:,:,:)
    real(c_double), allocatable :: fw(:,:,:,:), f1l(:,:,:), vx(:,:,:,:,:,:), v1l(:,:,:,:,:), vs(:,:,:,:,:), &
        e1l(:,:,:,:,:), es(:,:,:,:,:), rhs(:,:,:,:), xo(:,:,:,:), beta(:,:,:,:), alpha(:,:,:,:)
    real(c_double), allocatable :: mss(:,:,:,:,:), mss_c(:,:,:,:,:), uu(:,:,:,:,:,:), mlow(:,:,:,:), &
        mflux(:,:,:,:,:), trc(:,:,:,:), trh(:,:,:,:), tfct(:,:,:,:), tflux(:,:,:,:,:), tml(:,:,:,:)
    integer :: ti(3)
    real(c_double) :: tw(3)

    open(newunit=fu, file=trim(fin), access='stream', form='unformatted', status='old')
    read(fu) g, dt, grav, radius, omega
    n = 2 + 2**g
    allocate(h(n,n,10), hs(n,n,10), u(2,2,n,n,10), tracer(n,n,1,10))
    allocate(s1(n,n,10), s2(n,n,10), e1(3,n,n,10), vc(2,2,n,n,10), cc(2,n,n,10), hist(3,n,n,1,3,10))
    read(fu) h, hs, u, tracer, s1, s2, e1, vc, cc, hist
    close(fu)
    call swm_gpu_initialize(int(g), 1, dt, grav, radius, omega, h, hs, u, tracer)
    call initialize_wrap('swm')
    call initialize_multigrid('swm')
    allocate(area_inv(n,n,10), area(n,n,10), vw(3,2,2,n,n,10), lw(7,n,n,10), cai(2,n,n,10), sw(3,2,n,n,10), &
             sa(3,2,n,n,10), dc(3,2,n,n,10), nrm(3,3,2,n,n,10), tng(3,3,2,n,n,10), cmask(n,n,10))
    call swm_gpu_grid_array('area_inv', area_inv)
    call swm_gpu_grid_array('area', area)
    call swm_gpu_grid_array('v_weights', vw)
    call swm_gpu_grid_array('l_weights', lw)
    call swm_gpu_grid_array('corner_area_inv', cai)
    call swm_gpu_grid_array('s_weights', sw)
    call swm_gpu_grid_array('sub_area', sa)
    call swm_gpu_grid_array('dc', dc)
    call swm_gpu_grid_array('normal', nrm)
    call swm_gpu_grid_array('tangent', tng)
    call swm_gpu_grid_array('grid_center_mask', cmask)

    ! the operators of operators.F, panel by panel as the reference calls them
    allocate(o_curl(n,n,10), o_divv(n,n,10), o_dive(n,n,10), o_lap(n,n,10), o_kxv(2,2,n,n,10), &
             o_grad(2,2,n,n,10), o_avg(2,n,n,10), o_ke(n,n,10), o_vel(2,2,n,n,10), o_cmf(3,n,n,10), &
             o_umf(3,n,n,10), o_cvf(3,n,n,10), o_cdc(3,n,n,10))
    do k = 1, 10
      call curl_of_a_vector(1, n, n, 3, 2, 2, area_inv(:,:,k), vw(:,:,:,:,:,k), vc(:,:,:,:,k), o_curl(:,:,k))
      call div_of_a_vector(1, n, n, 3, 2, 2, area_inv(:,:,k), vw(:,:,:,:,:,k), vc(:,:,:,:,k), o_divv(:,:,k))
      call div_of_an_edge_flux(1, n, n, 3, area_inv(:,:,k), e1(:,:,:,k), o_dive(:,:,k))
      call laplacian_of_a_scalar(1, n, n, 7, lw(:,:,:,k), s1(:,:,k), o_lap(:,:,k))
      call k_cross_v(n, n, 2, 2, vc(:,:,:,:,k), o_kxv(:,:,:,:,k))
      call gradient(1, n, n, 3, 2, 2, cai(:,:,:,k), vw(:,:,:,:,:,k), s1(:,:,k), o_grad(:,:,:,:,k))
      call average_to_cell_corners(1, n, n, 3, 2, sw(:,:,:,:,k), s1(:,:,k), o_avg(:,:,:,k), .true.)
      call calculate_center_kinetic_energy(1, n, n, 3, 2, 2, sa(:,:,:,:,k), area_inv(:,:,k), cc(:,:,:,k), &
                                           o_ke(:,:,k))
      call calculate_velocity(1, n, n, 3, 2, 2, cai(:,:,:,k), vw(:,:,:,:,:,k), s1(:,:,k), s2(:,:,k), &
                              o_vel(:,:,:,:,k))
      call centered_mass_flux(1, n, n, 3, 2, 3, 2, 3, dc(:,:,:,:,k), nrm(:,:,:,:,:,k), cc(:,:,:,k), &
                              vc(:,:,:,:,k), s2(:,:,k), o_cmf(:,:,:,k))
      call upstream_mass_flux(1, n, n, 3, 2, 3, 2, 3, dc(:,:,:,:,k), nrm(:,:,:,:,:,k), vc(:,:,:,:,k), &
                              s2(:,:,k), o_umf(:,:,:,k))
      call centered_vorticity_flux(1, n, n, 3, e1(:,:,:,k), s2(:,:,k), s1(:,:,k), o_cvf(:,:,:,k))
      call centered_divergence_curl(1, n, n, 3, 2, 3, 2, 3, dc(:,:,:,:,k), tng(:,:,:,:,:,k), cc(:,:,:,k), &
                                    s1(:,:,k), vc(:,:,:,:,k), s2(:,:,k), o_cdc(:,:,:,k))
    end do

    ! wrap with keyword arguments
    allocate(fw(n,n,2,10), f1l(n,n,10), vx(2,2,n,n,1,10), v1l(2,2,n,n,10), vs(2,n,n,1,10), e1l(2,3,n,n,10), &
             es(3,n,n,1,10))
    fw(:,:,1,:) = s1
    fw(:,:,2,:) = s2
    f1l = s2
    vx(:,:,:,:,1,:) = vc
    v1l = vc
    vs(:,:,:,1,:) = cc
    e1l(1,:,:,:,:) = e1
    e1l(2,:,:,:,:) = 2.0_c_double * e1
    es(:,:,:,1,:) = e1
    call wrap('swm', face=fw, face_1L=f1l, vrtx=vx, vrtx_1L=v1l, vrtx_scalar=vs)
    call wrap('swm', edge_1L=e1l)
    call wrap('swm', l_HA_stencil=.true., edge_scalar=es)

    ! elliptic and one V-cycle
    allocate(rhs(n,n,2,10), xo(n,n,2,10), beta(n,n,2,10), alpha(n,n,2,10))
    rhs(:,:,1,:) = 1.0e-5_c_double * s1
    rhs(:,:,2,:) = 1.0e-6_c_double * s2
    xo = 0.0_c_double
    call elliptic('swm', 2, rhs, xo)
    beta = rhs
    alpha = 0.0_c_double
    call multigrid(2, beta, alpha)

    ! set_mass_fct and set_tracer_fct: level 1 in, level 2 out, the AB3
    ! history in permuted slots
    allocate(mss(n,n,1,2,10), mss_c(2,n,n,1,10), uu(2,2,n,n,1,10), mlow(n,n,1,10), mflux(3,n,n,1,10), &
             trc(n,n,1,10), trh(n,n,1,10), tfct(n,n,1,10), tflux(3,n,n,1,10), tml(n,n,1,10))
    mss = 0.0_c_double
    mss(:,:,1,1,:) = h
    do k = 1, 10
      call average_to_cell_corners(1, n, n, 3, 2, sw(:,:,:,:,k), h(:,:,k), mss_c(:,:,:,1,k), .true.)
    end do
    uu(:,:,:,:,1,:) = u
    ti = (/ 2, 3, 1 /)
    tw = (/ 23.0_c_double / 12.0_c_double, -16.0_c_double / 12.0_c_double, 5.0_c_double / 12.0_c_double /)
    call set_mass_fct('swm', n, n, 1, 10, 2, 3, 2, 3, 3, 2, 1, 3, 1, 2, ti, tw, dt, area_inv, area, dc, nrm, &
                      mss, mss_c, uu, mlow, mflux, hist)
    trc = tracer
    trh(:,:,1,:) = tracer(:,:,1,:) + 0.01_c_double * s1
    call set_tracer_fct('swm', n, n, 1, 10, 3, dt, area_inv, area, trc, trh, mss(:,:,:,1,:), mlow, &
                        mss(:,:,:,2,:), mflux, tfct, trc_flux_horz=tflux, msstrc_low_order=tml)
    total = global_real_sum(area, cmask, n * n * 10, 'swm')

    open(newunit=fu, file=trim(fout), access='stream', form='unformatted', status='replace')
    write(fu) o_curl, o_divv, o_dive, o_lap, o_kxv, o_grad, o_avg, o_ke, o_vel, o_cmf, o_umf, o_cvf, o_cdc
    write(fu) fw, f1l, vx, v1l, vs, e1l, es
    write(fu) rhs, xo, alpha
    write(fu) mss, mlow, mflux, hist, tfct, tflux, tml
    write(fu) total
    close(fu)
    call swm_gpu_finalize()
  end subroutine run_ops
end program swm_fortran_check
