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
!        working directory)
!=======================================================================
program swm_fortran_check
  use iso_c_binding
  use swm_mi355x
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
  do k = 1, nsteps
    call step_dt()
    call wrap_prognostics()
  end do
  if (trim(opt) == 'restart') call write_swm_restart()
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
end program swm_fortran_check
