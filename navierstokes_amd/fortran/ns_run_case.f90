! ns_run_case.f90 -- Fortran driver over the ISO_C_BINDING core: reads a case
! file, advances it with rk_euler-style calls (state round-trips through the
! caller's array every step, as in euler.f90:123-130) and writes U back.
!
! usage: ns_run_case <in.bin> <out.bin> <nsteps>
! in.bin (stream, native endian):
!   int32  imax, jmax, nk, D, flags
!   real8  dm, dn, dt, sigmaf, gamma, Re, Pr, Tref
!   real8  meshx(imax) meshy(jmax) sigmax(imax) sigmay(jmax) m(imax) n(jmax)
!          rhob(jmax) psib(jmax) Ub(nk,imax,jmax) U(nk,imax,jmax)
program ns_run_case
  use ns_shim_core
  implicit none
  integer(c_int32_t) :: hdr(5), rc
  real(c_double) :: sc(8)
  real(c_double), allocatable :: meshx(:), meshy(:), sigmax(:), sigmay(:), m(:), n(:), rhob(:), psib(:)
  real(c_double), allocatable :: Ub(:, :, :), U(:, :, :)
  type(ns_params) :: p
  character(len=512) :: fin, fout, arg
  integer :: nsteps, it, imax, jmax, nk, u_in, u_out

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  call get_command_argument(3, arg)
  read(arg, *) nsteps

  open(newunit=u_in, file=trim(fin), access='stream', form='unformatted', status='old')
  read(u_in) hdr
  read(u_in) sc
  imax = hdr(1)
  jmax = hdr(2)
  nk = hdr(3)
  allocate(meshx(imax), meshy(jmax), sigmax(imax), sigmay(jmax), m(imax), n(jmax), rhob(jmax), psib(jmax))
  allocate(Ub(nk, imax, jmax), U(nk, imax, jmax))
  read(u_in) meshx, meshy, sigmax, sigmay, m, n, rhob, psib, Ub, U
  close(u_in)

  rc = ns_default_params(p)
  p%imax = hdr(1)
  p%jmax = hdr(2)
  p%nk = hdr(3)
  p%D = hdr(4)
  p%flags = hdr(5)
  p%dm = sc(1)
  p%dn = sc(2)
  p%dt = sc(3)
  p%sigmaf = sc(4)
  p%gamma = sc(5)
  p%Re = sc(6)
  p%Pr = sc(7)
  p%Tref = sc(8)
  p%amplitude = 0.0_c_double
  call ns_shim_init(p, meshx, meshy, sigmax, sigmay, m, n, rhob, psib, Ub)
  do it = 0, nsteps - 1
    call ns_shim_step(U, it)
  end do
  call ns_shim_finalize()

  open(newunit=u_out, file=trim(fout), access='stream', form='unformatted', status='replace')
  write(u_out) U
  close(u_out)
  write(*, '(a,i0,a)') 'ns_run_case: ', nsteps, ' steps done'
end program ns_run_case
