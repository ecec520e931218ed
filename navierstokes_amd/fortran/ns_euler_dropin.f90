! ns_euler_dropin.f90 -- a driver shaped like the reference's euler.f90
! (euler.f90:118-130), built against the reference's own global.f90 and the
! drop-in modules (ns_dropin.f90).  It fills `module global` from a case file
! (what init_dados / init_variables do in the reference), then calls the
! drop-in routines the way the reference calls its own:
!
!   mode 1  the time loop: do iter = iter0, iter0+nsteps-1: rk_euler(U,t,iter,size(U,1))
!   mode 2  acousticfont(ct,iter0) then rhs_euler(dUdt,U,nk) (rk.f90:55-58)
!   mode 3  rhs_euler's body through its split pieces (rhs.f90:20-44):
!           TEMP = gamma p / rho; MOLECULAR_MI, k_coefficient; deriv(U,nk);
!           tau_ij(); EULERNONLINEAR(dUdt,U,nk), after acousticfont(ct,iter0)
!   mode 4  metricsX(dU,nk) and metricsY(dU,nk) of dU = U
!
! usage: ns_euler_dropin <in.bin> <out.bin>
! in.bin (stream, native endian):
!   int32  imax, jmax, D, mode, iter0, nsteps
!   real8  dm, dn, dt, sigmaf, gamma, Re, Pr, Tref, amplitude, wf, r0_acous,
!          m_pt_acous, n_pt_acous, ct
!   real8  meshx(imax) meshy(jmax) sigmax(imax) sigmay(jmax) m(imax) n(jmax)
!          rhob(jmax) psib(jmax) Ub(5,imax,jmax) U(5,imax,jmax)
! out.bin: mode 1 U; modes 2, 3 dUdt; mode 4 metricsX(U) then metricsY(U).
program ns_euler_dropin
  use global
  use rk
  use rhs
  use derivs
  use eqeuler
  use empirical_coefficients
  use diff
  use ns_shim_core, only: ns_shim_finalize
  implicit none
  integer(4) :: hdr(6)
  real(8) :: sc(14), ct_src, t
  real(kind=ip), allocatable :: U(:, :, :), dUdt(:, :, :), dU(:, :, :), T_inst(:, :, :)
  character(len=512) :: fin, fout
  integer :: u_in, u_out, mode, iter, iter0, nsteps

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open(newunit=u_in, file=trim(fin), access='stream', form='unformatted', status='old')
  read(u_in) hdr
  read(u_in) sc
  imax = hdr(1)
  jmax = hdr(2)
  D = hdr(3)
  mode = hdr(4)
  iter0 = hdr(5)
  nsteps = hdr(6)
  dm = sc(1)
  dn = sc(2)
  dt = sc(3)
  sigmaf = sc(4)
  gamma = sc(5)
  Re = sc(6)
  Pr = sc(7)
  Tref = sc(8)
  amplitude = sc(9)
  wf = sc(10)
  r0_acous = sc(11)
  m_pt_acous = sc(12)
  n_pt_acous = sc(13)
  ct_src = sc(14)
  ! init_dados / init_variables (init.f90:56-206, 224-284, 392-418)
  allocate(meshx(imax), meshy(jmax), sigmax(imax), sigmay(jmax), m(imax), n(jmax), rhob(jmax), psib(jmax))
  allocate(Ub(nVar, imax, jmax), U(nVar, imax, jmax), dUdt(nVar, imax, jmax), dU(nVar, imax, jmax))
  read(u_in) meshx, meshy, sigmax, sigmay, m, n, rhob, psib, Ub, U
  close(u_in)
  allocate(s_acou(imax, jmax), MI(imax, jmax), K1(imax, jmax), TEMP(1, imax, jmax), T_inst(1, imax, jmax))
  allocate(DTEMPDM(1, imax, jmax), DTEMPDN(1, imax, jmax), D2TEMPDM(1, imax, jmax), D2TEMPDN(1, imax, jmax))
  allocate(psiDiff(1, imax, jmax), D2psiFDM(1, imax, jmax), D2psiFDN(1, imax, jmax))
  allocate(TAU(3, imax, jmax), DTAUDM(3, imax, jmax), DTAUDN(3, imax, jmax))
  allocate(dUdm(nVar, imax, jmax), dUdn(nVar, imax, jmax))
  s_acou = 0.0d0
  call get_coefficients()

  select case (mode)
  case (1)  ! euler.f90:123-130
    do iter = iter0, iter0 + nsteps - 1
      t = (iter + 1) * dt
      call rk_euler(U, t, iter, size(U, 1))
    end do
  case (2)
    call acousticfont(ct_src, iter0)
    call rhs_euler(dUdt, U, size(U, 1))
  case (3)
    call acousticfont(ct_src, iter0)
    TEMP(1, :, :) = gamma * U(4, :, :) / U(1, :, :)
    T_inst = TEMP
    call molecular_mi(T_inst)
    call k_coefficient(T_inst)
    call deriv(U, size(U, 1))
    call tau_ij()
    call eulernonlinear(dUdt, U, size(U, 1))
  case (4)
    dU = U
    call metricsX(dU, nVar)
    dUdt = U
    call metricsY(dUdt, nVar)
  case default
    write(*, '(a,i0)') 'ns_euler_dropin: unknown mode ', mode
    error stop 2
  end select
  call ns_shim_finalize()

  open(newunit=u_out, file=trim(fout), access='stream', form='unformatted', status='replace')
  select case (mode)
  case (1)
    write(u_out) U
  case (2, 3)
    write(u_out) dUdt
  case (4)
    write(u_out) dU, dUdt
  end select
  close(u_out)
  write(*, '(a,i0,a)') 'ns_euler_dropin: mode ', mode, ' done'
end program ns_euler_dropin
