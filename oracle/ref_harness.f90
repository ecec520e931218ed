! ref_harness.f90 -- C-callable entry points into the parts of the reference
! (mathe1s/NavierStokes) that compile unmodified from their own sources:
! global.f90, coefficients.f90, cc.f90 and diffall.f90.  The recipe in
! oracle/Makefile compiles those files where they lie in the reference tree
! and links them with this harness into oracle/_ref/libnsref.so.
!
! TEST INFRASTRUCTURE: used only by tests/ to pin the oracle (ns_oracle.c)
! against the reference's own arithmetic.  Nothing in the product loads it.
module ref_harness
  use iso_c_binding
  use global
  use empirical_coefficients
  use cc, only: ccpml
  use diffall, only: ref_deropm => deropm, ref_deropn => deropn
  implicit none
contains

  ! global.f90 scalars the exported routines read
  subroutine nsref_setup(c_imax, c_jmax, c_dm, c_dn, c_gamma, c_tref) bind(C, name="nsref_setup")
    integer(c_int), value :: c_imax, c_jmax
    real(c_double), value :: c_dm, c_dn, c_gamma, c_tref
    imax = c_imax
    jmax = c_jmax
    dm = c_dm
    dn = c_dn
    gamma = c_gamma
    Tref = c_tref
    if (allocated(MI)) deallocate(MI)
    if (allocated(K1)) deallocate(K1)
    if (allocated(rhob)) deallocate(rhob)
    if (allocated(psib)) deallocate(psib)
    allocate(MI(imax, jmax), K1(imax, jmax), rhob(jmax), psib(jmax))
  end subroutine

  ! coefficients.f90:71-80 GET_COEFFICIENTS; out = A,B,C,D (mu), A,B,C,D (k), MI0, K0
  subroutine nsref_get_coefficients(out) bind(C, name="nsref_get_coefficients")
    real(c_double), intent(out) :: out(10)
    call GET_COEFFICIENTS()
    out = [A_MI, B_MI, C_MI, D_MI, A_K, B_K, C_K, D_K, MI0, K0]
  end subroutine

  ! coefficients.f90:27-59 MOLECULAR_MI, k_coefficient (GET_COEFFICIENTS first)
  subroutine nsref_transport(T, mi_out, k_out) bind(C, name="nsref_transport")
    real(c_double), intent(in) :: T(1, imax, jmax)
    real(c_double), intent(out) :: mi_out(imax, jmax), k_out(imax, jmax)
    real(kind=ip), allocatable :: tin(:, :, :)
    call GET_COEFFICIENTS()
    allocate(tin(1, imax, jmax))
    tin = T
    call MOLECULAR_MI(tin)
    call K_COEFFICIENT(tin)
    mi_out = MI
    k_out = K1
  end subroutine

  ! cc.f90:6-68 ccpml (the reference writes component 5, so nk must be 5)
  subroutine nsref_ccpml(U, nk, rb, pb) bind(C, name="nsref_ccpml")
    integer(c_int), value :: nk
    real(c_double), intent(inout) :: U(nk, imax, jmax)
    real(c_double), intent(in) :: rb(jmax), pb(jmax)
    rhob = rb
    psib = pb
    call ccpml(U, nk)
  end subroutine

  ! diffall.f90:5-45 / 195-232 deropm, deropn over the full range (no metrics)
  subroutine nsref_deropm(dU, U, nk) bind(C, name="nsref_deropm")
    integer(c_int), value :: nk
    real(c_double), intent(inout) :: dU(nk, imax, jmax)
    real(c_double), intent(inout) :: U(nk, imax, jmax)
    call ref_deropm(dU, U, 1, imax, 1, jmax, nk)
  end subroutine

  subroutine nsref_deropn(dU, U, nk) bind(C, name="nsref_deropn")
    integer(c_int), value :: nk
    real(c_double), intent(inout) :: dU(nk, imax, jmax)
    real(c_double), intent(inout) :: U(nk, imax, jmax)
    call ref_deropn(dU, U, 1, imax, 1, jmax, nk)
  end subroutine

end module ref_harness
