! ns_dropin.f90 -- source-level drop-in for the reference's hot-path modules.
!
! Built together with the reference's own global.f90 (the `module global`
! the driver and init.f90 fill), these modules replace rk.f90, filtering.f90,
! cc.f90, diff.f90, derivs.f90, eqeuler.f90, coefficients.f90 and rhs.f90:
! same module names, procedure names and dummy arguments, so euler.f90 keeps
! calling `rk_euler(U,t,iter,size(U,1))` and init.f90 `GET_COEFFICIENTS()`.
! Each hot-path call runs on the GPU through the C ABI (ns_capi.f90); the
! split pieces of rhs_euler (deriv, kdT, tau_ij, EULERNONLINEAR,
! MOLECULAR_MI, k_coefficient, metricsX/Y) read and write the same module
! arrays the reference's do.  The first call creates the device context from
! the module variables init_dados and init_variables set (profiles, base
! flow, scalars).
module ns_dropin_context
  use global
  use ns_shim_core
  implicit none
  ! (ct, iter) of the last acousticfont call: the reference's rhs_euler reads
  ! the s_acou that call left in module global (eqeuler.f90:84, after
  ! rk.f90:55), and the device recomputes that source from them.  Before any
  ! acousticfont call s_acou holds zeros, which (0, 0) reproduces:
  ! amplitude*sin(wf*(0+0)*dt) = 0.
  real(c_double), save :: ns_src_ct = 0.0_c_double
  integer(c_int32_t), save :: ns_src_iter = 0_c_int32_t
contains
  subroutine ns_dropin_ensure(nk)
    integer, intent(in) :: nk
    type(ns_params) :: p
    integer(c_int32_t) :: rc
    if (ns_shim_ready() .and. ns_ctx_nk == nk) return
    rc = ns_default_params(p)
    p%imax = imax
    p%jmax = jmax
    p%nk = nk
    p%D = D
    p%dm = dm
    p%dn = dn
    p%dt = dt
    p%sigmaf = sigmaf
    p%gamma = gamma
    p%Re = Re
    p%Pr = Pr
    p%Pe = Pe
    p%alpha_cond = alpha_cond
    p%Tref = Tref
    p%A_mi = A_MI
    p%B_mi = B_MI
    p%C_mi = C_MI
    p%D_mi = D_MI
    p%MI0 = MI0
    p%A_k = A_K
    p%B_k = B_K
    p%C_k = C_K
    p%D_k = D_K
    p%K0 = K0
    p%amplitude = amplitude
    p%wf = wf
    p%r0_acous = r0_acous
    p%m_pt_acous = m_pt_acous
    p%n_pt_acous = n_pt_acous
    p%precision = NS_PREC_FP64
    p%flags = 0
    p%device = 0
    call ns_shim_init(p, meshx, meshy, sigmax, sigmay, m, n, rhob, psib, Ub)
  end subroutine
end module ns_dropin_context

! rk.f90
module rk
  use global
  use ns_dropin_context
  implicit none
contains
  subroutine rk_euler(U, t, iter, nk)
    integer :: iter, nk
    real(kind=ip) :: t
    real(kind=ip), dimension(1:nk, 1:imax, 1:jmax) :: U
    call ns_dropin_ensure(nk)
    call ns_shim_step(U, iter)
  end subroutine

  subroutine bufferxy(U, nk)
    integer :: nk
    real(kind=ip), dimension(1:nk, 1:imax, 1:jmax) :: U
    call ns_dropin_ensure(nk)
    call ns_check(ns_bufferxy(ns_ctx_handle, U), 'ns_bufferxy')
  end subroutine

  subroutine rk_steep(alphat, betat, ct, U, dudt, w, nk)
    integer :: nk
    real(kind=ip) :: alphat, betat, ct
    real(kind=ip), dimension(nk, 1:imax, 1:jmax) :: U, dudt, w
    call ns_dropin_ensure(nk)
    call ns_check(ns_rk_steep(ns_ctx_handle, alphat, betat, U, dudt, w), 'ns_rk_steep')
  end subroutine

  subroutine acousticfont(ct, iter)
    integer :: iter
    real(kind=ip) :: ct
    call ns_dropin_ensure(nVar)
    call ns_check(ns_acousticfont(ns_ctx_handle, ct, int(iter, c_int32_t), s_acou), 'ns_acousticfont')
    ns_src_ct = ct
    ns_src_iter = int(iter, c_int32_t)
  end subroutine
end module rk

! filtering.f90
module filtering
  use global
  use ns_dropin_context
  implicit none
contains
  subroutine lddfilterx(U, nk)
    integer :: nk
    real(kind=ip), dimension(1:nk, 1:imax, 1:jmax) :: U
    call ns_dropin_ensure(nk)
    call ns_check(ns_filter_x(ns_ctx_handle, U), 'ns_filter_x')
  end subroutine

  subroutine lddfiltery(U, nk)
    integer :: nk
    real(kind=ip), dimension(1:nk, 1:imax, 1:jmax) :: U
    call ns_dropin_ensure(nk)
    call ns_check(ns_filter_y(ns_ctx_handle, U), 'ns_filter_y')
  end subroutine
end module filtering

! cc.f90
module cc
  use global
  use ns_dropin_context
  implicit none
contains
  subroutine ccpml(U, nk)
    integer :: nk
    real(kind=ip), dimension(1:nk, 1:imax, 1:jmax) :: U
    call ns_dropin_ensure(nk)
    call ns_check(ns_ccpml(ns_ctx_handle, U), 'ns_ccpml')
  end subroutine
end module cc

! diff.f90 (metrics applied inside, as the reference's deropm/deropn do)
module diff
  use global
  use ns_dropin_context
  implicit none
contains
  subroutine deropm(dU, U, ii, fi, ij, fj, nk)
    integer :: ii, fi, ij, fj, nk
    real(kind=ip), dimension(nk, 1:imax, 1:jmax) :: U, dU
    call ns_dropin_ensure(nVar)
    call ns_check(ns_deropm(ns_ctx_handle, int(nk, c_int32_t), dU, U), 'ns_deropm')
  end subroutine

  subroutine deropn(dU, U, ii, fi, ij, fj, nk)
    integer :: ii, fi, ij, fj, nk
    real(kind=ip), dimension(nk, 1:imax, 1:jmax) :: dU, U
    call ns_dropin_ensure(nVar)
    call ns_check(ns_deropn(ns_ctx_handle, int(nk, c_int32_t), dU, U), 'ns_deropn')
  end subroutine

  ! diff.f90:276 metricsY(dU,nk), :315 metricsX(dU,nk)
  subroutine metricsY(dU, nk)
    integer :: nk
    real(kind=ip), dimension(nk, 1:imax, 1:jmax), intent(inout) :: dU
    call ns_dropin_ensure(nVar)
    call ns_check(ns_metrics_y(ns_ctx_handle, int(nk, c_int32_t), dU), 'ns_metrics_y')
  end subroutine

  subroutine metricsX(dU, nk)
    integer :: nk
    real(kind=ip), dimension(nk, 1:imax, 1:jmax), intent(inout) :: dU
    call ns_dropin_ensure(nVar)
    call ns_check(ns_metrics_x(ns_ctx_handle, int(nk, c_int32_t), dU), 'ns_metrics_x')
  end subroutine
end module diff

! coefficients.f90 (the transport fits of the live "visc" module)
module empirical_coefficients
  use global
  use ns_dropin_context
  implicit none
contains
  ! coefficients.f90:27 MOLECULAR_MI(T_inst): MI from T
  subroutine molecular_mi(T_inst)
    real(kind=ip), dimension(1, 1:imax, 1:jmax) :: T_inst
    real(kind=ip), allocatable :: kk(:, :)
    call ns_dropin_ensure(nVar)
    allocate(kk(imax, jmax))
    call ns_check(ns_transport(ns_ctx_handle, T_inst, MI, kk), 'ns_transport')
  end subroutine

  ! coefficients.f90:46 k_coefficient(T_inst): K1 from T
  subroutine k_coefficient(T_inst)
    real(kind=ip), dimension(1, 1:imax, 1:jmax) :: T_inst
    real(kind=ip), allocatable :: mm(:, :)
    call ns_dropin_ensure(nVar)
    allocate(mm(imax, jmax))
    call ns_check(ns_transport(ns_ctx_handle, T_inst, mm, K1), 'ns_transport')
  end subroutine

  ! coefficients.f90:71-80 GET_COEFFICIENTS (one-time set-up on the host):
  ! the species fit, then its reference values at Tref
  subroutine get_coefficients()
    call specie_n()
    K0 = dexp(A_K * dlog(1.0d0 * Tref) + B_K / Tref + C_K / Tref / Tref + D_K)
    MI0 = dexp(A_MI * dlog(1.0d0 * Tref) + B_MI / Tref + C_MI / Tref / Tref + D_MI)
  end subroutine

  ! coefficients.f90:113-123 SPECIE_N: the N2 fit, from the library's
  ! parameter defaults (the same Svehla 1995 set)
  subroutine specie_n()
    type(ns_params) :: p
    call ns_check(ns_default_params(p), 'ns_default_params')
    A_MI = p%A_mi
    B_MI = p%B_mi
    C_MI = p%C_mi
    D_MI = p%D_mi
    A_K = p%A_k
    B_K = p%B_k
    C_K = p%C_k
    D_K = p%D_k
  end subroutine
end module empirical_coefficients

! derivs.f90
module derivs
  use global
  use ns_dropin_context
  implicit none
contains
  ! derivs.f90:9 deriv(U,nk): dUdm, dUdn, DTEMPDM/N (times K1), D2TEMPDM/N,
  ! psiDiff, D2psiFDM/N in module global, from U, TEMP and K1
  subroutine deriv(U, nk)
    integer :: nk
    real(kind=ip), dimension(1:nk, 1:imax, 1:jmax) :: U
    call ns_dropin_ensure(nVar)
    call ns_check(ns_deriv(ns_ctx_handle, int(nk, c_int32_t), U, TEMP, K1, dUdm, dUdn, DTEMPDM, DTEMPDN, &
                           D2TEMPDM, D2TEMPDN, psiDiff, D2psiFDM, D2psiFDN), 'ns_deriv')
  end subroutine

  ! derivs.f90:48 kdT(): DTEMPDM, DTEMPDN times K1
  subroutine kdT()
    call ns_dropin_ensure(nVar)
    call ns_check(ns_kdt(ns_ctx_handle, K1, DTEMPDM, DTEMPDN), 'ns_kdt')
  end subroutine
end module derivs

! eqeuler.f90
module eqeuler
  use global
  use ns_dropin_context
  implicit none
contains
  ! eqeuler.f90:28 tau_ij(): TAU, DTAUDM, DTAUDN from MI, dUdm, dUdn
  subroutine tau_ij()
    call ns_dropin_ensure(nVar)
    call ns_check(ns_tau_ij(ns_ctx_handle, int(size(dUdm, 1), c_int32_t), MI, dUdm, dUdn, TAU, DTAUDM, DTAUDN), &
                  'ns_tau_ij')
  end subroutine

  ! eqeuler.f90:53 EULERNONLINEAR(DUDT,U,NK) from the module fields above and s_acou
  subroutine eulernonlinear(dudt, U, nk)
    integer :: nk
    real(kind=ip), dimension(1:nk, 1:imax, 1:jmax) :: U, dudt
    call ns_dropin_ensure(nVar)
    call ns_check(ns_eulernonlinear(ns_ctx_handle, int(nk, c_int32_t), dudt, U, dUdm, dUdn, TAU, DTAUDM, DTAUDN, &
                                    D2TEMPDM, D2TEMPDN, D2psiFDM, D2psiFDN, s_acou), 'ns_eulernonlinear')
  end subroutine
end module eqeuler

! rhs.f90
module rhs
  use global
  use ns_dropin_context
  implicit none
contains
  subroutine rhs_euler(dUdt, U, nk)
    integer :: nk
    real(kind=ip), dimension(1:nk, 1:imax, 1:jmax) :: dUdt, U
    call ns_dropin_ensure(nk)
    call ns_check(ns_rhs(ns_ctx_handle, dUdt, U, ns_src_ct, ns_src_iter), 'ns_rhs')
  end subroutine
end module rhs
