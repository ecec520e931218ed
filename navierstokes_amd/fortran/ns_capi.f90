! ns_capi.f90 -- ISO_C_BINDING view of include/ns_api.h, plus the
! explicit-argument core the drop-in modules (ns_dropin.f90) forward to.
!
! Arrays are the reference's U(1:nk,1:imax,1:jmax) (real(8), column-major),
! which is exactly the C ABI's AoS host layout, so they pass without copies.
module ns_capi
  use iso_c_binding
  implicit none

  integer(c_int32_t), parameter :: NS_OK = 0
  integer(c_int32_t), parameter :: NS_PREC_FP64 = 0, NS_PREC_FP32_PERT = 1
  integer(c_int32_t), parameter :: NS_FLAG_INVISCID = 1, NS_FLAG_METRIC_EVERYWHERE = 2, NS_FLAG_PERIODIC_X = 4

  type, bind(C) :: ns_params
    integer(c_int32_t) :: imax, jmax, nk, D
    real(c_double) :: dm, dn, dt, sigmaf
    real(c_double) :: gamma, Re, Pr, Pe, alpha_cond, Tref
    real(c_double) :: A_mi, B_mi, C_mi, D_mi, MI0
    real(c_double) :: A_k, B_k, C_k, D_k, K0
    real(c_double) :: amplitude, wf, r0_acous, m_pt_acous, n_pt_acous
    integer(c_int32_t) :: precision
    integer(c_int32_t) :: flags
    integer(c_int32_t) :: device, reserved
  end type ns_params

  interface
    integer(c_int32_t) function ns_api_version() bind(C, name="ns_api_version")
      import :: c_int32_t
    end function
    type(c_ptr) function ns_last_error() bind(C, name="ns_last_error")
      import :: c_ptr
    end function
    integer(c_int32_t) function ns_default_params(p) bind(C, name="ns_default_params")
      import :: c_int32_t, ns_params
      type(ns_params), intent(out) :: p
    end function
    integer(c_int32_t) function ns_create(p, ctx) bind(C, name="ns_create")
      import :: c_int32_t, ns_params, c_ptr
      type(ns_params), intent(in) :: p
      type(c_ptr), intent(out) :: ctx
    end function
    integer(c_int32_t) function ns_destroy(ctx) bind(C, name="ns_destroy")
      import :: c_int32_t, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int32_t) function ns_set_profiles(ctx, meshx, meshy, sigmax, sigmay, m, n, rhob, psib) &
        bind(C, name="ns_set_profiles")
      import :: c_int32_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: meshx(*), meshy(*), sigmax(*), sigmay(*), m(*), n(*), rhob(*), psib(*)
    end function
    integer(c_int32_t) function ns_set_base(ctx, Ub) bind(C, name="ns_set_base")
      import :: c_int32_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: Ub(*)
    end function
    integer(c_int32_t) function ns_put_state(ctx, U) bind(C, name="ns_put_state")
      import :: c_int32_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: U(*)
    end function
    integer(c_int32_t) function ns_get_state(ctx, U) bind(C, name="ns_get_state")
      import :: c_int32_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: U(*)
    end function
    integer(c_int32_t) function ns_probe(ctx, npts, ij, out) bind(C, name="ns_probe")
      import :: c_int32_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: npts
      integer(c_int32_t), intent(in) :: ij(*)
      real(c_double), intent(out) :: out(*)
    end function
    integer(c_int32_t) function ns_rk_euler(ctx, iter) bind(C, name="ns_rk_euler")
      import :: c_int32_t, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: iter
    end function
    integer(c_int32_t) function ns_rk_steps(ctx, iter0, nsteps) bind(C, name="ns_rk_steps")
      import :: c_int32_t, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: iter0, nsteps
    end function
    integer(c_int32_t) function ns_deropm(ctx, nk, dU, U) bind(C, name="ns_deropm")
      import :: c_int32_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: nk
      real(c_double), intent(out) :: dU(*)
      real(c_double), intent(in) :: U(*)
    end function
    integer(c_int32_t) function ns_deropn(ctx, nk, dU, U) bind(C, name="ns_deropn")
      import :: c_int32_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: nk
      real(c_double), intent(out) :: dU(*)
      real(c_double), intent(in) :: U(*)
    end function
    integer(c_int32_t) function ns_rhs(ctx, dUdt, U, ct, iter) bind(C, name="ns_rhs")
      import :: c_int32_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: dUdt(*)
      real(c_double), intent(in) :: U(*)
      real(c_double), value :: ct
      integer(c_int32_t), value :: iter
    end function
    integer(c_int32_t) function ns_rk_steep(ctx, alphat, betat, U, dudt, w) bind(C, name="ns_rk_steep")
      import :: c_int32_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: alphat, betat
      real(c_double), intent(inout) :: U(*), w(*)
      real(c_double), intent(in) :: dudt(*)
    end function
    integer(c_int32_t) function ns_bufferxy(ctx, U) bind(C, name="ns_bufferxy")
      import :: c_int32_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(inout) :: U(*)
    end function
    integer(c_int32_t) function ns_filter_x(ctx, U) bind(C, name="ns_filter_x")
      import :: c_int32_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(inout) :: U(*)
    end function
    integer(c_int32_t) function ns_filter_y(ctx, U) bind(C, name="ns_filter_y")
      import :: c_int32_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(inout) :: U(*)
    end function
    integer(c_int32_t) function ns_ccpml(ctx, U) bind(C, name="ns_ccpml")
      import :: c_int32_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(inout) :: U(*)
    end function
    integer(c_int32_t) function ns_acousticfont(ctx, ct, iter, s) bind(C, name="ns_acousticfont")
      import :: c_int32_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: ct
      integer(c_int32_t), value :: iter
      real(c_double), intent(out) :: s(*)
    end function
    integer(c_int32_t) function ns_transport(ctx, T, MI, K1) bind(C, name="ns_transport")
      import :: c_int32_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: T(*)
      real(c_double), intent(out) :: MI(*), K1(*)
    end function
    integer(c_int32_t) function ns_deriv(ctx, nk, U, T, K1, dUdm, dUdn, dTdm, dTdn, d2Tdm, d2Tdn, psiDiff, &
                                         d2psidm, d2psidn) bind(C, name="ns_deriv")
      import :: c_int32_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: nk
      real(c_double), intent(in) :: U(*), T(*), K1(*)
      real(c_double), intent(out) :: dUdm(*), dUdn(*), dTdm(*), dTdn(*), d2Tdm(*), d2Tdn(*)
      real(c_double), intent(out) :: psiDiff(*), d2psidm(*), d2psidn(*)
    end function
    integer(c_int32_t) function ns_kdt(ctx, K1, dTdm, dTdn) bind(C, name="ns_kdt")
      import :: c_int32_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: K1(*)
      real(c_double), intent(inout) :: dTdm(*), dTdn(*)
    end function
    integer(c_int32_t) function ns_tau_ij(ctx, nk, MI, dUdm, dUdn, tau, dtaudm, dtaudn) bind(C, name="ns_tau_ij")
      import :: c_int32_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: nk
      real(c_double), intent(in) :: MI(*), dUdm(*), dUdn(*)
      real(c_double), intent(out) :: tau(*), dtaudm(*), dtaudn(*)
    end function
    integer(c_int32_t) function ns_eulernonlinear(ctx, nk, dUdt, U, dUdm, dUdn, tau, dtaudm, dtaudn, d2Tdm, &
                                                  d2Tdn, d2psidm, d2psidn, s_acou) bind(C, name="ns_eulernonlinear")
      import :: c_int32_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: nk
      real(c_double), intent(out) :: dUdt(*)
      real(c_double), intent(in) :: U(*), dUdm(*), dUdn(*), tau(*), dtaudm(*), dtaudn(*)
      real(c_double), intent(in) :: d2Tdm(*), d2Tdn(*), d2psidm(*), d2psidn(*), s_acou(*)
    end function
    integer(c_int32_t) function ns_metrics_x(ctx, nk, dU) bind(C, name="ns_metrics_x")
      import :: c_int32_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: nk
      real(c_double), intent(inout) :: dU(*)
    end function
    integer(c_int32_t) function ns_metrics_y(ctx, nk, dU) bind(C, name="ns_metrics_y")
      import :: c_int32_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: nk
      real(c_double), intent(inout) :: dU(*)
    end function
  end interface

contains

  ! error stop with the library's message when a call fails
  subroutine ns_check(rc, what)
    integer(c_int32_t), intent(in) :: rc
    character(len=*), intent(in) :: what
    character(kind=c_char), pointer :: msg(:)
    character(len=512) :: text
    integer :: i
    if (rc == NS_OK) return
    call c_f_pointer(ns_last_error(), msg, [512])
    text = ' '
    do i = 1, 512
      if (msg(i) == c_null_char) exit
      text(i:i) = msg(i)
    end do
    write(*, '(a,a,a,i0,a,a)') 'ns_api: ', what, ' failed (', rc, '): ', trim(text)
    error stop 1
  end subroutine ns_check

end module ns_capi


! One resident context for the whole program (the reference keeps one set of
! module globals).  U crosses the host/device boundary on every call of the
! per-step routines so that the caller's array stays authoritative, as in the
! reference; ns_shim_steps keeps the state on the GPU for many steps.
module ns_shim_core
  use ns_capi
  implicit none
  type(c_ptr), save :: ns_ctx_handle = c_null_ptr
  integer, save :: ns_ctx_nk = -1

contains

  logical function ns_shim_ready()
    ns_shim_ready = c_associated(ns_ctx_handle)
  end function

  subroutine ns_shim_init(p, meshx, meshy, sigmax, sigmay, m, n, rhob, psib, Ub)
    type(ns_params), intent(in) :: p
    real(c_double), intent(in) :: meshx(*), meshy(*), sigmax(*), sigmay(*), m(*), n(*), rhob(*), psib(*)
    real(c_double), intent(in) :: Ub(*)
    if (c_associated(ns_ctx_handle)) call ns_check(ns_destroy(ns_ctx_handle), 'ns_destroy')
    call ns_check(ns_create(p, ns_ctx_handle), 'ns_create')
    call ns_check(ns_set_profiles(ns_ctx_handle, meshx, meshy, sigmax, sigmay, m, n, rhob, psib), &
                  'ns_set_profiles')
    call ns_check(ns_set_base(ns_ctx_handle, Ub), 'ns_set_base')
    ns_ctx_nk = p%nk
  end subroutine

  subroutine ns_shim_finalize()
    if (c_associated(ns_ctx_handle)) call ns_check(ns_destroy(ns_ctx_handle), 'ns_destroy')
    ns_ctx_handle = c_null_ptr
  end subroutine

  ! rk.f90:12 rk_euler(U,t,iter,nk): U in, one LDDRK6 step on the GPU, U out
  subroutine ns_shim_step(U, iter)
    real(c_double), intent(inout) :: U(*)
    integer, intent(in) :: iter
    call ns_check(ns_put_state(ns_ctx_handle, U), 'ns_put_state')
    call ns_check(ns_rk_euler(ns_ctx_handle, int(iter, c_int32_t)), 'ns_rk_euler')
    call ns_check(ns_get_state(ns_ctx_handle, U), 'ns_get_state')
  end subroutine

  ! euler.f90:123-130 for nsteps steps with the state resident on the GPU
  subroutine ns_shim_steps(U, iter0, nsteps)
    real(c_double), intent(inout) :: U(*)
    integer, intent(in) :: iter0, nsteps
    call ns_check(ns_put_state(ns_ctx_handle, U), 'ns_put_state')
    call ns_check(ns_rk_steps(ns_ctx_handle, int(iter0, c_int32_t), int(nsteps, c_int32_t)), 'ns_rk_steps')
    call ns_check(ns_get_state(ns_ctx_handle, U), 'ns_get_state')
  end subroutine

end module ns_shim_core
