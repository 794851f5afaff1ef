! iso_c_binding interfaces to the host-layout context of the fused nest step
! (include/tgx_step.h).  Compile with the model's flags (-fdefault-real-8: the
! model's real is C double).
module tgx_ng_c
  use iso_c_binding, only: c_int, c_double, c_ptr, c_char, c_null_char, c_f_pointer, c_associated
  implicit none

  integer(c_int), parameter :: TGX_SCHEME_AUTO = 0, TGX_SCHEME_THOMAS = 1, TGX_SCHEME_PCR = 2

  ! tgx_ng_grid, member for member
  type, bind(C) :: tgx_ng_grid
    integer(c_int) :: nk, nlon, nlat, ld
    integer(c_int) :: nstep_sub, calc_helium, scheme
    real(c_double) :: dz, dlamda, dphi, dtx2inv, shapiro, expzmid, expzmid_inv
    real(c_double) :: colfac, opdiffcap, difhor, pshelb, ar_glbm1, ar_glbm2
    real(c_double) :: b(9), fb(3)
    type(c_ptr) :: expz, difk, dift, xmue, zpint, zpmid
    type(c_ptr) :: racs, cs, tanphi, cor, glat
    type(c_ptr) :: capture
  end type tgx_ng_grid

  interface
    type(c_ptr) function tgx_ng_ctx_create(g) bind(C, name="tgx_ng_ctx_create")
      import :: c_ptr, tgx_ng_grid
      type(tgx_ng_grid), intent(in) :: g
    end function tgx_ng_ctx_create

    integer(c_int) function tgx_ng_ctx_destroy(ctx) bind(C, name="tgx_ng_ctx_destroy")
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
    end function tgx_ng_ctx_destroy

    integer(c_int) function tgx_ng_ctx_put(ctx, name, host) bind(C, name="tgx_ng_ctx_put")
      import :: c_ptr, c_int, c_char, c_double
      type(c_ptr), value :: ctx
      character(kind=c_char), intent(in) :: name(*)
      real(c_double), intent(in) :: host(*)
    end function tgx_ng_ctx_put

    integer(c_int) function tgx_ng_ctx_put_out(ctx, name, host) bind(C, name="tgx_ng_ctx_put_out")
      import :: c_ptr, c_int, c_char, c_double
      type(c_ptr), value :: ctx
      character(kind=c_char), intent(in) :: name(*)
      real(c_double), intent(in) :: host(*)
    end function tgx_ng_ctx_put_out

    integer(c_int) function tgx_ng_ctx_get(ctx, name, host) bind(C, name="tgx_ng_ctx_get")
      import :: c_ptr, c_int, c_char, c_double
      type(c_ptr), value :: ctx
      character(kind=c_char), intent(in) :: name(*)
      real(c_double), intent(inout) :: host(*)
    end function tgx_ng_ctx_get

    integer(c_int) function tgx_ng_ctx_run(ctx, routine) bind(C, name="tgx_ng_ctx_run")
      import :: c_ptr, c_int, c_char
      type(c_ptr), value :: ctx
      character(kind=c_char), intent(in) :: routine(*)
    end function tgx_ng_ctx_run

    type(c_ptr) function tgx_last_error_c() bind(C, name="tgx_last_error")
      import :: c_ptr
    end function tgx_last_error_c
  end interface

contains

  ! The library's last error message as a Fortran string.
  function tgx_error_message() result(msg)
    character(len=512) :: msg
    character(kind=c_char), pointer :: s(:)
    type(c_ptr) :: p
    integer :: i
    msg = ' '
    p = tgx_last_error_c()
    if (.not. c_associated(p)) return
    call c_f_pointer(p, s, [len(msg)])
    do i = 1, len(msg)
      if (s(i) == c_null_char) exit
      msg(i:i) = s(i)
    enddo
  end function tgx_error_message
end module tgx_ng_c
