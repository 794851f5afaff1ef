! iso_c_binding interfaces to libtgx (include/tgx.h).  Compile with the
! model's flags (-fdefault-real-8: the model's real is C double).
module tgx_c
  use iso_c_binding, only: c_double, c_int, c_char, c_ptr
  implicit none
  interface
    subroutine tgx_trsolv_f64(a, b, c, f, x, lev0, lev1, k1, k2, lon0, lon1, lonmax, lat, &
                              idebug) bind(C, name="tgx_trsolv_f64")
      import :: c_double, c_int
      real(c_double), intent(in) :: a(*), b(*), c(*), f(*)
      real(c_double), intent(inout) :: x(*)
      integer(c_int), value :: lev0, lev1, k1, k2, lon0, lon1, lonmax, lat, idebug
    end subroutine tgx_trsolv_f64

    subroutine tgx_trsolv_ng_f64(a, b, c, f, x, nk, nlond, nlatd, k2) &
        bind(C, name="tgx_trsolv_ng_f64")
      import :: c_double, c_int
      real(c_double), intent(in) :: a(*), b(*), c(*), f(*)
      real(c_double), intent(inout) :: x(*)
      integer(c_int), value :: nk, nlond, nlatd, k2
    end subroutine tgx_trsolv_ng_f64

    subroutine tgx_blktri_f64(a, b, c, f, beta, gamma, y, x, lev0, lev1, lon0, lon1, lonmax) &
        bind(C, name="tgx_blktri_f64")
      import :: c_double, c_int
      real(c_double), intent(in) :: a(*), b(*), c(*), f(*)
      real(c_double), intent(inout) :: beta(*), gamma(*), y(*), x(*)
      integer(c_int), value :: lev0, lev1, lon0, lon1, lonmax
    end subroutine tgx_blktri_f64

    subroutine tgx_ringfilter_f64(f, lev0, lev1, order, name, lat, nlon, nlat, dlamda, nlat_filter, &
                                  chunk) bind(C, name="tgx_ringfilter_f64")
      import :: c_double, c_int, c_char
      real(c_double), intent(inout) :: f(*)
      integer(c_int), value :: lev0, lev1, order, lat, nlon, nlat, nlat_filter
      character(kind=c_char), intent(in) :: name(*)
      real(c_double), value :: dlamda
      integer(c_int), intent(in) :: chunk(*)
    end subroutine tgx_ringfilter_f64

    integer(c_int) function tgx_last_status() bind(C, name="tgx_last_status")
      import :: c_int
    end function tgx_last_status
  end interface
end module tgx_c
