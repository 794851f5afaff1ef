! Drop-in replacements for the reference's global-grid column solvers with
! the same names and arguments (trsolv.F:2-3, duv.F:801-820); they forward to
! libtgx.  Link these instead of trsolv.F and blktri from duv.F.
subroutine trsolv(a, b, c, f, x, lev0, lev1, k1, k2, lon0, lon1, lonmax, lat, idebug)
  use tgx_c, only: tgx_trsolv_f64
  implicit none
  integer, intent(in) :: lev0, lev1, k1, k2, lon0, lon1, lonmax, lat, idebug
  real, dimension(lev0:lev1, lon0:lon1), intent(in) :: a, b, c, f
  real, dimension(lev0:lev1, lon0:lon1), intent(inout) :: x
  call tgx_trsolv_f64(a, b, c, f, x, lev0, lev1, k1, k2, lon0, lon1, lonmax, lat, idebug)
end subroutine trsolv

! blktri reads nlonp4 from params_module in the reference (duv.F:806); the
! shim takes it the same way when built inside the model, or from the
! optional module variable below when built stand-alone.
subroutine blktri(a, b, c, f, beta, gamma, y, x, lev0, lev1, lon0, lon1)
  use tgx_c, only: tgx_blktri_f64
#ifdef TGX_IN_MODEL
  use params_module, only: nlonp4
#endif
  implicit none
#ifndef TGX_IN_MODEL
  integer, parameter :: nlonp4 = TGX_NLONP4
#endif
  integer, intent(in) :: lev0, lev1, lon0, lon1
  real, dimension(2, 2, lev0:lev1, lon0:lon1), intent(in) :: a, b, c
  real, dimension(2, lev0:lev1, lon0:lon1), intent(in) :: f
  real, dimension(2, 2, lev0:lev1, lon0:lon1), intent(inout) :: beta, gamma
  real, dimension(2, lev0:lev1, lon0:lon1), intent(inout) :: x, y
  call tgx_blktri_f64(a, b, c, f, beta, gamma, y, x, lev0, lev1, lon0, lon1, nlonp4)
end subroutine blktri
