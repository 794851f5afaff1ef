! C binding onto the reference's own interp_module (compiled from its source
! where it lies in the reference tree), so the oracle's interp1d/interp2d/
! interp3d can be checked bit for bit against the compiled reference.  Test
! infrastructure only.  Arrays in Fortran layout.
subroutine tgx_ref_interp3d(z, nz, x, nx, y, ny, zp, nzp, xp, nxp, yp, nyp, fp, zlog, f) &
    bind(C, name="tgx_ref_interp3d")
  use iso_c_binding, only: c_double, c_int
  use interp_module, only: interp3d
  implicit none
  integer(c_int), value :: nz, nx, ny, nzp, nxp, nyp, zlog
  real(c_double), intent(in) :: z(nz), x(nx), y(ny), zp(nzp), xp(nxp), yp(nyp), fp(nzp, nxp, nyp)
  real(c_double), intent(out) :: f(nz, nx, ny)
  f = interp3d(z, x, y, zp, xp, yp, fp, zlog /= 0)
end subroutine tgx_ref_interp3d

subroutine tgx_ref_interp2d(x, nx, y, ny, xp, nxp, yp, nyp, fp, f) bind(C, name="tgx_ref_interp2d")
  use iso_c_binding, only: c_double, c_int
  use interp_module, only: interp2d
  implicit none
  integer(c_int), value :: nx, ny, nxp, nyp
  real(c_double), intent(in) :: x(nx), y(ny), xp(nxp), yp(nyp), fp(nxp, nyp)
  real(c_double), intent(out) :: f(nx, ny)
  f = interp2d(x, y, xp, yp, fp)
end subroutine tgx_ref_interp2d

subroutine tgx_ref_interp1d(x, nx, xp, nxp, fp, f) bind(C, name="tgx_ref_interp1d")
  use iso_c_binding, only: c_double, c_int
  use interp_module, only: interp1d
  implicit none
  integer(c_int), value :: nx, nxp
  real(c_double), intent(in) :: x(nx), xp(nxp), fp(nxp)
  real(c_double), intent(out) :: f(nx)
  f = interp1d(x, xp, fp)
end subroutine tgx_ref_interp1d
