! C binding onto the reference's own matutil_module (compiled from its source
! where it lies in the reference tree), so the oracle's matinv3/matinv2 can be
! checked bit for bit against the compiled reference.  Test infrastructure only.
subroutine tgx_ref_matinv3(a, b) bind(C, name="tgx_ref_matinv3")
  use iso_c_binding, only: c_double
  use matutil_module, only: matinv3
  implicit none
  real(c_double), intent(in) :: a(3,3)
  real(c_double), intent(out) :: b(3,3)
  b = matinv3(a)
end subroutine tgx_ref_matinv3

subroutine tgx_ref_matinv2(a, b) bind(C, name="tgx_ref_matinv2")
  use iso_c_binding, only: c_double
  use matutil_module, only: matinv2
  implicit none
  real(c_double), intent(in) :: a(2,2)
  real(c_double), intent(out) :: b(2,2)
  b = matinv2(a)
end subroutine tgx_ref_matinv2
