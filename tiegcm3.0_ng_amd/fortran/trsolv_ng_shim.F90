! Drop-in replacement for trsolv_ng (trsolv_ng.F90:1-28) with the same name
! and arguments.  The array bounds live in model module state, so this file
! is compiled inside the model build (params_module, fields_ng_module); it
! passes the dimensions to libtgx by value.
subroutine trsolv_ng(a, b, c, f, x, k2, i_ng)
  use params_module, only: nlevp1_ng
  use fields_ng_module, only: flds
  use tgx_c, only: tgx_trsolv_ng_f64
  implicit none
  integer, intent(in) :: k2, i_ng
  real, dimension(nlevp1_ng(i_ng), flds(i_ng)%lond0:flds(i_ng)%lond1, &
                  flds(i_ng)%latd0:flds(i_ng)%latd1), intent(in) :: a, b, c, f
  real, dimension(nlevp1_ng(i_ng), flds(i_ng)%lond0:flds(i_ng)%lond1, &
                  flds(i_ng)%latd0:flds(i_ng)%latd1), intent(inout) :: x
  call tgx_trsolv_ng_f64(a, b, c, f, x, nlevp1_ng(i_ng), &
                         flds(i_ng)%lond1 - flds(i_ng)%lond0 + 1, &
                         flds(i_ng)%latd1 - flds(i_ng)%latd0 + 1, k2)
end subroutine trsolv_ng
