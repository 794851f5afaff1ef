! Drop-in replacement for ringfilter (ringfilter.F:1-111) with the same name
! and arguments.  It reads the module state the reference reads (params_module
! nlon, nlat; cons_module dlamda, nlat_filter, chunk_array), so it is compiled
! inside the model build, and filters the latitude row on the GPU.  An order
! the filter cannot take ends in shutdown, as the reference's does.
subroutine ringfilter(f, lev0, lev1, order, name, lat)
  use params_module, only: nlon, nlonp4, nlat
  use cons_module, only: dlamda, nlat_filter, chunk_array
  use iso_c_binding, only: c_null_char, c_int
  use tgx_c, only: tgx_ringfilter_f64, tgx_last_status
  implicit none
  integer, intent(in) :: lev0, lev1, order, lat
  real, dimension(nlonp4, lev0:lev1), intent(inout) :: f
  character(len=*), intent(in) :: name
  integer(c_int) :: chunk(max(nlat_filter, 1))
  external :: shutdown
  chunk = 0
  if (nlat_filter > 0) chunk(1:nlat_filter) = chunk_array(1:nlat_filter)
  call tgx_ringfilter_f64(f, lev0, lev1, order, trim(name)//c_null_char, lat, nlon, nlat, dlamda, &
                          nlat_filter, chunk)
  if (tgx_last_status() /= 0) call shutdown('ringfilter: filtering order too high or bad arguments')
end subroutine ringfilter
