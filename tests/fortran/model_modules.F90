! Test stand-ins for the model modules that the Fortran drop-ins read
! (params_module, fields_ng_module, init_module, input_module, cons_module,
! lbc, ar_module, char_module) and the model's shutdown.  The global-grid
! members the ringfilter drop-in reads (nlon, nlat, dlamda, the chunk table)
! are variables here, parameters or allocated in the model.  They hold only the
! members the drop-ins and the test driver use, under the model's names and
! shapes (fields_ng.F90:12-60 and :163-166, advance.F:219-222), so that
! tiegcm3.0_ng_amd/fortran/ng_shims.F90 and trsolv_ng_shim.F90 compile and
! run here exactly as they would inside the model build.
module params_module
  implicit none
  integer :: nlon = 0, nlonp4 = 0, nlat = 0   ! global grid (ringfilter)
  integer, parameter :: mx_ng = 5
  integer :: n_ng = 1
  integer, dimension(mx_ng) :: nlevp1_ng = 0, nlat_ng = 0, nlon_ng = 0
  real, dimension(:, :), allocatable :: zpmid_ng, zpint_ng, glat_ng
end module params_module

module fields_ng_module
  use params_module, only: mx_ng
  implicit none
  type fields
    integer :: lon0, lon1, lat0, lat1, lond0, lond1, latd0, latd1
    logical, dimension(4) :: is_bndry
    real, dimension(:), allocatable :: racs, cs, tanphi, cor
    real, dimension(:, :), allocatable :: bx, by, bz, bmod2, rlatm, dipmag, sndec, csdec, chi
    real, dimension(:, :), allocatable :: t_lbc, u_lbc, v_lbc, flx_he
    real, dimension(:, :, :, :), allocatable :: &
      tn, un, vn, o2, o1, he, op, n2d, n4s, no, ar, ti, te, ne, w, o2p, z, &
      tn_nm, un_nm, vn_nm, o2_nm, o1_nm, he_nm, op_nm, n2d_nm, n4s_nm, no_nm, ar_nm, &
      mbar, barm, xnmbar, xnmbari, scht, schti
    real, dimension(:, :, :, :, :), allocatable :: lon_b, lat_b
    real, dimension(:, :, :, :), allocatable :: op_lon_b, op_lat_b
    real, dimension(:, :), allocatable :: tlbc, ulbc, vlbc, tlbc_nm, ulbc_nm, vlbc_nm
    real, dimension(:, :, :), allocatable :: &
      cp, kt, km, ui, vi, wi, xiop2p, xiop2d, nplus, n2p, nop, lxx, lyy, lxy, lyx, &
      qji_tn, cool_implicit, cool_explicit, hdt, hdu, hdv, hdo2, hdo1, hdhe, n2, Fe, Fn, &
      rk1, rk2, rk3, ra1, ra2, ra3, beta1, beta3, beta5, beta8, beta9, beta9n, beta17, &
      rk19, rk20, rk25, rkm12, rj, qtef, qtotal, qop2p, qop2d, qo2p, qop
  end type fields
  type(fields), dimension(mx_ng), target :: flds
  integer, dimension(mx_ng) :: itp = 1, itc = 2
  real, dimension(mx_ng) :: shapiro, dtx2inv, dphi, dlamda, dz, expzmid, expzmid_inv
  real, dimension(:, :), allocatable :: expz
  real, dimension(:, :, :), allocatable :: difk, dift, xmue
  character(len=*), dimension(*), parameter :: bndry = &
    (/'TN   ', 'UN   ', 'VN   ', 'O2   ', 'O1   ', 'HE   ', 'OP   ', 'N2D  ', 'N4S  ', 'NO   ', 'AR   ', 'OMEGA'/)
  integer, parameter :: nbnd = size(bndry)
end module fields_ng_module

module init_module
  implicit none
  integer :: iday = 1
end module init_module

module input_module
  implicit none
  integer :: nstep_sub = 1, calc_helium = 1
  real :: colfac = 1.5, opdiffcap = 0.
end module input_module

module cons_module
  implicit none
  integer, parameter :: difhor = 1
  real :: dlamda = 0.                        ! global grid (ringfilter)
  integer :: nlat_filter = 0
  integer, dimension(:), allocatable :: chunk_array
end module cons_module

module lbc
  implicit none
  real :: b(3, 3), fb(3)
  real, parameter :: pshelb = 0.1154e-5
end module lbc

module ar_module
  implicit none
  real, dimension(:), allocatable :: ar_glbm
end module ar_module

module char_module
  implicit none
contains
  pure function find_index(name, list) result(idx)
    character(len=*), intent(in) :: name
    character(len=*), dimension(:), intent(in) :: list
    integer :: idx
    do idx = 1, size(list)
      if (trim(list(idx)) == trim(name)) return
    enddo
    idx = 0
  end function find_index
end module char_module

subroutine shutdown(msg)
  implicit none
  character(len=*), intent(in) :: msg
  write(6, '(a)') 'shutdown: '//trim(msg)
  error stop 3
end subroutine shutdown
