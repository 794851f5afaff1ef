! Drop-ins for the nest column routines, with the reference's names and
! arguments: dt_ng (dt_ng.F90:1), duv_ng (duv_ng.F90:1), oplus_ng
! (oplus_ng.F90:1), comp_ng (comp_ng.F90:1) and the minor species
! comp_n2d_ng, comp_n4s_ng, comp_no_ng, comp_ar_ng (comp_*_ng.F90:1, each
! with its minor_ng call).  Like the routines they replace they read the
! model's module state (params_module, fields_ng_module, ...), so this file
! is compiled inside the model build, after those modules; drop the replaced
! routines from the source list.
!
! Each call copies the inputs its routine reads into the nest's device
! context (include/tgx_step.h), runs the routine there and copies the
! results back into the arguments.  The context of nest i_ng is made at the
! first call from the module state and remade when iday changes (difk, dift
! and xmue depend on it).  The minor species run fused: the first of the
! four calls of a step (comp_n2d_ng in dynamics_ng.F90:147-165) solves all
! four, using this step's N2D for NO and N4S as the reference does, and
! each species' call returns its pair; a second call for a species in the
! same step solves again.  comp_o2o_ng stays in the model: comp_ng's
! drop-in forms the same fs terms itself from the same inputs.  The nest
! must run on one task (all four sides domain edges), as the fused step
! assumes; library failures go to the model's shutdown.
module tgx_ng_shim_module
  use iso_c_binding, only: c_ptr, c_null_ptr, c_associated, c_loc, c_null_char, c_int, c_double
  use tgx_ng_c
  use params_module, only: mx_ng
  implicit none
  private
  public :: ng_context, ng_put, ng_put_out, ng_get, ng_run, ng_minor_species

  type(c_ptr), save :: ctx(mx_ng) = c_null_ptr
  integer, save :: ctx_day(mx_ng) = -1
  integer, save :: minors_step(mx_ng) = -1
  logical, save :: minors_open(4, mx_ng) = .false.

contains

  function ng_context(i_ng) result(c)
    use params_module, only: nlevp1_ng, zpint_ng, zpmid_ng, glat_ng
    use fields_ng_module, only: flds, shapiro, dtx2inv, dlamda, dphi, dz, expzmid, expzmid_inv, &
      expz, difk, dift, xmue
    use init_module, only: iday
    use input_module, only: nstep_sub, colfac, opdiffcap, calc_helium
    use cons_module, only: difhor
    use lbc, only: b, fb, pshelb
    use ar_module, only: ar_glbm
    integer, intent(in) :: i_ng
    type(c_ptr) :: c
    type(tgx_ng_grid) :: g
    real(c_double), dimension(:), allocatable, target :: h_expz, h_difk, h_dift, h_xmue, h_zpint, h_zpmid
    real(c_double), dimension(:), allocatable, target :: h_racs, h_cs, h_tanphi, h_cor, h_glat
    integer :: nk, l, m, latd0, latd1
    external :: shutdown

    if (c_associated(ctx(i_ng)) .and. ctx_day(i_ng) == iday) then
      c = ctx(i_ng)
      return
    endif
    if (c_associated(ctx(i_ng))) call check(tgx_ng_ctx_destroy(ctx(i_ng)), 'ctx_destroy')
    ctx(i_ng) = c_null_ptr
    if (.not. all(flds(i_ng)%is_bndry)) &
      call shutdown('tgx: the fused nest step needs the nest on one task (all four sides domain edges)')

    nk = nlevp1_ng(i_ng)
    latd0 = flds(i_ng)%latd0
    latd1 = flds(i_ng)%latd1
    g%nk = nk
    g%nlon = flds(i_ng)%lon1 - flds(i_ng)%lon0 + 1
    g%nlat = flds(i_ng)%lat1 - flds(i_ng)%lat0 + 1
    g%ld = 0
    g%nstep_sub = nstep_sub
    g%calc_helium = calc_helium
    g%scheme = TGX_SCHEME_THOMAS
    g%dz = dz(i_ng)
    g%dlamda = dlamda(i_ng)
    g%dphi = dphi(i_ng)
    g%dtx2inv = dtx2inv(i_ng)
    g%shapiro = shapiro(i_ng)
    g%expzmid = expzmid(i_ng)
    g%expzmid_inv = expzmid_inv(i_ng)
    g%colfac = colfac
    g%opdiffcap = opdiffcap
    g%difhor = difhor
    g%pshelb = pshelb
    g%ar_glbm1 = ar_glbm(1)
    g%ar_glbm2 = ar_glbm(2)
    do l = 1, 3
      do m = 1, 3
        g%b(3 * (l - 1) + m) = b(l, m)
      enddo
    enddo
    g%fb = fb
    h_expz = expz(i_ng, 1:nk)
    h_difk = difk(i_ng, 1:nk, iday)
    h_dift = dift(i_ng, 1:nk, iday)
    h_xmue = xmue(i_ng, 1:nk, iday)
    h_zpint = zpint_ng(i_ng, 1:nk)
    h_zpmid = zpmid_ng(i_ng, 1:nk)
    h_racs = flds(i_ng)%racs(latd0:latd1)
    h_cs = flds(i_ng)%cs(latd0:latd1)
    h_tanphi = flds(i_ng)%tanphi(latd0:latd1)
    h_cor = flds(i_ng)%cor(latd0:latd1)
    h_glat = glat_ng(i_ng, latd0:latd1)
    g%expz = c_loc(h_expz)
    g%difk = c_loc(h_difk)
    g%dift = c_loc(h_dift)
    g%xmue = c_loc(h_xmue)
    g%zpint = c_loc(h_zpint)
    g%zpmid = c_loc(h_zpmid)
    g%racs = c_loc(h_racs)
    g%cs = c_loc(h_cs)
    g%tanphi = c_loc(h_tanphi)
    g%cor = c_loc(h_cor)
    g%glat = c_loc(h_glat)
    g%capture = c_null_ptr
    ctx(i_ng) = tgx_ng_ctx_create(g)
    if (.not. c_associated(ctx(i_ng))) call check(1_c_int, 'ctx_create')
    ctx_day(i_ng) = iday
    minors_step(i_ng) = -1
    c = ctx(i_ng)
  end function ng_context

  subroutine check(rc, what)
    integer(c_int), intent(in) :: rc
    character(*), intent(in) :: what
    external :: shutdown
    if (rc /= 0) then
      write(6, '(a)') 'tgx '//what//': '//trim(tgx_error_message())
      call shutdown('tgx: '//what//' failed')
    endif
  end subroutine check

  subroutine ng_put(c, name, f)
    type(c_ptr), intent(in) :: c
    character(*), intent(in) :: name
    real, intent(in) :: f(*)
    call check(tgx_ng_ctx_put(c, name//c_null_char, f), 'put '//name)
  end subroutine ng_put

  subroutine ng_put_out(c, name, f)
    type(c_ptr), intent(in) :: c
    character(*), intent(in) :: name
    real, intent(in) :: f(*)
    call check(tgx_ng_ctx_put_out(c, name//c_null_char, f), 'put_out '//name)
  end subroutine ng_put_out

  subroutine ng_get(c, name, f)
    type(c_ptr), intent(in) :: c
    character(*), intent(in) :: name
    real, intent(inout) :: f(*)
    call check(tgx_ng_ctx_get(c, name//c_null_char, f), 'get '//name)
  end subroutine ng_get

  subroutine ng_run(c, routine)
    type(c_ptr), intent(in) :: c
    character(*), intent(in) :: routine
    call check(tgx_ng_ctx_run(c, routine//c_null_char), 'run '//routine)
  end subroutine ng_run

  ! The lateral data of field `fname` (a name of bndry) at step istep.
  subroutine put_lateral(c, i_ng, istep, fname, cname)
    use fields_ng_module, only: flds, bndry
    use char_module, only: find_index
    type(c_ptr), intent(in) :: c
    integer, intent(in) :: i_ng, istep
    character(*), intent(in) :: fname, cname
    integer :: idx
    idx = find_index(fname, bndry)
    call ng_put(c, 'lonb_'//cname, flds(i_ng)%lon_b(:, :, :, istep, idx))
    call ng_put(c, 'latb_'//cname, flds(i_ng)%lat_b(:, :, :, istep, idx))
  end subroutine put_lateral

  ! Species isp (1 N2D, 2 N4S, 3 NO, 4 AR) of the fused minor solve of step
  ! istep into f, fnm; solves the four species first when this step's
  ! result for isp has been taken already or not yet computed.
  subroutine ng_minor_species(i_ng, istep, isp, f, fnm)
    use fields_ng_module, only: flds, itp, itc
    integer, intent(in) :: i_ng, istep, isp
    real, intent(inout) :: f(*), fnm(*)
    character(len=3), parameter :: sp(4) = ['n2d', 'n4s', 'no ', 'ar ']
    character(len=3), parameter :: bn(4) = ['N2D', 'N4S', 'NO ', 'AR ']
    type(c_ptr) :: c
    integer :: t, q
    c = ng_context(i_ng)
    if (minors_step(i_ng) /= istep .or. .not. minors_open(isp, i_ng)) then
      t = itp(i_ng)
      call ng_put(c, 'tn', flds(i_ng)%tn(:, :, :, t))
      call ng_put(c, 'o2', flds(i_ng)%o2(:, :, :, t))
      call ng_put(c, 'o1', flds(i_ng)%o1(:, :, :, t))
      call ng_put(c, 'he', flds(i_ng)%he(:, :, :, t))
      call ng_put(c, 'op', flds(i_ng)%op(:, :, :, t))
      call ng_put(c, 'ne', flds(i_ng)%ne(:, :, :, t))
      call ng_put(c, 'o2p', flds(i_ng)%o2p(:, :, :, t))
      call ng_put(c, 'mbar', flds(i_ng)%mbar(:, :, :, t))
      call ng_put(c, 'barm', flds(i_ng)%barm(:, :, :, t))
      call ng_put(c, 'xnmbar', flds(i_ng)%xnmbar(:, :, :, t))
      call ng_put(c, 'xnmbari', flds(i_ng)%xnmbari(:, :, :, t))
      call ng_put(c, 'n2d', flds(i_ng)%n2d(:, :, :, t))
      call ng_put(c, 'n2d_nm', flds(i_ng)%n2d_nm(:, :, :, t))
      call ng_put(c, 'n4s', flds(i_ng)%n4s(:, :, :, t))
      call ng_put(c, 'n4s_nm', flds(i_ng)%n4s_nm(:, :, :, t))
      call ng_put(c, 'no', flds(i_ng)%no(:, :, :, t))
      call ng_put(c, 'no_nm', flds(i_ng)%no_nm(:, :, :, t))
      call ng_put(c, 'ar', flds(i_ng)%ar(:, :, :, t))
      call ng_put(c, 'ar_nm', flds(i_ng)%ar_nm(:, :, :, t))
      call ng_put(c, 'w_upd', flds(i_ng)%w(:, :, :, itc(i_ng)))
      call ng_put(c, 'n2', flds(i_ng)%n2)
      call ng_put(c, 'n2p', flds(i_ng)%n2p)
      call ng_put(c, 'nplus', flds(i_ng)%nplus)
      call ng_put(c, 'nop', flds(i_ng)%nop)
      call ng_put(c, 'qtef', flds(i_ng)%qtef)
      call ng_put(c, 'qop2p', flds(i_ng)%qop2p)
      call ng_put(c, 'qop2d', flds(i_ng)%qop2d)
      call ng_put(c, 'qop', flds(i_ng)%qop)
      call ng_put(c, 'qo2p', flds(i_ng)%qo2p)
      call ng_put(c, 'rj', flds(i_ng)%rj)
      call ng_put(c, 'rkm12', flds(i_ng)%rkm12)
      call ng_put(c, 'rk1', flds(i_ng)%rk1)
      call ng_put(c, 'rk2', flds(i_ng)%rk2)
      call ng_put(c, 'rk3', flds(i_ng)%rk3)
      call ng_put(c, 'rk19', flds(i_ng)%rk19)
      call ng_put(c, 'rk20', flds(i_ng)%rk20)
      call ng_put(c, 'rk25', flds(i_ng)%rk25)
      call ng_put(c, 'ra1', flds(i_ng)%ra1)
      call ng_put(c, 'ra2', flds(i_ng)%ra2)
      call ng_put(c, 'ra3', flds(i_ng)%ra3)
      call ng_put(c, 'beta1', flds(i_ng)%beta1)
      call ng_put(c, 'beta3', flds(i_ng)%beta3)
      call ng_put(c, 'beta5', flds(i_ng)%beta5)
      call ng_put(c, 'beta8', flds(i_ng)%beta8)
      call ng_put(c, 'beta9', flds(i_ng)%beta9)
      call ng_put(c, 'beta9n', flds(i_ng)%beta9n)
      call ng_put(c, 'beta17', flds(i_ng)%beta17)
      call ng_put(c, 'tlbc', flds(i_ng)%tlbc)
      do q = 1, 4
        call put_lateral(c, i_ng, istep, trim(bn(q)), trim(sp(q)))
      enddo
      call ng_run(c, 'minors')
      minors_step(i_ng) = istep
      minors_open(:, i_ng) = .true.
    endif
    call ng_get(c, trim(sp(isp)), f)
    call ng_get(c, trim(sp(isp))//'_nm', fnm)
    minors_open(isp, i_ng) = .false.
  end subroutine ng_minor_species
end module tgx_ng_shim_module

subroutine dt_ng(tn_upd, tn_nm_upd, tlbc, tlbc_nm, istep, i_ng)
  use params_module, only: nlevp1_ng
  use fields_ng_module, only: flds, itp, itc, bndry
  use char_module, only: find_index
  use tgx_ng_shim_module
  use iso_c_binding, only: c_ptr
  implicit none
  integer, intent(in) :: istep, i_ng
  real, dimension(nlevp1_ng(i_ng), flds(i_ng)%lond0:flds(i_ng)%lond1, flds(i_ng)%latd0:flds(i_ng)%latd1), &
    intent(out) :: tn_upd, tn_nm_upd
  real, dimension(flds(i_ng)%lond0:flds(i_ng)%lond1, flds(i_ng)%latd0:flds(i_ng)%latd1), intent(inout) :: tlbc
  real, dimension(flds(i_ng)%lond0:flds(i_ng)%lond1, flds(i_ng)%latd0:flds(i_ng)%latd1), intent(out) :: tlbc_nm
  type(c_ptr) :: c
  integer :: t
  c = ng_context(i_ng)
  t = itp(i_ng)
  call ng_put(c, 'tn', flds(i_ng)%tn(:, :, :, t))
  call ng_put(c, 'tn_nm', flds(i_ng)%tn_nm(:, :, :, t))
  call ng_put(c, 'un', flds(i_ng)%un(:, :, :, t))
  call ng_put(c, 'vn', flds(i_ng)%vn(:, :, :, t))
  call ng_put(c, 'o1', flds(i_ng)%o1(:, :, :, t))
  call ng_put(c, 'mbar', flds(i_ng)%mbar(:, :, :, t))
  call ng_put(c, 'xnmbar', flds(i_ng)%xnmbar(:, :, :, t))
  call ng_put(c, 'scht', flds(i_ng)%scht(:, :, :, t))
  call ng_put(c, 'schti', flds(i_ng)%schti(:, :, :, t))
  call ng_put(c, 'w_upd', flds(i_ng)%w(:, :, :, itc(i_ng)))
  call ng_put(c, 'cp', flds(i_ng)%cp)
  call ng_put(c, 'kt', flds(i_ng)%kt)
  call ng_put(c, 'km', flds(i_ng)%km)
  call ng_put(c, 'hdt', flds(i_ng)%hdt)
  call ng_put(c, 'qji_tn', flds(i_ng)%qji_tn)
  call ng_put(c, 'cool_imp', flds(i_ng)%cool_implicit)
  call ng_put(c, 'cool_exp', flds(i_ng)%cool_explicit)
  call ng_put(c, 'qtotal', flds(i_ng)%qtotal)
  call ng_put(c, 'rkm12', flds(i_ng)%rkm12)
  call ng_put(c, 't_lbc', flds(i_ng)%t_lbc)
  call ng_put(c, 'tlbc', tlbc)
  call ng_put(c, 'ulbc', flds(i_ng)%ulbc)
  call ng_put(c, 'vlbc', flds(i_ng)%vlbc)
  call ng_put(c, 'lonb_tn', flds(i_ng)%lon_b(:, :, :, istep, find_index('TN', bndry)))
  call ng_put(c, 'latb_tn', flds(i_ng)%lat_b(:, :, :, istep, find_index('TN', bndry)))
  call ng_run(c, 'dt')
  call ng_get(c, 'tn', tn_upd)
  call ng_get(c, 'tn_nm', tn_nm_upd)
  call ng_get(c, 'tlbc', tlbc)
  call ng_get(c, 'tlbc_nm', tlbc_nm)
end subroutine dt_ng

subroutine duv_ng(un_upd, unm_upd, vn_upd, vnm_upd, ulbc, vlbc, ulbc_nm, vlbc_nm, istep, i_ng)
  use params_module, only: nlevp1_ng
  use fields_ng_module, only: flds, itp, itc, bndry
  use char_module, only: find_index
  use tgx_ng_shim_module
  use iso_c_binding, only: c_ptr
  implicit none
  integer, intent(in) :: istep, i_ng
  real, dimension(nlevp1_ng(i_ng), flds(i_ng)%lond0:flds(i_ng)%lond1, flds(i_ng)%latd0:flds(i_ng)%latd1), &
    intent(out) :: un_upd, unm_upd, vn_upd, vnm_upd
  real, dimension(flds(i_ng)%lond0:flds(i_ng)%lond1, flds(i_ng)%latd0:flds(i_ng)%latd1), intent(inout) :: ulbc, vlbc
  real, dimension(flds(i_ng)%lond0:flds(i_ng)%lond1, flds(i_ng)%latd0:flds(i_ng)%latd1), intent(out) :: &
    ulbc_nm, vlbc_nm
  type(c_ptr) :: c
  integer :: t, iu, iv
  c = ng_context(i_ng)
  t = itp(i_ng)
  call ng_put(c, 'tn', flds(i_ng)%tn(:, :, :, t))
  call ng_put(c, 'tn_nm', flds(i_ng)%tn_nm(:, :, :, t))
  call ng_put(c, 'un', flds(i_ng)%un(:, :, :, t))
  call ng_put(c, 'vn', flds(i_ng)%vn(:, :, :, t))
  call ng_put(c, 'un_nm', flds(i_ng)%un_nm(:, :, :, t))
  call ng_put(c, 'vn_nm', flds(i_ng)%vn_nm(:, :, :, t))
  call ng_put(c, 'mbar', flds(i_ng)%mbar(:, :, :, t))
  call ng_put(c, 'scht', flds(i_ng)%scht(:, :, :, t))
  call ng_put(c, 'schti', flds(i_ng)%schti(:, :, :, t))
  call ng_put(c, 'z', flds(i_ng)%z(:, :, :, t))
  call ng_put(c, 'w_upd', flds(i_ng)%w(:, :, :, itc(i_ng)))
  call ng_put(c, 'hdu', flds(i_ng)%hdu)
  call ng_put(c, 'hdv', flds(i_ng)%hdv)
  call ng_put(c, 'ui', flds(i_ng)%ui)
  call ng_put(c, 'vi', flds(i_ng)%vi)
  call ng_put(c, 'lxx', flds(i_ng)%lxx)
  call ng_put(c, 'lyy', flds(i_ng)%lyy)
  call ng_put(c, 'lxy', flds(i_ng)%lxy)
  call ng_put(c, 'lyx', flds(i_ng)%lyx)
  call ng_put(c, 'km', flds(i_ng)%km)
  call ng_put(c, 'u_lbc', flds(i_ng)%u_lbc)
  call ng_put(c, 'v_lbc', flds(i_ng)%v_lbc)
  call ng_put(c, 'ulbc', ulbc)
  call ng_put(c, 'vlbc', vlbc)
  iu = find_index('UN', bndry)
  iv = find_index('VN', bndry)
  call ng_put(c, 'lonb_un', flds(i_ng)%lon_b(:, :, :, istep, iu))
  call ng_put(c, 'latb_un', flds(i_ng)%lat_b(:, :, :, istep, iu))
  call ng_put(c, 'lonb_vn', flds(i_ng)%lon_b(:, :, :, istep, iv))
  call ng_put(c, 'latb_vn', flds(i_ng)%lat_b(:, :, :, istep, iv))
  ! this step's TN (dt_ng) and Fe, Fn (oplus_ng)
  call ng_put_out(c, 'tn', flds(i_ng)%tn(:, :, :, itc(i_ng)))
  call ng_put_out(c, 'Fe', flds(i_ng)%Fe)
  call ng_put_out(c, 'Fn', flds(i_ng)%Fn)
  call ng_run(c, 'duv')
  call ng_get(c, 'un', un_upd)
  call ng_get(c, 'un_nm', unm_upd)
  call ng_get(c, 'vn', vn_upd)
  call ng_get(c, 'vn_nm', vnm_upd)
  call ng_get(c, 'ulbc', ulbc)
  call ng_get(c, 'vlbc', vlbc)
  call ng_get(c, 'ulbc_nm', ulbc_nm)
  call ng_get(c, 'vlbc_nm', vlbc_nm)
end subroutine duv_ng

! One O+ sub-step (dynamics_ng.F90:103-119 calls it nstep_sub times per
! step); istep counts sub-steps and selects the lateral data block.
subroutine oplus_ng(op, optm1, opout, optm1out, xiop2p, xiop2d, Fe, Fn, istep, i_ng)
  use params_module, only: nlevp1_ng
  use fields_ng_module, only: flds, itp
  use tgx_ng_shim_module
  use iso_c_binding, only: c_ptr
  implicit none
  integer, intent(in) :: istep, i_ng
  real, dimension(nlevp1_ng(i_ng), flds(i_ng)%lond0:flds(i_ng)%lond1, flds(i_ng)%latd0:flds(i_ng)%latd1), &
    intent(in) :: op, optm1
  real, dimension(nlevp1_ng(i_ng), flds(i_ng)%lond0:flds(i_ng)%lond1, flds(i_ng)%latd0:flds(i_ng)%latd1), &
    intent(out) :: opout, optm1out, xiop2p, xiop2d, Fe, Fn
  type(c_ptr) :: c
  integer :: t
  c = ng_context(i_ng)
  t = itp(i_ng)
  call ng_put(c, 'tn', flds(i_ng)%tn(:, :, :, t))
  call ng_put(c, 'te', flds(i_ng)%te(:, :, :, t))
  call ng_put(c, 'ti', flds(i_ng)%ti(:, :, :, t))
  call ng_put(c, 'o2', flds(i_ng)%o2(:, :, :, t))
  call ng_put(c, 'o1', flds(i_ng)%o1(:, :, :, t))
  call ng_put(c, 'he', flds(i_ng)%he(:, :, :, t))
  call ng_put(c, 'n2d', flds(i_ng)%n2d(:, :, :, t))
  call ng_put(c, 'ne', flds(i_ng)%ne(:, :, :, t))
  call ng_put(c, 'un', flds(i_ng)%un(:, :, :, t))
  call ng_put(c, 'vn', flds(i_ng)%vn(:, :, :, t))
  call ng_put(c, 'w', flds(i_ng)%w(:, :, :, t))
  call ng_put(c, 'xnmbar', flds(i_ng)%xnmbar(:, :, :, t))
  call ng_put(c, 'scht', flds(i_ng)%scht(:, :, :, t))
  call ng_put(c, 'n2', flds(i_ng)%n2)
  call ng_put(c, 'ui', flds(i_ng)%ui)
  call ng_put(c, 'vi', flds(i_ng)%vi)
  call ng_put(c, 'wi', flds(i_ng)%wi)
  call ng_put(c, 'qop2p', flds(i_ng)%qop2p)
  call ng_put(c, 'qop2d', flds(i_ng)%qop2d)
  call ng_put(c, 'qop', flds(i_ng)%qop)
  call ng_put(c, 'rk1', flds(i_ng)%rk1)
  call ng_put(c, 'rk2', flds(i_ng)%rk2)
  call ng_put(c, 'rk19', flds(i_ng)%rk19)
  call ng_put(c, 'rk20', flds(i_ng)%rk20)
  call ng_put(c, 'rk25', flds(i_ng)%rk25)
  call ng_put(c, 'bx', flds(i_ng)%bx)
  call ng_put(c, 'by', flds(i_ng)%by)
  call ng_put(c, 'bz', flds(i_ng)%bz)
  call ng_put(c, 'bmod2', flds(i_ng)%bmod2)
  call ng_put(c, 'dipmag', flds(i_ng)%dipmag)
  call ng_put(c, 'sndec', flds(i_ng)%sndec)
  call ng_put(c, 'csdec', flds(i_ng)%csdec)
  call ng_put(c, 'chi', flds(i_ng)%chi)
  call ng_put(c, 'rlatm', flds(i_ng)%rlatm)
  call ng_put(c, 'op', op)
  call ng_put(c, 'op_nm', optm1)
  call ng_put(c, 'lonb_op_sub', flds(i_ng)%op_lon_b(:, :, :, istep))
  call ng_put(c, 'latb_op_sub', flds(i_ng)%op_lat_b(:, :, :, istep))
  call ng_run(c, 'oplus_sub')
  call ng_get(c, 'op', opout)
  call ng_get(c, 'op_nm', optm1out)
  call ng_get(c, 'xiop2p', xiop2p)
  call ng_get(c, 'xiop2d', xiop2d)
  call ng_get(c, 'Fe', Fe)
  call ng_get(c, 'Fn', Fn)
end subroutine oplus_ng

subroutine comp_ng(o2_upd, o2nm_upd, o1_upd, o1nm_upd, he_upd, henm_upd, flx_he, istep, i_ng)
  use params_module, only: nlevp1_ng
  use fields_ng_module, only: flds, itp, bndry
  use char_module, only: find_index
  use tgx_ng_shim_module
  use iso_c_binding, only: c_ptr
  implicit none
  integer, intent(in) :: istep, i_ng
  real, dimension(nlevp1_ng(i_ng), flds(i_ng)%lond0:flds(i_ng)%lond1, flds(i_ng)%latd0:flds(i_ng)%latd1), &
    intent(out) :: o2_upd, o2nm_upd, o1_upd, o1nm_upd, he_upd, henm_upd
  real, dimension(flds(i_ng)%lond0:flds(i_ng)%lond1, flds(i_ng)%latd0:flds(i_ng)%latd1), intent(inout) :: flx_he
  type(c_ptr) :: c
  integer :: t, q
  character(len=2), parameter :: sp(3) = ['o2', 'o1', 'he']
  character(len=2), parameter :: bn(3) = ['O2', 'O1', 'HE']
  c = ng_context(i_ng)
  t = itp(i_ng)
  call ng_put(c, 'tn', flds(i_ng)%tn(:, :, :, t))
  call ng_put(c, 'o2', flds(i_ng)%o2(:, :, :, t))
  call ng_put(c, 'o2_nm', flds(i_ng)%o2_nm(:, :, :, t))
  call ng_put(c, 'o1', flds(i_ng)%o1(:, :, :, t))
  call ng_put(c, 'o1_nm', flds(i_ng)%o1_nm(:, :, :, t))
  call ng_put(c, 'he', flds(i_ng)%he(:, :, :, t))
  call ng_put(c, 'he_nm', flds(i_ng)%he_nm(:, :, :, t))
  call ng_put(c, 'w', flds(i_ng)%w(:, :, :, t))
  call ng_put(c, 'mbar', flds(i_ng)%mbar(:, :, :, t))
  call ng_put(c, 'xnmbar', flds(i_ng)%xnmbar(:, :, :, t))
  call ng_put(c, 'hdo2', flds(i_ng)%hdo2)
  call ng_put(c, 'hdo1', flds(i_ng)%hdo1)
  call ng_put(c, 'hdhe', flds(i_ng)%hdhe)
  ! comp_o2o_ng's inputs: the drop-in forms fs itself
  call ng_put(c, 'op', flds(i_ng)%op(:, :, :, t))
  call ng_put(c, 'ne', flds(i_ng)%ne(:, :, :, t))
  call ng_put(c, 'o2p', flds(i_ng)%o2p(:, :, :, t))
  call ng_put(c, 'n2d', flds(i_ng)%n2d(:, :, :, t))
  call ng_put(c, 'n4s', flds(i_ng)%n4s(:, :, :, t))
  call ng_put(c, 'no', flds(i_ng)%no(:, :, :, t))
  call ng_put(c, 'n2p', flds(i_ng)%n2p)
  call ng_put(c, 'nplus', flds(i_ng)%nplus)
  call ng_put(c, 'nop', flds(i_ng)%nop)
  call ng_put(c, 'qo2p', flds(i_ng)%qo2p)
  call ng_put(c, 'qop', flds(i_ng)%qop)
  call ng_put(c, 'rj', flds(i_ng)%rj)
  call ng_put(c, 'rkm12', flds(i_ng)%rkm12)
  call ng_put(c, 'rk1', flds(i_ng)%rk1)
  call ng_put(c, 'rk3', flds(i_ng)%rk3)
  call ng_put(c, 'ra1', flds(i_ng)%ra1)
  call ng_put(c, 'ra2', flds(i_ng)%ra2)
  call ng_put(c, 'beta1', flds(i_ng)%beta1)
  call ng_put(c, 'beta3', flds(i_ng)%beta3)
  call ng_put(c, 'beta8', flds(i_ng)%beta8)
  ! flds%tlbc as dt_ng of this step left it (the library's t_lbc slot)
  call ng_put(c, 't_lbc', flds(i_ng)%tlbc)
  call ng_put(c, 'flx_he', flx_he)
  do q = 1, 3
    call ng_put(c, 'lonb_'//sp(q), flds(i_ng)%lon_b(:, :, :, istep, find_index(bn(q), bndry)))
    call ng_put(c, 'latb_'//sp(q), flds(i_ng)%lat_b(:, :, :, istep, find_index(bn(q), bndry)))
  enddo
  call ng_run(c, 'comp')
  call ng_get(c, 'o2', o2_upd)
  call ng_get(c, 'o2_nm', o2nm_upd)
  call ng_get(c, 'o1', o1_upd)
  call ng_get(c, 'o1_nm', o1nm_upd)
  call ng_get(c, 'he', he_upd)
  call ng_get(c, 'he_nm', henm_upd)
  call ng_get(c, 'flx_he', flx_he)
end subroutine comp_ng

subroutine comp_n2d_ng(n2d_out, n2d_nm1_out, istep, i_ng)
  use params_module, only: nlevp1_ng
  use fields_ng_module, only: flds
  use tgx_ng_shim_module, only: ng_minor_species
  implicit none
  integer, intent(in) :: istep, i_ng
  real, dimension(nlevp1_ng(i_ng), flds(i_ng)%lond0:flds(i_ng)%lond1, flds(i_ng)%latd0:flds(i_ng)%latd1), &
    intent(out) :: n2d_out, n2d_nm1_out
  call ng_minor_species(i_ng, istep, 1, n2d_out, n2d_nm1_out)
end subroutine comp_n2d_ng

subroutine comp_n4s_ng(n4s_out, n4s_nm1_out, istep, i_ng)
  use params_module, only: nlevp1_ng
  use fields_ng_module, only: flds
  use tgx_ng_shim_module, only: ng_minor_species
  implicit none
  integer, intent(in) :: istep, i_ng
  real, dimension(nlevp1_ng(i_ng), flds(i_ng)%lond0:flds(i_ng)%lond1, flds(i_ng)%latd0:flds(i_ng)%latd1), &
    intent(out) :: n4s_out, n4s_nm1_out
  call ng_minor_species(i_ng, istep, 2, n4s_out, n4s_nm1_out)
end subroutine comp_n4s_ng

subroutine comp_no_ng(no_out, no_nm1_out, istep, i_ng)
  use params_module, only: nlevp1_ng
  use fields_ng_module, only: flds
  use tgx_ng_shim_module, only: ng_minor_species
  implicit none
  integer, intent(in) :: istep, i_ng
  real, dimension(nlevp1_ng(i_ng), flds(i_ng)%lond0:flds(i_ng)%lond1, flds(i_ng)%latd0:flds(i_ng)%latd1), &
    intent(out) :: no_out, no_nm1_out
  call ng_minor_species(i_ng, istep, 3, no_out, no_nm1_out)
end subroutine comp_no_ng

subroutine comp_ar_ng(ar_out, arnm_out, istep, i_ng)
  use params_module, only: nlevp1_ng
  use fields_ng_module, only: flds
  use tgx_ng_shim_module, only: ng_minor_species
  implicit none
  integer, intent(in) :: istep, i_ng
  real, dimension(nlevp1_ng(i_ng), flds(i_ng)%lond0:flds(i_ng)%lond1, flds(i_ng)%latd0:flds(i_ng)%latd1), &
    intent(out) :: ar_out, arnm_out
  call ng_minor_species(i_ng, istep, 4, ar_out, arnm_out)
end subroutine comp_ar_ng
