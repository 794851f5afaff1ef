! Test driver for the nest drop-ins: fills the stand-in model modules from raw
! files in a directory (the model's own layout, written by the test), runs
! one nest step through the drop-ins in the order of dynamics_ng.F90:99-233
! (the O+ sub-cycles, the four minor species, dt_ng, duv_ng, comp_ng) and
! writes every result field back.  Also checks trsolv_ng's drop-in on the
! dt-shaped system of the same grid (mode 2 of the directory's run.txt);
! mode 3 is mode 1 with the nest's west side not a domain edge; mode 4 runs
! the ringfilter drop-in on every latitude of a global field.
program ng_driver
  use params_module, only: mx_ng, n_ng, nlevp1_ng, nlon_ng, nlat_ng, zpint_ng, zpmid_ng, glat_ng
  use fields_ng_module
  use init_module, only: iday
  use input_module, only: nstep_sub, calc_helium, colfac, opdiffcap
  use lbc, only: b, fb
  use ar_module, only: ar_glbm
  use char_module, only: find_index
  implicit none
  character(len=512) :: dir
  integer :: u, nk, nlon, nlat, nlond, nlatd, istep, isub, itp_sub, itc_sub, itmp, i, mode
  real :: rv(30)
  real, allocatable :: op_i(:, :, :, :), op_nm_i(:, :, :, :), lvl(:)
  real, allocatable :: a(:, :, :), bb(:, :, :), c(:, :, :), f(:, :, :), x(:, :, :)
  character(len=8), parameter :: f4(34) = [character(len=8) :: 'tn', 'un', 'vn', 'o2', 'o1', 'he', 'op', &
    'n2d', 'n4s', 'no', 'ar', 'ti', 'te', 'ne', 'w', 'o2p', 'z', 'tn_nm', 'un_nm', 'vn_nm', 'o2_nm', 'o1_nm', &
    'he_nm', 'op_nm', 'n2d_nm', 'n4s_nm', 'no_nm', 'ar_nm', 'mbar', 'barm', 'xnmbar', 'xnmbari', 'scht', 'schti']
  character(len=8), parameter :: lat_names(10) = [character(len=8) :: 'tn', 'un', 'vn', 'o2', 'o1', 'he', &
    'n2d', 'n4s', 'no', 'ar']
  character(len=8), parameter :: bnd_names(10) = [character(len=8) :: 'TN', 'UN', 'VN', 'O2', 'O1', 'HE', &
    'N2D', 'N4S', 'NO', 'AR']

  call get_command_argument(1, dir)
  open(newunit=u, file=trim(dir)//'/run.txt', status='old')
  read(u, *) mode, nk, nlon, nlat, nstep_sub, calc_helium
  read(u, *) rv(1:24)
  close(u)
  nlond = nlon + 4
  nlatd = nlat + 4
  n_ng = 1
  nlevp1_ng(1) = nk
  nlon_ng(1) = nlon
  nlat_ng(1) = nlat
  iday = 1
  dz(1) = rv(1); dlamda(1) = rv(2); dphi(1) = rv(3); dtx2inv(1) = rv(4); shapiro(1) = rv(5)
  expzmid(1) = rv(6); expzmid_inv(1) = rv(7); colfac = rv(8); opdiffcap = rv(9)
  allocate(ar_glbm(nk))
  ar_glbm = 0.
  ar_glbm(1:2) = rv(10:11)
  b = reshape(rv(12:20), [3, 3], order=[2, 1])   ! rv holds b row-major
  fb = rv(21:23)

  flds(1)%lon0 = 1; flds(1)%lon1 = nlon; flds(1)%lat0 = 1; flds(1)%lat1 = nlat
  flds(1)%lond0 = -1; flds(1)%lond1 = nlon + 2; flds(1)%latd0 = -1; flds(1)%latd1 = nlat + 2
  flds(1)%is_bndry = .true.
  if (mode == 3) flds(1)%is_bndry(1) = .false.   ! a nest split over tasks: refused

  if (mode == 4) then
    call ring_rows()
    stop
  endif
  if (mode == 2) then
    ! trsolv_ng on a system of this grid's shape, x written back
    allocate(a(nk, -1:nlon + 2, -1:nlat + 2), bb(nk, -1:nlon + 2, -1:nlat + 2), c(nk, -1:nlon + 2, -1:nlat + 2))
    allocate(f(nk, -1:nlon + 2, -1:nlat + 2), x(nk, -1:nlon + 2, -1:nlat + 2))
    call rd('a', a, size(a)); call rd('b', bb, size(bb)); call rd('c', c, size(c))
    call rd('f', f, size(f)); call rd('x', x, size(x))
    call trsolv_ng(a, bb, c, f, x, nk - 1, 1)
    call wr('xout', x, size(x))
    stop
  endif

  allocate(lvl(nk))
  allocate(expz(1, nk), difk(1, nk, 1), dift(1, nk, 1), xmue(1, nk, 1))
  allocate(zpint_ng(1, nk), zpmid_ng(1, nk), glat_ng(1, -1:nlat + 2))
  call rd('expz', lvl, nk); expz(1, :) = lvl
  call rd('difk', lvl, nk); difk(1, :, 1) = lvl
  call rd('dift', lvl, nk); dift(1, :, 1) = lvl
  call rd('xmue', lvl, nk); xmue(1, :, 1) = lvl
  call rd('zpint', lvl, nk); zpint_ng(1, :) = lvl
  call rd('zpmid', lvl, nk); zpmid_ng(1, :) = lvl
  call rd('glat', glat_ng(1, :), nlatd)
  allocate(flds(1)%racs(-1:nlat + 2), flds(1)%cs(-1:nlat + 2), flds(1)%tanphi(-1:nlat + 2), flds(1)%cor(-1:nlat + 2))
  call rd('racs', flds(1)%racs, nlatd)
  call rd('cs', flds(1)%cs, nlatd)
  call rd('tanphi', flds(1)%tanphi, nlatd)
  call rd('cor', flds(1)%cor, nlatd)

  call alloc_fields()
  do i = 1, size(f4)
    call rd4(trim(f4(i)))
  enddo
  call rd('w_upd', flds(1)%w(:, :, :, 2), nk * nlond * nlatd)
  call rd('n2', flds(1)%n2, size(flds(1)%n2))
  call rd('cp', flds(1)%cp, size(flds(1)%cp)); call rd('kt', flds(1)%kt, size(flds(1)%kt))
  call rd('km', flds(1)%km, size(flds(1)%km)); call rd('ui', flds(1)%ui, size(flds(1)%ui))
  call rd('vi', flds(1)%vi, size(flds(1)%vi)); call rd('wi', flds(1)%wi, size(flds(1)%wi))
  call rd('lxx', flds(1)%lxx, size(flds(1)%lxx)); call rd('lyy', flds(1)%lyy, size(flds(1)%lyy))
  call rd('lxy', flds(1)%lxy, size(flds(1)%lxy)); call rd('lyx', flds(1)%lyx, size(flds(1)%lyx))
  call rd('qji_tn', flds(1)%qji_tn, size(flds(1)%qji_tn))
  call rd('cool_imp', flds(1)%cool_implicit, size(flds(1)%cool_implicit))
  call rd('cool_exp', flds(1)%cool_explicit, size(flds(1)%cool_explicit))
  call rd('hdt', flds(1)%hdt, size(flds(1)%hdt)); call rd('hdu', flds(1)%hdu, size(flds(1)%hdu))
  call rd('hdv', flds(1)%hdv, size(flds(1)%hdv)); call rd('hdo2', flds(1)%hdo2, size(flds(1)%hdo2))
  call rd('hdo1', flds(1)%hdo1, size(flds(1)%hdo1)); call rd('hdhe', flds(1)%hdhe, size(flds(1)%hdhe))
  call rd('rk1', flds(1)%rk1, size(flds(1)%rk1)); call rd('rk2', flds(1)%rk2, size(flds(1)%rk2))
  call rd('rk3', flds(1)%rk3, size(flds(1)%rk3)); call rd('ra1', flds(1)%ra1, size(flds(1)%ra1))
  call rd('ra2', flds(1)%ra2, size(flds(1)%ra2)); call rd('ra3', flds(1)%ra3, size(flds(1)%ra3))
  call rd('beta1', flds(1)%beta1, size(flds(1)%beta1)); call rd('beta3', flds(1)%beta3, size(flds(1)%beta3))
  call rd('beta5', flds(1)%beta5, size(flds(1)%beta5)); call rd('beta8', flds(1)%beta8, size(flds(1)%beta8))
  call rd('beta9', flds(1)%beta9, size(flds(1)%beta9)); call rd('beta9n', flds(1)%beta9n, size(flds(1)%beta9n))
  call rd('beta17', flds(1)%beta17, size(flds(1)%beta17)); call rd('rk19', flds(1)%rk19, size(flds(1)%rk19))
  call rd('rk20', flds(1)%rk20, size(flds(1)%rk20)); call rd('rk25', flds(1)%rk25, size(flds(1)%rk25))
  call rd('rkm12', flds(1)%rkm12, size(flds(1)%rkm12)); call rd('rj', flds(1)%rj, size(flds(1)%rj))
  call rd('qtef', flds(1)%qtef, size(flds(1)%qtef)); call rd('qtotal', flds(1)%qtotal, size(flds(1)%qtotal))
  call rd('qop2p', flds(1)%qop2p, size(flds(1)%qop2p)); call rd('qop2d', flds(1)%qop2d, size(flds(1)%qop2d))
  call rd('qo2p', flds(1)%qo2p, size(flds(1)%qo2p)); call rd('qop', flds(1)%qop, size(flds(1)%qop))
  call rd('n2p', flds(1)%n2p, size(flds(1)%n2p)); call rd('nplus', flds(1)%nplus, size(flds(1)%nplus))
  call rd('nop', flds(1)%nop, size(flds(1)%nop))
  call rd('t_lbc', flds(1)%t_lbc, size(flds(1)%t_lbc)); call rd('u_lbc', flds(1)%u_lbc, size(flds(1)%u_lbc))
  call rd('v_lbc', flds(1)%v_lbc, size(flds(1)%v_lbc)); call rd('flx_he', flds(1)%flx_he, size(flds(1)%flx_he))
  call rd('tlbc', flds(1)%tlbc, size(flds(1)%tlbc)); call rd('ulbc', flds(1)%ulbc, size(flds(1)%ulbc))
  call rd('vlbc', flds(1)%vlbc, size(flds(1)%vlbc))
  call rd('bx', flds(1)%bx, size(flds(1)%bx)); call rd('by', flds(1)%by, size(flds(1)%by))
  call rd('bz', flds(1)%bz, size(flds(1)%bz)); call rd('bmod2', flds(1)%bmod2, size(flds(1)%bmod2))
  call rd('dipmag', flds(1)%dipmag, size(flds(1)%dipmag)); call rd('sndec', flds(1)%sndec, size(flds(1)%sndec))
  call rd('csdec', flds(1)%csdec, size(flds(1)%csdec)); call rd('chi', flds(1)%chi, size(flds(1)%chi))
  call rd('rlatm', flds(1)%rlatm, size(flds(1)%rlatm))
  istep = 1
  do i = 1, size(lat_names)
    call rd('lonb_'//trim(lat_names(i)), flds(1)%lon_b(:, :, :, istep, find_index(bnd_names(i), bndry)), &
            nk * 2 * nlatd)
    call rd('latb_'//trim(lat_names(i)), flds(1)%lat_b(:, :, :, istep, find_index(bnd_names(i), bndry)), &
            nk * nlond * 2)
  enddo
  call rd('lonb_op', flds(1)%op_lon_b(:, :, :, 1:nstep_sub), nk * 2 * nlatd * nstep_sub)
  call rd('latb_op', flds(1)%op_lat_b(:, :, :, 1:nstep_sub), nk * nlond * 2 * nstep_sub)

  ! dynamics_ng.F90:99-121: the O+ sub-cycles alternate the two time levels
  itp_sub = itp(1)
  itc_sub = itc(1)
  allocate(op_i, source=flds(1)%op)
  allocate(op_nm_i, source=flds(1)%op_nm)
  do isub = 1, nstep_sub
    call oplus_ng(op_i(:, :, :, itp_sub), op_nm_i(:, :, :, itp_sub), op_i(:, :, :, itc_sub), &
                  op_nm_i(:, :, :, itc_sub), flds(1)%xiop2p, flds(1)%xiop2d, flds(1)%Fe, flds(1)%Fn, &
                  (istep - 1) * nstep_sub + isub, 1)
    if (isub /= nstep_sub) then
      itmp = itp_sub
      itp_sub = itc_sub
      itc_sub = itmp
    endif
  enddo
  flds(1)%op(:, :, :, itc(1)) = op_i(:, :, :, itc_sub)
  flds(1)%op_nm(:, :, :, itc(1)) = op_nm_i(:, :, :, itc_sub)
  ! :147-165
  call comp_n2d_ng(flds(1)%n2d(:, :, :, itc(1)), flds(1)%n2d_nm(:, :, :, itc(1)), istep, 1)
  call comp_n4s_ng(flds(1)%n4s(:, :, :, itc(1)), flds(1)%n4s_nm(:, :, :, itc(1)), istep, 1)
  call comp_no_ng(flds(1)%no(:, :, :, itc(1)), flds(1)%no_nm(:, :, :, itc(1)), istep, 1)
  call comp_ar_ng(flds(1)%ar(:, :, :, itc(1)), flds(1)%ar_nm(:, :, :, itc(1)), istep, 1)
  ! :200-233
  call dt_ng(flds(1)%tn(:, :, :, itc(1)), flds(1)%tn_nm(:, :, :, itc(1)), flds(1)%tlbc, flds(1)%tlbc_nm, istep, 1)
  call duv_ng(flds(1)%un(:, :, :, itc(1)), flds(1)%un_nm(:, :, :, itc(1)), flds(1)%vn(:, :, :, itc(1)), &
              flds(1)%vn_nm(:, :, :, itc(1)), flds(1)%ulbc, flds(1)%vlbc, flds(1)%ulbc_nm, flds(1)%vlbc_nm, &
              istep, 1)
  call comp_ng(flds(1)%o2(:, :, :, itc(1)), flds(1)%o2_nm(:, :, :, itc(1)), flds(1)%o1(:, :, :, itc(1)), &
               flds(1)%o1_nm(:, :, :, itc(1)), flds(1)%he(:, :, :, itc(1)), flds(1)%he_nm(:, :, :, itc(1)), &
               flds(1)%flx_he, istep, 1)

  do i = 1, size(f4)
    select case (trim(f4(i)))
    case ('tn', 'un', 'vn', 'o2', 'o1', 'he', 'op', 'n2d', 'n4s', 'no', 'ar', 'tn_nm', 'un_nm', 'vn_nm', &
          'o2_nm', 'o1_nm', 'he_nm', 'op_nm', 'n2d_nm', 'n4s_nm', 'no_nm', 'ar_nm')
      call wr4(trim(f4(i)))
    end select
  enddo
  call wr('xiop2p', flds(1)%xiop2p, size(flds(1)%xiop2p)); call wr('xiop2d', flds(1)%xiop2d, size(flds(1)%xiop2d))
  call wr('Fe', flds(1)%Fe, size(flds(1)%Fe)); call wr('Fn', flds(1)%Fn, size(flds(1)%Fn))
  call wr('tlbc', flds(1)%tlbc, size(flds(1)%tlbc)); call wr('tlbc_nm', flds(1)%tlbc_nm, size(flds(1)%tlbc_nm))
  call wr('ulbc', flds(1)%ulbc, size(flds(1)%ulbc)); call wr('vlbc', flds(1)%vlbc, size(flds(1)%vlbc))
  call wr('ulbc_nm', flds(1)%ulbc_nm, size(flds(1)%ulbc_nm)); call wr('vlbc_nm', flds(1)%vlbc_nm, size(flds(1)%vlbc_nm))
  call wr('flx_he', flds(1)%flx_he, size(flds(1)%flx_he))

contains

  ! mode 4: the ringfilter drop-in on every latitude of rows.bin, f(nlon+4,
  ! nk, nlat) (run.txt: nstep_sub = order, calc_helium = 1 for 'OPLUS',
  ! dlamda; chunk.txt: the chunk table)
  subroutine ring_rows()
    use params_module, only: g_nlon => nlon, g_nlonp4 => nlonp4, g_nlat => nlat
    use cons_module, only: c_dlamda => dlamda, nlat_filter, chunk_array
    real, allocatable :: rows(:, :, :)
    integer :: lat, uu
    g_nlon = nlon
    g_nlonp4 = nlon + 4
    g_nlat = nlat
    c_dlamda = rv(2)
    open(newunit=uu, file=trim(dir)//'/chunk.txt', status='old')
    read(uu, *) nlat_filter
    allocate(chunk_array(nlat_filter))
    read(uu, *) chunk_array
    close(uu)
    allocate(rows(nlon + 4, nk, nlat))
    call rd('rows', rows, size(rows))
    do lat = 1, nlat
      if (calc_helium == 1) then
        call ringfilter(rows(:, :, lat), 1, nk, nstep_sub, 'OPLUS', lat)
      else
        call ringfilter(rows(:, :, lat), 1, nk, nstep_sub, 'TN', lat)
      endif
    enddo
    call wr('rows', rows, size(rows))
  end subroutine ring_rows

  subroutine alloc_fields()
    integer :: l0, l1, j0, j1
    l0 = -1; l1 = nlon + 2; j0 = -1; j1 = nlat + 2
    associate (d => flds(1))
      allocate(d%tn(nk, l0:l1, j0:j1, 2))
      allocate(d%un, d%vn, d%o2, d%o1, d%he, d%op, d%n2d, d%n4s, d%no, d%ar, d%ti, d%te, d%ne, d%w, d%o2p, &
               d%z, d%tn_nm, d%un_nm, d%vn_nm, d%o2_nm, d%o1_nm, d%he_nm, d%op_nm, d%n2d_nm, d%n4s_nm, &
               d%no_nm, d%ar_nm, d%mbar, d%barm, d%xnmbar, d%xnmbari, d%scht, d%schti, mold=d%tn)
      allocate(d%cp(nk, l0:l1, j0:j1))
      allocate(d%kt, d%km, d%ui, d%vi, d%wi, d%xiop2p, d%xiop2d, d%nplus, d%n2p, d%nop, d%lxx, d%lyy, &
               d%lxy, d%lyx, d%qji_tn, d%cool_implicit, d%cool_explicit, d%hdt, d%hdu, d%hdv, d%hdo2, &
               d%hdo1, d%hdhe, d%n2, d%Fe, d%Fn, d%rk1, d%rk2, d%rk3, d%ra1, d%ra2, d%ra3, d%beta1, &
               d%beta3, d%beta5, d%beta8, d%beta9, d%beta9n, d%beta17, d%rk19, d%rk20, d%rk25, d%rkm12, &
               d%rj, d%qtef, d%qtotal, d%qop2p, d%qop2d, d%qo2p, d%qop, mold=d%cp)
      allocate(d%bx(l0:l1, j0:j1))
      allocate(d%by, d%bz, d%bmod2, d%rlatm, d%dipmag, d%sndec, d%csdec, d%chi, d%t_lbc, d%u_lbc, d%v_lbc, &
               d%flx_he, d%tlbc, d%ulbc, d%vlbc, d%tlbc_nm, d%ulbc_nm, d%vlbc_nm, mold=d%bx)
      allocate(d%lon_b(nk, 1:2, j0:j1, 0:1, nbnd), d%lat_b(nk, l0:l1, 3:4, 0:1, nbnd))
      allocate(d%op_lon_b(nk, 1:2, j0:j1, 0:nstep_sub), d%op_lat_b(nk, l0:l1, 3:4, 0:nstep_sub))
      d%lon_b = 0.
      d%lat_b = 0.
      d%op_lon_b = 0.
      d%op_lat_b = 0.
      d%tlbc_nm = 0.
      d%ulbc_nm = 0.
      d%vlbc_nm = 0.
    end associate
  end subroutine alloc_fields

  subroutine rd4(name)
    character(len=*), intent(in) :: name
    real, pointer :: p(:, :, :, :)
    p => field4(name)
    p = 0.
    call rd(name, p(:, :, :, itp(1)), nk * nlond * nlatd)
  end subroutine rd4

  subroutine wr4(name)
    character(len=*), intent(in) :: name
    real, pointer :: p(:, :, :, :)
    p => field4(name)
    call wr(name, p(:, :, :, itc(1)), nk * nlond * nlatd)
  end subroutine wr4

  function field4(name) result(p)
    character(len=*), intent(in) :: name
    real, pointer :: p(:, :, :, :)
    select case (name)
    case ('tn'); p => flds(1)%tn
    case ('un'); p => flds(1)%un
    case ('vn'); p => flds(1)%vn
    case ('o2'); p => flds(1)%o2
    case ('o1'); p => flds(1)%o1
    case ('he'); p => flds(1)%he
    case ('op'); p => flds(1)%op
    case ('n2d'); p => flds(1)%n2d
    case ('n4s'); p => flds(1)%n4s
    case ('no'); p => flds(1)%no
    case ('ar'); p => flds(1)%ar
    case ('ti'); p => flds(1)%ti
    case ('te'); p => flds(1)%te
    case ('ne'); p => flds(1)%ne
    case ('w'); p => flds(1)%w
    case ('o2p'); p => flds(1)%o2p
    case ('z'); p => flds(1)%z
    case ('tn_nm'); p => flds(1)%tn_nm
    case ('un_nm'); p => flds(1)%un_nm
    case ('vn_nm'); p => flds(1)%vn_nm
    case ('o2_nm'); p => flds(1)%o2_nm
    case ('o1_nm'); p => flds(1)%o1_nm
    case ('he_nm'); p => flds(1)%he_nm
    case ('op_nm'); p => flds(1)%op_nm
    case ('n2d_nm'); p => flds(1)%n2d_nm
    case ('n4s_nm'); p => flds(1)%n4s_nm
    case ('no_nm'); p => flds(1)%no_nm
    case ('ar_nm'); p => flds(1)%ar_nm
    case ('mbar'); p => flds(1)%mbar
    case ('barm'); p => flds(1)%barm
    case ('xnmbar'); p => flds(1)%xnmbar
    case ('xnmbari'); p => flds(1)%xnmbari
    case ('scht'); p => flds(1)%scht
    case ('schti'); p => flds(1)%schti
    case default
      write(6, '(a)') 'ng_driver: unknown field '//name
      error stop 2
    end select
  end function field4

  subroutine rd(name, v, n)
    character(len=*), intent(in) :: name
    integer, intent(in) :: n
    real, intent(out) :: v(n)
    integer :: uu
    open(newunit=uu, file=trim(dir)//'/'//name//'.bin', access='stream', form='unformatted', status='old')
    read(uu) v
    close(uu)
  end subroutine rd

  subroutine wr(name, v, n)
    character(len=*), intent(in) :: name
    integer, intent(in) :: n
    real, intent(in) :: v(n)
    integer :: uu
    open(newunit=uu, file=trim(dir)//'/out_'//name//'.bin', access='stream', form='unformatted', &
         status='replace')
    write(uu) v
    close(uu)
  end subroutine wr
end program ng_driver
