! Test driver for the Fortran drop-ins: reads a system from raw files, calls
! trsolv / blktri exactly as the model does (trsolv per latitude, dt.F:330;
! blktri per latitude, duv.F:375) and writes x.  Used by the GPU test suite.
program shim_driver
  implicit none
  integer :: lev0, lev1, lon0, lon1, lonmax, nlev, ncol, mode, u
  real, allocatable :: a(:,:), b(:,:), c(:,:), f(:,:), x(:,:)
  real, allocatable :: a2(:,:,:,:), b2(:,:,:,:), c2(:,:,:,:), f2(:,:,:), x2(:,:,:), wk(:,:,:,:), wy(:,:,:)
  character(len=512) :: dir
  call get_command_argument(1, dir)
  open(newunit=u, file=trim(dir)//'/dims.txt', status='old')
  read(u, *) mode, lev0, lev1, lon0, lon1, lonmax
  close(u)
  nlev = lev1 - lev0 + 1
  ncol = lon1 - lon0 + 1
  if (mode == 1) then
    allocate(a(lev0:lev1, lon0:lon1), b(lev0:lev1, lon0:lon1), c(lev0:lev1, lon0:lon1))
    allocate(f(lev0:lev1, lon0:lon1), x(lev0:lev1, lon0:lon1))
    call rd(a, 'a'); call rd(b, 'b'); call rd(c, 'c'); call rd(f, 'f'); call rd(x, 'x')
    call trsolv(a, b, c, f, x, lev0, lev1, lev0, lev1 - 1, lon0, lon1, lonmax, 1, 0)
    call wr(x, 'xout')
  else
    allocate(a2(2,2,lev0:lev1,lon0:lon1), b2(2,2,lev0:lev1,lon0:lon1), c2(2,2,lev0:lev1,lon0:lon1))
    allocate(f2(2,lev0:lev1,lon0:lon1), x2(2,lev0:lev1,lon0:lon1), wk(2,2,lev0:lev1,lon0:lon1))
    allocate(wy(2,lev0:lev1,lon0:lon1))
    call rd4(a2, 'a'); call rd4(b2, 'b'); call rd4(c2, 'c'); call rd3(f2, 'f'); call rd3(x2, 'x')
    call blktri(a2, b2, c2, f2, wk, wk, wy, x2, lev0, lev1, lon0, lon1)
    call wr3(x2, 'xout')
  endif
contains
  subroutine rd(v, nm)
    real, intent(out) :: v(:,:)
    character(*), intent(in) :: nm
    integer :: uu
    open(newunit=uu, file=trim(dir)//'/'//nm//'.bin', access='stream', form='unformatted', status='old')
    read(uu) v
    close(uu)
  end subroutine rd
  subroutine wr(v, nm)
    real, intent(in) :: v(:,:)
    character(*), intent(in) :: nm
    integer :: uu
    open(newunit=uu, file=trim(dir)//'/'//nm//'.bin', access='stream', form='unformatted', status='replace')
    write(uu) v
    close(uu)
  end subroutine wr
  subroutine rd4(v, nm)
    real, intent(out) :: v(:,:,:,:)
    character(*), intent(in) :: nm
    integer :: uu
    open(newunit=uu, file=trim(dir)//'/'//nm//'.bin', access='stream', form='unformatted', status='old')
    read(uu) v
    close(uu)
  end subroutine rd4
  subroutine rd3(v, nm)
    real, intent(out) :: v(:,:,:)
    character(*), intent(in) :: nm
    integer :: uu
    open(newunit=uu, file=trim(dir)//'/'//nm//'.bin', access='stream', form='unformatted', status='old')
    read(uu) v
    close(uu)
  end subroutine rd3
  subroutine wr3(v, nm)
    real, intent(in) :: v(:,:,:)
    character(*), intent(in) :: nm
    integer :: uu
    open(newunit=uu, file=trim(dir)//'/'//nm//'.bin', access='stream', form='unformatted', status='replace')
    write(uu) v
    close(uu)
  end subroutine wr3
end program shim_driver
