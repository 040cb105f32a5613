!=====================================================================
! A Fortran program that uses the satt_host module the way the reference's
! Fortran caller would: it reads a src_rec table (the reference's CSV
! columns, station and event names as CHARACTER(30), src/src_rec.f90:24-48),
! takes the model from raw files, and runs forward_simulate over the table's
! own arrays through satt_forward_simulate_host, which matches the names to
! build the station list and the gathers (src/src_rec.f90:162-217,252-283).
!
!   satt_host_example04 table.csv model.bin out.bin precision
!
! model.bin (stream, native doubles): nx, ny, nperiod as doubles, then
! xgrids(nx), ygrids(ny), svel, m11, m12, m22 (nperiod,nx,ny).
! out.bin: status, chi, then tt_fwd(npath), adj_s(nperiod,nx,ny).
!=====================================================================
program satt_host_example04
  use, intrinsic :: iso_c_binding
  use satt_host, only: satt_forward_simulate_host, satt_host_release
  implicit none
  character(len=1024) :: line, csv, modf, outf, arg
  character(len=30), allocatable :: evtname(:), staname(:)
  real(c_double), allocatable :: tt(:), stla(:), stlo(:), evla(:), evlo(:), per(:), wt(:)
  real(c_double), allocatable :: xg(:), yg(:), periods(:), svel(:,:,:), m11(:,:,:)
  real(c_double), allocatable :: m12(:,:,:), m22(:,:,:), tt_fwd(:), adj_s(:,:,:), dens(:,:,:)
  character(len=64) :: fields(32), head(32)
  real(c_double) :: hdr(3), chi
  integer :: u, ios, npath, nf, nh, i, k, nx, ny, np, prec, st
  integer :: c_tt, c_stla, c_stlo, c_evla, c_evlo, c_evt, c_sta, c_per, c_w

  call get_command_argument(1, csv)
  call get_command_argument(2, modf)
  call get_command_argument(3, outf)
  call get_command_argument(4, arg)
  read(arg, *) prec

  ! the table: header, then one path per line
  open(newunit=u, file=trim(csv), status='old', action='read')
  read(u, '(a)') line
  call split(line, head, nh)
  c_tt = col('tt'); c_stla = col('stla'); c_stlo = col('stlo'); c_evla = col('evla')
  c_evlo = col('evlo'); c_evt = col('evtname'); c_sta = col('staname'); c_per = col('period')
  c_w = col('weight')
  npath = 0
  do
    read(u, '(a)', iostat=ios) line
    if (ios /= 0) exit
    if (len_trim(line) > 0) npath = npath + 1
  end do
  allocate(evtname(npath), staname(npath), tt(npath), stla(npath), stlo(npath), evla(npath), &
           evlo(npath), per(npath), wt(npath))
  rewind(u)
  read(u, '(a)') line
  do i = 1, npath
    read(u, '(a)') line
    call split(line, fields, nf)
    read(fields(c_tt), *) tt(i)
    read(fields(c_stla), *) stla(i)
    read(fields(c_stlo), *) stlo(i)
    read(fields(c_evla), *) evla(i)
    read(fields(c_evlo), *) evlo(i)
    read(fields(c_per), *) per(i)
    read(fields(c_w), *) wt(i)
    evtname(i) = fields(c_evt)
    staname(i) = fields(c_sta)
  end do
  close(u)

  ! the sorted unique periods (SrcRec%periods)
  allocate(periods(npath))
  np = 0
  do i = 1, npath
    if (np == 0) then
      np = 1; periods(1) = per(i)
    else if (all(abs(periods(1:np) - per(i)) > 0)) then
      np = np + 1; periods(np) = per(i)
    end if
  end do
  call sort(periods(1:np))

  open(newunit=u, file=trim(modf), access='stream', form='unformatted', status='old')
  read(u) hdr
  nx = nint(hdr(1)); ny = nint(hdr(2))
  if (nint(hdr(3)) /= np) stop 'model periods do not match the table'
  allocate(xg(nx), yg(ny), svel(np,nx,ny), m11(np,nx,ny), m12(np,nx,ny), m22(np,nx,ny))
  read(u) xg, yg, svel, m11, m12, m22
  close(u)

  allocate(tt_fwd(npath), adj_s(np,nx,ny), dens(np,nx,ny))
  call satt_forward_simulate_host(2, xg, yg, periods(1:np), svel, m11, m12, m22, &
                                  evtname, staname, evla, evlo, stla, stlo, per, tt, wt, &
                                  .true., .true., tt_fwd, chi, adj_s, dens, st, prec)
  open(newunit=u, file=trim(outf), access='stream', form='unformatted', status='replace')
  write(u) real(st, c_double), chi, tt_fwd, adj_s
  close(u)
  call satt_host_release()
  if (st < 0) stop 1

contains

  integer function col(name)
    character(len=*), intent(in) :: name
    integer :: k
    col = 0
    do k = 1, nh
      if (trim(adjustl(head(k))) == name) col = k
    end do
    if (col == 0) stop 'missing column'
  end function col

  subroutine split(s, out, n)
    character(len=*), intent(in) :: s
    character(len=64), intent(out) :: out(:)
    integer, intent(out) :: n
    integer :: a, b
    n = 0
    a = 1
    do
      b = index(s(a:), ',')
      n = n + 1
      if (b == 0) then
        out(n) = adjustl(s(a:len_trim(s)))
        exit
      end if
      out(n) = adjustl(s(a:a + b - 2))
      a = a + b
    end do
  end subroutine split

  subroutine sort(v)
    real(c_double), intent(inout) :: v(:)
    integer :: i, j
    real(c_double) :: t
    do i = 2, size(v)
      t = v(i)
      j = i - 1
      do while (j >= 1)
        if (v(j) <= t) exit
        v(j + 1) = v(j)
        j = j - 1
      end do
      v(j + 1) = t
    end do
  end subroutine sort
end program satt_host_example04
