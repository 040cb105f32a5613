!=====================================================================
! C-callable wrapper of satt_forward_simulate_host for harnesses that cannot
! pass Fortran CHARACTER arrays: names are given as integer ids and turned
! into the reference's fixed-length names ('S' followed by six digits).
!=====================================================================
subroutine satt_fhost_forward_simulate(dim, nx, ny, nperiod, xgrids, ygrids, periods, svel, &
                                       m11, m12, m22, npath, evt_id, sta_id, evla, evlo, &
                                       stla, stlo, periods_all, tt, weight, istotable, isadj, &
                                       tt_fwd, chi, adj_s, adj_density, status, precision) &
                                       bind(C, name="satt_fhost_forward_simulate")
  use, intrinsic :: iso_c_binding
  use satt_host, only: satt_forward_simulate_host
  implicit none
  integer(c_int), value :: dim, nx, ny, nperiod, npath, istotable, isadj, precision
  real(c_double), intent(in) :: xgrids(nx), ygrids(ny), periods(nperiod)
  real(c_double), intent(in) :: svel(nperiod,nx,ny), m11(nperiod,nx,ny), m12(nperiod,nx,ny)
  real(c_double), intent(in) :: m22(nperiod,nx,ny)
  integer(c_int), intent(in) :: evt_id(npath), sta_id(npath)
  real(c_double), intent(in) :: evla(npath), evlo(npath), stla(npath), stlo(npath)
  real(c_double), intent(in) :: periods_all(npath), tt(npath), weight(npath)
  real(c_double), intent(out) :: tt_fwd(npath), chi
  real(c_double), intent(out) :: adj_s(nperiod,nx,ny), adj_density(nperiod,nx,ny)
  integer(c_int), intent(out) :: status
  character(len=30), allocatable :: evtname(:), staname(:)
  integer :: i, st
  allocate(evtname(npath), staname(npath))
  do i = 1, npath
    write(evtname(i), '(a,i6.6)') 'S', evt_id(i)
    write(staname(i), '(a,i6.6)') 'S', sta_id(i)
  end do
  call satt_forward_simulate_host(dim, xgrids, ygrids, periods, svel, m11, m12, m22, &
                                  evtname, staname, evla, evlo, stla, stlo, periods_all, tt, &
                                  weight, istotable /= 0, isadj /= 0, tt_fwd, chi, adj_s, &
                                  adj_density, st, precision)
  status = st
end subroutine satt_fhost_forward_simulate

subroutine satt_fhost_linesearch(dim, nx, ny, nperiod, xgrids, ygrids, periods, svel, m11, &
                                 m12, m22, npath, evt_id, sta_id, evla, evlo, stla, stlo, &
                                 periods_all, tt, weight, ker_s, chi0, maxshrink, &
                                 max_sub_niter, updatemax, chi, nsub, status) &
                                 bind(C, name="satt_fhost_linesearch")
  use, intrinsic :: iso_c_binding
  use satt_host, only: satt_linesearch_host
  implicit none
  integer(c_int), value :: dim, nx, ny, nperiod, npath, max_sub_niter
  real(c_double), intent(in) :: xgrids(nx), ygrids(ny), periods(nperiod)
  real(c_double), intent(inout) :: svel(nperiod,nx,ny)
  real(c_double), intent(in) :: m11(nperiod,nx,ny), m12(nperiod,nx,ny), m22(nperiod,nx,ny)
  integer(c_int), intent(in) :: evt_id(npath), sta_id(npath)
  real(c_double), intent(in) :: evla(npath), evlo(npath), stla(npath), stlo(npath)
  real(c_double), intent(in) :: periods_all(npath), tt(npath), weight(npath)
  real(c_double), intent(in) :: ker_s(nperiod,nx,ny)
  real(c_double), value :: chi0, maxshrink
  real(c_double), intent(inout) :: updatemax
  real(c_double), intent(out) :: chi
  integer(c_int), intent(out) :: nsub, status
  character(len=30), allocatable :: evtname(:), staname(:)
  integer :: i, st, ns
  allocate(evtname(npath), staname(npath))
  do i = 1, npath
    write(evtname(i), '(a,i6.6)') 'S', evt_id(i)
    write(staname(i), '(a,i6.6)') 'S', sta_id(i)
  end do
  call satt_linesearch_host(dim, xgrids, ygrids, periods, svel, m11, m12, m22, evtname, &
                            staname, evla, evlo, stla, stlo, periods_all, tt, weight, ker_s, &
                            chi0, maxshrink, max_sub_niter, updatemax, chi, ns, st)
  nsub = ns
  status = st
end subroutine satt_fhost_linesearch

!=====================================================================
! C-callable wrappers of the model-side host routines (for harnesses).
!=====================================================================
subroutine satt_fhost_topo_metric(nx, ny, nperiod, xgrids, ygrids, z, vel, periods, &
                                  wavelen_factor, m11, m12, m22, topo_angle) &
                                  bind(C, name="satt_fhost_topo_metric")
  use, intrinsic :: iso_c_binding
  use satt_host, only: satt_get_topo_metric
  implicit none
  integer(c_int), value :: nx, ny, nperiod
  real(c_double), intent(in) :: xgrids(nx), ygrids(ny), z(nx,ny), vel(nperiod), periods(nperiod)
  real(c_double), value :: wavelen_factor
  real(c_double), intent(out) :: m11(nperiod,nx,ny), m12(nperiod,nx,ny), m22(nperiod,nx,ny)
  real(c_double), intent(out) :: topo_angle(nperiod,nx,ny)
  call satt_get_topo_metric(xgrids, ygrids, z, vel, periods, wavelen_factor, m11, m12, m22, &
                            topo_angle)
end subroutine satt_fhost_topo_metric

subroutine satt_fhost_steepest_descent(nx, ny, nperiod, xgrids, ygrids, adj_s, sigma_2d, &
                                       misfit_now, misfit_prev, maxshrink, updatemax, svel, &
                                       ker_s) bind(C, name="satt_fhost_steepest_descent")
  use, intrinsic :: iso_c_binding
  use satt_host, only: satt_steepest_descent_host
  implicit none
  integer(c_int), value :: nx, ny, nperiod
  real(c_double), intent(in) :: xgrids(nx), ygrids(ny), adj_s(nperiod,nx,ny)
  real(c_double), value :: sigma_2d, misfit_now, misfit_prev, maxshrink
  real(c_double), intent(inout) :: updatemax, svel(nperiod,nx,ny)
  real(c_double), intent(out) :: ker_s(nperiod,nx,ny)
  call satt_steepest_descent_host(xgrids, ygrids, adj_s, sigma_2d, misfit_now, misfit_prev, &
                                  maxshrink, updatemax, svel, ker_s)
end subroutine satt_fhost_steepest_descent

subroutine satt_fhost_smooth(nx, ny, data, lons, lats, sigma, smdata) &
                             bind(C, name="satt_fhost_smooth")
  use, intrinsic :: iso_c_binding
  use satt_host, only: satt_gaussian_smooth_geo_2
  implicit none
  integer(c_int), value :: nx, ny
  real(c_double), intent(in) :: data(nx,ny), lons(nx), lats(ny)
  real(c_double), value :: sigma
  real(c_double), intent(out) :: smdata(nx,ny)
  smdata = satt_gaussian_smooth_geo_2(data, lons, lats, sigma)
end subroutine satt_fhost_smooth
