!=====================================================================
! satt_host: the Fortran host side of the MI355X path.
!
!  - satt_host_ctx()   one device context per process (device from the
!                      SATT_DEVICE environment variable, default 0)
!  - satt_host_check() maps status codes to the reference's behaviour:
!                      "source on the boundary" stops the run as
!                      src/eikon_solver_2d.f90:863-867 does; reaching MaxIter
!                      is silent as in :1117-1119,1436-1438
!  - satt_forward_simulate_host()  the per-(source, period) loop of
!                      att_tomo_2d%forward_simulate (src/tomo2d.f90:158-201)
!                      and att_acqui%forward_simulate (src/acqui.f90:91-161)
!                      over the reference's own arrays: the grid axes, the
!                      period-major model svel/m11/m12/m22(nperiod,nx,ny) and
!                      the src_rec path columns.  It builds the station list,
!                      the isrcs work list and the receiver gathers in the
!                      reference's orders (src/src_rec.f90:162-217,252-283,
!                      338-369) and runs every item in one device call.
!  - satt_gaussian_smooth_geo_2(), satt_gradient_2_geo()  the reference's
!                      gaussian_smooth_geo_2 and gradient_2_geo
!                      (src/shared/utils.f90:1194,1118) with their arguments
!  - satt_get_topo_metric()  the per-period metric loop of att_grid%get_topo
!                      (src/grid.f90:114-123,137-141) over every period
!  - satt_steepest_descent_host()  att_tomo_2d%optimize + steepest_descent
!                      (src/tomo2d.f90:203-251), the misfit-increase shrink
!                      of the step length included
!  - satt_linesearch_host()  att_tomo_2d%linesearch (src/tomo2d.f90:253-288)
!                      with prepare_fwd_linesearch (src/acqui_2d.f90:173-183):
!                      the trial models and the forward-only runs on the
!                      device, over the same arrays as forward_simulate
!  - satt_combine_kernels_host()  att_acqui%combine_kernels
!                      (src/acqui.f90:197-245) over its arrays
!  - satt_post_processing_host()  att_acqui%post_processing_for_kernel
!                      (src/acqui.f90:247-304): preconditioning and both
!                      inversion-grid projections on the device
!  - satt_fwdsurf3d_host()  fwdsurf3d_decomp (src/surfker/surfker.f90:200-222)
!                      over the whole grid: the velocity maps of
!                      att_grid%fwdsurf (src/grid.f90:150-163)
!=====================================================================
module satt_host
  use, intrinsic :: iso_c_binding
  use satt_cabi
  implicit none
  private
  public :: satt_host_ctx, satt_host_check, satt_host_release, satt_forward_simulate_host
  public :: satt_gaussian_smooth_geo_2, satt_gradient_2_geo, satt_get_topo_metric
  public :: satt_steepest_descent_host, satt_fwdsurf3d_host, satt_linesearch_host
  public :: satt_combine_kernels_host, satt_post_processing_host

  type(c_ptr), save :: ctx = c_null_ptr

  ! The C-ABI batch of one forward_simulate call and the arrays it points to
  ! (the work list and the receiver gathers built from the path columns).
  type :: host_batch
    integer(c_int), allocatable :: item_pidx(:), rec_off(:), idx_path(:)
    real(c_double), allocatable :: item_srcx(:), item_srcy(:), recx(:), recy(:)
    real(c_double), allocatable :: rtt(:), rweight(:)
    type(satt_batch) :: b
  end type host_batch

contains

  function satt_host_ctx() result(c)
    type(c_ptr) :: c
    integer(c_int) :: st, dev
    character(len=16) :: env
    integer :: l, ios
    if (.not. c_associated(ctx)) then
      dev = 0
      call get_environment_variable('SATT_DEVICE', env, length=l)
      if (l > 0) then
        read(env(1:l), *, iostat=ios) dev
        if (ios /= 0) dev = 0
      end if
      st = satt_create(ctx, dev)
      if (st /= SATT_OK) then
        write(*,*) 'satt: no usable GPU (status ', st, ')'
        stop 1
      end if
    end if
    c = ctx
  end function satt_host_ctx

  subroutine satt_host_release()
    if (c_associated(ctx)) call satt_destroy(ctx)
    ctx = c_null_ptr
  end subroutine satt_host_release

  subroutine satt_host_check(st, where)
    integer(c_int), intent(in) :: st
    character(len=*), intent(in) :: where
    if (st >= 0) return
    if (st == SATT_E_SRC_ON_BOUNDARY) then
      write(*,*) 'source on the boundary, mesh error'
    else
      write(*,*) 'satt: ', where, ': ', satt_error_message(ctx)
    end if
    stop 1
  end subroutine satt_host_check

  !> forward_simulate over the reference's arrays.  Names are CHARACTER
  !> arrays like SrcRec%evtname/staname; periods(nperiod) is the sorted unique
  !> list (SrcRec%periods).  dim = 2 (tomo2d) or 3 (acqui).  Outputs are fully
  !> overwritten: tt_fwd(npath) (zero where no item maps, as local_tt), chi,
  !> adj_s(nperiod,nx,ny) and, for dim 3, adj_density(nperiod,nx,ny).
  subroutine satt_forward_simulate_host(dim, xgrids, ygrids, periods, svel, m11, m12, m22, &
                                        evtname, staname, evla, evlo, stla, stlo, &
                                        periods_all, tt, weight, istotable, isadj, &
                                        tt_fwd, chi, adj_s, adj_density, status, precision)
    integer, intent(in) :: dim
    real(c_double), intent(in), target, contiguous :: xgrids(:), ygrids(:), periods(:)
    real(c_double), intent(in), target, contiguous :: svel(:,:,:), m11(:,:,:), m12(:,:,:)
    real(c_double), intent(in), target, contiguous :: m22(:,:,:)
    character(len=*), intent(in) :: evtname(:), staname(:)
    real(c_double), intent(in) :: evla(:), evlo(:), stla(:), stlo(:), periods_all(:)
    real(c_double), intent(in) :: tt(:), weight(:)
    logical, intent(in) :: istotable, isadj
    real(c_double), intent(out), target, contiguous :: tt_fwd(:)
    real(c_double), intent(out) :: chi
    real(c_double), intent(out), target, contiguous :: adj_s(:,:,:), adj_density(:,:,:)
    integer, intent(out), optional :: status
    integer, intent(in), optional :: precision   ! SATT_PREC_FP64 (default) or SATT_PREC_MIXED

    type(host_batch), target :: hb
    real(c_double), target :: chi_c(1)
    integer(c_int) :: prec
    type(satt_result) :: r
    integer(c_int) :: st

    call build_batch(dim, xgrids, ygrids, periods, svel, m11, m12, m22, evtname, staname, &
                     evla, evlo, stla, stlo, periods_all, tt, weight, hb)
    r%tt_fwd = c_loc(tt_fwd)
    r%chi = c_loc(chi_c)
    r%adj_s = c_loc(adj_s)
    if (dim == 3) r%adj_density = c_loc(adj_density)
    prec = SATT_PREC_FP64
    if (present(precision)) prec = int(precision, c_int)
    st = satt_forward_simulate(satt_host_ctx(), hb%b, merge(1_c_int, 0_c_int, istotable), &
                               merge(1_c_int, 0_c_int, isadj), prec, r)
    if (present(status)) then
      status = st
    else
      call satt_host_check(st, 'forward_simulate')
    end if
    chi = chi_c(1)
  end subroutine satt_forward_simulate_host

  !> att_tomo_2d%linesearch (src/tomo2d.f90:253-288) over forward_simulate's
  !> arrays: chi0 is acqui%misfits(acqui%iter), ker_s the smoothed kernel of
  !> optimize, updatemax the step length (shrunk in place, :275-278).  On
  !> return svel is the last model tried (acqui%svel = acqui%ag%svel,
  !> :285-287), chi its misfit and nsub the sub-iterations run.  The trial
  !> models and the forward-only runs stay on the device.
  subroutine satt_linesearch_host(dim, xgrids, ygrids, periods, svel, m11, m12, m22, &
                                  evtname, staname, evla, evlo, stla, stlo, periods_all, &
                                  tt, weight, ker_s, chi0, maxshrink, max_sub_niter, &
                                  updatemax, chi, nsub, status, precision)
    integer, intent(in) :: dim
    real(c_double), intent(in), target, contiguous :: xgrids(:), ygrids(:), periods(:)
    real(c_double), intent(inout), target, contiguous :: svel(:,:,:)
    real(c_double), intent(in), target, contiguous :: m11(:,:,:), m12(:,:,:), m22(:,:,:)
    character(len=*), intent(in) :: evtname(:), staname(:)
    real(c_double), intent(in) :: evla(:), evlo(:), stla(:), stlo(:), periods_all(:)
    real(c_double), intent(in) :: tt(:), weight(:)
    real(c_double), intent(in), contiguous :: ker_s(:,:,:)
    real(c_double), intent(in) :: chi0, maxshrink
    integer, intent(in) :: max_sub_niter
    real(c_double), intent(inout) :: updatemax
    real(c_double), intent(out) :: chi
    integer, intent(out) :: nsub
    integer, intent(out), optional :: status
    integer, intent(in), optional :: precision   ! SATT_PREC_FP64 (default) or SATT_PREC_MIXED

    type(host_batch), target :: hb
    real(c_double), allocatable :: svel_out(:,:,:)
    type(c_ptr) :: plan
    integer(c_int) :: st, ns, prec

    call build_batch(dim, xgrids, ygrids, periods, svel, m11, m12, m22, evtname, staname, &
                     evla, evlo, stla, stlo, periods_all, tt, weight, hb)
    allocate(svel_out(size(svel,1), size(svel,2), size(svel,3)))
    chi = chi0
    nsub = 0
    plan = c_null_ptr
    prec = SATT_PREC_FP64
    if (present(precision)) prec = int(precision, c_int)
    st = satt_plan_create(satt_host_ctx(), hb%b, prec, plan)
    if (st >= 0) then
      st = satt_plan_linesearch(plan, svel, ker_s, chi0, maxshrink, &
                                int(max_sub_niter, c_int), updatemax, svel_out, chi, ns)
      if (st >= 0) then
        svel = svel_out
        nsub = ns
      end if
    end if
    if (c_associated(plan)) call satt_plan_destroy(plan)
    if (present(status)) then
      status = st
    else
      call satt_host_check(st, 'linesearch')
    end if
  end subroutine satt_linesearch_host

  ! The isrcs work list and the receiver gathers of forward_simulate's path
  ! columns, and the C-ABI batch pointing at them and at the model.
  subroutine build_batch(dim, xgrids, ygrids, periods, svel, m11, m12, m22, evtname, staname, &
                         evla, evlo, stla, stlo, periods_all, tt, weight, hb)
    integer, intent(in) :: dim
    real(c_double), intent(in), target, contiguous :: xgrids(:), ygrids(:), periods(:)
    real(c_double), intent(in), target, contiguous :: svel(:,:,:), m11(:,:,:), m12(:,:,:)
    real(c_double), intent(in), target, contiguous :: m22(:,:,:)
    character(len=*), intent(in) :: evtname(:), staname(:)
    real(c_double), intent(in) :: evla(:), evlo(:), stla(:), stlo(:), periods_all(:)
    real(c_double), intent(in) :: tt(:), weight(:)
    type(host_batch), intent(inout), target :: hb

    integer :: npath, nsta, i, j, k, np, nitems, n, p
    character(len=len(staname)), allocatable :: names(:)
    real(c_double), allocatable :: slat(:), slon(:)
    integer, allocatable :: evt_sta(:), rec_sta(:), item_pidx_f(:), item_sta(:), plist(:)
    logical :: seen

    npath = size(evtname)
    np = size(periods)
    ! station list: receivers in order of first appearance, then events not
    ! yet listed (get_sta, src/src_rec.f90:180-217)
    allocate(names(2 * npath), slat(2 * npath), slon(2 * npath))
    nsta = 0
    do i = 1, npath
      if (find_name(names, nsta, staname(i)) == 0) then
        nsta = nsta + 1; names(nsta) = staname(i); slat(nsta) = stla(i); slon(nsta) = stlo(i)
      end if
    end do
    do i = 1, npath
      if (find_name(names, nsta, evtname(i)) == 0) then
        nsta = nsta + 1; names(nsta) = evtname(i); slat(nsta) = evla(i); slon(nsta) = evlo(i)
      end if
    end do
    allocate(evt_sta(npath), rec_sta(npath))
    do i = 1, npath
      evt_sta(i) = find_name(names, nsta, evtname(i))
      rec_sta(i) = find_name(names, nsta, staname(i))
    end do
    ! work list isrcs (scatter_src_gather / get_periods_by_src,
    ! src/src_rec.f90:338-369, :162-179): stations in order, each event's
    ! periods in order of first appearance
    allocate(item_pidx_f(npath), item_sta(npath), plist(np))
    nitems = 0
    do j = 1, nsta
      n = 0
      do i = 1, npath
        if (evt_sta(i) /= j) cycle
        seen = .false.
        do k = 1, n
          if (periods(plist(k)) == periods_all(i)) seen = .true.
        end do
        if (seen) cycle
        n = n + 1
        plist(n) = period_index(periods, periods_all(i))
      end do
      do k = 1, n
        nitems = nitems + 1
        item_pidx_f(nitems) = plist(k)
        item_sta(nitems) = j
      end do
    end do
    ! receiver gathers in path order (get_evt_gather, src/src_rec.f90:252-283)
    allocate(hb%item_pidx(nitems), hb%item_srcx(nitems), hb%item_srcy(nitems))
    allocate(hb%rec_off(nitems + 1), hb%recx(npath), hb%recy(npath), hb%rtt(npath))
    allocate(hb%rweight(npath), hb%idx_path(npath))
    hb%rec_off(1) = 0
    n = 0
    do k = 1, nitems
      p = item_pidx_f(k)
      hb%item_pidx(k) = p - 1
      hb%item_srcx(k) = slon(item_sta(k))
      hb%item_srcy(k) = slat(item_sta(k))
      do i = 1, npath
        if (evt_sta(i) == item_sta(k) .and. periods_all(i) == periods(p)) then
          n = n + 1
          hb%recx(n) = slon(rec_sta(i)); hb%recy(n) = slat(rec_sta(i))
          hb%rtt(n) = tt(i); hb%rweight(n) = weight(i); hb%idx_path(n) = i - 1
        end if
      end do
      hb%rec_off(k + 1) = n
    end do

    hb%b%dim = dim
    hb%b%nx = size(xgrids); hb%b%ny = size(ygrids); hb%b%nperiod = np
    hb%b%xgrids = c_loc(xgrids); hb%b%ygrids = c_loc(ygrids)
    hb%b%svel = c_loc(svel); hb%b%m11 = c_loc(m11); hb%b%m12 = c_loc(m12)
    hb%b%m22 = c_loc(m22)
    hb%b%nitems = nitems
    hb%b%item_pidx = c_loc(hb%item_pidx)
    hb%b%item_srcx = c_loc(hb%item_srcx); hb%b%item_srcy = c_loc(hb%item_srcy)
    hb%b%rec_off = c_loc(hb%rec_off)
    hb%b%recx = c_loc(hb%recx); hb%b%recy = c_loc(hb%recy)
    hb%b%tt = c_loc(hb%rtt); hb%b%weight = c_loc(hb%rweight)
    hb%b%idx_path = c_loc(hb%idx_path)
    hb%b%npath = npath
  end subroutine build_batch

  !> gaussian_smooth_geo_2(data, lons, lats, sigma) (src/shared/utils.f90:1194)
  function satt_gaussian_smooth_geo_2(data, lons, lats, sigma) result(smdata)
    real(c_double), intent(in) :: data(:,:), lons(:), lats(:), sigma
    real(c_double), allocatable :: smdata(:,:)
    real(c_double) :: sig(1), src(size(data,1), size(data,2))
    allocate(smdata(size(data,1), size(data,2)))
    sig(1) = sigma
    src = data
    call satt_host_check(satt_smooth_geo2(satt_host_ctx(), lons, lats, size(lons), size(lats), &
                                          1, 1, src, sig, smdata), 'gaussian_smooth_geo_2')
  end function satt_gaussian_smooth_geo_2

  !> gradient_2_geo(f, lon, lat, tx, ty) (src/shared/utils.f90:1118)
  subroutine satt_gradient_2_geo(f, lon, lat, tx, ty)
    real(c_double), intent(in) :: f(:,:), lon(:), lat(:)
    real(c_double), allocatable, intent(out) :: tx(:,:), ty(:,:)
    real(c_double) :: src(size(f,1), size(f,2))
    allocate(tx(size(f,1), size(f,2)), ty(size(f,1), size(f,2)))
    src = f
    call satt_host_check(satt_gradient_geo2(satt_host_ctx(), lon, lat, size(lon), size(lat), &
                                            1, src, tx, ty), 'gradient_2_geo')
  end subroutine satt_gradient_2_geo

  !> The metric of att_grid%get_topo for every period (src/grid.f90:114-123,
  !> 137-141): z is the topography on the model grid (km, as after
  !> att_topo%grid_topo), vel(nperiod) the 1-D model's velocities from
  !> fwdsurf1d; m11/m12/m22/topo_angle are (nperiod,nx,ny).
  subroutine satt_get_topo_metric(xgrids, ygrids, z, vel, periods, wavelen_factor, &
                                  m11, m12, m22, topo_angle)
    real(c_double), intent(in) :: xgrids(:), ygrids(:), z(:,:), vel(:), periods(:)
    real(c_double), intent(in) :: wavelen_factor
    real(c_double), intent(out) :: m11(:,:,:), m12(:,:,:), m22(:,:,:), topo_angle(:,:,:)
    real(c_double) :: sigma(size(periods))
    integer :: ip
    do ip = 1, size(periods)
      sigma(ip) = satt_topo_sigma(vel(ip), periods(ip), wavelen_factor)
    end do
    call satt_host_check(satt_topo_metric(satt_host_ctx(), xgrids, ygrids, size(xgrids), &
                                          size(ygrids), z, size(periods), sigma, m11, m12, &
                                          m22, topo_angle), 'get_topo')
  end subroutine satt_get_topo_metric

  !> optimize + steepest_descent (src/tomo2d.f90:203-251): the step length
  !> shrinks by maxshrink when the misfit rose (:238-246), then
  !> svel = svel*(1 + updatemax*ker_s/maxval(abs(ker_s))).
  subroutine satt_steepest_descent_host(xgrids, ygrids, adj_s, sigma_2d, misfit_now, &
                                        misfit_prev, maxshrink, updatemax, svel, ker_s)
    real(c_double), intent(in) :: xgrids(:), ygrids(:), adj_s(:,:,:), sigma_2d
    real(c_double), intent(in) :: misfit_now, misfit_prev, maxshrink
    real(c_double), intent(inout) :: updatemax, svel(:,:,:)
    real(c_double), intent(out) :: ker_s(:,:,:)
    if (misfit_now > misfit_prev) updatemax = updatemax*maxshrink
    call satt_host_check(satt_steepest_descent(satt_host_ctx(), xgrids, ygrids, size(xgrids), &
                                               size(ygrids), size(adj_s,1), adj_s, sigma_2d, &
                                               updatemax, svel, ker_s), 'steepest_descent')
  end subroutine satt_steepest_descent_host

  !> fwdsurf3d_decomp(vel, ..., iwave, igr, tRc, depz, svel) for every column:
  !> svel(nperiod,nx,ny) phase (igr 0) or group velocities of Love (iwave 1)
  !> or Rayleigh (iwave 2) waves.  A column without a root stops the run, as
  !> the reference's surfdisp96 does.
  subroutine satt_fwdsurf3d_host(vel, iwave, igr, tRc, depz, svel)
    real(c_double), intent(in) :: vel(:,:,:), tRc(:), depz(:)
    integer, intent(in) :: iwave, igr
    real(c_double), intent(out) :: svel(:,:,:)
    integer(c_int) :: cs(size(vel,1)*size(vel,2)), st
    st = satt_fwdsurf3d(satt_host_ctx(), vel, size(vel,1), size(vel,2), size(vel,3), depz, &
                        size(tRc), tRc, iwave, igr, svel, cs)
    if (st == SATT_E_DISPERSION) then
      write(*,*) 'WARNING:improper initial value in disper - no zero found'
      stop 'Dispersion calculation failed'
    end if
    call satt_host_check(st, 'fwdsurf3d')
  end subroutine satt_fwdsurf3d_host

  !> combine_kernels over att_acqui's arrays: ker and ker_density
  !> (nker,nx,ny,nz) accumulate -adj*sen over the periods.
  subroutine satt_combine_kernels_host(adj_s, adj_density, sen_vs, sen_vp, sen_rho, vs3d, &
                                       vp3d, use_alpha_beta_rho, rho_scaling, ker, ker_density)
    real(c_double), intent(in), contiguous :: adj_s(:,:,:), adj_density(:,:,:)
    real(c_double), intent(in), contiguous :: sen_vs(:,:,:,:), sen_vp(:,:,:,:), sen_rho(:,:,:,:)
    real(c_double), intent(in), contiguous :: vs3d(:,:,:), vp3d(:,:,:)
    logical, intent(in) :: use_alpha_beta_rho, rho_scaling
    real(c_double), intent(inout), contiguous :: ker(:,:,:,:), ker_density(:,:,:,:)
    call satt_host_check(satt_combine_kernels(satt_host_ctx(), size(sen_vs,1), size(sen_vs,2), &
                                              size(sen_vs,3), size(sen_vs,4), adj_s, &
                                              adj_density, sen_vs, sen_vp, sen_rho, vs3d, vp3d, &
                                              merge(1_c_int, 0_c_int, use_alpha_beta_rho), &
                                              merge(1_c_int, 0_c_int, rho_scaling), ker, &
                                              ker_density), 'combine_kernels')
  end subroutine satt_combine_kernels_host

  !> post_processing_for_kernel: every component of ker (nker,nx,ny,nz)
  !> preconditioned (kdensity_coe > 0, in place), projected to the inversion
  !> grids xinv(nxinv,nset), yinv, zinv and back, and added to gradient_s as
  !> update/nset/chi0*weight.
  subroutine satt_post_processing_host(xgrids, ygrids, zgrids, xinv, yinv, zinv, ker, &
                                       ker_density, kdensity_coe, chi0, weight, gradient_s)
    real(c_double), intent(in), contiguous :: xgrids(:), ygrids(:), zgrids(:)
    real(c_double), intent(in), contiguous :: xinv(:,:), yinv(:,:), zinv(:,:)
    real(c_double), intent(inout), contiguous :: ker(:,:,:,:), gradient_s(:,:,:,:)
    real(c_double), intent(in), contiguous :: ker_density(:,:,:,:)
    real(c_double), intent(in) :: kdensity_coe, chi0, weight
    call satt_host_check(satt_kernel_postproc(satt_host_ctx(), size(ker,1), size(ker,2), &
                                              size(ker,3), size(ker,4), xgrids, ygrids, zgrids, &
                                              size(xinv,2), size(xinv,1), size(yinv,1), &
                                              size(zinv,1), xinv, yinv, zinv, ker, ker_density, &
                                              kdensity_coe, chi0, weight, gradient_s), &
                         'post_processing_for_kernel')
  end subroutine satt_post_processing_host

  integer function find_name(names, n, s)
    character(len=*), intent(in) :: names(:), s
    integer, intent(in) :: n
    integer :: k
    find_name = 0
    do k = 1, n
      if (names(k) == s) then
        find_name = k
        return
      end if
    end do
  end function find_name

  integer function period_index(periods, p)
    real(c_double), intent(in) :: periods(:), p
    integer :: k
    period_index = 0
    do k = 1, size(periods)
      if (periods(k) == p) then
        period_index = k
        return
      end if
    end do
  end function period_index

end module satt_host
