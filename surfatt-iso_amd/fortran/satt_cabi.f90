!=====================================================================
! satt_cabi: ISO_C_BINDING interface to libsatt.so (include/satt.h).
!
! The derived types mirror satt_batch / satt_result field for field; the
! functions are the C entry points with Fortran-side argument intents.
!=====================================================================
module satt_cabi
  use, intrinsic :: iso_c_binding
  implicit none

  integer(c_int), parameter :: SATT_OK = 0, SATT_W_MAXITER = 1
  integer(c_int), parameter :: SATT_E_ARG = -1, SATT_E_SRC_ON_BOUNDARY = -2
  integer(c_int), parameter :: SATT_E_REC_OUT_OF_GRID = -3, SATT_E_NOMEM = -4
  integer(c_int), parameter :: SATT_E_HIP = -5, SATT_E_NODEV = -6, SATT_E_TEAM_WAIT = -7
  integer(c_int), parameter :: SATT_E_DISPERSION = -8
  integer(c_int), parameter :: SATT_PREC_FP64 = 0, SATT_PREC_MIXED = 1
  integer(c_int), parameter :: SATT_LAUNCH_BATCHED = 0, SATT_LAUNCH_PER_SOLVE = 1

  type, bind(C) :: satt_batch
    integer(c_int) :: dim, nx, ny, nperiod
    type(c_ptr) :: xgrids, ygrids
    type(c_ptr) :: svel, m11, m12, m22
    integer(c_int) :: nitems
    type(c_ptr) :: item_pidx, item_srcx, item_srcy
    type(c_ptr) :: rec_off, recx, recy, tt, weight, idx_path
    integer(c_int) :: npath
  end type satt_batch

  type, bind(C) :: satt_result
    type(c_ptr) :: tt_fwd = c_null_ptr, chi = c_null_ptr, adj_s = c_null_ptr
    type(c_ptr) :: adj_density = c_null_ptr, item_chi = c_null_ptr, iters = c_null_ptr
    type(c_ptr) :: item_status = c_null_ptr, weight_out = c_null_ptr
    type(c_ptr) :: iters_fp32 = c_null_ptr
  end type satt_result

  interface
    integer(c_int) function satt_create(ctx, device) bind(C, name="satt_create")
      import :: c_int, c_ptr
      type(c_ptr), intent(out) :: ctx
      integer(c_int), value :: device
    end function satt_create

    subroutine satt_destroy(ctx) bind(C, name="satt_destroy")
      import :: c_ptr
      type(c_ptr), value :: ctx
    end subroutine satt_destroy

    type(c_ptr) function satt_last_error(ctx) bind(C, name="satt_last_error")
      import :: c_ptr
      type(c_ptr), value :: ctx
    end function satt_last_error

    integer(c_int) function satt_fwd_lonlat(ctx, xx_deg, yy_deg, nx, ny, spha, sphb, sphc, &
                                            fun, x0_deg, y0_deg, T, iters) &
                                            bind(C, name="satt_fwd_lonlat")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: nx, ny
      real(c_double), intent(in) :: xx_deg(*), yy_deg(*), spha(*), sphb(*), sphc(*), fun(*)
      real(c_double), value :: x0_deg, y0_deg
      real(c_double), intent(out) :: T(*)
      integer(c_int), intent(out) :: iters
    end function satt_fwd_lonlat

    integer(c_int) function satt_adj_lonlat(ctx, xx_deg, yy_deg, nx, ny, spha, sphb, sphc, &
                                            T, xrec_deg, yrec_deg, src_adj, nr, Ta, iters) &
                                            bind(C, name="satt_adj_lonlat")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: nx, ny, nr
      real(c_double), intent(in) :: xx_deg(*), yy_deg(*), spha(*), sphb(*), sphc(*), T(*)
      real(c_double), intent(in) :: xrec_deg(*), yrec_deg(*), src_adj(*)
      real(c_double), intent(out) :: Ta(*)
      integer(c_int), intent(out) :: iters
    end function satt_adj_lonlat

    integer(c_int) function satt_mask(ctx, xx, yy, nx, ny, table, x0, y0) &
                                      bind(C, name="satt_mask")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: nx, ny
      real(c_double), intent(in) :: xx(*), yy(*)
      real(c_double), intent(inout) :: table(*)
      real(c_double), value :: x0, y0
    end function satt_mask

    integer(c_int) function satt_interp2(ctx, x, y, nx, ny, v, xq, yq, nq, vq) &
                                         bind(C, name="satt_interp2")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: nx, ny, nq
      real(c_double), intent(in) :: x(*), y(*), v(*), xq(*), yq(*)
      real(c_double), intent(out) :: vq(*)
    end function satt_interp2

    integer(c_int) function satt_smooth_geo2(ctx, lons, lats, nx, ny, nfield, ndata, data, &
                                             sigma, out) bind(C, name="satt_smooth_geo2")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: nx, ny, nfield, ndata
      real(c_double), intent(in) :: lons(*), lats(*), data(*), sigma(*)
      real(c_double), intent(out) :: out(*)
    end function satt_smooth_geo2

    integer(c_int) function satt_gradient_geo2(ctx, lons, lats, nx, ny, nfield, f, tx, ty) &
                                               bind(C, name="satt_gradient_geo2")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: nx, ny, nfield
      real(c_double), intent(in) :: lons(*), lats(*), f(*)
      real(c_double), intent(out) :: tx(*), ty(*)
    end function satt_gradient_geo2

    integer(c_int) function satt_topo_metric(ctx, lons, lats, nx, ny, z, nperiod, sigma, &
                                             m11, m12, m22, angle) &
                                             bind(C, name="satt_topo_metric")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: nx, ny, nperiod
      real(c_double), intent(in) :: lons(*), lats(*), z(*), sigma(*)
      real(c_double), intent(out) :: m11(*), m12(*), m22(*), angle(*)
    end function satt_topo_metric

    integer(c_int) function satt_fwdsurf3d(ctx, vs3d, nx, ny, nz, depz, nperiod, periods, &
                                           iwave, igr, svel, col_status) &
                                           bind(C, name="satt_fwdsurf3d")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: nx, ny, nz, nperiod, iwave, igr
      real(c_double), intent(in) :: vs3d(*), depz(*), periods(*)
      real(c_double), intent(out) :: svel(*)
      integer(c_int), intent(out) :: col_status(*)
    end function satt_fwdsurf3d

    integer(c_int) function satt_combine_kernels(ctx, nperiod, nx, ny, nz, adj_s, adj_density, &
                                                 sen_vs, sen_vp, sen_rho, vs3d, vp3d, &
                                                 use_alpha_beta_rho, rho_scaling, ker, &
                                                 ker_density) bind(C, name="satt_combine_kernels")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: nperiod, nx, ny, nz, use_alpha_beta_rho, rho_scaling
      real(c_double), intent(in) :: adj_s(*), adj_density(*), sen_vs(*), sen_vp(*), sen_rho(*)
      real(c_double), intent(in) :: vs3d(*), vp3d(*)
      real(c_double), intent(inout) :: ker(*), ker_density(*)
    end function satt_combine_kernels

    integer(c_int) function satt_inv_grid_iso(ctx, xinv, yinv, zinv, nxinv, nyinv, nzinv, &
                                              nset, adj, nx, ny, nz, gk, vel_x, vel_y, vel_z) &
                                              bind(C, name="satt_inv_grid_iso")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: nxinv, nyinv, nzinv, nset, nx, ny, nz
      real(c_double), intent(in) :: xinv(*), yinv(*), zinv(*), adj(*), vel_x(*), vel_y(*)
      real(c_double), intent(in) :: vel_z(*)
      real(c_double), intent(out) :: gk(*)
    end function satt_inv_grid_iso

    integer(c_int) function satt_inv2fwd_iso(ctx, xinv, yinv, zinv, nxinv, nyinv, nzinv, nset, &
                                             nx, ny, nz, gk, vel_x, vel_y, vel_z, update) &
                                             bind(C, name="satt_inv2fwd_iso")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: nxinv, nyinv, nzinv, nset, nx, ny, nz
      real(c_double), intent(in) :: xinv(*), yinv(*), zinv(*), gk(*), vel_x(*), vel_y(*)
      real(c_double), intent(in) :: vel_z(*)
      real(c_double), intent(out) :: update(*)
    end function satt_inv2fwd_iso

    integer(c_int) function satt_kernel_postproc(ctx, nker, nx, ny, nz, vel_x, vel_y, vel_z, &
                                                 nset, nxinv, nyinv, nzinv, xinv, yinv, zinv, &
                                                 ker, ker_density, kdensity_coe, chi0, weight, &
                                                 gradient_s) bind(C, name="satt_kernel_postproc")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: nker, nx, ny, nz, nset, nxinv, nyinv, nzinv
      real(c_double), intent(in) :: vel_x(*), vel_y(*), vel_z(*), xinv(*), yinv(*), zinv(*)
      real(c_double), intent(inout) :: ker(*), gradient_s(*)
      real(c_double), intent(in) :: ker_density(*)
      real(c_double), value :: kdensity_coe, chi0, weight
    end function satt_kernel_postproc

    real(c_double) function satt_topo_sigma(vel, period, wavelen_factor) &
                                            bind(C, name="satt_topo_sigma")
      import :: c_double
      real(c_double), value :: vel, period, wavelen_factor
    end function satt_topo_sigma

    integer(c_int) function satt_steepest_descent(ctx, lons, lats, nx, ny, nperiod, adj_s, &
                                                  sigma_2d, updatemax, svel, ker_s) &
                                                  bind(C, name="satt_steepest_descent")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: nx, ny, nperiod
      real(c_double), intent(in) :: lons(*), lats(*), adj_s(*)
      real(c_double), value :: sigma_2d, updatemax
      real(c_double), intent(inout) :: svel(*)
      real(c_double), intent(out) :: ker_s(*)
    end function satt_steepest_descent

    integer(c_int) function satt_forward_simulate(ctx, batch, istotable, isadj, precision, &
                                                  result) bind(C, name="satt_forward_simulate")
      import :: c_int, c_ptr, satt_batch, satt_result
      type(c_ptr), value :: ctx
      type(satt_batch), intent(in) :: batch
      integer(c_int), value :: istotable, isadj, precision
      type(satt_result), intent(inout) :: result
    end function satt_forward_simulate

    integer(c_int) function satt_plan_create(ctx, batch, precision, plan) &
                                             bind(C, name="satt_plan_create")
      import :: c_int, c_ptr, satt_batch
      type(c_ptr), value :: ctx
      type(satt_batch), intent(in) :: batch
      integer(c_int), value :: precision
      type(c_ptr), intent(out) :: plan
    end function satt_plan_create

    integer(c_int) function satt_plan_set_model(plan, svel, m11, m12, m22) &
                                                bind(C, name="satt_plan_set_model")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: plan
      real(c_double), intent(in) :: svel(*), m11(*), m12(*), m22(*)
    end function satt_plan_set_model

    integer(c_int) function satt_plan_run(plan, istotable, isadj, launch_form) &
                                          bind(C, name="satt_plan_run")
      import :: c_int, c_ptr
      type(c_ptr), value :: plan
      integer(c_int), value :: istotable, isadj, launch_form
    end function satt_plan_run

    integer(c_int) function satt_plan_results(plan, result) bind(C, name="satt_plan_results")
      import :: c_int, c_ptr, satt_result
      type(c_ptr), value :: plan
      type(satt_result), intent(inout) :: result
    end function satt_plan_results

    integer(c_int) function satt_plan_linesearch(plan, svel, ker_s, chi0, maxshrink, &
                                                 max_sub_niter, updatemax, svel_out, chi_out, &
                                                 nsub) bind(C, name="satt_plan_linesearch")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: plan
      real(c_double), intent(in) :: svel(*), ker_s(*)
      real(c_double), value :: chi0, maxshrink
      integer(c_int), value :: max_sub_niter
      real(c_double), intent(inout) :: updatemax
      real(c_double), intent(out) :: svel_out(*), chi_out
      integer(c_int), intent(out) :: nsub
    end function satt_plan_linesearch

    subroutine satt_plan_destroy(plan) bind(C, name="satt_plan_destroy")
      import :: c_ptr
      type(c_ptr), value :: plan
    end subroutine satt_plan_destroy
  end interface

contains

  !> The context's last error message as a Fortran string.
  function satt_error_message(ctx) result(msg)
    type(c_ptr), intent(in) :: ctx
    character(len=:), allocatable :: msg
    character(kind=c_char), pointer :: p(:)
    type(c_ptr) :: cp
    integer :: n
    cp = satt_last_error(ctx)
    call c_f_pointer(cp, p, [512])
    n = 0
    do while (n < 512)
      if (p(n + 1) == c_null_char) exit
      n = n + 1
    end do
    allocate(character(len=n) :: msg)
    msg = transfer(p(1:n), msg)
  end function satt_error_message

end module satt_cabi
