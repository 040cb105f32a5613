!=====================================================================
! Link-level drop-ins for the reference's solver file: the same external
! subroutine names and argument lists as src/eikon_solver_2d.f90, so a
! reference build that links this object (and libsatt.so) instead of
! eikon_solver_2d.f90 runs its solves on the GPU with no source change at
! the call sites (src/measadj.f90:68-70,85-89,99-103), and its inversion-grid
! projections too (src/acqui.f90:274-279).
!=====================================================================

! src/eikon_solver_2d.f90:764  FSM_UW_PS_lonlat_2d
subroutine FSM_UW_PS_lonlat_2d(xx_deg, yy_deg, nx, ny, spha, sphb, sphc, T, fun, x0_deg, y0_deg, u)
  use, intrinsic :: iso_c_binding
  use satt_cabi, only: satt_fwd_lonlat
  use satt_host, only: satt_host_ctx, satt_host_check
  implicit none
  integer :: nx, ny
  double precision :: xx_deg(nx), yy_deg(ny), x0_deg, y0_deg
  double precision :: spha(nx,ny), sphb(nx,ny), sphc(nx,ny), T(nx,ny), fun(nx,ny), u(nx,ny)
  integer(c_int) :: st, iters
  ! u only feeds an error norm the reference computes and never uses (:1099-1104)
  st = satt_fwd_lonlat(satt_host_ctx(), xx_deg, yy_deg, nx, ny, spha, sphb, sphc, fun, &
                       x0_deg, y0_deg, T, iters)
  call satt_host_check(st, 'FSM_UW_PS_lonlat_2d')
end subroutine FSM_UW_PS_lonlat_2d

! src/eikon_solver_2d.f90:1276  FSM_O1_JSE_lonlat_2d
subroutine FSM_O1_JSE_lonlat_2d(xx_deg, yy_deg, nx, ny, spha, sphb, sphc, T, Ta, &
                                xrec_deg, yrec_deg, sourceAdj, nr)
  use, intrinsic :: iso_c_binding
  use satt_cabi, only: satt_adj_lonlat
  use satt_host, only: satt_host_ctx, satt_host_check
  implicit none
  integer :: nx, ny, nr
  double precision :: xx_deg(nx), yy_deg(ny), xrec_deg(nr), yrec_deg(nr), sourceAdj(nr)
  double precision :: spha(nx,ny), sphb(nx,ny), sphc(nx,ny), T(nx,ny), Ta(nx,ny)
  integer(c_int) :: st, iters
  st = satt_adj_lonlat(satt_host_ctx(), xx_deg, yy_deg, nx, ny, spha, sphb, sphc, T, &
                       xrec_deg, yrec_deg, sourceAdj, nr, Ta, iters)
  call satt_host_check(st, 'FSM_O1_JSE_lonlat_2d')
end subroutine FSM_O1_JSE_lonlat_2d

! src/eikon_solver_2d.f90:1449  mask
subroutine mask(xx, yy, nx, ny, TableAdj, x0, y0)
  use, intrinsic :: iso_c_binding
  use satt_cabi, only: satt_mask
  use satt_host, only: satt_host_ctx, satt_host_check
  implicit none
  integer :: nx, ny
  double precision :: xx(nx), yy(ny), TableAdj(nx,ny), x0, y0
  call satt_host_check(satt_mask(satt_host_ctx(), xx, yy, nx, ny, TableAdj, x0, y0), 'mask')
end subroutine mask

! src/eikon_solver_2d.f90:1538  inv_grid_iso
subroutine inv_grid_iso(xinv, yinv, zinv, nxinv, nyinv, nzinv, nset, adj, nx, ny, nz_3d, gk, &
                        vel_x, vel_y, vel_z_3d)
  use, intrinsic :: iso_c_binding
  use satt_cabi, only: satt_inv_grid_iso
  use satt_host, only: satt_host_ctx, satt_host_check
  implicit none
  integer :: nxinv, nyinv, nzinv, nset, nx, ny, nz_3d
  double precision :: xinv(nxinv, nset), yinv(nyinv, nset), zinv(nzinv, nset)
  double precision :: adj(nx, ny, nz_3d), gk(nset*nxinv*nyinv*nzinv)
  double precision :: vel_x(nx), vel_y(ny), vel_z_3d(nz_3d)
  call satt_host_check(satt_inv_grid_iso(satt_host_ctx(), xinv, yinv, zinv, nxinv, nyinv, &
                                         nzinv, nset, adj, nx, ny, nz_3d, gk, vel_x, vel_y, &
                                         vel_z_3d), 'inv_grid_iso')
end subroutine inv_grid_iso

! src/eikon_solver_2d.f90:1613  inv2fwd_iso
subroutine inv2fwd_iso(xinv, yinv, zinv, nxinv, nyinv, nzinv, nset, nx, ny, nz_3d, gk, &
                       vel_x, vel_y, vel_z_3d, update)
  use, intrinsic :: iso_c_binding
  use satt_cabi, only: satt_inv2fwd_iso
  use satt_host, only: satt_host_ctx, satt_host_check
  implicit none
  integer :: nxinv, nyinv, nzinv, nset, nx, ny, nz_3d
  double precision :: xinv(nxinv, nset), yinv(nyinv, nset), zinv(nzinv, nset)
  double precision :: gk(nset*nxinv*nyinv*nzinv), update(nx, ny, nz_3d)
  double precision :: vel_x(nx), vel_y(ny), vel_z_3d(nz_3d)
  call satt_host_check(satt_inv2fwd_iso(satt_host_ctx(), xinv, yinv, zinv, nxinv, nyinv, &
                                        nzinv, nset, nx, ny, nz_3d, gk, vel_x, vel_y, vel_z_3d, &
                                        update), 'inv2fwd_iso')
end subroutine inv2fwd_iso
