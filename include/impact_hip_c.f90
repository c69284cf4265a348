! Fortran interface to libimpact_hip (include/impact_hip.h) through iso_c_binding.
!
! This is the module a Fortran host such as IMPACT adds to reach the library:
! the C entry points with their C argument types.  A second, host-specific
! module re-exposes the reference's own subroutine names on top of it (see
! the drop-in module in the integration notes); the reference interfaces
! each entry point replaces are listed in include/impact_hip.h.
module impact_hip_c
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: IH_ABI_VERSION = 4
  integer(c_int), parameter :: IH_OK = 0

  ! fields owned by a case (enum ih_field)
  integer(c_int), parameter :: IH_VEL1 = 0, IH_VEL2 = 1, IH_VEL3 = 2
  integer(c_int), parameter :: IH_NL1 = 3, IH_NL2 = 4, IH_NL3 = 5
  integer(c_int), parameter :: IH_RHS1 = 6, IH_RHS2 = 7, IH_RHS3 = 8
  integer(c_int), parameter :: IH_WORK1 = 9, IH_WORK2 = 10, IH_WORK3 = 11
  integer(c_int), parameter :: IH_PRE = 12, IH_RES = 13, IH_GPRE = 14, IH_PP = 15, IH_COMP = 16, IH_DIG = 17
  integer(c_int), parameter :: IH_RR = 18, IH_RH = 19, IH_PK = 20, IH_AP = 21, IH_AR = 22
  integer(c_int), parameter :: IH_TMP1 = 23, IH_TMP2 = 24, IH_TMP3 = 25, IH_TMP4 = 26
  integer(c_int), parameter :: IH_Z1 = 27, IH_Z2 = 28
  integer(c_int), parameter :: IH_SYNC_TO_HOST = 0, IH_SYNC_TO_DEVICE = 1

  type, bind(C) :: ih_grid_desc
     integer(c_int) :: version
     integer(c_int) :: N(3)
     integer(c_int) :: bc(6)
     real(c_double) :: L(3)
     real(c_double) :: stretch(3)
  end type ih_grid_desc

  interface
     ! ih_abi_version() (Fortran names are case-blind: IH_ABI_VERSION is the constant above)
     integer(c_int) function ih_get_abi_version() bind(C, name="ih_abi_version")
       import :: c_int
     end function ih_get_abi_version

     integer(c_int) function ih_last_error(buf, n) bind(C, name="ih_last_error")
       import :: c_int, c_char, c_size_t
       character(kind=c_char) :: buf(*)
       integer(c_size_t), value :: n
     end function ih_last_error

     integer(c_int) function ih_case_create(desc, device, ctx) bind(C, name="ih_case_create")
       import :: c_int, c_ptr, ih_grid_desc
       type(ih_grid_desc), intent(in) :: desc
       integer(c_int), value :: device
       type(c_ptr), intent(out) :: ctx
     end function ih_case_create

     subroutine ih_case_destroy(ctx) bind(C, name="ih_case_destroy")
       import :: c_ptr
       type(c_ptr), value :: ctx
     end subroutine ih_case_destroy

     integer(c_int) function ih_case_generate_coeffs(ctx) bind(C, name="ih_case_generate_coeffs")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function ih_case_generate_coeffs

     integer(c_int) function ih_field_upload(ctx, f, host) bind(C, name="ih_field_upload")
       import :: c_int, c_double, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), value :: f
       real(c_double), intent(in) :: host(*)
     end function ih_field_upload

     integer(c_int) function ih_field_download(ctx, f, host) bind(C, name="ih_field_download")
       import :: c_int, c_double, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), value :: f
       real(c_double), intent(out) :: host(*)
     end function ih_field_download

     integer(c_int) function ih_jor(ctx, n_sweeps, omega, ss, nn, rel, bb) bind(C, name="ih_jor")
       import :: c_int, c_double, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), value :: n_sweeps
       real(c_double), value :: omega
       integer(c_int), intent(in) :: ss(3), nn(3)
       integer(c_int), value :: rel, bb
     end function ih_jor

     ! BiCGstab_orig with the reference's preconditioner argument (0 none, 1 multigrid V cycle)
     integer(c_int) function ih_bicgstab_precond(ctx, eps, n_it_max, ss, nn, bb, phi, preconditioner, st) &
          bind(C, name="ih_bicgstab_precond")
       import :: c_int, c_double, c_ptr
       type(c_ptr), value :: ctx
       real(c_double), value :: eps
       integer(c_int), value :: n_it_max
       integer(c_int), intent(in) :: ss(3), nn(3)
       integer(c_int), value :: bb, phi, preconditioner
       type(c_ptr), value :: st
     end function ih_bicgstab_precond

     integer(c_int) function ih_mg_configure(ctx, n_grids_limit, n_relax_down, n_relax_up, n_relax_bottom) &
          bind(C, name="ih_mg_configure")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), value :: n_grids_limit, n_relax_down, n_relax_up, n_relax_bottom
     end function ih_mg_configure

     ! multigridV(.TRUE., 1, bb, phi) on resident fields
     integer(c_int) function ih_multigrid_v(ctx, bb, phi) bind(C, name="ih_multigrid_v")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), value :: bb, phi
     end function ih_multigrid_v

     integer(c_int) function ih_device_jacobi_solver(ctx, grid_id, omega, SS1, SS2, SS3, NN1, NN2, NN3, rel, bb) &
          bind(C, name="ih_device_jacobi_solver")
       import :: c_int, c_double, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), value :: grid_id, SS1, SS2, SS3, NN1, NN2, NN3
       real(c_double), value :: omega
       real(c_double) :: rel(*), bb(*)
     end function ih_device_jacobi_solver

     integer(c_int) function ih_device_gradient(ctx, dimension_id, phi, grad) bind(C, name="ih_device_gradient")
       import :: c_int, c_double, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), value :: dimension_id
       real(c_double) :: phi(*), grad(*)
     end function ih_device_gradient

     integer(c_int) function ih_device_divergence2(ctx, phi3, div) bind(C, name="ih_device_divergence2")
       import :: c_int, c_double, c_ptr
       type(c_ptr), value :: ctx
       real(c_double) :: phi3(*), div(*)
     end function ih_device_divergence2

     integer(c_int) function ih_device_nonlinear_upwind(ctx, component_id, pp, vel_c, nl_c) &
          bind(C, name="ih_device_nonlinear_upwind")
       import :: c_int, c_double, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), value :: component_id
       real(c_double) :: pp(*), vel_c(*), nl_c(*)
     end function ih_device_nonlinear_upwind

     integer(c_int) function ih_device_helmholtz(ctx, component_id, multL, vel_c, helm_c) &
          bind(C, name="ih_device_helmholtz")
       import :: c_int, c_double, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), value :: component_id
       real(c_double), value :: multL
       real(c_double) :: vel_c(*), helm_c(*)
     end function ih_device_helmholtz

     integer(c_int) function ih_device_axpby(ctx, a, b, x, y, z) bind(C, name="ih_device_axpby")
       import :: c_int, c_double, c_ptr
       type(c_ptr), value :: ctx
       real(c_double), value :: a, b
       real(c_double) :: x(*), y(*), z(*)
     end function ih_device_axpby

     integer(c_int) function ih_device_interpolate(ctx, dimension_id, phi, inter) bind(C, name="ih_device_interpolate")
       import :: c_int, c_double, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), value :: dimension_id
       real(c_double) :: phi(*), inter(*)
     end function ih_device_interpolate

     integer(c_int) function ih_device_interpolate2(ctx, dimension_id, phi, inter) &
          bind(C, name="ih_device_interpolate2")
       import :: c_int, c_double, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), value :: dimension_id
       real(c_double) :: phi(*), inter(*)
     end function ih_device_interpolate2

     ! residency registry: field f mirrors the host array (identified by its address)
     integer(c_int) function ih_bind_field(ctx, f, host) bind(C, name="ih_bind_field")
       import :: c_int, c_double, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), value :: f
       real(c_double) :: host(*)
     end function ih_bind_field

     integer(c_int) function ih_sync(ctx, f, dir) bind(C, name="ih_sync")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), value :: f, dir
     end function ih_sync

     integer(c_int) function ih_sync_all(ctx, dir) bind(C, name="ih_sync_all")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), value :: dir
     end function ih_sync_all

     integer(c_int) function ih_transfer_stats(ctx, out) bind(C, name="ih_transfer_stats")
       import :: c_int, c_int64_t, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int64_t), intent(out) :: out(4)
     end function ih_transfer_stats
  end interface
end module impact_hip_c
