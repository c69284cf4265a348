! Drop-in replacements for IMPACT's device_* host subroutines on the
! projection-step path: the same external subroutine names and argument lists
! as the reference's host files, each a thin call through iso_c_binding
! (include/impact_hip_c.f90) into libimpact_hip.
!
!   device_jacobi_solver      src/jacobi_solver_host_v4.f90:5
!   device_gradient           src/grad_host.f90:39
!   device_divergence2        src/divergence2_host.f90:6
!   device_Nonlinear_upwind   src/Nonlinear_upwind_host.f90:6
!   device_helmholtz          src/helmholtz_explicit_host.f90:6
!   device_axpby              src/axpby_host.f90:7
!   device_interpolate        src/interpolate2_pre_vel_host.f90:7
!   device_interpolate2       src/interpolate2_vel_pre_host.f90:6
!
! The reference subroutines take the grid sizes, the halo bounds, pp and multL
! from its modules (mod_dims, mod_vars).  Here module impact_hip_state holds
! them; inside IMPACT the USE of impact_hip_state is replaced by those modules
! and impact_hip_init/impact_hip_finish by the GPU_malloc/GPU_memfree calls.
module impact_hip_state
  use iso_c_binding
  use impact_hip_c
  implicit none
  type(c_ptr), save :: ih_ctx = c_null_ptr
  integer, save :: N1 = 0, N2 = 0, N3 = 0
  integer, save :: NN(3, 1) = 0                    ! grid sizes per multigrid level (finest only)
  integer, parameter :: b1L = -3, b1U = 3, b2L = -3, b2U = 3, b3L = -3, b3U = 3
  real(8), save :: multL = 0d0                     ! Helmholtz factor of the current substep
  real(8), allocatable, save :: pp(:, :, :)        ! interpolated velocity of the upwind term
contains
  ! GPU_malloc (src/mod_mem_GPU.f90:152): every field of the block in HBM, coefficients generated
  subroutine impact_hip_init(N, bc, L, stretch, device)
    integer, intent(in) :: N(3), bc(6), device
    real(8), intent(in) :: L(3), stretch(3)
    type(ih_grid_desc) :: d
    d%version = IH_ABI_VERSION
    d%N = int(N, c_int)
    d%bc = int(bc, c_int)
    d%L = L
    d%stretch = stretch
    call ih_check(ih_case_create(d, int(device, c_int), ih_ctx))
    call ih_check(ih_case_generate_coeffs(ih_ctx))
    N1 = N(1)
    N2 = N(2)
    N3 = N(3)
    NN(:, 1) = N
    if (allocated(pp)) deallocate(pp)
    allocate(pp(b1L:(N1 + b1U), b2L:(N2 + b2U), b3L:(N3 + b3U)))
    pp = 0d0
  end subroutine impact_hip_init

  ! GPU_memfree (src/mod_mem_GPU.f90:437)
  subroutine impact_hip_finish()
    if (c_associated(ih_ctx)) call ih_case_destroy(ih_ctx)
    ih_ctx = c_null_ptr
    if (allocated(pp)) deallocate(pp)
  end subroutine impact_hip_finish

  ! a failed call prints the library's message and stops, as the reference's
  ! cuda_err_check does (src/grad_host.f90:7)
  subroutine ih_check(rc)
    integer(c_int), intent(in) :: rc
    character(kind=c_char) :: msg(512)
    integer(c_int) :: n, q
    if (rc /= IH_OK) then
       n = ih_last_error(msg, int(512, c_size_t))
       write(*, '(a)', advance='no') 'impact_hip: '
       do q = 1, min(n, 512)
          write(*, '(a)', advance='no') msg(q)
       end do
       write(*, *)
       stop 1
    end if
  end subroutine ih_check
end module impact_hip_state

subroutine device_jacobi_solver(grid_id, omega, SS1, SS2, SS3, NN1, NN2, NN3, rel, bb)
  use impact_hip_state
  implicit none
  integer, intent(in) :: grid_id, SS1, SS2, SS3, NN1, NN2, NN3
  real(8), intent(in) :: omega
  real(8), intent(inout) :: rel(b1L:(NN(1, grid_id) + b1U), b2L:(NN(2, grid_id) + b2U), b3L:(NN(3, grid_id) + b3U))
  real(8), intent(inout) :: bb(b1L:(NN(1, grid_id) + b1U), b2L:(NN(2, grid_id) + b2U), b3L:(NN(3, grid_id) + b3U))
  call ih_check(ih_device_jacobi_solver(ih_ctx, int(grid_id, c_int), omega, int(SS1, c_int), int(SS2, c_int), &
       int(SS3, c_int), int(NN1, c_int), int(NN2, c_int), int(NN3, c_int), rel, bb))
end subroutine device_jacobi_solver

subroutine device_gradient(dimension_id, phi, grad)
  use impact_hip_state
  implicit none
  integer, intent(in) :: dimension_id
  real(8), intent(inout) :: phi(b1L:(N1 + b1U), b2L:(N2 + b2U), b3L:(N3 + b3U))
  real(8), intent(inout) :: grad(b1L:(N1 + b1U), b2L:(N2 + b2U), b3L:(N3 + b3U))
  call ih_check(ih_device_gradient(ih_ctx, int(dimension_id, c_int), phi, grad))
end subroutine device_gradient

subroutine device_divergence2(phi, div)
  use impact_hip_state
  implicit none
  real(8), intent(inout) :: phi(b1L:(N1 + b1U), b2L:(N2 + b2U), b3L:(N3 + b3U), 3)
  real(8), intent(out) :: div(b1L:(N1 + b1U), b2L:(N2 + b2U), b3L:(N3 + b3U))
  call ih_check(ih_device_divergence2(ih_ctx, phi, div))
end subroutine device_divergence2

subroutine device_Nonlinear_upwind(component_id, vel_component, nl_component)
  use impact_hip_state
  implicit none
  integer, intent(in) :: component_id
  real(8), intent(in) :: vel_component(b1L:(N1 + b1U), b2L:(N2 + b2U), b3L:(N3 + b3U))
  real(8), intent(inout) :: nl_component(b1L:(N1 + b1U), b2L:(N2 + b2U), b3L:(N3 + b3U))
  call ih_check(ih_device_nonlinear_upwind(ih_ctx, int(component_id, c_int), pp, vel_component, nl_component))
end subroutine device_Nonlinear_upwind

subroutine device_helmholtz(component_id, vel_component, helm_component)
  use impact_hip_state
  implicit none
  integer, intent(in) :: component_id
  real(8), intent(in) :: vel_component(b1L:(N1 + b1U), b2L:(N2 + b2U), b3L:(N3 + b3U))
  real(8), intent(inout) :: helm_component(b1L:(N1 + b1U), b2L:(N2 + b2U), b3L:(N3 + b3U))
  call ih_check(ih_device_helmholtz(ih_ctx, int(component_id, c_int), multL, vel_component, helm_component))
end subroutine device_helmholtz

subroutine device_axpby(a, b, x, y, z)
  use impact_hip_state
  implicit none
  real(8), intent(in) :: a, b
  real(8), intent(in) :: x(b1L:(N1 + b1U), b2L:(N2 + b2U), b3L:(N3 + b3U))
  real(8), intent(in) :: y(b1L:(N1 + b1U), b2L:(N2 + b2U), b3L:(N3 + b3U))
  real(8), intent(inout) :: z(b1L:(N1 + b1U), b2L:(N2 + b2U), b3L:(N3 + b3U))
  call ih_check(ih_device_axpby(ih_ctx, a, b, x, y, z))
end subroutine device_axpby

subroutine device_interpolate(dimension_id, phi_mat, inter_mat)
  use impact_hip_state
  implicit none
  integer, intent(in) :: dimension_id
  real(8), intent(inout) :: phi_mat(b1L:(N1 + b1U), b2L:(N2 + b2U), b3L:(N3 + b3U))
  real(8), intent(out) :: inter_mat(b1L:(N1 + b1U), b2L:(N2 + b2U), b3L:(N3 + b3U))
  call ih_check(ih_device_interpolate(ih_ctx, int(dimension_id, c_int), phi_mat, inter_mat))
end subroutine device_interpolate

subroutine device_interpolate2(dimension_id, phi_mat, inter_mat)
  use impact_hip_state
  implicit none
  integer, intent(in) :: dimension_id
  real(8), intent(inout) :: phi_mat(b1L:(N1 + b1U), b2L:(N2 + b2U), b3L:(N3 + b3U))
  real(8), intent(inout) :: inter_mat(b1L:(N1 + b1U), b2L:(N2 + b2U), b3L:(N3 + b3U))
  call ih_check(ih_device_interpolate2(ih_ctx, int(dimension_id, c_int), phi_mat, inter_mat))
end subroutine device_interpolate2
