! A Fortran host calling libimpact_hip the way IMPACT's host calls its device
! wrappers: the drop-in subroutines of include/impact_hip_host.f90 (the
! reference's names and argument lists) on arrays with the reference's bounds
! (-3:N+3 per direction), plus the resident calls of include/impact_hip_c.f90.
!
! Usage: dropin_driver <output directory>
! Grid 17x13x11, walls, L = (2, 1, 1), tanh 1.5 in y and z.  Inputs are the
! exactly representable pattern f(i,j,k) = mod(7(i+3) + 13(j+3) + 29(k+3) + s, 97)/32 - 1.5
! for seeds s; every result array is written to <dir>/<name>.bin as raw
! doubles in Fortran (i fastest) order.
program dropin_driver
  use iso_c_binding
  use impact_hip_c
  use impact_hip_state
  implicit none
  real(8), allocatable :: phi(:,:,:), grad(:,:,:), vel(:,:,:,:), div(:,:,:), inter(:,:,:)
  real(8), allocatable :: x(:,:,:), y(:,:,:), z(:,:,:), rel(:,:,:), bb(:,:,:), work(:,:,:,:), nl(:,:,:,:)
  character(len=1024) :: outdir
  integer :: m, q, d
  external :: device_gradient, device_divergence2, device_interpolate, device_interpolate2, device_axpby
  external :: device_Nonlinear_upwind, device_helmholtz, device_jacobi_solver

  call get_command_argument(1, outdir)
  if (len_trim(outdir) == 0) then
     write(*,*) 'usage: dropin_driver <output directory>'
     stop 2
  end if
  if (ih_get_abi_version() /= IH_ABI_VERSION) then
     write(*,*) 'impact_hip: ABI version mismatch'
     stop 1
  end if

  ! GPU_malloc
  call impact_hip_init((/ 17, 13, 11 /), (/ 1, 1, 1, 1, 1, 1 /), (/ 2d0, 1d0, 1d0 /), (/ 0d0, 1.5d0, 1.5d0 /), 0)

  allocate(phi(b1L:N1+b1U, b2L:N2+b2U, b3L:N3+b3U), grad(b1L:N1+b1U, b2L:N2+b2U, b3L:N3+b3U))
  allocate(div(b1L:N1+b1U, b2L:N2+b2U, b3L:N3+b3U), inter(b1L:N1+b1U, b2L:N2+b2U, b3L:N3+b3U))
  allocate(vel(b1L:N1+b1U, b2L:N2+b2U, b3L:N3+b3U, 1:3), work(b1L:N1+b1U, b2L:N2+b2U, b3L:N3+b3U, 1:3))
  allocate(nl(b1L:N1+b1U, b2L:N2+b2U, b3L:N3+b3U, 1:3))
  allocate(x(b1L:N1+b1U, b2L:N2+b2U, b3L:N3+b3U), y(b1L:N1+b1U, b2L:N2+b2U, b3L:N3+b3U))
  allocate(z(b1L:N1+b1U, b2L:N2+b2U, b3L:N3+b3U))
  allocate(rel(b1L:N1+b1U, b2L:N2+b2U, b3L:N3+b3U), bb(b1L:N1+b1U, b2L:N2+b2U, b3L:N3+b3U))

  call fill(phi, 1)
  call fill(grad, 2)
  call device_gradient(2, phi, grad)
  call dump('gradient2', grad)

  do m = 1, 3
     call fill(vel(:,:,:,m), 10 + m)
  end do
  call fill(div, 3)
  call device_divergence2(vel, div)
  call dump('divergence2', div)

  call fill(phi, 4)
  call fill(inter, 5)
  call device_interpolate2(3, phi, inter)
  call dump('interpolate2_3', inter)

  call fill(x, 6)
  call fill(y, 7)
  z = 0d0
  call device_axpby(1d0, -1d0, x, y, z)
  call dump('axpby', z)

  ! the upwind nonlinear term and the explicit Helmholtz step per component,
  ! as rhs_vel drives them: pp = I_q(work_d), then the upwind term of component
  ! (q, d), then the Helmholtz term of q
  do m = 1, 3
     call fill(vel(:,:,:,m), 49 + m)
     call fill(work(:,:,:,m), 52 + m)
     call fill(nl(:,:,:,m), 55 + m)
  end do
  multL = 1d0 / 137d0
  do q = 1, 3
     do d = 1, 3
        call device_interpolate(q, work(:,:,:,d), pp)
        call device_Nonlinear_upwind(3 * (q - 1) + d, vel(:,:,:,q), nl(:,:,:,q))
     end do
     call device_helmholtz(q, vel(:,:,:,q), nl(:,:,:,q))
  end do
  call dump('nl1', nl(:,:,:,1))
  call dump('nl2', nl(:,:,:,2))
  call dump('nl3', nl(:,:,:,3))

  call fill(rel, 20)
  call fill(bb, 21)
  call device_jacobi_solver(1, 0.8d0, 1, 1, 1, N1, N2, N3, rel, bb)
  call dump('jacobi', rel)

  ! resident mode: upload, 5 JOR sweeps on the pressure range, download
  call fill(rel, 8)
  call fill(bb, 9)
  call ih_check(ih_field_upload(ih_ctx, IH_PRE, rel))
  call ih_check(ih_field_upload(ih_ctx, IH_RES, bb))
  call ih_check(ih_jor(ih_ctx, 5_c_int, 0.8d0, (/ 1_c_int, 1_c_int, 1_c_int /), (/ int(N1, c_int), &
       int(N2, c_int), int(N3, c_int) /), IH_PRE, IH_RES))
  call ih_check(ih_field_download(ih_ctx, IH_PRE, rel))
  call dump('jor5', rel)

  ! GPU_memfree
  call impact_hip_finish()
  write(*,*) 'dropin_driver: done'

contains

  subroutine fill(a, s)
    real(8), intent(out) :: a(-3:, -3:, -3:)
    integer, intent(in) :: s
    integer :: i, j, k
    do k = lbound(a, 3), ubound(a, 3)
       do j = lbound(a, 2), ubound(a, 2)
          do i = lbound(a, 1), ubound(a, 1)
             a(i, j, k) = dble(mod(7 * (i + 3) + 13 * (j + 3) + 29 * (k + 3) + s, 97)) / 32d0 - 1.5d0
          end do
       end do
    end do
  end subroutine fill

  subroutine dump(name, a)
    character(len=*), intent(in) :: name
    real(8), intent(in) :: a(:, :, :)
    integer :: u
    open(newunit=u, file=trim(outdir)//'/'//name//'.bin', access='stream', form='unformatted', status='replace')
    write(u) a
    close(u)
  end subroutine dump

end program dropin_driver
