! A Fortran host calling libimpact_hip the way IMPACT's host calls its device
! wrappers: arrays with the reference's bounds (-3:N+3 per direction), REAL(8),
! passed to the reference-signature entry points through iso_c_binding.
!
! Usage: dropin_driver <output directory>
! Grid 17x13x11, walls, L = (2, 1, 1), tanh 1.5 in y and z.  Inputs are the
! exactly representable pattern f(i,j,k) = mod(7(i+3) + 13(j+3) + 29(k+3) + s, 97)/32 - 1.5
! for seeds s; every result array is written to <dir>/<name>.bin as raw
! doubles in Fortran (i fastest) order.
program dropin_driver
  use iso_c_binding
  use impact_hip_c
  implicit none
  integer, parameter :: N1 = 17, N2 = 13, N3 = 11
  real(8), allocatable :: phi(:,:,:), grad(:,:,:), vel(:,:,:,:), div(:,:,:), inter(:,:,:)
  real(8), allocatable :: x(:,:,:), y(:,:,:), z(:,:,:), rel(:,:,:), bb(:,:,:)
  type(ih_grid_desc) :: d
  type(c_ptr) :: ctx
  character(len=1024) :: outdir
  integer :: m

  call get_command_argument(1, outdir)
  if (len_trim(outdir) == 0) then
     write(*,*) 'usage: dropin_driver <output directory>'
     stop 2
  end if
  if (ih_get_abi_version() /= IH_ABI_VERSION) then
     write(*,*) 'impact_hip: ABI version mismatch'
     stop 1
  end if

  allocate(phi(-3:N1+3,-3:N2+3,-3:N3+3), grad(-3:N1+3,-3:N2+3,-3:N3+3), div(-3:N1+3,-3:N2+3,-3:N3+3))
  allocate(inter(-3:N1+3,-3:N2+3,-3:N3+3), vel(-3:N1+3,-3:N2+3,-3:N3+3,1:3))
  allocate(x(-3:N1+3,-3:N2+3,-3:N3+3), y(-3:N1+3,-3:N2+3,-3:N3+3), z(-3:N1+3,-3:N2+3,-3:N3+3))
  allocate(rel(-3:N1+3,-3:N2+3,-3:N3+3), bb(-3:N1+3,-3:N2+3,-3:N3+3))

  d%version = IH_ABI_VERSION
  d%N = (/ N1, N2, N3 /)
  d%bc = (/ 1, 1, 1, 1, 1, 1 /)
  d%L = (/ 2d0, 1d0, 1d0 /)
  d%stretch = (/ 0d0, 1.5d0, 1.5d0 /)
  call check(ih_case_create(d, 0_c_int, ctx))
  call check(ih_case_generate_coeffs(ctx))

  ! device_gradient (src/grad_host.f90:39)
  call fill(phi, 1)
  call fill(grad, 2)
  call check(ih_device_gradient(ctx, 2_c_int, phi, grad))
  call dump('gradient2', grad)

  ! device_divergence2 (src/divergence2_host.f90:6)
  do m = 1, 3
     call fill(vel(:,:,:,m), 10 + m)
  end do
  call fill(div, 3)
  call check(ih_device_divergence2(ctx, vel, div))
  call dump('divergence2', div)

  ! device_interpolate2 (src/interpolate2_vel_pre_host.f90)
  call fill(phi, 4)
  call fill(inter, 5)
  call check(ih_device_interpolate2(ctx, 3_c_int, phi, inter))
  call dump('interpolate2_3', inter)

  ! device_axpby (src/axpby_host.f90:5): z = x - y
  call fill(x, 6)
  call fill(y, 7)
  z = 0d0
  call check(ih_device_axpby(ctx, 1d0, -1d0, x, y, z))
  call dump('axpby', z)

  ! resident mode: upload, 5 JOR sweeps on the pressure range, download
  call fill(rel, 8)
  call fill(bb, 9)
  call check(ih_field_upload(ctx, IH_PRE, rel))
  call check(ih_field_upload(ctx, IH_RES, bb))
  call check(ih_jor(ctx, 5_c_int, 0.8d0, (/ 1_c_int, 1_c_int, 1_c_int /), (/ N1, N2, N3 /), IH_PRE, IH_RES))
  call check(ih_field_download(ctx, IH_PRE, rel))
  call dump('jor5', rel)

  call ih_case_destroy(ctx)
  write(*,*) 'dropin_driver: done'

contains

  subroutine check(rc)
    integer(c_int), intent(in) :: rc
    character(kind=c_char) :: msg(512)
    integer(c_int) :: n, q
    if (rc /= IH_OK) then
       n = ih_last_error(msg, int(512, c_size_t))
       write(*,'(a)', advance='no') 'impact_hip: '
       do q = 1, min(n, 512)
          write(*,'(a)', advance='no') msg(q)
       end do
       write(*,*)
       stop 1
    end if
  end subroutine check

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
