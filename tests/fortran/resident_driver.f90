! A Fortran host running one projection substep through the reference's
! device_* names (include/impact_hip_host.f90) twice from the same inputs:
! first as the reference's wrappers work (every array crosses PCIe on every
! call), then with the module arrays bound to resident fields (ih_bind_field)
! and one coherence point at the end (ih_sync_all).  The chain: velocity to the
! pressure grid (device_interpolate2), the upwind nonlinear term and the
! explicit Helmholtz step per component (device_interpolate,
! device_Nonlinear_upwind, device_helmholtz), rhs = u - dt a nl (device_axpby),
! div rhs (device_divergence2), the pressure by device_jacobi_solver, and
! u = rhs - G p (device_gradient, device_axpby).
!
! Usage: resident_driver <output directory>
! Grid 17x13x11, walls, L = (2, 1, 1), tanh 1.5 in y and z; inputs the pattern
! mod(7(i+3) + 13(j+3) + 29(k+3) + s, 97)/32 - 1.5.  Writes <pass>_<name>.bin
! (raw doubles, Fortran order) for both passes and transfers.txt with the
! whole-field uploads and downloads of each pass.
program resident_driver
  use iso_c_binding
  use impact_hip_c
  use impact_hip_state
  implicit none
  real(8), allocatable :: vel(:,:,:,:), work(:,:,:,:), nl(:,:,:,:), rhs(:,:,:,:)
  real(8), allocatable :: pre(:,:,:), div(:,:,:), gpre(:,:,:)
  real(8), parameter :: dta = -1d0 / 64d0
  character(len=1024) :: outdir
  integer(c_int64_t) :: t0(4), t1(4), t2(4)
  integer :: m, u
  external :: device_gradient, device_divergence2, device_interpolate, device_interpolate2, device_axpby
  external :: device_Nonlinear_upwind, device_helmholtz, device_jacobi_solver

  call get_command_argument(1, outdir)
  if (len_trim(outdir) == 0) then
     write(*,*) 'usage: resident_driver <output directory>'
     stop 2
  end if
  call impact_hip_init((/ 17, 13, 11 /), (/ 1, 1, 1, 1, 1, 1 /), (/ 2d0, 1d0, 1d0 /), (/ 0d0, 1.5d0, 1.5d0 /), 0)
  allocate(vel(b1L:N1+b1U, b2L:N2+b2U, b3L:N3+b3U, 1:3), work(b1L:N1+b1U, b2L:N2+b2U, b3L:N3+b3U, 1:3))
  allocate(nl(b1L:N1+b1U, b2L:N2+b2U, b3L:N3+b3U, 1:3), rhs(b1L:N1+b1U, b2L:N2+b2U, b3L:N3+b3U, 1:3))
  allocate(pre(b1L:N1+b1U, b2L:N2+b2U, b3L:N3+b3U), div(b1L:N1+b1U, b2L:N2+b2U, b3L:N3+b3U))
  allocate(gpre(b1L:N1+b1U, b2L:N2+b2U, b3L:N3+b3U))
  multL = 1d0 / 137d0

  ! pass 1: the reference's wrappers (nothing bound)
  call init_state()
  call ih_check(ih_transfer_stats(ih_ctx, t0))
  call substep()
  call ih_check(ih_transfer_stats(ih_ctx, t1))
  call dump_all('plain')

  ! pass 2: the module arrays bound to resident fields, one coherence point
  call init_state()
  do m = 1, 3
     call ih_check(ih_bind_field(ih_ctx, IH_VEL1 + m - 1, vel(:,:,:,m)))
     call ih_check(ih_bind_field(ih_ctx, IH_WORK1 + m - 1, work(:,:,:,m)))
     call ih_check(ih_bind_field(ih_ctx, IH_NL1 + m - 1, nl(:,:,:,m)))
     call ih_check(ih_bind_field(ih_ctx, IH_RHS1 + m - 1, rhs(:,:,:,m)))
  end do
  call ih_check(ih_bind_field(ih_ctx, IH_PP, pp))
  call ih_check(ih_bind_field(ih_ctx, IH_PRE, pre))
  call ih_check(ih_bind_field(ih_ctx, IH_RES, div))
  call ih_check(ih_bind_field(ih_ctx, IH_GPRE, gpre))
  call substep()
  call ih_check(ih_sync_all(ih_ctx, IH_SYNC_TO_HOST))
  call ih_check(ih_transfer_stats(ih_ctx, t2))
  call dump_all('bound')

  open(newunit=u, file=trim(outdir)//'/transfers.txt', status='replace')
  write(u, *) t1(1) - t0(1), t1(2) - t0(2)
  write(u, *) t2(1) - t1(1), t2(2) - t1(2)
  close(u)
  call impact_hip_finish()
  write(*,*) 'resident_driver: done'

contains

  subroutine init_state()
    integer :: q
    do q = 1, 3
       call fill(vel(:,:,:,q), 60 + q)
       call fill(work(:,:,:,q), 63 + q)
       call fill(nl(:,:,:,q), 66 + q)
       call fill(rhs(:,:,:,q), 69 + q)
    end do
    call fill(pp, 73)
    call fill(div, 74)
    call fill(gpre, 75)
    pre = 0d0
  end subroutine init_state

  subroutine substep()
    integer :: q, d
    do d = 1, 3
       call device_interpolate2(d, vel(:,:,:,d), work(:,:,:,d))
    end do
    do q = 1, 3
       do d = 1, 3
          call device_interpolate(q, work(:,:,:,d), pp)
          call device_Nonlinear_upwind(3 * (q - 1) + d, vel(:,:,:,q), nl(:,:,:,q))
       end do
       call device_helmholtz(q, vel(:,:,:,q), nl(:,:,:,q))
    end do
    do q = 1, 3
       call device_axpby(1d0, dta, vel(:,:,:,q), nl(:,:,:,q), rhs(:,:,:,q))
    end do
    call device_divergence2(rhs, div)
    call device_jacobi_solver(1, 0.8d0, 1, 1, 1, N1, N2, N3, pre, div)
    do q = 1, 3
       call device_gradient(q, pre, gpre)
       call device_axpby(1d0, -1d0, rhs(:,:,:,q), gpre, vel(:,:,:,q))
    end do
  end subroutine substep

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

  subroutine dump_all(tag)
    character(len=*), intent(in) :: tag
    integer :: q
    character(len=1) :: c
    do q = 1, 3
       write(c, '(i1)') q
       call dump(tag//'_vel'//c, vel(:,:,:,q))
       call dump(tag//'_nl'//c, nl(:,:,:,q))
       call dump(tag//'_rhs'//c, rhs(:,:,:,q))
    end do
    call dump(tag//'_pre', pre)
    call dump(tag//'_div', div)
  end subroutine dump_all

  subroutine dump(name, a)
    character(len=*), intent(in) :: name
    real(8), intent(in) :: a(:, :, :)
    integer :: v
    open(newunit=v, file=trim(outdir)//'/'//name//'.bin', access='stream', form='unformatted', status='replace')
    write(v) a
    close(v)
  end subroutine dump

end program resident_driver
