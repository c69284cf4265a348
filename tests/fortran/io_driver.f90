! A Fortran host writing and reading IMPACT's HDF5 files through libimpact_io
! (include/impact_io_c.f90), in the order of IMPACT's time loop: write_xdmf_xml
! and write_fields (src/mod_timeint.f90:199-201), write_restart (:605), then
! read_restart (:106) into cleared fields.
!
! Usage: io_driver <output directory>
! Grid 17x13x11, walls, L = (2, 1, 1), tanh 1.5 in y and z; velocity m and the
! pressure hold the pattern f(i,j,k) = mod(7(i+3) + 13(j+3) + 29(k+3) + s, 97)/32 - 1.5
! with s = 60 + m and 64.  Writes the HDF5 and XDMF files into the directory,
! the fields read back as <dir>/back_vel<m>.bin, back_pre.bin (raw doubles,
! i fastest) and the run state after the calls as state_i.bin / state_r.bin.
program io_driver
  use iso_c_binding
  use impact_hip_c
  use impact_hip_state
  use impact_io_c
  implicit none
  real(8), allocatable :: vel(:,:,:,:), pre(:,:,:), zero(:,:,:)
  character(len=1024) :: outdir
  character(len=:), allocatable :: prefix
  type(ih_io_state) :: st, st2
  integer(c_int) :: sizes(3), strides(9)
  integer :: m, u

  call get_command_argument(1, outdir)
  if (len_trim(outdir) == 0) then
     write(*,*) 'usage: io_driver <output directory>'
     stop 2
  end if
  if (ih_io_get_abi_version() /= IH_IO_ABI_VERSION) then
     write(*,*) 'impact_io: ABI version mismatch'
     stop 1
  end if
  prefix = trim(outdir) // '/' // c_null_char

  call impact_hip_init((/ 17, 13, 11 /), (/ 1, 1, 1, 1, 1, 1 /), (/ 2d0, 1d0, 1d0 /), (/ 0d0, 1.5d0, 1.5d0 /), 0)
  allocate(vel(b1L:N1+b1U, b2L:N2+b2U, b3L:N3+b3U, 1:3), pre(b1L:N1+b1U, b2L:N2+b2U, b3L:N3+b3U))
  allocate(zero(b1L:N1+b1U, b2L:N2+b2U, b3L:N3+b3U))
  zero = 0d0
  do m = 1, 3
     call fill(vel(:,:,:,m), 60 + m)
     call ih_check(ih_field_upload(ih_ctx, IH_VEL1 + m - 1, vel(:,:,:,m)))
  end do
  call fill(pre, 64)
  call ih_check(ih_field_upload(ih_ctx, IH_PRE, pre))

  st = ih_io_state(IH_IO_ABI_VERSION, 5, 2, 100, 1, 0, 0, 0, 1, 1, 1, &
                   0.5d0, 1d-3, 0.6d0, 0.7d0, 0.1d0, 0.2d0, 500d0, 0.6d0, 0.25d0, 1d-6, 1d0, 2d0, 3d0, 0.3d0, 0.02d0)
  sizes = (/ 1, 1, 0 /)
  strides = (/ 1, 1, 1, 2, 2, 2, 1, 1, 1 /)
  call io_check(ih_io_write_xdmf(ih_ctx, prefix, sizes, strides, st))
  call io_check(ih_io_write_fields(ih_ctx, prefix, sizes, strides, st))
  call io_check(ih_io_write_restart(ih_ctx, prefix, st))

  do m = 1, 3
     call ih_check(ih_field_upload(ih_ctx, IH_VEL1 + m - 1, zero))
  end do
  call ih_check(ih_field_upload(ih_ctx, IH_PRE, zero))
  st2 = st
  st2%time = 0d0
  st2%timestep = 0
  st2%dtime_out_vect = 0.05d0   ! the configured intervals of the resumed run
  st2%dtime_out_scal = 0.4d0
  call io_check(ih_io_read_restart(ih_ctx, prefix, 5_c_int, st2))
  do m = 1, 3
     vel(:,:,:,m) = -1d0
     call ih_check(ih_field_download(ih_ctx, IH_VEL1 + m - 1, vel(:,:,:,m)))
     call dump('back_vel' // achar(48 + m), vel(:,:,:,m))
  end do
  call ih_check(ih_field_download(ih_ctx, IH_PRE, pre))
  call dump('back_pre', pre)

  open(newunit=u, file=trim(outdir)//'/state_i.bin', access='stream', form='unformatted', status='replace')
  write(u) st%write_count, st%write_out_vect, st2%restart, st2%write_count, st2%timestep
  close(u)
  open(newunit=u, file=trim(outdir)//'/state_r.bin', access='stream', form='unformatted', status='replace')
  write(u) st%time_out_vect, st2%time, st2%dtime, st2%time_out_vect, st2%time_out_scal, st2%p_end_systole_out
  close(u)

  call impact_hip_finish()
  write(*,*) 'io_driver: done'

contains

  subroutine io_check(rc)
    integer(c_int), intent(in) :: rc
    character(kind=c_char) :: buf(512)
    integer :: q
    if (rc /= 0) then
       q = ih_io_last_error(buf, 512_c_size_t)
       write(*,*) 'impact_io error ', rc, ': ', (buf(q), q = 1, 200)
       stop 1
    end if
  end subroutine io_check

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
    integer :: v
    open(newunit=v, file=trim(outdir)//'/'//name//'.bin', access='stream', form='unformatted', status='replace')
    write(v) a
    close(v)
  end subroutine dump

end program io_driver
