! Fortran interface to libimpact_io (include/impact_io.h) through iso_c_binding:
! HDF5 field output and restart of one block.  IMPACT's time loop calls these
! where it calls write_xdmf_xml, write_fields, write_restart and read_restart
! (src/mod_timeint.f90:106, 199-201, 605), with ih_io_state carrying the module
! variables those routines read and write.  Strings are passed NUL-terminated
! (trim(s)//c_null_char).
module impact_io_c
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: IH_IO_ABI_VERSION = 1

  ! ih_io_state: the run state stored as dataset attributes (logicals as 0/1)
  type, bind(C) :: ih_io_state
     integer(c_int) :: version
     integer(c_int) :: restart, write_count, timestep
     integer(c_int) :: write_out_vect, write_out_scal, new_dtime
     integer(c_int) :: Euler_yes, twostep_yes, mapping_yes, upwind_yes
     real(c_double) :: time, dtime
     real(c_double) :: time_out_vect, time_out_scal, dtime_out_vect, dtime_out_scal
     real(c_double) :: Re, CFL, thetaL, epsU
     real(c_double) :: p_end_systole_in, p_end_systole_out, p_end_systole_out_2
     real(c_double) :: U_ref, L_ref
  end type ih_io_state

  interface
     integer(c_int) function ih_io_get_abi_version() bind(C, name="ih_io_abi_version")
       import :: c_int
     end function ih_io_get_abi_version

     integer(c_int) function ih_io_last_error(buf, n) bind(C, name="ih_io_last_error")
       import :: c_int, c_char, c_size_t
       character(kind=c_char), intent(out) :: buf(*)
       integer(c_size_t), value :: n
     end function ih_io_last_error

     ! write_hdf (src/mod_inout.f90:503-694)
     integer(c_int) function ih_io_write_hdf(ctx, filename, dsetname, f, vel_dir, stride, st) &
          bind(C, name="ih_io_write_hdf")
       import :: c_int, c_ptr, c_char, ih_io_state
       type(c_ptr), value :: ctx
       character(kind=c_char), intent(in) :: filename(*), dsetname(*)
       integer(c_int), value :: f, vel_dir
       integer(c_int), intent(in) :: stride(3)
       type(ih_io_state), intent(in) :: st
     end function ih_io_write_hdf

     ! read2_hdf (src/mod_inout.f90:1452-1720)
     integer(c_int) function ih_io_read_hdf(ctx, filename, dsetname, f, vel_dir, st) bind(C, name="ih_io_read_hdf")
       import :: c_int, c_ptr, c_char, ih_io_state
       type(c_ptr), value :: ctx
       character(kind=c_char), intent(in) :: filename(*), dsetname(*)
       integer(c_int), value :: f, vel_dir
       type(ih_io_state), intent(inout) :: st
     end function ih_io_read_hdf

     ! write_restart (src/mod_inout.f90:326-359)
     integer(c_int) function ih_io_write_restart(ctx, prefix, st) bind(C, name="ih_io_write_restart")
       import :: c_int, c_ptr, c_char, ih_io_state
       type(c_ptr), value :: ctx
       character(kind=c_char), intent(in) :: prefix(*)
       type(ih_io_state), intent(in) :: st
     end function ih_io_write_restart

     ! read_restart (src/mod_inout.f90:371-491)
     integer(c_int) function ih_io_read_restart(ctx, prefix, restart, st) bind(C, name="ih_io_read_restart")
       import :: c_int, c_ptr, c_char, ih_io_state
       type(c_ptr), value :: ctx
       character(kind=c_char), intent(in) :: prefix(*)
       integer(c_int), value :: restart
       type(ih_io_state), intent(inout) :: st
     end function ih_io_read_restart

     ! write_fields (src/mod_inout.f90:78-314): sizes = (write_large, write_med, write_small)
     integer(c_int) function ih_io_write_fields(ctx, prefix, sizes, strides, st) bind(C, name="ih_io_write_fields")
       import :: c_int, c_ptr, c_char, ih_io_state
       type(c_ptr), value :: ctx
       character(kind=c_char), intent(in) :: prefix(*)
       integer(c_int), intent(in) :: sizes(3), strides(9)
       type(ih_io_state), intent(inout) :: st
     end function ih_io_write_fields

     ! write_xdmf_xml (src/mod_inout.f90:3113-3159)
     integer(c_int) function ih_io_write_xdmf(ctx, prefix, sizes, strides, st) bind(C, name="ih_io_write_xdmf")
       import :: c_int, c_ptr, c_char, ih_io_state
       type(c_ptr), value :: ctx
       character(kind=c_char), intent(in) :: prefix(*)
       integer(c_int), intent(in) :: sizes(3), strides(9)
       type(ih_io_state), intent(in) :: st
     end function ih_io_write_xdmf
  end interface
end module impact_io_c
