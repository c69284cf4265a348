/* HDF5 field output, restart and XDMF of the IMPACT projection step
 * (src/mod_inout.f90), as libimpact_io.so over libimpact_hip.so and HDF5 1.10.
 *
 * One block (NB1 = NB2 = NB3 = 1, M = N): file offsets are zero and a file's
 * extent is the written component's index range, as filespace_props gives
 * them for the first block (src/mod_inout.f90:2713-2815).  Files are written
 * serially (the reference's collective MPI-IO property lists have no role on
 * one rank).  Names are the reference's: "<prefix><filename>.h5" with prefix a
 * directory ("out/") or "".
 *
 * Reference interface each entry point replaces:
 *   ih_io_write_hdf      write_hdf        src/mod_inout.f90:503-694
 *   ih_io_read_hdf       read2_hdf        src/mod_inout.f90:1452-1720
 *   ih_io_write_restart  write_restart    src/mod_inout.f90:326-359
 *   ih_io_read_restart   read_restart     src/mod_inout.f90:371-491
 *   ih_io_write_fields   write_fields     src/mod_inout.f90:78-314 (3-D branch)
 *   ih_io_write_xdmf     write_xdmf_xml   src/mod_inout.f90:3113-3159, 3162-3290
 */
#ifndef IMPACT_IO_H
#define IMPACT_IO_H

#include "impact_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

#define IH_IO_ABI_VERSION 1

/* The run state write_hdf stores as attributes of every dataset
 * (src/mod_inout.f90:594-656) and read2_hdf / read_restart read back.
 * Logicals are 0/1 (write_hdf_infoLOG stores them as integers). */
typedef struct {
    int version; /* IH_IO_ABI_VERSION */
    int restart, write_count, timestep;
    int write_out_vect, write_out_scal, new_dtime;
    int Euler_yes, twostep_yes, mapping_yes, upwind_yes;
    double time, dtime;
    double time_out_vect, time_out_scal, dtime_out_vect, dtime_out_scal;
    double Re, CFL, thetaL, epsU;
    double p_end_systole_in, p_end_systole_out, p_end_systole_out_2;
    double U_ref, L_ref;
} ih_io_state;

int ih_io_abi_version(void);
/* Text of the last error of this thread (the I/O layer's, else the case library's). */
int ih_io_last_error(char *buf, size_t n);

/* write_hdf(filename, dsetname, S, N, vel_dir, stride, f): field f over the
 * range of vel_dir (0: pressure S?p..N?p; m = 1..3: velocity SmB..NmB) with
 * stride, one dataset, the run state as its attributes, and the coordinates of
 * the written points as datasets VectorX/Y/Z.  Rejected as the reference
 * rejects them: a stride other than 1 along vel_dir, and (N-1) not divisible
 * by the stride. */
int ih_io_write_hdf(ih_case *c, const char *filename, const char *dsetname, ih_field f, int vel_dir,
                    const int stride[3], const ih_io_state *st);

/* read2_hdf(filename, dsetname, S, N, vel_dir, f): the dataset into f over
 * the overlap of its S?w..M?w attributes with the range of vel_dir.  The
 * attributes read2_hdf reads (time, dtime, time_out_vect, time_out_scal,
 * timestep, write_count, write_out_vect, write_out_scal, new_dtime) go to *st
 * when st is not NULL; its other members are left as they are. */
int ih_io_read_hdf(ih_case *c, const char *filename, const char *dsetname, ih_field f, int vel_dir,
                   ih_io_state *st);

/* write_restart: <prefix>velX_restart<nnn>.h5 (dataset velX_restart), velY,
 * velZ from IH_VEL1..3 over SmB..NmB, pre_restart<nnn>.h5 from IH_PRE over
 * S?p..N?p; nnn = st->restart in three digits. */
int ih_io_write_restart(ih_case *c, const char *prefix, const ih_io_state *st);

/* read_restart: the four files of restart number `restart` into IH_VEL1..3 and
 * IH_PRE, the run state into *st.  On entry st->dtime_out_vect and
 * st->dtime_out_scal are the configured output intervals and st->dtime_out*
 * of the files are replaced by them: time_out_* = (time_out_* - old dtime_out_*)
 * + dtime_out_*, moved to time + dtime when not after time
 * (src/mod_inout.f90:452-487). */
int ih_io_read_restart(ih_case *c, const char *prefix, int restart, ih_io_state *st);

/* write_fields, 3-D branch, without lambda2 and forces: for each size q in
 * {large, med, small} with sizes[q] != 0 and stride strides[3q..3q+2],
 * velX/Y/Z_<size>_<ccccc>.h5 (velocity m after exchange(m, m, vel_m),
 * interpolated to the pressure grid with cIup/cIvp/cIwp) and pre_<size>_<ccccc>.h5,
 * ccccc = st->write_count.  IH_TMP1 is the interpolation's scratch field (the
 * reference uses nl), and IH_VEL1..3 leave with refreshed ghost layers.  Then,
 * as write_fields ends: write_count + 1, time_out_vect + dtime_out_vect,
 * write_out_vect = 0. */
int ih_io_write_fields(ih_case *c, const char *prefix, const int sizes[3], const int strides[9], ih_io_state *st);

/* write_xdmf_xml: <prefix>IMPACT_xdmf3d_<size>_<ccccc>.xmf for each requested
 * size, pointing at the files write_fields writes with the same write_count
 * (call it first, as the time loop does, src/mod_timeint.f90:199-201).  The
 * unscaled form (scale_output_yes false), no lambda2 or force attributes. */
int ih_io_write_xdmf(ih_case *c, const char *prefix, const int sizes[3], const int strides[9],
                     const ih_io_state *st);

#ifdef __cplusplus
}
#endif

#endif /* IMPACT_IO_H */
