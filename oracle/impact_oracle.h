/*
 * CPU oracle for the IMPACT projection-step hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product links, loads or calls this
 * code; only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * use it, and only as the checker (or as the timed CPU baseline).
 *
 * It restates, loop by loop and in the same floating-point operation order,
 * the reference Fortran CPU path (single MPI block, walls or periodic BCs):
 *   - index ranges        src/mod_setup.f90:465-700, 828-867
 *   - coordinates         src/usr_geometry.f90:93-157, src/mod_geometry.f90:37-430,
 *                         src/usr_func.f90:1964-2000 (coord_tanh)
 *   - FD coefficients     src/mod_coeffs.f90:37-120, 200-596, 598-772
 *   - Poisson stencil     src/mod_coeffs.f90:1048-1258 (get_stencil -> cdg)
 *   - ghost fill          src/mod_exchange.f90:47-294, 565-817, 1273-1480
 *   - stencils            src/mod_diff.f90:130-215, 309-414, 549-762, 784-1424,
 *                         1436-1472, 1567-1648, 1740-1825
 *   - JOR relaxation      src/mod_laplace.f90:466-597, src/jacobi_solver_host_v4.f90:283-420
 *   - BiCGstab            src/mod_solvers.f90:1509-1935, 2319-2515;
 *                         operator src/mod_laplace.f90:38-197
 *   - multigrid V cycle   src/mod_solvers.f90:353-1141, src/mod_laplace.f90:200-790,
 *                         src/mod_coeffs.f90:1683-2035 (orc_mg.c)
 *   - projection / RK3    src/mod_solvers.f90:64-346, src/mod_rhs.f90:35-420,
 *                         src/mod_timeint.f90:275-377, src/mod_lib.f90:507-575
 *   - IBM volume mask     src/mod_sharp_interface_ibm.f90:585-640
 *   - CT labels, forcing  src/usr_func.f90:187-555, 574-651, 753-760, 893-945 (orc_ct.c)
 *
 * Arrays are Fortran-ordered with halo 3: phi(-3:N1+3,-3:N2+3,-3:N3+3),
 * element (i,j,k) at (i+3) + (j+3)*nx + (k+3)*nx*ny, nx = N1+7, ny = N2+7.
 * Coefficient arrays c(lo:hi,0:N) at (ii-lo) + (hi-lo+1)*i.
 * cdg?(-1:1,1:N) at (ii+1) + 3*(i-1).
 *
 * Compiled with -ffp-contract=off so no multiply-add is fused.
 */
#ifndef IMPACT_ORACLE_H
#define IMPACT_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_case orc_case;

/* Create a single-block case.  N[d] grid points per direction (IMPACT N1..N3),
 * bc[2*d+0/1] lower/upper BC (-1 periodic, 1 wall), L[d] extents,
 * stretch[d] = 0 for an equidistant grid, else the tanh factor ii0L=ii0U. */
orc_case *orc_case_create(const int N[3], const int bc[6], const double L[3], const double stretch[3]);
void orc_case_destroy(orc_case *c);

/* Field size in doubles, (N1+7)(N2+7)(N3+7). */
int64_t orc_field_size(const orc_case *c);

/* Index ranges: out[0..2]=S?p, [3..5]=N?p, then for m=1..3: S1m B..(9 values
 * SB[m][d]), NB (9), SI (9), NI (9), then SNF lower (3) and upper (3). */
void orc_ranges(const orc_case *c, int out[48]);

/* Coordinates, (-3:N+3) each: which = 0 x?p, 1 x?u  (direction d). */
void orc_coords(const orc_case *c, int d, int which, double *out);

/* Coefficient export by name ("cGp1", "cDu1", "cIpu", "cIup", "cNu1U", "cNp2D",
 * "cu11", "cp22", "cdg1", ...).  Returns number of doubles written, or -1. */
int orc_coeff(const orc_case *c, const char *name, double *out);

/* One diff_coeffs_exact table lookup; returns 1 if the table has the entry. */
int orc_exact_table(int abl, int n_coeff, double deltaX, double *cc);

/* ---------------- ghost fill ---------------- */
void orc_exchange(const orc_case *c, int dir, int vel_dir, double *phi);
void orc_exchange_all_all(const orc_case *c, double *vel /* 3 fields, contiguous */);
void orc_exchange_relax(const orc_case *c, double *phi);

/* ---------------- stencils ---------------- */
void orc_divergence2(const orc_case *c, double *vel3, double *div);
void orc_gradient(const orc_case *c, int m, double *phi, double *grad);
void orc_interpolate2_vel_pre(const orc_case *c, int m, double *phi, double *inter);
void orc_interpolate2_pre_vel(const orc_case *c, int m, double *phi, double *inter);
void orc_interpolate_vel(const orc_case *c, double *vel3, double *work3);
/* nonlinear (upwind) into nl3; reads work3 (exchanged), vel3 */
void orc_nonlinear(const orc_case *c, double *vel3, double *work3, double *nl3, double *pp);
void orc_helmholtz_explicit(const orc_case *c, double multL, double *vel3, double *nl3);
void orc_product_div_grad(const orc_case *c, double *phi, double *lap, double *dig);

/* ---------------- reductions ---------------- */
/* mode 0: sequential k,j,i loop order (reference order);
 * mode 1: the fixed two-level tree the device kernels use. */
void orc_set_reduction_mode(orc_case *c, int mode);
/* threads of the OpenMP build's point loops (n > 0 sets them); returns the count in use (1 when serial) */
int orc_set_threads(int n);
double orc_dot(const orc_case *c, const int ss[3], const int nn[3], const double *a, const double *b);

/* ---------------- Poisson ---------------- */
void orc_jor(const orc_case *c, int n_sweeps, double omega, const int ss[3], const int nn[3],
             double *rel, const double *bb, double *comp);
double orc_residual_cdg(const orc_case *c, const int ss[3], const int nn[3], double *rel, const double *bb);
typedef struct {
    int iterations;
    double norm_inf_initial;
    double norm_inf_final;
    int exit_reason; /* 0 converged, 1 max iterations, 2 breakdown */
} orc_bicgstab_stats;
void orc_bicgstab(const orc_case *c, double eps, int n_it_max, const int ss[3], const int nn[3],
                  const double *bb, double *phi, double *work /* 7 fields */, orc_bicgstab_stats *st);
/* Mixed precision (the device's FP32 smoothing; no reference counterpart):
 * BiCGstab in its right-preconditioned form with M = `sweeps` FP32 JOR sweeps
 * (omega) on cdg from zero, and FP32 JOR as defect correction of rel. */
void orc_bicgstab_mixed(const orc_case *c, double eps, int n_it_max, const int ss[3], const int nn[3],
                        const double *bb, double *phi, int sweeps, double omega, orc_bicgstab_stats *st);
void orc_jor_mixed(const orc_case *c, int n_sweeps, double omega, const int ss[3], const int nn[3], double *rel,
                   const double *bb);
/* The same right-preconditioned BiCGstab with the sweeps of M in FP64 (the
 * like-for-like FP64 baseline of orc_bicgstab_mixed; no reference counterpart). */
void orc_bicgstab_jor64(const orc_case *c, double eps, int n_it_max, const int ss[3], const int nn[3],
                        const double *bb, double *phi, int sweeps, double omega, orc_bicgstab_stats *st);

/* ---------------- multigrid V cycle (preconditioner 1) ---------------- */
/* Hierarchy, coefficients and work arrays of every level; NULL for BCs other
 * than walls or periodic pairs.  n_down, n_up, n_bottom: relaxations per level
 * (prog/config.txt:108-110), n_grids_limit (:116); omega is the reference's 0.98. */
typedef struct orc_mg orc_mg;
orc_mg *orc_mg_create(const orc_case *c, int n_grids_limit, int n_down, int n_up, int n_bottom);
void orc_mg_destroy(orc_mg *m);
int orc_mg_grids(const orc_mg *m);
/* NN[3*(g-1)+d] for g = 1..G; returns G */
int orc_mg_levels(const orc_mg *m, int *NN);
/* "cdg1".."cdg3" (3*N), "cR1".. (3*N, levels >= 2), "cI1".. (2*N, levels <= G-1),
 * "xp1".., "xu1".. (N+7): level g's tables; returns the number written or -1 */
int orc_mg_coeff(const orc_mg *m, int g, const char *name, double *out);
/* elements of a level-g array, (N1+7)(N2+7)(N3+7) */
int64_t orc_mg_field_size(const orc_mg *m, int g);
/* multigridV(.TRUE., 1, bb, phi): phi = M bb; bb's edges of wall pairs are averaged in place */
void orc_mg_vcycle(orc_mg *m, double *bb, double *phi);
/* single operations (test hooks): coarse = R (bb - A x) from level g to g+1;
 * fine += P coarse from level g+1 to g; relaxation on level g */
void orc_mg_restrict_residual(orc_mg *m, int g, double *bb, double *x, double *coarse);
void orc_mg_interpolate_add(orc_mg *m, int g, double *coarse, double *fine);
void orc_mg_relax(orc_mg *m, int g, int init, int n_relax, const double *bb, double *rel);
/* Mixed precision (no reference counterpart): the same cycle in binary32 on
 * b32 = (float) bb, phi = (double) result; bb's edges are averaged in FP64 too */
void orc_mg_vcycle_f32(orc_mg *m, double *bb, double *phi);
void orc_bicgstab_mg_f32(const orc_case *c, orc_mg *mg, double eps, int n_it_max, const int ss[3], const int nn[3],
                         const double *bb, double *phi, orc_bicgstab_stats *st);
/* BiCGstab_orig with preconditioner 1 */
void orc_bicgstab_mg(const orc_case *c, orc_mg *mg, double eps, int n_it_max, const int ss[3], const int nn[3],
                     const double *bb, double *phi, orc_bicgstab_stats *st);

/* ---------------- CT label volume and forcing ---------------- */
/* read_vtk_3D's file reading: returns np and the values (malloc'd), or -1 */
int orc_vtk_read(const char *path, int dims[3], double **values);
/* convert_waveform (CT-based): t_out, q_out in reference units; ref = velocity, length, time */
void orc_waveform_convert(int npts, const double *t_ms, const double *q, int M1, double *t_out, double *q_out,
                          double ref[3]);
/* read_vtk_3D's placement (x/y transpose, offsets), inlet / outlet detection and labels
 * 7, 5, 9, 11 into lab (uint8, field layout); info: inlet, outlet, outlet-2 cells, jet
 * voxels, k_start, k_start_2; centre: inlet x, y, outlet x, y.  0 or -1 */
int orc_ct_build(const orc_case *c, const int dims[3], const double *x, int x_offset, int y_offset, int z_offset,
                 double reference_velocity, double reference_length, uint8_t *lab, int info[6], double centre[4]);
/* Q(t) of the converted waveform after the pulse shift; *found = 0 if no interval holds it */
double orc_ct_inflow(int npts, const double *t_conv, const double *q_conv, double time, int *found);
/* apply_CT_forcing (CT branch) on vel (3 fields, contiguous) */
void orc_ct_forcing(const orc_case *c, const uint8_t *lab, int npts, const double *t_conv, const double *q_conv,
                    double time, double *vel);

/* CT state of the two apply_CT_forcing branches (src/usr_func.f90:909-1048) and
 * of the label-driven pressure of the GPU Jacobi solver: the module variables
 * the reference keeps (CT_based_aorta, idealized_aorta, the waveform as read and
 * converted, the reference scales, inlet_voxels_count, k_start_index,
 * p_end_systole_in/out/out_2, pre_inlet/outlet/outlet_2). */
typedef struct {
    int ct_based, idealized;
    int npts;
    const double *t_ms, *q_raw;   /* waveform as read (ms, m/s) */
    const double *t_conv, *q_conv; /* converted (reference units) */
    double ref[3];                 /* reference velocity, length, time */
    int inlet_voxels;              /* inlet_voxels_count */
    int k_start;                   /* k_start_index */
    double wk_state[3];            /* p_end_systole_in, _out, _out_2 */
    double pres[3];                /* pre_inlet, pre_outlet, pre_outlet_2 (nondimensional) */
} orc_ct_state;

/* convert_waveform with the idealized-aorta override of the scales when idealized != 0 */
void orc_waveform_convert_branch(int npts, const double *t_ms, const double *q, int M1, int idealized,
                                 double *t_out, double *q_out, double ref[3]);
/* get_pressure_and_flow_rate_WKM (branch 1 inflow, 2 outflow, 3 outflow 2); state = p_end_systole_* */
void orc_wk_pressure_flow(int branch, double time, const double ref[3], double state[3], double *p, double *q);
/* get_pressure_waveform_echo_WKM */
void orc_wk_echo(int npts, const double *t_ms, const double *q_raw, const double *t_conv, double time,
                 double reference_velocity, double *pre_inlet, double *pre_outlet);
/* apply_CT_forcing, the branches st->ct_based and st->idealized in the reference's order; updates st->pres */
void orc_ct_apply_forcing(const orc_case *c, const uint8_t *lab, orc_ct_state *st, double time, double *vel);
/* set_pressure_BC_kernel as pressure_bc_launcher launches it (plane k_start_index-4, see orc_ct.c) */
void orc_set_pressure_bc(const orc_case *c, const uint8_t *lab, int type1, int type2, int type3, int k_start_index,
                         double p_in, double p_out, double p_out2, double *rel);

/* device_jacobi_solver (src/jacobi_solver_host_v4.f90:5-717): JOR sweeps on
 * [ss, nn] with the stopping test every 20 sweeps; on substep 3 with CT labels
 * (lab != NULL) the Dirichlet pressure before every sweep and before the exit
 * (types 1 and 2, values pres * (aRK + bRK) * dtime).  The sum err_tot follows
 * the case's reduction mode.  Returns 0 when the test stopped it (or err_tot is
 * 0), 1 at max_iter; *sweeps = sweeps done. */
int orc_device_jacobi_solver(const orc_case *c, double omega, const int ss[3], const int nn[3], double *rel,
                             const double *bb, const uint8_t *lab, int k_start, int substep, double dtime,
                             const double pres[3], int max_iter, int *sweeps);

/* ---------------- projection step ---------------- */
typedef struct {
    double Re, dtime;
    int jor_sweeps;
    double jor_omega;
    double bicg_rel_tol;   /* stop when ||r||_inf < rel_tol * ||b||_inf */
    int bicg_max_it;
    int level_pressure;    /* 1: subtract mean pressure after each step */
    int inflow_kind;       /* 0: none, 1: parabola(y,z), 2: pulsatile plug inside mask */
    double inflow_amp;
    const uint8_t *mask3;  /* optional: 3 masks (u,v,w grids), 1 = fluid; NULL = none */
    int jor_fp32;          /* 1: the JOR sweeps as FP32 defect correction */
    int precond_sweeps;    /* > 0: BiCGstab preconditioned by this many FP32 JOR sweeps */
    int precond_poisson;   /* 1: BiCGstab preconditioned by the multigrid V cycle (precond_sweeps 0) */
    int mg_fp32;           /* with precond_poisson 1: the cycle in FP32 (mixed precision) */
    const uint8_t *ct_lab; /* optional: CT labels (orc_ct_build); apply_CT_forcing at the step's end time */
    int ct_npts;
    const double *ct_t, *ct_q; /* converted waveform */
    orc_ct_state *ct_state; /* optional (with ct_lab): both forcing branches, the Windkessel state and pressures */
    int pressure_bc;        /* 1 (with ct_state): set_pressure_BC before each JOR sweep of substep 3 and after the last */
    int precond_fp64;       /* with precond_sweeps > 0: the preconditioner's sweeps in FP64 */
} orc_step_params;

typedef struct {
    int bicg_iterations[3];
    double bicg_final[3];
    double div_max_after;  /* max |div u| after the last substep */
} orc_step_stats;

/* Boundary values bcXY used by rhs_vel: filled from the inflow model at subtime. */
void orc_timestep(orc_case *c, const orc_step_params *p, int step_index, double time,
                  double *vel3, double *pre, double *nl3, orc_step_stats *st);

/* IBM volume mask of a stenosed tube along x (axis (yc,zc), radius profile r(x_i)
 * given per i on the three staggered grids).  mask3 = 3 uint8 fields. */
void orc_tube_mask(const orc_case *c, double yc, double zc, const double *r_u /* (-3:N1+3) at x1u */,
                   const double *r_p /* at x1p */, uint8_t *mask3);
void orc_apply_mask(const orc_case *c, const uint8_t *mask3, double *vel3);
void orc_level_pressure(const orc_case *c, double *pre);
/* inflow amplitude at time t: kind 1 constant amp, kind 2 amp*(1 + 0.5 sin 2pi t + 0.2 sin(4pi t+0.7) + 0.1 sin(6pi t+1.3)) */
double orc_inflow_amplitude(int kind, double amp, double t);
void orc_fill_corners(const orc_case *c, double *phi);

#ifdef __cplusplus
}
#endif
#endif
