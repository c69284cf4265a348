/*
 * impact_hip -- MI355X (gfx950) implementation of IMPACT's projection-step hot
 * path behind a plain C ABI.
 *
 * A case (ih_case) owns one grid block and all of its fields resident in HBM.
 * Fields are addressed by id; host arrays cross the boundary in the reference's
 * Fortran layout phi(-3:N1+3,-3:N2+3,-3:N3+3) (REAL(8), column-major, dense).
 * All work of a case is ordered on the case's own HIP stream.  Calls on distinct
 * cases may run concurrently from different threads; calls on one case must not.
 *
 * Every function returns IH_OK (0) or a negative status; ih_last_error() returns
 * the message of the last failure on the calling thread.
 *
 * Reference interfaces replaced (paths relative to the reference tree):
 *   device_jacobi_solver     src/jacobi_solver_host_v4.f90:5-40   -> ih_device_jacobi_solver, ih_jor
 *   device_gradient          src/grad_host.f90:39-50              -> ih_device_gradient, ih_gradient
 *   device_divergence2       src/divergence2_host.f90:6-20        -> ih_device_divergence2, ih_divergence2
 *   device_Nonlinear_upwind  src/Nonlinear_upwind_host.f90:6-30   -> ih_device_nonlinear_upwind
 *   device_helmholtz         src/helmholtz_explicit_host.f90:6-30 -> ih_device_helmholtz
 *   device_axpby             src/axpby_host.f90:5-30              -> ih_device_axpby
 *   device_interpolate       src/interpolate2_pre_vel_host.f90    -> ih_device_interpolate
 *   device_interpolate2      src/interpolate2_vel_pre_host.f90    -> ih_device_interpolate2
 *   GPU_malloc / GPU_memcopy src/mod_mem_GPU.f90:152-434          -> ih_case_create, ih_case_set_coeffs
 *   GPU_memfree              src/mod_mem_GPU.f90:437-610          -> ih_case_destroy
 *   launcher ABI, e.g. jor_launcher_ src/jor_kernel_launcher.cu:12 (replaced wholesale)
 */
#ifndef IMPACT_HIP_H
#define IMPACT_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define IH_ABI_VERSION 4

enum {
    IH_OK = 0,
    IH_EINVAL = -22,
    IH_ENOMEM = -12,
    IH_EHIP = -1000,
    IH_ESTATE = -1001,
    IH_ENOTCONV = -1002 /* an iteration stopped at its limit without meeting its test (results are kept) */
};

/* Field ids.  Vectors occupy three consecutive ids. */
typedef enum {
    IH_VEL1 = 0, IH_VEL2, IH_VEL3,
    IH_NL1, IH_NL2, IH_NL3,
    IH_RHS1, IH_RHS2, IH_RHS3,
    IH_WORK1, IH_WORK2, IH_WORK3,
    IH_PRE, IH_RES, IH_GPRE, IH_PP, IH_COMP, IH_DIG,
    IH_RR, IH_RH, IH_PK, IH_AP, IH_AR,
    IH_TMP1, IH_TMP2, IH_TMP3, IH_TMP4,
    IH_Z1, IH_Z2,   /* preconditioned search directions of the mixed-precision BiCGstab */
    IH_NFIELDS
} ih_field;

typedef struct ih_case ih_case;

/* Grid of one block: N[d] points per direction (IMPACT N1..N3; for a single
 * block N = M), bc[2d], bc[2d+1] the lower/upper BC: -1 periodic, 1 wall.
 * L[d] extents and stretch[d] (0 equidistant, else the tanh factor) are used
 * only by ih_case_generate_coeffs. */
typedef struct {
    int version;
    int N[3];
    int bc[6];
    double L[3];
    double stretch[3];
} ih_grid_desc;

/* Coefficient arrays exactly as the reference keeps them (mod_vars.f90:280-370):
 * c(lo:hi, 0:N) column-major.  cdg?(-1:1, 1:N) for the finest grid. */
typedef struct {
    const double *cGp[3];   /* cGp1..3   (-2:3, 0:N)  gradient                 */
    const double *cD[3];    /* cDu1,cDv2,cDw3 (-3:2, 0:N) divergence           */
    const double *cIp[3];   /* cIpu,cIpv,cIpw (-2:3, 0:N) pressure -> velocity  */
    const double *cIu[3];   /* cIup,cIvp,cIwp (-3:2, 0:N) velocity -> pressure  */
    const double *cNpD[3];  /* cNp1D..   (-3:3, 0:N)  upwind, pressure grid     */
    const double *cNpU[3];
    const double *cNuD[3];  /* cNu1D,cNv2D,cNw3D (-3:3, 0:N) velocity grid      */
    const double *cNuU[3];
    const double *cp2[3];   /* cp11,cp22,cp33 (-3:3, 0:N)                       */
    const double *cu2[3];   /* cu11,cv22,cw33 (-3:3, 0:N)                       */
    const double *cdg[3];   /* cdg1..3 (-1:1, 1:N)                              */
} ih_coeff_set;

/* ---- library / case lifetime ---------------------------------------------- */
int  ih_abi_version(void);
int  ih_last_error(char *buf, size_t n);
int  ih_device_count(int *count);
int  ih_case_create(const ih_grid_desc *g, int device, ih_case **out);
void ih_case_destroy(ih_case *c);
/* Upload coefficients computed by the host (GPU_memcopy's role). */
int  ih_case_set_coeffs(ih_case *c, const ih_coeff_set *cs);
/* Compute coordinates and coefficients on the host from g->L, g->stretch
 * (FD_coeffs + get_stencil with mapping, src/mod_coeffs.f90:37-120,1048-1258)
 * and upload them. */
int  ih_case_generate_coeffs(ih_case *c);
/* Export one generated coefficient array by reference name ("cGp1", "cdg2", ...);
 * *n receives the element count.  out may be NULL to query the size. */
int  ih_case_get_coeff(ih_case *c, const char *name, double *out, int64_t *n);
/* The grid the case was created with. */
int  ih_case_get_grid(const ih_case *c, ih_grid_desc *out);
/* Coordinates x?p (which = 0) or x?u (which = 1) of direction d, (-3:N+3). */
int  ih_case_get_coords(ih_case *c, int d, int which, double *out);
/* Host-only forms of the two above (no device needed): coefficients and
 * coordinates the library would generate for the grid g. */
int  ih_grid_coeff(const ih_grid_desc *g, const char *name, double *out, int64_t *n);
int  ih_grid_coords(const ih_grid_desc *g, int d, int which, double *out);
/* Index ranges in the layout of the reference's mod_vars (48 ints, see docs). */
int  ih_case_get_ranges(ih_case *c, int out[48]);
/* Elements of one host field, (N1+7)(N2+7)(N3+7). */
int64_t ih_field_elems(const ih_case *c);
/* Bytes of HBM held by the case. */
int64_t ih_case_bytes(const ih_case *c);

/* ---- field transfer (Fortran-dense host arrays) -------------------------- */
int ih_field_upload(ih_case *c, ih_field f, const double *host);
int ih_field_download(ih_case *c, ih_field f, double *host);
int ih_field_fill(ih_case *c, ih_field f, double value);
int ih_field_copy(ih_case *c, ih_field dst, ih_field src);
/* The strided box lo, lo+stride, .. <= hi (reference indices, within -3..N+3)
 * of field f as a dense host array, x fastest: the hyperslab that write_hdf
 * writes and read2_hdf reads (src/mod_inout.f90:570-575, 657-667, 1452-1720).
 * Gathered and scattered on the device: only the box crosses the host link. */
int ih_field_get_box(ih_case *c, ih_field f, const int lo[3], const int hi[3], const int stride[3], double *host);
int ih_field_set_box(ih_case *c, ih_field f, const int lo[3], const int hi[3], const int stride[3],
                     const double *host);
int ih_synchronize(ih_case *c);

/* ---- device-resident operators ------------------------------------------ */
/* exchange(dir, vel_dir, phi): src/mod_exchange.f90:47-294 as an on-device fill. */
int ih_exchange(ih_case *c, int dir, int vel_dir, ih_field f);
int ih_exchange_relax(ih_case *c, ih_field f);
int ih_divergence2(ih_case *c, ih_field u, ih_field v, ih_field w, ih_field div);
int ih_gradient(ih_case *c, int m, ih_field phi, ih_field grad);
int ih_interpolate2_vel_pre(ih_case *c, int m, ih_field phi, ih_field inter);
int ih_interpolate2_pre_vel(ih_case *c, int m, ih_field phi, ih_field inter);
/* work1..3 = interpolated velocities + ghost fill (interpolate_vel). */
int ih_interpolate_vel(ih_case *c);
/* component_id 1..9 as in the reference (1,4,7 assign; others accumulate). */
int ih_nonlinear_upwind(ih_case *c, int component_id, ih_field pp, ih_field vel_c, ih_field nl_c);
int ih_helmholtz(ih_case *c, int component, double multL, ih_field vel_c, ih_field nl_c);
int ih_axpby(ih_case *c, double a, double b, ih_field x, ih_field y, ih_field z);
int ih_product_div_grad(ih_case *c, ih_field phi, ih_field lap);
int ih_dot(ih_case *c, const int ss[3], const int nn[3], ih_field a, ih_field b, double *out);
/* ih_dot without allocation.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Same result, bit for
 * bit, as ih_dot; the block partials and the sum live in one fixed 2 MiB device
 * array per GPU instead of the case's partial buffer, so a box may need at most
 * 262112 reduction partials (16 x 16 columns x 32-plane chunks) or the call
 * fails with IH_EINVAL.  Calls that overlap on one GPU race on that array. */
int ih_dot_noalloc(ih_case *c, const int ss[3], const int nn[3], ih_field a, ih_field b, double *out);

typedef struct {
    int iterations;
    double norm_inf_initial;
    double norm_inf_final;
    int exit_reason; /* 0 converged, 1 max iterations, 2 breakdown */
} ih_bicgstab_stats;

/* JOR on the compact cdg operator, out of place, n_sweeps sweeps. */
int ih_jor(ih_case *c, int n_sweeps, double omega, const int ss[3], const int nn[3], ih_field rel, ih_field bb);
/* max |bb - A_cdg rel| over [ss, nn]. */
int ih_residual_cdg(ih_case *c, const int ss[3], const int nn[3], ih_field rel, ih_field bb, double *out);
/* BiCGstab_orig (preconditioner 0) on product_div_grad; stop at ||r||_inf < eps. */
int ih_bicgstab(ih_case *c, double eps, int n_it_max, const int ss[3], const int nn[3], ih_field bb, ih_field phi,
                ih_bicgstab_stats *st);
/* BiCGstab_orig in its right-preconditioned form (z1 = M pp, Ap = A z1, ...,
 * z2 = M rr, Ar = A z2, phi += omega z2 + alpha z1; src/mod_solvers.f90:1684-1693,
 * 1750-1752, 1895-1899) with M = `sweeps` JOR sweeps (omega) on the cdg operator
 * in FP32 from zero.  Residuals, inner products and the solution update stay
 * FP64; only the smoothing inside M is FP32.  The reference preconditions with
 * a multigrid V cycle instead. */
int ih_bicgstab_mixed(ih_case *c, double eps, int n_it_max, const int ss[3], const int nn[3], ih_field bb,
                      ih_field phi, int sweeps, double omega, ih_bicgstab_stats *st);
/* No counterpart in the reference: the same right-preconditioned BiCGstab with
 * M = `sweeps` JOR sweeps from zero in FP64 -- the like-for-like FP64 baseline
 * of ih_bicgstab_mixed, which isolates what the FP32 sweeps gain.  Uses TMP3,
 * TMP4 and DIG as work fields. */
int ih_bicgstab_jor64(ih_case *c, double eps, int n_it_max, const int ss[3], const int nn[3], ih_field bb,
                      ih_field phi, int sweeps, double omega, ih_bicgstab_stats *st);

/* BiCGstab_orig with the reference's preconditioner argument: 0 none (as
 * ih_bicgstab), 1 the multigrid V cycle multigridV (src/mod_solvers.f90:1684-1693,
 * 1750-1752; z1 = M pp, z2 = M rr).  The cycle averages the edges of wall pairs
 * of pp and rr in place, as the reference's restriction does. */
int ih_bicgstab_precond(ih_case *c, double eps, int n_it_max, const int ss[3], const int nn[3], ih_field bb,
                        ih_field phi, int preconditioner, ih_bicgstab_stats *st);

/* ---- multigrid V cycle (preconditioner 1) --------------------------------- */
/* Levels and relaxations per level (prog/config.txt:108-116: 15, 4, 4, 4, the
 * defaults).  The cycle follows multigridV (src/mod_solvers.f90:456-585) with
 * Jacobi smoothing (omega 0.98, src/mod_laplace.f90:470) on each level's
 * pressure range; the hierarchy halves a direction while N-1 is even and N >= 5
 * (src/mod_setup.f90:96-143).  Walls or periodic pairs in every direction. */
int ih_mg_configure(ih_case *c, int n_grids_limit, int n_relax_down, int n_relax_up, int n_relax_bottom);
/* number of levels and their sizes NN[3*(g-1)+d], g = 1..n_grids (NN holds 45 ints) */
int ih_mg_levels(ih_case *c, int *n_grids, int *NN);
/* level tables: "cdg1".."cdg3" (3N), "cR1".."cR3" (3N, levels >= 2),
 * "cI1".."cI3" (2N, levels below the coarsest); out may be NULL to query n */
int ih_mg_get_coeff(ih_case *c, int level, const char *name, double *out, int64_t *n);
/* phi = M bb, one V cycle multigridV(.TRUE., 1, bb, phi); bb's edges of wall pairs are averaged */
int ih_multigrid_v(ih_case *c, ih_field bb, ih_field phi);
/* Mixed precision (no counterpart in the reference): the same cycle in FP32 on
 * (float) bb, phi = (double) of its result; bb's edges are averaged in FP64
 * as the FP64 cycle does.  ih_bicgstab_mg_f32 is BiCGstab_orig with it as M;
 * residuals, inner products and updates stay FP64. */
int ih_multigrid_v_f32(ih_case *c, ih_field bb, ih_field phi);
int ih_bicgstab_mg_f32(ih_case *c, double eps, int n_it_max, const int ss[3], const int nn[3], ih_field bb,
                       ih_field phi, ih_bicgstab_stats *st);

int ih_fill_corners(ih_case *c, ih_field f);
int ih_level_pressure(ih_case *c, ih_field pre);

/* ---- immersed-boundary mask ---------------------------------------------- */
/* Stenosed tube along x: radius per x-index on the u grid (r_u) and on the
 * pressure grid (r_p), arrays (-3:N1+3).  The mask is built on the host and
 * uploaded (uint8, 1 = fluid). */
int ih_set_tube_mask(ih_case *c, double yc, double zc, const double *r_u, const double *r_p);
int ih_get_mask(ih_case *c, uint8_t *host3);
int ih_apply_mask(ih_case *c);

/* ---- CT geometry (src/usr_func.f90:187-651, 909-945) --------------------- */
typedef struct {
    int inlet_cells, outlet_cells, outlet2_cells; /* fluid cells found on the image's first / last plane */
    int inlet_jet_voxels;                         /* inlet cells within the jet radius (label 5 at k_start-1) */
    int k_start, k_start2;                        /* z_offset + 1, z_offset + image depth */
    double inlet_centre[2], outlet_centre[2];     /* mean x, y of the inlet / outlet cells */
    double reference_velocity, reference_length, reference_time;
} ih_ct_info;
/* read_vtk_3D: a legacy ASCII VTK label image placed at the offsets with the
 * reference's x/y transpose (image axis 1 -> grid y, axis 2 -> grid x), the
 * rest of the grid label 2; inlet cells (label 1, left half of the first
 * plane) get ten planes of 7 below them, 5 within the jet radius; outlet cells
 * (right half) ten planes of 9, the last plane's cells ten planes of 11 above.
 * The echo waveform (t in ms, velocity in m/s) is converted to reference units
 * as convert_waveform does for a CT image.  Labels live on the device. */
int ih_ct_load_vtk(ih_case *c, const char *path, int x_offset, int y_offset, int z_offset, const double *wave_t_ms,
                   const double *wave_q, int npts, ih_ct_info *info);
/* labels as uint8 (N1+7)(N2+7)(N3+7), Fortran element order */
int ih_ct_get_labels(ih_case *c, uint8_t *host);
/* the inflow velocity the forcing uses at `time` (reference units); found = 0 outside the waveform */
int ih_ct_inflow(ih_case *c, double time, double *q, int *found);
/* The two branches of apply_CT_forcing (src/usr_func.f90:909-1048): the
 * reference's switches CT_based_aorta and idealized_aorta (src/mod_vars.f90:260;
 * its compile-time default is the idealized aorta alone).  Set before
 * ih_ct_load_vtk: the idealized aorta also overrides the reference scales of
 * convert_waveform (0.2 m, 2.7 m/s, src/usr_func.f90:637-640), which the jet
 * radius of the labels depends on.  The library's default is the CT branch alone. */
int ih_ct_configure(ih_case *c, int ct_based, int idealized);
/* apply_CT_forcing on VEL1..3 at `time` (after the time update, as
 * src/mod_timeint.f90:405-416 calls it), the configured branches in the
 * reference's order:
 *   CT-based:  get_pressure_waveform_echo_WKM (:654-750) sets pre_inlet and
 *              pre_outlet; solid and label-7 points to rest; label 5 to
 *              (0, 0, Q(t)) of the waveform;
 *   idealized: get_pressure_and_flow_rate_WKM (:762-855) for the inflow (label
 *              5 present: pre_inlet, q_inlet), the outflow (label 9: pre_outlet)
 *              and the second outflow (label 11: pre_outlet_2); solid and label 7
 *              to rest; label 5 to (0, 0, q_inlet / (inlet_voxels_count h^2)),
 *              h = L1/(M1-1).
 * The Windkessel keeps p_end_systole_* between calls as the reference's module
 * variables do (zero after ih_ct_load_vtk). */
int ih_ct_forcing(ih_case *c, double time);
/* The module state the forcing leaves: pres = pre_inlet, pre_outlet,
 * pre_outlet_2 (nondimensional) and wk = p_end_systole_in, _out, _out_2 (mmHg).
 * set: for a host that computes them itself, or a restart. */
int ih_ct_get_state(ih_case *c, double pres[3], double wk[3]);
int ih_ct_set_state(ih_case *c, const double pres[3], const double wk[3]);
/* set_pressure_BC_kernel (src/set_pressure_BC_kernel.cu:1-42) on field rel, as
 * pressure_bc_launcher (src/set_pressure_BC_kernel_launcher.cu:12-57) launches
 * it: over the points its grid reaches (1..b*t per direction, b = (N-1)/8 blocks
 * of t = (N-1)/b threads; none when a direction has no block or a block would
 * exceed 1024 threads), label-5 points (type1 = 1) and label-9 points
 * (type2 = 1) of plane k_start_index - 4 take p_in and p_out, label-11 points
 * (type3 = 1) take p_out2.  The kernel compares its storage offset k+4 with
 * k_start_index - 1; the array starts at index -3, so that is Fortran plane
 * k_start_index - 4.  Needs the CT labels. */
int ih_set_pressure_bc(ih_case *c, ih_field rel, int type1, int type2, int type3, double p_in, double p_out,
                       double p_out2);
/* The module variables device_jacobi_solver reads besides its arguments: the RK
 * substep (1..3; 0 = outside a time step) and dtime.  With CT labels loaded,
 * substep 3 applies set_pressure_BC (types 1, 1, 0) with the pressures of the
 * last forcing times (aRK + bRK) dtime before every sweep
 * (src/jacobi_solver_host_v4.f90:302-305). */
int ih_case_set_substep(ih_case *c, int substep, double dtime);

/* ---- projection time step ------------------------------------------------ */
typedef struct {
    double Re, dtime;
    int jor_sweeps;
    double jor_omega;
    double bicg_rel_tol;
    int bicg_max_it;
    int level_pressure;
    int inflow_kind; /* 0 none, 1 parabola 16y(1-y)z(1-z), 2 pulsatile plug in the mask */
    double inflow_amp;
    int use_mask;
    int jor_fp32;   /* 1: JOR sweeps in FP32 (residual and update kept FP64) */
    int precond_sweeps; /* > 0: BiCGstab right-preconditioned by this many FP32 JOR sweeps */
    int precond_poisson; /* with precond_sweeps 0: BiCGstab's preconditioner, 0 none, 1 multigrid V cycle */
    int mg_fp32;         /* with precond_poisson 1: the cycle in FP32 (an FP64-cycle solve redoes a failure) */
    int use_ct;          /* 1: apply_CT_forcing with the loaded CT labels at the step's end time */
    int pressure_bc;     /* 1 (with CT labels, FP64 sweeps): set_pressure_BC before every JOR sweep of
                            substep 3 and after the last, as device_jacobi_solver does */
    int precond_fp64;    /* with precond_sweeps > 0: the preconditioner's sweeps in FP64 (ih_bicgstab_jor64) */
} ih_step_params;

typedef struct {
    int bicg_iterations[3];
    double bicg_final[3];
    double div_max_after;
} ih_step_stats;

/* One time step (3 RK substeps) of the fields VEL, PRE, NL in place.
 * stats may be NULL; when given, div_max_after costs one extra divergence.
 * div_max_after is max |div u| over the pressure range at the end of the step:
 * after the mask and the CT forcing, which set velocities in the wall and the
 * jet and so leave divergence at those interfaces; without them it is what
 * the projection leaves. */
int ih_timestep(ih_case *c, const ih_step_params *p, double time, ih_step_stats *st);

/* ---- reference-signature entry points (host arrays in, host arrays out) --- */
/* device_jacobi_solver: JOR sweeps with the reference's stopping test every 20
 * sweeps (|log(err_tot_old / err_tot)| against its threshold, err_tot the sum
 * of |rel - rel_old| over the box in the fixed reduction order) and, on
 * substep 3 with CT labels, the Dirichlet pressure (ih_case_set_substep).  An
 * err_tot of 0 (a fixed point) stops.  After 200000 sweeps it stops with
 * IH_ENOTCONV, rel holding the last iterate. */
int ih_device_jacobi_solver(ih_case *c, int grid_id, double omega, int SS1, int SS2, int SS3, int NN1, int NN2,
                            int NN3, double *rel, double *bb);
int ih_device_gradient(ih_case *c, int dimension_id, double *phi, double *grad);
int ih_device_divergence2(ih_case *c, double *phi3, double *div);
int ih_device_nonlinear_upwind(ih_case *c, int component_id, double *pp, double *vel_component,
                               double *nl_component);
int ih_device_helmholtz(ih_case *c, int component_id, double multL, double *vel_component,
                        double *helm_component);
int ih_device_axpby(ih_case *c, double a, double b, double *x, double *y, double *z);
int ih_device_interpolate(ih_case *c, int dimension_id, double *phi, double *inter);
int ih_device_interpolate2(ih_case *c, int dimension_id, double *phi, double *inter);

/* Residency registry of the calls above.  Without it every host array crosses
 * PCIe both ways on every call, as in the reference's *_host.f90 wrappers
 * (e.g. src/helmholtz_explicit_host.f90:204-214, 386).  ih_bind_field makes
 * field f the device mirror of the host array `host` (the reference's module
 * array, identified by its address): a call given that array uploads it only
 * when the host copy is newer, works on f in place, and leaves a result on the
 * device until ih_sync copies it back.  The host copy starts as the newer one.
 * host = NULL unbinds (a device copy newer than the host's is copied back
 * first).  TMP1..TMP4 and COMP are the calls' staging fields and cannot be
 * bound; one host array binds to one field.  Resident calls (ih_jor,
 * ih_timestep, ...) on a bound field do not change its state: follow them with
 * ih_sync as needed. */
int ih_bind_field(ih_case *c, ih_field f, double *host);
enum { IH_SYNC_TO_HOST = 0, IH_SYNC_TO_DEVICE = 1 };
/* IH_SYNC_TO_HOST: copy f back if the device copy is newer; IH_SYNC_TO_DEVICE:
 * the host changed the array, upload it now.  f must be bound. */
int ih_sync(ih_case *c, ih_field f, int dir);
int ih_sync_all(ih_case *c, int dir);
/* Whole-field transfers of the case so far: uploads, downloads, bytes up, bytes down. */
int ih_transfer_stats(ih_case *c, int64_t out[4]);

/* ---- batched relaxation of many cases ------------------------------------ */
/* Many cases (uneven grids, one device) relaxed together: each sweep is one
 * launch over the tiles of all cases.  No reference counterpart: the reference
 * relaxes one block per call (device_jacobi_solver); this serves callers that
 * advance many patient geometries at once. */
typedef struct ih_batch ih_batch;
int  ih_batch_create(ih_case *const *cases, int n, ih_batch **out);
void ih_batch_destroy(ih_batch *b);
/* Times every launch shape -- threads per tile 128/256/512, z chunk 8/12/16/24/32,
 * layout 0 (one flat grid over all tiles) or 1 (one row of blocks per case),
 * group budget all/192/96/48 MB (ih_batch_set_groups) --
 * on `sweeps` sweeps without changing PRE, keeps the fastest and reports it
 * with its time per sweep. */
int  ih_batch_autotune(ih_batch *b, int sweeps, int *tpb, int *zchunk, int *layout, double *ms_per_sweep);
/* Sets the launch shape directly (tpb 128/256/512, zchunk 1..64, layout 0/1). */
int  ih_batch_set_shape(ih_batch *b, int tpb, int zchunk, int layout);
/* Cases are relaxed in groups (consecutive cases whose rel, out and b fields
 * together fit `bytes`; 0: one group), all sweeps of a group before the next,
 * so that a group can be swept from the device's last-level cache.  Results
 * do not depend on the grouping.  ih_batch_autotune also chooses it. */
int  ih_batch_set_groups(ih_batch *b, int64_t bytes);
int  ih_batch_get_groups(ih_batch *b, int64_t *bytes, int *count);
/* n_sweeps JOR sweeps (omega) of PRE on RES in every case, each bit-identical to
 * ih_jor on that case.  fp32 != 0: FP32 defect correction instead (r = RES -
 * A_cdg PRE in FP64, n_sweeps FP32 sweeps on A e = r from e = 0, PRE += e).
 * use_graph != 0: the sweeps are captured once into a hipGraph and replayed. */
int  ih_batch_jor(ih_batch *b, int n_sweeps, double omega, int fp32, int use_graph);
/* One projection time step of every case (ih_timestep's sequence,
 * src/mod_timeint.f90:275-377, per case), each case's fields bit-identical to
 * its own ih_timestep.  The stencil phases run case by case on the batch's
 * stream; the Poisson solves are shared launches: the JOR sweeps of all cases
 * (as ih_batch_jor), then BiCGstab -- plain, or right-preconditioned by
 * p->precond_sweeps FP32 JOR sweeps -- with every kernel of an iteration one
 * launch over the tiles of all cases, each case stopping on its own test, the
 * iterations replayed from a captured hipGraph.  st: one entry per case, or
 * NULL.  Not available here: precond_poisson 1, pressure_bc, precond_fp64. */
int  ih_batch_timestep(ih_batch *b, const ih_step_params *p, double time, ih_step_stats *st);

/* ---- serving: a trace of simulation requests ------------------------------ */
/* No reference counterpart (the reference runs one simulation per MPI job).
 * Requests arrive over time; each is a stenosed tube (walls, L = 2 x 1 x 1,
 * tanh 1.5 in y and z, axis at (0.5, 0.5), cosine constriction of half-length
 * 0.4 at x = 1) started from plug flow at rest pressure and advanced `steps`
 * time steps with `step`.  Requests are admitted first come first served while
 * the resident cases fit in hbm_budget and fewer than max_active run; each runs
 * on its own case and stream from its own host thread, so the kernels of the
 * admitted requests overlap on the device.  Every result is independent of the
 * mix it ran in: checksum_* are <u,u> and <p,p> over the pressure range, in the
 * fixed reduction order. */
typedef struct {
    double arrival_s;   /* after the start of the trace (scaled by time_scale) */
    int N[3];           /* grid points */
    int steps;
    double radius;      /* tube radius */
    double stenosis;    /* fraction of the area removed at the throat, 0..1 */
} ih_request;

typedef struct {
    double admit_s, start_s, finish_s; /* after the start of the trace */
    double latency_s;                  /* finish - arrival */
    int bicg_iterations;               /* over all steps */
    double checksum_vel, checksum_pre;
    int status;                        /* IH_OK, IH_ENOMEM if it can never fit, or the failing call's status */
} ih_request_result;

typedef struct {
    int device;
    int64_t hbm_budget;  /* bytes the resident cases may hold together */
    int max_active;      /* simulations running at once */
    double time_scale;   /* arrival times are multiplied by this (1 = as given) */
    ih_step_params step;
    int progress_every;  /* > 0: one line on stderr after every that many finished requests */
} ih_serve_config;

typedef struct {
    double makespan_s;         /* first arrival to last finish (set-up excluded) */
    int64_t peak_bytes;        /* largest sum of the running requests' slot bytes */
    int peak_active;
    double setup_s;            /* building the slot pool, before the trace starts */
    int64_t pool_bytes;        /* HBM held by the pool (at most hbm_budget) */
    int pool_slots;
    double admit_wait_max_s;   /* arrival to admission */
    double admit_wait_mean_s;
} ih_serve_summary;

/* HBM a case of grid g holds (fields, FP32 buffers, mask, tables). */
int64_t ih_grid_bytes(const ih_grid_desc *g);
/* The same for a case that runs steps with p: the multigrid levels and their
 * FP32 copies are added when p selects preconditioner 1. */
int64_t ih_grid_bytes_step(const ih_grid_desc *g, const ih_step_params *p);
/* Runs the trace to completion; out[q] belongs to req[q].  Before the trace
 * starts, the trace's grids get a pool of resident slots (cases with their
 * coefficients, multigrid levels and tuned launch shapes), one more per size
 * class in turn while the pool fits hbm_budget, at most max_active or the
 * class's request count per class; serving then allocates nothing.  A request
 * is admitted when a slot of its size is free and fewer than max_active run;
 * a request whose size got no slot fails with IH_ENOMEM.  Returns IH_OK or
 * the status of a failed request. */
int ih_serve(const ih_serve_config *cfg, const ih_request *req, int n, ih_request_result *out,
             ih_serve_summary *summary);

/* ---- timing of the case's stream (for benchmarks) ------------------------ */
int ih_timer_start(ih_case *c);
int ih_timer_stop(ih_case *c, double *ms);
/* Per-kernel-family launch statistics collected with HIP events when enabled. */
int ih_kernel_timing_enable(ih_case *c, int on);
int ih_kernel_timing_get(ih_case *c, const char *family, double *total_ms, int64_t *launches);
/* `launches` launches of one kernel back to back between two events on the
 * case's stream: "jor" (an FP64 sweep over the pressure range), "jor_f32" (an
 * FP32 sweep of the mixed-precision preconditioner, on the case's float
 * fields), "apply" (the fused operator alone), the five kernels of a BiCGstab
 * iteration -- "p_update", "apply_w" (operator with <rh, Ap>), "s_update",
 * "apply_self" (operator with <r, Ar>, <Ar, Ar>), "x_update" -- on the case's
 * Krylov fields (their results are meaningless), and the substep stencils
 * "nl_helm_rhs0/1/2", "divergence2", "interp_vel_pre3", "grad_correct0". */
int ih_bench_kernel(ih_case *c, const char *name, int launches, double *ms_per_launch);

#ifdef __cplusplus
}
#endif
#endif
