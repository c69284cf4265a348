/*
 * Oracle (test infrastructure only): CT label-volume ingest, inlet/outlet
 * detection and the CT forcing of the velocity.  Build with -ffp-contract=off.
 *
 *   read_vtk_3D            src/usr_func.f90:187-555 (legacy VTK, ASCII, list-directed
 *                          reads; labels placed with the x/y transpose and offsets)
 *   read_velocity_waveform src/usr_func.f90:574-596
 *   convert_waveform       src/usr_func.f90:599-651 (CT-based branch)
 *   get_pulse_number_and_length, get_shifted_time_in_pulse  src/usr_func.f90:753-760, 893-903
 *   apply_CT_forcing       src/usr_func.f90:909-945 (CT_based_aorta branch)
 *
 * Labels are stored as uint8 on the case's index space (-3..N+3 per
 * direction, the field layout); the reference keeps them as REAL(8) values
 * (0, 1, 2 from the image, 5, 7, 9, 11 written here), compared with == and
 * with 0.5 / 2.0 thresholds, so integer labels 0..255 carry the same
 * decisions.  Points of the velocity ranges outside 1..N (the wall plane 0 of
 * the component-1 range) read label 2 (no forcing); the reference reads
 * outside its ct_grid(1:M1,...) there.
 */
#include "orc_internal.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define LBL(c, lab, i, j, k) (lab)[ORC_IDX(c, i, j, k)]

/* read one whitespace-separated token */
static int token(FILE *f, char *buf, size_t n)
{
    int ch;
    size_t k = 0;
    do {
        ch = fgetc(f);
    } while (ch != EOF && (ch == ' ' || ch == '\t' || ch == '\n' || ch == '\r' || ch == ','));
    if (ch == EOF) return 0;
    while (ch != EOF && !(ch == ' ' || ch == '\t' || ch == '\n' || ch == '\r' || ch == ',')) {
        if (k + 1 < n) buf[k++] = (char)ch;
        ch = fgetc(f);
    }
    buf[k] = 0;
    return 1;
}

static int skip_line(FILE *f)
{
    int ch;
    while ((ch = fgetc(f)) != EOF && ch != '\n') {
    }
    return ch != EOF;
}

/* read_vtk_3D: the header as the reference reads it (two lines, the format,
 * DATASET, DIMENSIONS d1 d2 d3, SPACING, ORIGIN, POINT_DATA np, SCALARS, LOOKUP_TABLE),
 * then np values.  Returns the number of values (dims out) or -1. */
int orc_vtk_read(const char *path, int dims[3], double **values)
{
    FILE *f = fopen(path, "r");
    if (!f) return -1;
    char t[256];
    int ok = skip_line(f) && skip_line(f);           /* version, title */
    ok = ok && token(f, t, sizeof t) && !strcmp(t, "ASCII");
    ok = ok && token(f, t, sizeof t) && token(f, t, sizeof t); /* DATASET <type> */
    ok = ok && token(f, t, sizeof t) && !strcmp(t, "DIMENSIONS");
    for (int d = 0; ok && d < 3; ++d) ok = token(f, t, sizeof t) && (dims[d] = atoi(t)) > 0;
    ok = ok && token(f, t, sizeof t) && !strcmp(t, "SPACING") && skip_line(f);
    ok = ok && token(f, t, sizeof t) && !strcmp(t, "ORIGIN") && skip_line(f);
    int np = 0;
    ok = ok && token(f, t, sizeof t) && !strcmp(t, "POINT_DATA") && token(f, t, sizeof t) && (np = atoi(t)) > 0;
    ok = ok && token(f, t, sizeof t) && !strcmp(t, "SCALARS") && skip_line(f);
    ok = ok && token(f, t, sizeof t) && !strcmp(t, "LOOKUP_TABLE") && skip_line(f);
    ok = ok && np == dims[0] * dims[1] * dims[2];
    if (!ok) {
        fclose(f);
        return -1;
    }
    double *x = malloc(sizeof(double) * (size_t)np);
    for (int q = 0; q < np; ++q) {
        if (!token(f, t, sizeof t)) {
            free(x);
            fclose(f);
            return -1;
        }
        x[q] = strtod(t, NULL);
    }
    fclose(f);
    *values = x;
    return np;
}

/* convert_waveform (CT-based: CT pixel width 0.0004 m, upsampling 1): reference
 * velocity = max Q / 1.0, reference length = 0.0004 / (1/M1), reference time =
 * length / velocity; times from ms to reference units, velocities by the
 * reference velocity */
void orc_waveform_convert(int npts, const double *t_ms, const double *q, int M1, double *t_out, double *q_out,
                          double ref[3])
{
    const double CT_pixel_width = 0.0004;
    const double upsampling_factor = 1.0;
    const double CT_pixel_width_star = (1.0 / M1) * upsampling_factor;
    const double maximum_inflow_velocity_star = 1.0;
    double vmax = q[0];
    for (int i = 1; i < npts; ++i) vmax = q[i] > vmax ? q[i] : vmax;
    const double reference_velocity = vmax / maximum_inflow_velocity_star;
    const double reference_length = CT_pixel_width / CT_pixel_width_star;
    const double reference_time = reference_length / reference_velocity;
    for (int i = 0; i < npts; ++i) {
        t_out[i] = t_ms[i] / (1000.0 * reference_time);
        q_out[i] = q[i] / reference_velocity;
    }
    ref[0] = reference_velocity;
    ref[1] = reference_length;
    ref[2] = reference_time;
}

/* read_vtk_3D's placement, detection and labelling.  info: [0] inlet count,
 * [1] outlet count, [2] outlet-2 count, [3] inlet voxels within the jet,
 * [4] k_start_index, [5] k_start_index_2; centre: inlet x, y, outlet x, y.
 * Returns 0, or -1 if a detected list is empty or the labels leave the grid. */
int orc_ct_build(const orc_case *c, const int dims[3], const double *x, int x_offset, int y_offset, int z_offset,
                 double reference_velocity, double reference_length, uint8_t *lab, int info[6], double centre[4])
{
    const int M1 = c->N[0], M2 = c->N[1], M3 = c->N[2];
    if (dims[1] + y_offset > M1 || dims[0] + x_offset > M2 || dims[2] + z_offset > M3 || z_offset < 10 ||
        z_offset + dims[2] + 10 > M3)
        return -1;
    /* grid_set = 2 over 1..M, then the image with the x/y transpose */
    memset(lab, 0, (size_t)c->ntot);
    for (int k = 1; k <= M3; ++k)
        for (int j = 1; j <= M2; ++j)
            for (int i = 1; i <= M1; ++i) LBL(c, lab, i, j, k) = 2;
    for (int i = 1; i <= dims[0]; ++i)
        for (int j = 1; j <= dims[1]; ++j)
            for (int k = 1; k <= dims[2]; ++k) {
                const double v = x[(size_t)(k - 1) * dims[1] * dims[0] + (size_t)(j - 1) * dims[0] + (i - 1)];
                if (!(v >= 0.0 && v <= 255.0 && v == floor(v))) return -1;
                LBL(c, lab, j + y_offset, i + x_offset, k + z_offset) = (uint8_t)v;
            }
    const double *y1p = c->xp[0], *y2p = c->xp[1];
#define FLUID(i, j, k) (LBL(c, lab, i, j, k) >= 1 && LBL(c, lab, i, j, k) < 2) /* 0.5 <= v < 2.0 */
    const int half = M1 / 2; /* floor(M1/2.) */
    const int ks = z_offset + 1, ks2 = z_offset + dims[2];
    int ni = 0, no = 0, no2 = 0;
    double xi = 0.0, yi = 0.0, xo = 0.0, yo = 0.0;
    int *ii = malloc(sizeof(int) * (size_t)M1 * M2 * 2), *io = malloc(sizeof(int) * (size_t)M1 * M2 * 2),
        *io2 = malloc(sizeof(int) * (size_t)M1 * M2 * 2);
    for (int j = 1; j <= M2; ++j)
        for (int i = 1; i <= half; ++i)
            if (FLUID(i, j, ks)) {
                xi = xi + y1p[i + 3];
                yi = yi + y2p[j + 3];
                ii[2 * ni] = i;
                ii[2 * ni + 1] = j;
                ++ni;
            }
    for (int j = 1; j <= M2; ++j)
        for (int i = half; i <= M1; ++i)
            if (FLUID(i, j, ks)) {
                xo = xo + y1p[i + 3];
                yo = yo + y2p[j + 3];
                io[2 * no] = i;
                io[2 * no + 1] = j;
                ++no;
            }
    for (int j = 1; j <= M2; ++j)
        for (int i = 1; i <= M1; ++i)
            if (FLUID(i, j, ks2)) {
                io2[2 * no2] = i;
                io2[2 * no2 + 1] = j;
                ++no2;
            }
#undef FLUID
    if (ni == 0 || no == 0) {
        free(ii);
        free(io);
        free(io2);
        return -1;
    }
    xi = xi / ni;
    yi = yi / ni;
    xo = xo / no;
    yo = yo / no;
    const double R_inlet = sqrt(0.0004 / (4.0 * atan(1.0) * reference_velocity)) / reference_length;
    int inlet_voxels = 0;
    for (int q = 0; q < ni; ++q)
        for (int k = -10; k <= -1; ++k) {
            const int i = ii[2 * q], j = ii[2 * q + 1];
            LBL(c, lab, i, j, ks + k) = 7;
            const double dx = y1p[i + 3] - xi, dy = y2p[j + 3] - yi;
            if (sqrt(dx * dx + dy * dy) <= R_inlet) {
                LBL(c, lab, i, j, ks + k) = 5;
                if (k == -1) ++inlet_voxels;
            }
        }
    for (int q = 0; q < no; ++q)
        for (int k = -10; k <= -1; ++k) LBL(c, lab, io[2 * q], io[2 * q + 1], ks + k) = 9;
    for (int q = 0; q < no2; ++q)
        for (int k = 1; k <= 10; ++k) LBL(c, lab, io2[2 * q], io2[2 * q + 1], ks2 + k) = 11;
    info[0] = ni;
    info[1] = no;
    info[2] = no2;
    info[3] = inlet_voxels;
    info[4] = ks;
    info[5] = ks2;
    centre[0] = xi;
    centre[1] = yi;
    centre[2] = xo;
    centre[3] = yo;
    free(ii);
    free(io);
    free(io2);
    return 0;
}

/* the inflow velocity at time (reference units): the pulse-shifted time, then
 * the last interval [t_i, t_i+1] of the converted waveform holding it gives Q_i */
double orc_ct_inflow(int npts, const double *t_conv, const double *q_conv, double time, int *found)
{
    const double pulse_length = t_conv[npts - 1];
    const double pulse_number = floor(time / pulse_length) + 1.0;
    const double st = time - (pulse_number - 1.0) * pulse_length;
    double v = 0.0;
    *found = 0;
    for (int i = 0; i + 1 < npts; ++i)
        if (st >= t_conv[i] && st <= t_conv[i + 1]) {
            v = q_conv[i] * 1.0; /* flow_rate_correction_factor */
            *found = 1;
        }
    return v;
}

/* apply_CT_forcing, CT-based branch: over the component-1 velocity range
 * S11B..N11B, S21B..N21B, S31B..N31B, all three components: solid (< 0.5) and
 * label 7 set to 0; label 5 gets (0, 0, Q(t)) */
void orc_ct_forcing(const orc_case *c, const uint8_t *lab, int npts, const double *t_conv, const double *q_conv,
                    double time, double *vel)
{
    const int64_t n = c->ntot;
    int found = 0;
    const double q = orc_ct_inflow(npts, t_conv, q_conv, time, &found);
    for (int k = c->SB[0][2]; k <= c->NB[0][2]; ++k)
        for (int j = c->SB[0][1]; j <= c->NB[0][1]; ++j)
            for (int i = c->SB[0][0]; i <= c->NB[0][0]; ++i) {
                const int inside = i >= 1 && i <= c->N[0] && j >= 1 && j <= c->N[1] && k >= 1 && k <= c->N[2];
                const int L = inside ? LBL(c, lab, i, j, k) : 2;
                const int64_t x = ORC_IDX(c, i, j, k);
                if (L < 1 || L == 7) {
                    vel[x] = 0.0;
                    vel[n + x] = 0.0;
                    vel[2 * n + x] = 0.0;
                } else if (L == 5 && found) {
                    vel[2 * n + x] = q;
                    vel[n + x] = 0.0;
                    vel[x] = 0.0;
                }
            }
}

/* ------------------------------------------------------------------------ */
/* Windkessel models, both branches of apply_CT_forcing and the label-driven */
/* Dirichlet pressure of the GPU Jacobi solver.                              */
/* ------------------------------------------------------------------------ */

/* blood properties, src/mod_vars.f90:235-236 */
#define ORC_BLOOD_NU 0.0000035
#define ORC_BLOOD_RHO 1060.

/* convert_waveform (src/usr_func.f90:599-651) with the idealized-aorta
 * override of the reference scales (:637-640: length 0.2, velocity 2.7) when
 * idealized != 0; the converted waveform uses the final scales. */
void orc_waveform_convert_branch(int npts, const double *t_ms, const double *q, int M1, int idealized,
                                 double *t_out, double *q_out, double ref[3])
{
    const double CT_pixel_width = 0.0004;
    const double upsampling_factor = 1.0;
    const double CT_pixel_width_star = (1.0 / M1) * upsampling_factor;
    const double maximum_inflow_velocity_star = 1.0;
    double vmax = q[0];
    for (int i = 1; i < npts; ++i) vmax = q[i] > vmax ? q[i] : vmax;
    double reference_velocity = vmax / maximum_inflow_velocity_star;
    double reference_length = CT_pixel_width / CT_pixel_width_star;
    if (idealized) {
        reference_length = 0.2;
        reference_velocity = 2.7;
    }
    const double reference_time = reference_length / reference_velocity;
    for (int i = 0; i < npts; ++i) {
        t_out[i] = t_ms[i] / (1000.0 * reference_time);
        q_out[i] = q[i] / reference_velocity;
    }
    ref[0] = reference_velocity;
    ref[1] = reference_length;
    ref[2] = reference_time;
}

/* get_pressure_and_flow_rate_WKM (src/usr_func.f90:762-855): the reference's
 * closed form of a two-element Windkessel driven by a half-sine systolic flow
 * (T_beat 1 s, T_systole 0.4 s, p_diastole 80 mmHg), statement by statement.
 * The physiological time is time * reference_time (not wrapped into a beat).
 * state[branch-1] is p_end_systole_in / _out / _out_2: written on every call in
 * systole, read in diastole.  p and q are returned nondimensional. */
void orc_wk_pressure_flow(int branch, double time, const double ref[3], double state[3], double *p_out, double *q_out)
{
    const double pi = 4. * atan(1.);
    const double T_beat = 1.;
    const double T_systole = 0.4 * T_beat;
    const double p_diastole = 80.;
    double CWK, RWK, Q_peak_systole;
    if (branch == 1) {
        CWK = 1.5496 / 1000000.;
        RWK = 0.9758 * 1000000.;
        Q_peak_systole = 0.0004;
    } else if (branch == 2) {
        CWK = 1.5 / 1000000.;
        RWK = 1.12 * 1000000.;
        Q_peak_systole = 0.0004 * 0.85;
    } else {
        CWK = 0.059 / 1000000.;
        RWK = 19.68 * 1000000.;
        Q_peak_systole = 0.0004 * 0.05;
    }
    const double reference_velocity = ref[0], reference_length = ref[1], reference_time = ref[2];
    const double time_phys = time * reference_time;
    double p, q;
    if (time_phys <= T_systole) {
        const double den = pow(T_systole, 2.) + pow(CWK, 2.) * pow(pi, 2.) * pow(RWK, 2.);
        p = (p_diastole + Q_peak_systole * T_systole * RWK * CWK * pi * RWK / den) * exp(-(1. * time_phys) / (RWK * CWK)) -
            (T_systole * Q_peak_systole * RWK * (CWK * pi * RWK * cos(pi * time_phys / T_systole)) -
             T_systole * sin(pi * time_phys / T_systole)) / den;
        q = Q_peak_systole * sin(pi * (time_phys) / T_systole);
        state[branch - 1] = p;
    } else {
        p = state[branch - 1] * exp(-(1. * (time_phys - T_systole)) / (RWK * CWK));
        q = 0.;
    }
    p = p / (ORC_BLOOD_RHO * pow(reference_velocity, 2.)) * 133.322;
    q = q / (pow(reference_length, 3.) / reference_time);
    *p_out = p;
    *q_out = q;
}

/* get_pressure_waveform_echo_WKM (src/usr_func.f90:654-750): from 80 mmHg, one
 * step p = p (1 - dt/(C R)) - dt/C q_in for every waveform interval i (1..n-1)
 * whose converted times hold the pulse-shifted time, dt the interval in s and
 * q_in = Q_i * q_max / max Q; inlet (R 0.78e6, C 2.166e-6, q_max 4e-4) and
 * outlet (R 0.898e6, C 2.066e-6, q_max 0.85 * 4e-4); nondimensional. */
void orc_wk_echo(int npts, const double *t_ms, const double *q_raw, const double *t_conv, double time,
                 double reference_velocity, double *pre_inlet, double *pre_outlet)
{
    const double pulse_length = t_conv[npts - 1];
    const double pulse_number = floor(time / pulse_length) + 1.0;
    const double shifted = time - (pulse_number - 1.0) * pulse_length;
    const double cardiac_output_factor = 1.0;
    double vel_max = q_raw[0];
    for (int i = 1; i < npts; ++i) vel_max = q_raw[i] > vel_max ? q_raw[i] : vel_max;
    for (int side = 0; side < 2; ++side) {
        const double R = side == 0 ? 0.78 * 1000000. : 0.898 * 1000000.;
        const double Cw = side == 0 ? 2.166 / 1000000. : 2.066 / 1000000.;
        const double q_max = side == 0 ? 0.0004 * cardiac_output_factor : 0.0004 * 0.85 * cardiac_output_factor;
        const double inlet_area = q_max / vel_max;
        double pre = 80.;
        for (int i = 0; i + 1 < npts; ++i)
            if (shifted >= t_conv[i] && shifted <= t_conv[i + 1]) {
                const double dt = t_ms[i + 1] / 1000. - t_ms[i] / 1000.;
                const double q_in = q_raw[i] * inlet_area;
                pre = pre * (1. - dt / (Cw * R)) - dt / Cw * q_in;
            }
        pre = pre / (ORC_BLOOD_RHO * pow(reference_velocity, 2.)) * 133.322;
        if (side == 0)
            *pre_inlet = pre;
        else
            *pre_outlet = pre;
    }
}

/* apply_CT_forcing (src/usr_func.f90:909-1048), both branches in the
 * reference's order over the component-1 velocity range.  CT-based branch
 * (st->ct_based): the echo Windkessel pressures, then solid / label 7 to zero
 * and label 5 to (0, 0, Q(t)).  Idealized-aorta branch (st->idealized): solid /
 * label 7 to zero; label 5 to (0, 0, q_inlet / (inlet_voxels_count h^2)),
 * h = L1/(M1-1), with the inflow Windkessel's pressure and flow; labels 9 and
 * 11 set the outflow Windkesssel pressures (each model evaluated when a point
 * of its label lies in the range, as the per-point calls of the reference). */
void orc_ct_apply_forcing(const orc_case *c, const uint8_t *lab, orc_ct_state *st, double time, double *vel)
{
    const int64_t n = c->ntot;
    if (st->ct_based) {
        orc_wk_echo(st->npts, st->t_ms, st->q_raw, st->t_conv, time, st->ref[0], &st->pres[0], &st->pres[1]);
        orc_ct_forcing(c, lab, st->npts, st->t_conv, st->q_conv, time, vel);
    }
    if (!st->idealized) return;
    int has5 = 0, has9 = 0, has11 = 0;
    for (int k = c->SB[0][2]; k <= c->NB[0][2]; ++k)
        for (int j = c->SB[0][1]; j <= c->NB[0][1]; ++j)
            for (int i = c->SB[0][0]; i <= c->NB[0][0]; ++i) {
                const int inside = i >= 1 && i <= c->N[0] && j >= 1 && j <= c->N[1] && k >= 1 && k <= c->N[2];
                const int L = inside ? LBL(c, lab, i, j, k) : 2;
                has5 |= L == 5;
                has9 |= L == 9;
                has11 |= L == 11;
            }
    double q_inlet = 0.0, q_dummy;
    if (has5) orc_wk_pressure_flow(1, time, st->ref, st->wk_state, &st->pres[0], &q_inlet);
    if (has9) orc_wk_pressure_flow(2, time, st->ref, st->wk_state, &st->pres[1], &q_dummy);
    if (has11) orc_wk_pressure_flow(3, time, st->ref, st->wk_state, &st->pres[2], &q_dummy);
    const double h = c->L[0] / (double)(c->N[0] - 1);
    const double w_in = q_inlet / ((double)st->inlet_voxels * (h * h));
    for (int k = c->SB[0][2]; k <= c->NB[0][2]; ++k)
        for (int j = c->SB[0][1]; j <= c->NB[0][1]; ++j)
            for (int i = c->SB[0][0]; i <= c->NB[0][0]; ++i) {
                const int inside = i >= 1 && i <= c->N[0] && j >= 1 && j <= c->N[1] && k >= 1 && k <= c->N[2];
                const int L = inside ? LBL(c, lab, i, j, k) : 2;
                const int64_t x = ORC_IDX(c, i, j, k);
                if (L < 1 || L == 7) {
                    vel[x] = 0.0;
                    vel[n + x] = 0.0;
                    vel[2 * n + x] = 0.0;
                } else if (L == 5) {
                    vel[2 * n + x] = w_in;
                    vel[n + x] = 0.;
                    vel[x] = 0.;
                }
            }
}

/* set_pressure_BC_kernel (src/set_pressure_BC_kernel.cu:1-42) as
 * pressure_bc_launcher launches it (src/set_pressure_BC_kernel_launcher.cu:16-57):
 * blocks (n-1)/8 and threads (n-1)/blocks per direction (n = N+7 the padded
 * extent), so thread (i, j, k) reaches Fortran point (i+1, j+1, k+1) for
 * i < blocks*threads; no point is written when a direction has no block or a
 * block would exceed 1024 threads (the launch fails).  The kernel compares the
 * storage offset k+4 with k_start_index-1; the array starts at index -3, so
 * offset k+4 is Fortran k+1 and the written plane is k_start_index-4.
 * Label 5 (type1) and label 9 (type2) points of that plane get p_in and p_out;
 * label 11 points (type3) anywhere in the covered box get p_out2. */
void orc_set_pressure_bc(const orc_case *c, const uint8_t *lab, int type1, int type2, int type3, int k_start_index,
                         double p_in, double p_out, double p_out2, double *rel)
{
    int cover[3], threads = 1;
    for (int d = 0; d < 3; ++d) {
        const int m = c->N[d] - 1; /* (N+7) - 8 */
        const int blocks = m / 8;
        if (blocks < 1) return;
        const int t = m / blocks;
        threads *= t;
        cover[d] = blocks * t;
    }
    if (threads > 1024) return;
    for (int k = 1; k <= cover[2]; ++k)
        for (int j = 1; j <= cover[1]; ++j)
            for (int i = 1; i <= cover[0]; ++i) {
                const int L = LBL(c, lab, i, j, k);
                const int64_t x = ORC_IDX(c, i, j, k);
                const int plane = (k + 3) == k_start_index - 1;
                if (type1 == 1 && L == 5 && plane) rel[x] = p_in;
                if (type2 == 1 && L == 9 && plane) rel[x] = p_out;
                if (type3 == 1 && L == 11) rel[x] = p_out2;
            }
}
