/*
 * Oracle (test infrastructure only): Poisson solvers and the projection step.
 * Build with -ffp-contract=off.
 */
#include "orc_internal.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define A(p, i, j, k) (p)[ORC_IDX(c, i, j, k)]

/* RK3 coefficients, src/mod_vars.f90:774-775 (evaluated in double precision,
 * as the reference's CPU builds do with -fdefault-real-8). */
static const double aRK[3] = {8.0 / 15.0, 5.0 / 12.0, 3.0 / 4.0};
static const double bRK[3] = {0.0, -17.0 / 60.0, -5.0 / 12.0};

/* ------------------------------------------------------------------------ */
/* JOR on the compact cdg operator: src/mod_laplace.f90:514-518 (out of      */
/* place into comp, then copied back over the relaxed range), ghost layer    */
/* refreshed before every sweep as src/jacobi_solver_host_v4.f90:298 does.   */
/* ------------------------------------------------------------------------ */
void orc_jor(const orc_case *c, int n_sweeps, double omega, const int ss[3], const int nn[3], double *rel,
             const double *bb, double *comp)
{
    const double *g1 = c->cdg[0], *g2 = c->cdg[1], *g3 = c->cdg[2];
    for (int r = 0; r < n_sweeps; ++r) {
        orc_exchange_relax(c, rel);
        ORC_PAR
        for (int k = ss[2]; k <= nn[2]; ++k)
            for (int j = ss[1]; j <= nn[1]; ++j)
                for (int i = ss[0]; i <= nn[0]; ++i) {
                    double t = A(bb, i, j, k);
                    t = t - CDG(g1, -1, i) * A(rel, i - 1, j, k);
                    t = t - CDG(g1, 1, i) * A(rel, i + 1, j, k);
                    t = t - CDG(g2, -1, j) * A(rel, i, j - 1, k);
                    t = t - CDG(g2, 1, j) * A(rel, i, j + 1, k);
                    t = t - CDG(g3, -1, k) * A(rel, i, j, k - 1);
                    t = t - CDG(g3, 1, k) * A(rel, i, j, k + 1);
                    const double diag = (CDG(g1, 0, i) + CDG(g2, 0, j)) + CDG(g3, 0, k);
                    A(comp, i, j, k) = omega * t / diag + (1.0 - omega) * A(rel, i, j, k);
                }
        ORC_PAR
        for (int k = ss[2]; k <= nn[2]; ++k)
            for (int j = ss[1]; j <= nn[1]; ++j)
                for (int i = ss[0]; i <= nn[0]; ++i) A(rel, i, j, k) = A(comp, i, j, k);
    }
    orc_exchange_relax(c, rel);
}

typedef struct {
    const double *a, *b;
} absdiff_ctx;
static double absdiff_pt(const void *ctx, int64_t idx)
{
    const absdiff_ctx *d = (const absdiff_ctx *)ctx;
    return fabs(d->a[idx] - d->b[idx]);
}

/* the Dirichlet pressure of device_jacobi_solver's substep 3
 * (src/jacobi_solver_host_v4.f90:302-305): type flags 1, 1, 0 and the
 * pressures times (aRK + bRK) * dtime */
static void jacobi_pressure_bc(const orc_case *c, const uint8_t *lab, int k_start, int substep, double dtime,
                               const double pres[3], double *rel)
{
    if (!lab || substep != 3) return;
    const double f = aRK[substep - 1] + bRK[substep - 1];
    orc_set_pressure_bc(c, lab, 1, 1, 0, k_start, pres[0] * f * dtime, pres[1] * f * dtime, pres[2] * f * dtime,
                        rel);
}

/* device_jacobi_solver (src/jacobi_solver_host_v4.f90:5-717), in the CPU
 * relaxation's out-of-place form: each trip fills the relaxation ghosts,
 * applies the substep-3 pressure, and every 20th trip (iter mod 20 == 1)
 * compares err_tot = sum |rel - rel_old| over [ss, nn] with the previous one:
 * |log(err_tot_old / err_tot)| <= 20 mode_j 8/3 0.8 3.1415^2 (1/M1^2 + 1/M2^2 +
 * 1/M3^2) stops (:357-414); otherwise rel_old = rel and one sweep follows.
 * rel_old and err_tot_old are undefined at the reference's entry, so its first
 * test compares with undefined values; here the first test only records them.
 * A later err_tot of 0 (a fixed point, where the reference's log is undefined
 * and it would never stop) stops. */
int orc_device_jacobi_solver(const orc_case *c, double omega, const int ss[3], const int nn[3], double *rel,
                             const double *bb, const uint8_t *lab, int k_start, int substep, double dtime,
                             const double pres[3], int max_iter, int *sweeps)
{
    const int64_t n = c->ntot;
    double *rel_old = calloc((size_t)n, sizeof(double));
    double *comp = calloc((size_t)n, sizeof(double));
    const int mode_j = 20;
    const int M1 = c->N[0], M2 = c->N[1], M3 = c->N[2];
    const double thresh = 20. * mode_j * 8. * 1. / 3.0 * 0.8 * 3.1415 * 3.1415 *
                          (1. / (double)(M1 * M1) + 1. / (double)(M2 * M2) + 1. / (double)(M3 * M3));
    double err_tot_old = 0.0;
    int done = 0, status = 1;
    for (int iter = 1; iter <= max_iter; ++iter) {
        orc_exchange_relax(c, rel);
        jacobi_pressure_bc(c, lab, k_start, substep, dtime, pres, rel);
        if (iter % mode_j == 1) {
            orc_exchange_relax(c, rel);
            const absdiff_ctx d = {rel, rel_old};
            const double err_tot = orc_reduce_sum(c, ss, nn, absdiff_pt, &d);
            if (iter > 1 && (err_tot == 0.0 || fabs(log(err_tot_old / err_tot)) <= thresh)) {
                status = 0;
                break;
            }
            memcpy(rel_old, rel, sizeof(double) * (size_t)n);
            err_tot_old = err_tot;
        }
        orc_jor(c, 1, omega, ss, nn, rel, bb, comp);
        ++done;
    }
    if (status == 1) jacobi_pressure_bc(c, lab, k_start, substep, dtime, pres, rel);
    orc_exchange_relax(c, rel);
    *sweeps = done;
    free(rel_old);
    free(comp);
    return status;
}

/* max |bb - A_cdg rel|: src/jacobi_solver_host_v4.f90:357-371 */
double orc_residual_cdg(const orc_case *c, const int ss[3], const int nn[3], double *rel, const double *bb)
{
    const double *g1 = c->cdg[0], *g2 = c->cdg[1], *g3 = c->cdg[2];
    orc_exchange_relax(c, rel);
    double m = 0.0;
    for (int k = ss[2]; k <= nn[2]; ++k)
        for (int j = ss[1]; j <= nn[1]; ++j)
            for (int i = ss[0]; i <= nn[0]; ++i) {
                double t = A(bb, i, j, k);
                t = t - CDG(g1, -1, i) * A(rel, i - 1, j, k);
                t = t - CDG(g1, 1, i) * A(rel, i + 1, j, k);
                t = t - CDG(g2, -1, j) * A(rel, i, j - 1, k);
                t = t - CDG(g2, 1, j) * A(rel, i, j + 1, k);
                t = t - CDG(g3, -1, k) * A(rel, i, j, k - 1);
                t = t - CDG(g3, 1, k) * A(rel, i, j, k + 1);
                t = t - ((CDG(g1, 0, i) + CDG(g2, 0, j)) + CDG(g3, 0, k)) * A(rel, i, j, k);
                const double a = fabs(t);
                m = (a > m || a != a) ? a : m; /* a NaN wins */
            }
    return m;
}

/* ------------------------------------------------------------------------ */
/* Mixed precision (no reference counterpart; the device's FP32 smoothing):  */
/* the JOR loop above in binary32, with cdg, omega and the right-hand side   */
/* rounded to float, out of place, periodic neighbours read across the       */
/* range (the values exchange_relax would place in the ghost layer).  e and  */
/* tmp are float fields whose ghost cells hold 0; returns the buffer that    */
/* holds the result.                                                         */
/* ------------------------------------------------------------------------ */
static float *jor_f32(const orc_case *c, int n_sweeps, float omega, const int ss[3], const int nn[3], float *e,
                      float *tmp, const float *b)
{
    const double *g1 = c->cdg[0], *g2 = c->cdg[1], *g3 = c->cdg[2];
    const int p0 = c->bc[0][0] == -1, p1 = c->bc[1][0] == -1, p2 = c->bc[2][0] == -1;
    for (int r = 0; r < n_sweeps; ++r) {
        ORC_PAR
        for (int k = ss[2]; k <= nn[2]; ++k)
            for (int j = ss[1]; j <= nn[1]; ++j)
                for (int i = ss[0]; i <= nn[0]; ++i) {
                    const int im = (p0 && i == c->Sp[0]) ? c->Np[0] : i - 1;
                    const int ip = (p0 && i == c->Np[0]) ? c->Sp[0] : i + 1;
                    const int jm = (p1 && j == c->Sp[1]) ? c->Np[1] : j - 1;
                    const int jp = (p1 && j == c->Np[1]) ? c->Sp[1] : j + 1;
                    const int km = (p2 && k == c->Sp[2]) ? c->Np[2] : k - 1;
                    const int kp = (p2 && k == c->Np[2]) ? c->Sp[2] : k + 1;
                    float t = A(b, i, j, k);
                    t = t - (float)CDG(g1, -1, i) * A(e, im, j, k);
                    t = t - (float)CDG(g1, 1, i) * A(e, ip, j, k);
                    t = t - (float)CDG(g2, -1, j) * A(e, i, jm, k);
                    t = t - (float)CDG(g2, 1, j) * A(e, i, jp, k);
                    t = t - (float)CDG(g3, -1, k) * A(e, i, j, km);
                    t = t - (float)CDG(g3, 1, k) * A(e, i, j, kp);
                    const float diag = ((float)CDG(g1, 0, i) + (float)CDG(g2, 0, j)) + (float)CDG(g3, 0, k);
                    A(tmp, i, j, k) = omega * t / diag + (1.0f - omega) * A(e, i, j, k);
                }
        float *sw = e;
        e = tmp;
        tmp = sw;
    }
    return e;
}

/* defect correction: r = bb - A_cdg rel (FP64), A e = r by FP32 sweeps from
 * e = 0, rel += e (FP64); ghost layer of rel refreshed before and after */
void orc_jor_mixed(const orc_case *c, int n_sweeps, double omega, const int ss[3], const int nn[3], double *rel,
                   const double *bb)
{
    const double *g1 = c->cdg[0], *g2 = c->cdg[1], *g3 = c->cdg[2];
    float *b = calloc((size_t)c->ntot, sizeof(float));
    float *e = calloc((size_t)c->ntot, sizeof(float));
    float *tmp = calloc((size_t)c->ntot, sizeof(float));
    orc_exchange_relax(c, rel);
    for (int k = ss[2]; k <= nn[2]; ++k)
        for (int j = ss[1]; j <= nn[1]; ++j)
            for (int i = ss[0]; i <= nn[0]; ++i) {
                double t = A(bb, i, j, k);
                t = t - CDG(g1, -1, i) * A(rel, i - 1, j, k);
                t = t - CDG(g1, 1, i) * A(rel, i + 1, j, k);
                t = t - CDG(g2, -1, j) * A(rel, i, j - 1, k);
                t = t - CDG(g2, 1, j) * A(rel, i, j + 1, k);
                t = t - CDG(g3, -1, k) * A(rel, i, j, k - 1);
                t = t - CDG(g3, 1, k) * A(rel, i, j, k + 1);
                t = t - ((CDG(g1, 0, i) + CDG(g2, 0, j)) + CDG(g3, 0, k)) * A(rel, i, j, k);
                A(b, i, j, k) = (float)t;
            }
    const float *res = jor_f32(c, n_sweeps, (float)omega, ss, nn, e, tmp, b);
    for (int k = ss[2]; k <= nn[2]; ++k)
        for (int j = ss[1]; j <= nn[1]; ++j)
            for (int i = ss[0]; i <= nn[0]; ++i) A(rel, i, j, k) = A(rel, i, j, k) + (double)A(res, i, j, k);
    orc_exchange_relax(c, rel);
    free(b);
    free(e);
    free(tmp);
}

/* z = M v, M = n_sweeps FP32 JOR sweeps on A_cdg e = (float)v from e = 0 */
static void precond(const orc_case *c, int n_sweeps, float omega, const int ss[3], const int nn[3], const double *v,
                    double *z, float *b, float *e, float *tmp)
{
    for (int k = ss[2]; k <= nn[2]; ++k)
        for (int j = ss[1]; j <= nn[1]; ++j)
            for (int i = ss[0]; i <= nn[0]; ++i) A(b, i, j, k) = (float)A(v, i, j, k);
    memset(e, 0, sizeof(float) * (size_t)c->ntot);
    const float *res = jor_f32(c, n_sweeps, omega, ss, nn, e, tmp, b);
    for (int k = ss[2]; k <= nn[2]; ++k)
        for (int j = ss[1]; j <= nn[1]; ++j)
            for (int i = ss[0]; i <= nn[0]; ++i) A(z, i, j, k) = (double)A(res, i, j, k);
}

/* z = M v with the sweeps in FP64 (the like-for-like FP64 form of precond):
 * n_sweeps JOR sweeps on A_cdg e = v from e = 0 */
static void precond64(const orc_case *c, int n_sweeps, double omega, const int ss[3], const int nn[3],
                      const double *v, double *z, double *e, double *tmp)
{
    memset(e, 0, sizeof(double) * (size_t)c->ntot);
    orc_jor(c, n_sweeps, omega, ss, nn, e, v, tmp);
    for (int k = ss[2]; k <= nn[2]; ++k)
        for (int j = ss[1]; j <= nn[1]; ++j)
            for (int i = ss[0]; i <= nn[0]; ++i) A(z, i, j, k) = A(e, i, j, k);
}

static double absmax(const orc_case *c, const int ss[3], const int nn[3], const double *a)
{
    double m = 0.0;
    for (int k = ss[2]; k <= nn[2]; ++k)
        for (int j = ss[1]; j <= nn[1]; ++j)
            for (int i = ss[0]; i <= nn[0]; ++i) {
                const double v = fabs(A(a, i, j, k));
                m = (v > m || v != v) ? v : m; /* a NaN wins */
            }
    return m;
}

/* ------------------------------------------------------------------------ */
/* BiCGstab_orig, preconditioner 0, init_yes = .FALSE., problem_type 2:      */
/* src/mod_solvers.f90:1509-1935 with product_scalar (:2319-2354) and       */
/* status_iteration (:2482-2515).                                            */
/* ------------------------------------------------------------------------ */
static void bicgstab_run(const orc_case *c, double eps, int n_it_max, const int ss[3], const int nn[3],
                         const double *bb, double *phi, double *work, int sweeps, double pre_omega,
                         orc_mg *mg, int mg_fp32, int pre64, orc_bicgstab_stats *st)
{
    const int64_t n = c->ntot;
    double *own = NULL;
    if (!work) {
        own = calloc((size_t)7 * n, sizeof(double));
        work = own;
    }
    double *rr = work, *rh = work + n, *pk = work + 2 * n, *Ap = work + 3 * n, *Ar = work + 4 * n,
           *dig = work + 5 * n;
    /* right preconditioner (src/mod_solvers.f90:1684-1693, 1750-1752, 1895-1899) */
    double *z1 = NULL, *z2 = NULL;
    float *fb = NULL, *fe = NULL, *ft = NULL;
    const float omf = (float)pre_omega;
    const int pre = sweeps > 0 || mg;
    if (pre) {
        z1 = calloc((size_t)n, sizeof(double));
        z2 = calloc((size_t)n, sizeof(double));
    }
    double *de = NULL, *dt = NULL;
    if (sweeps > 0 && pre64) {
        de = calloc((size_t)n, sizeof(double));
        dt = calloc((size_t)n, sizeof(double));
    } else if (sweeps > 0) {
        fb = calloc((size_t)n, sizeof(float));
        fe = calloc((size_t)n, sizeof(float));
        ft = calloc((size_t)n, sizeof(float));
    }

    orc_product_div_grad(c, phi, rr, dig);
    for (int k = ss[2]; k <= nn[2]; ++k)
        for (int j = ss[1]; j <= nn[1]; ++j)
            for (int i = ss[0]; i <= nn[0]; ++i) {
                A(rr, i, j, k) = A(bb, i, j, k) - A(rr, i, j, k);
                A(rh, i, j, k) = A(rr, i, j, k);
            }
    double norm_inf = absmax(c, ss, nn, rr);
    if (st) {
        st->norm_inf_initial = norm_inf;
        st->exit_reason = 0;
    }

    int counter = 0;
    double rhr = 0.0, rhr_prev, alpha = 0.0, beta, omega = 0.0;
    for (;;) {
        int exit_yes = 0;
        if (norm_inf < eps)
            exit_yes = 1;
        if (!exit_yes && counter == n_it_max) {
            exit_yes = 1;
            if (st) st->exit_reason = 1;
        }
        if (counter < 1 && norm_inf != 0.0)
            exit_yes = 0; /* number_poisson == 1: at least one iteration */
        if (!(norm_inf == norm_inf) || norm_inf > 1.0e300) {
            /* non-finite residual: stop as a breakdown (the reference would run to n_it_max) */
            exit_yes = 1;
            if (st) st->exit_reason = 2;
        }
        if (exit_yes)
            break;
        counter = counter + 1;

        rhr_prev = rhr;
        rhr = orc_dot(c, ss, nn, rr, rh);
        if (fabs(rhr) == 0.0) {
            for (int k = ss[2]; k <= nn[2]; ++k)
                for (int j = ss[1]; j <= nn[1]; ++j)
                    for (int i = ss[0]; i <= nn[0]; ++i) A(rh, i, j, k) = A(rr, i, j, k);
            rhr = orc_dot(c, ss, nn, rr, rh);
        }
        if (counter >= 2) {
            if (omega == 0.0 || rhr_prev == 0.0) {
                if (st) st->exit_reason = 2;
                break;
            }
            beta = (alpha / omega) * (rhr / rhr_prev);
            omega = -beta * omega;
            ORC_PAR
            for (int k = ss[2]; k <= nn[2]; ++k)
                for (int j = ss[1]; j <= nn[1]; ++j)
                    for (int i = ss[0]; i <= nn[0]; ++i)
                        A(pk, i, j, k) = A(rr, i, j, k) + beta * A(pk, i, j, k) + omega * A(Ap, i, j, k);
        } else {
            for (int k = ss[2]; k <= nn[2]; ++k)
                for (int j = ss[1]; j <= nn[1]; ++j)
                    for (int i = ss[0]; i <= nn[0]; ++i) A(pk, i, j, k) = A(rr, i, j, k);
        }
        if (mg) {
            /* multigridV(.TRUE.,1,pp,z1) (src/mod_solvers.f90:1688) */
            if (mg_fp32)
                orc_mg_vcycle_f32(mg, pk, z1);
            else
                orc_mg_vcycle(mg, pk, z1);
            orc_product_div_grad(c, z1, Ap, dig);
        } else if (sweeps > 0) {
            if (pre64)
                precond64(c, sweeps, pre_omega, ss, nn, pk, z1, de, dt);
            else
                precond(c, sweeps, omf, ss, nn, pk, z1, fb, fe, ft);
            orc_product_div_grad(c, z1, Ap, dig);
        } else {
            orc_product_div_grad(c, pk, Ap, dig);
        }
        const double rhAp = orc_dot(c, ss, nn, Ap, rh);
        if (fabs(rhAp) == 0.0) {
            if (st) st->exit_reason = 2;
            break;
        }
        alpha = rhr / rhAp;
        ORC_PAR
        for (int k = ss[2]; k <= nn[2]; ++k)
            for (int j = ss[1]; j <= nn[1]; ++j)
                for (int i = ss[0]; i <= nn[0]; ++i) A(rr, i, j, k) = A(rr, i, j, k) - alpha * A(Ap, i, j, k);
        if (mg) {
            if (mg_fp32)
                orc_mg_vcycle_f32(mg, rr, z2);
            else
                orc_mg_vcycle(mg, rr, z2); /* src/mod_solvers.f90:1751 */
            orc_product_div_grad(c, z2, Ar, dig);
        } else if (sweeps > 0) {
            if (pre64)
                precond64(c, sweeps, pre_omega, ss, nn, rr, z2, de, dt);
            else
                precond(c, sweeps, omf, ss, nn, rr, z2, fb, fe, ft);
            orc_product_div_grad(c, z2, Ar, dig);
        } else {
            orc_product_div_grad(c, rr, Ar, dig);
        }
        const double rAr = orc_dot(c, ss, nn, rr, Ar);
        const double ArAr = orc_dot(c, ss, nn, Ar, Ar);
        if (fabs(ArAr) == 0.0) {
            if (st) st->exit_reason = 2;
            break;
        }
        omega = rAr / ArAr;
        norm_inf = 0.0;
        ORC_PAR_MAX_NORM
        for (int k = ss[2]; k <= nn[2]; ++k)
            for (int j = ss[1]; j <= nn[1]; ++j)
                for (int i = ss[0]; i <= nn[0]; ++i) {
                    if (pre)
                        A(phi, i, j, k) = A(phi, i, j, k) + omega * A(z2, i, j, k) + alpha * A(z1, i, j, k);
                    else
                        A(phi, i, j, k) = A(phi, i, j, k) + omega * A(rr, i, j, k) + alpha * A(pk, i, j, k);
                    A(rr, i, j, k) = A(rr, i, j, k) - omega * A(Ar, i, j, k);
                    const double a = fabs(A(rr, i, j, k));
                    norm_inf = (a > norm_inf) ? a : norm_inf;
                }
    }
    if (st) {
        st->iterations = counter;
        st->norm_inf_final = norm_inf;
    }
    free(own);
    free(z1);
    free(z2);
    free(de);
    free(dt);
    free(fb);
    free(fe);
    free(ft);
}

void orc_bicgstab(const orc_case *c, double eps, int n_it_max, const int ss[3], const int nn[3], const double *bb,
                  double *phi, double *work, orc_bicgstab_stats *st)
{
    bicgstab_run(c, eps, n_it_max, ss, nn, bb, phi, work, 0, 0.0, NULL, 0, 0, st);
}

void orc_bicgstab_mixed(const orc_case *c, double eps, int n_it_max, const int ss[3], const int nn[3],
                        const double *bb, double *phi, int sweeps, double omega, orc_bicgstab_stats *st)
{
    bicgstab_run(c, eps, n_it_max, ss, nn, bb, phi, NULL, sweeps, omega, NULL, 0, 0, st);
}

/* BiCGstab_orig with preconditioner 1 (multigridV, src/mod_solvers.f90:1684-1693,
 * 1750-1752); bb is not changed (the cycle averages edges of its own copies of
 * pp and rr only) */
void orc_bicgstab_jor64(const orc_case *c, double eps, int n_it_max, const int ss[3], const int nn[3],
                        const double *bb, double *phi, int sweeps, double omega, orc_bicgstab_stats *st)
{
    bicgstab_run(c, eps, n_it_max, ss, nn, bb, phi, NULL, sweeps, omega, NULL, 0, 1, st);
}

void orc_bicgstab_mg(const orc_case *c, orc_mg *mg, double eps, int n_it_max, const int ss[3], const int nn[3],
                     const double *bb, double *phi, orc_bicgstab_stats *st)
{
    bicgstab_run(c, eps, n_it_max, ss, nn, bb, phi, NULL, 0, 0.0, mg, 0, 0, st);
}

/* the same with the FP32 cycle (orc_mg_vcycle_f32) as M */
void orc_bicgstab_mg_f32(const orc_case *c, orc_mg *mg, double eps, int n_it_max, const int ss[3], const int nn[3],
                         const double *bb, double *phi, orc_bicgstab_stats *st)
{
    bicgstab_run(c, eps, n_it_max, ss, nn, bb, phi, NULL, 0, 0.0, mg, 1, 0, st);
}

/* ------------------------------------------------------------------------ */
/* IBM volume mask: src/mod_sharp_interface_ibm.f90:585-640 pattern --       */
/* outside sqrt((y-yc)**2+(z-zc)**2) <= r the velocity is zeroed.  Each      */
/* component is evaluated at its own staggered position over its B range.   */
/* ------------------------------------------------------------------------ */
void orc_tube_mask(const orc_case *c, double yc, double zc, const double *r_u, const double *r_p, uint8_t *mask3)
{
    for (int64_t q = 0; q < 3 * c->ntot; ++q) mask3[q] = 1;
    for (int q = 0; q < 3; ++q) {
        uint8_t *mk = mask3 + (int64_t)q * c->ntot;
        for (int k = c->SB[q][2]; k <= c->NB[q][2]; ++k)
            for (int j = c->SB[q][1]; j <= c->NB[q][1]; ++j)
                for (int i = c->SB[q][0]; i <= c->NB[q][0]; ++i) {
                    const double y = (q == 1) ? c->xu[1][j + 3] : c->xp[1][j + 3];
                    const double z = (q == 2) ? c->xu[2][k + 3] : c->xp[2][k + 3];
                    const double r = (q == 0) ? r_u[i + 3] : r_p[i + 3];
                    const double dy = y - yc, dz = z - zc;
                    mk[ORC_IDX(c, i, j, k)] = (sqrt(dy * dy + dz * dz) <= r) ? 1 : 0;
                }
    }
}

void orc_apply_mask(const orc_case *c, const uint8_t *mask3, double *vel)
{
    for (int q = 0; q < 3; ++q) {
        const uint8_t *mk = mask3 + (int64_t)q * c->ntot;
        double *v = vel + (int64_t)q * c->ntot;
        for (int k = c->SB[q][2]; k <= c->NB[q][2]; ++k)
            for (int j = c->SB[q][1]; j <= c->NB[q][1]; ++j)
                for (int i = c->SB[q][0]; i <= c->NB[q][0]; ++i)
                    if (!mk[ORC_IDX(c, i, j, k)]) A(v, i, j, k) = 0.0;
    }
}

/* level_pressure: src/mod_lib.f90:507-530 */
typedef struct {
    const double *p;
} pre_ctx;
static double pre_pt(const void *ctx, int64_t idx) { return ((const pre_ctx *)ctx)->p[idx]; }

void orc_level_pressure(const orc_case *c, double *pre)
{
    const int ss[3] = {c->Sp[0], c->Sp[1], c->Sp[2]};
    const int nn[3] = {c->Np[0], c->Np[1], c->Np[2]};
    pre_ctx pc = {pre};
    double pre0 = orc_reduce_sum(c, ss, nn, pre_pt, &pc);
    const int dim1 = c->bc[0][0] == -1 ? c->N[0] - 1 : c->N[0];
    const int dim2 = c->bc[1][0] == -1 ? c->N[1] - 1 : c->N[1];
    const int dim3 = c->bc[2][0] == -1 ? c->N[2] - 1 : c->N[2];
    pre0 = pre0 / (double)dim1 / (double)dim2 / (double)dim3;
    for (int k = ss[2]; k <= nn[2]; ++k)
        for (int j = ss[1]; j <= nn[1]; ++j)
            for (int i = ss[0]; i <= nn[0]; ++i) A(pre, i, j, k) = A(pre, i, j, k) - pre0;
}

/* fill_corners: src/mod_lib.f90:549-575 */
void orc_fill_corners(const orc_case *c, double *phi)
{
    const int N1 = c->N[0], N2 = c->N[1], N3 = c->N[2];
    const int w1L = c->bc[0][0] > 0, w1U = c->bc[0][1] > 0;
    const int w2L = c->bc[1][0] > 0, w2U = c->bc[1][1] > 0;
    const int w3L = c->bc[2][0] > 0, w3U = c->bc[2][1] > 0;
    for (int k = 1; k <= N3; ++k) {
        if (w1L && w2L) A(phi, 1, 1, k) = (A(phi, 2, 1, k) + A(phi, 1, 2, k) + A(phi, 2, 2, k)) / 3.0;
        if (w1L && w2U) A(phi, 1, N2, k) = (A(phi, 2, N2, k) + A(phi, 1, N2 - 1, k) + A(phi, 2, N2 - 1, k)) / 3.0;
        if (w1U && w2L) A(phi, N1, 1, k) = (A(phi, N1 - 1, 1, k) + A(phi, N1, 2, k) + A(phi, N1 - 1, 2, k)) / 3.0;
        if (w1U && w2U)
            A(phi, N1, N2, k) = (A(phi, N1 - 1, N2, k) + A(phi, N1, N2 - 1, k) + A(phi, N1 - 1, N2 - 1, k)) / 3.0;
    }
    for (int j = 1; j <= N2; ++j) {
        if (w1L && w3L) A(phi, 1, j, 1) = (A(phi, 2, j, 1) + A(phi, 1, j, 2) + A(phi, 2, j, 2)) / 3.0;
        if (w1L && w3U) A(phi, 1, j, N3) = (A(phi, 2, j, N3) + A(phi, 1, j, N3 - 1) + A(phi, 2, j, N3 - 1)) / 3.0;
        if (w1U && w3L) A(phi, N1, j, 1) = (A(phi, N1 - 1, j, 1) + A(phi, N1, j, 2) + A(phi, N1 - 1, j, 2)) / 3.0;
        if (w1U && w3U)
            A(phi, N1, j, N3) = (A(phi, N1 - 1, j, N3) + A(phi, N1, j, N3 - 1) + A(phi, N1 - 1, j, N3 - 1)) / 3.0;
    }
    for (int i = 1; i <= N1; ++i) {
        if (w2L && w3L) A(phi, i, 1, 1) = (A(phi, i, 2, 1) + A(phi, i, 1, 2) + A(phi, i, 2, 2)) / 3.0;
        if (w2L && w3U) A(phi, i, 1, N3) = (A(phi, i, 2, N3) + A(phi, i, 1, N3 - 1) + A(phi, i, 2, N3 - 1)) / 3.0;
        if (w2U && w3L) A(phi, i, N2, 1) = (A(phi, i, N2 - 1, 1) + A(phi, i, N2, 2) + A(phi, i, N2 - 1, 2)) / 3.0;
        if (w2U && w3U)
            A(phi, i, N2, N3) = (A(phi, i, N2 - 1, N3) + A(phi, i, N2, N3 - 1) + A(phi, i, N2 - 1, N3 - 1)) / 3.0;
    }
}

/* ------------------------------------------------------------------------ */
/* boundary values for the Dirichlet planes of rhs (src/mod_rhs.f90:213-241):*/
/* every wall value is 0 except bc11 (u on the x-faces), which carries the   */
/* inflow profile of the case at both x = 0 and x = L1.                      */
/* ------------------------------------------------------------------------ */
double orc_inflow_amplitude(int kind, double amp, double t)
{
    const double pi = 3.14159265358979323846;
    if (kind == 2)
        return amp * (1.0 + 0.5 * sin(2.0 * pi * t) + 0.2 * sin(4.0 * pi * t + 0.7) + 0.1 * sin(6.0 * pi * t + 1.3));
    return amp;
}

static double bc11_value(const orc_case *c, const orc_step_params *p, double t, int side, int j, int k)
{
    if (p->inflow_kind == 1) {
        const double y = c->xp[1][j + 3], z = c->xp[2][k + 3];
        return p->inflow_amp * (16.0 * y * (1.0 - y) * z * (1.0 - z));
    }
    if (p->inflow_kind == 2) {
        const int i = side == 0 ? 0 : c->N[0];
        const uint8_t fluid = p->mask3 ? p->mask3[ORC_IDX(c, i, j, k)] : 1;
        return fluid ? orc_inflow_amplitude(2, p->inflow_amp, t) : 0.0;
    }
    return 0.0;
}

static void set_plane(const orc_case *c, double *f, int q, int e, int plane, double value_or_nan, const orc_step_params *p,
                      double t, int side)
{
    int lo[3], hi[3];
    for (int g = 0; g < 3; ++g) {
        lo[g] = c->SB[q][g];
        hi[g] = c->NB[q][g];
    }
    lo[e] = hi[e] = plane;
    for (int k = lo[2]; k <= hi[2]; ++k)
        for (int j = lo[1]; j <= hi[1]; ++j)
            for (int i = lo[0]; i <= hi[0]; ++i)
                A(f, i, j, k) = (q == 0 && e == 0) ? bc11_value(c, p, t, side, j, k) : value_or_nan;
}

/* "RB einsetzen", src/mod_rhs.f90:213-241: the order of the planes decides the
 * edge values (the last plane written wins). */
static void apply_bc_planes(const orc_case *c, double *rhs, const orc_step_params *p, double t)
{
    static const int order[3][3] = {{1, 2, 0}, {0, 2, 1}, {0, 1, 2}};
    for (int q = 0; q < 3; ++q) {
        double *f = rhs + (int64_t)q * c->ntot;
        for (int o = 0; o < 3; ++o) {
            const int e = order[q][o];
            for (int side = 0; side < 2; ++side) {
                if (!(c->bc[e][side] > 0))
                    continue;
                const int plane = side == 0 ? (e == q ? 0 : 1) : c->N[e];
                set_plane(c, f, q, e, plane, 0.0, p, t, side);
            }
        }
    }
}

static void range_op(const orc_case *c, int q, double *dst, const double *src, double mult, int kind)
{
    for (int k = c->SB[q][2]; k <= c->NB[q][2]; ++k)
        for (int j = c->SB[q][1]; j <= c->NB[q][1]; ++j)
            for (int i = c->SB[q][0]; i <= c->NB[q][0]; ++i) {
                const int64_t x = ORC_IDX(c, i, j, k);
                if (kind == 0)
                    dst[x] = src[x];
                else
                    dst[x] = dst[x] - mult * src[x];
            }
}

/* one time step = 3 RK substeps: src/mod_timeint.f90:275-377 with rhs_vel
 * (src/mod_rhs.f90:35-241) and explicit (src/mod_solvers.f90:64-346); after the
 * step the mask (apply_CT_forcing pattern, :416) and level_pressure (:552). */
void orc_timestep(orc_case *c, const orc_step_params *p, int step_index, double time, double *vel, double *pre,
                  double *nl, orc_step_stats *st)
{
    (void)step_index;
    const int64_t n = c->ntot;
    double *work = calloc((size_t)3 * n, sizeof(double));
    double *rhs = calloc((size_t)3 * n, sizeof(double));
    double *res = calloc((size_t)n, sizeof(double));
    double *gpre = calloc((size_t)n, sizeof(double));
    double *pp = calloc((size_t)n, sizeof(double));
    double *comp = calloc((size_t)n, sizeof(double));
    /* preconditioner 1 with the reference configuration's cycle (prog/config.txt:108-116) */
    orc_mg *mg = p->precond_poisson == 1 ? orc_mg_create(c, 15, 4, 4, 4) : NULL;
    double *bw = calloc((size_t)7 * n, sizeof(double));
    const int ss[3] = {c->Sp[0], c->Sp[1], c->Sp[2]};
    const int nn[3] = {c->Np[0], c->Np[1], c->Np[2]};
    const double dtime = p->dtime;

    for (int s = 0; s < 3; ++s) {
        double subtime;
        if (s == 0)
            subtime = time + dtime * aRK[0];
        else if (s == 1)
            subtime = time + dtime * (aRK[0] + aRK[1] + bRK[1]);
        else
            subtime = time + dtime;

        orc_exchange_all_all(c, vel);
        orc_interpolate_vel(c, vel, work);

        /* rhs_vel */
        for (int q = 0; q < 3; ++q) range_op(c, q, rhs + q * n, vel + q * n, 0.0, 0);
        if (s != 0) {
            const double mult = dtime * bRK[s];
            for (int q = 0; q < 3; ++q) range_op(c, q, rhs + q * n, nl + q * n, mult, 1);
        }
        orc_nonlinear(c, vel, work, nl, pp);
        orc_helmholtz_explicit(c, 1.0 / p->Re, vel, nl);
        {
            const double mult = dtime * aRK[s];
            for (int q = 0; q < 3; ++q) range_op(c, q, rhs + q * n, nl + q * n, mult, 1);
        }
        apply_bc_planes(c, rhs, p, subtime);

        /* pressure as the increment's initial guess */
        for (int k = ss[2]; k <= nn[2]; ++k)
            for (int j = ss[1]; j <= nn[1]; ++j)
                for (int i = ss[0]; i <= nn[0]; ++i) A(pre, i, j, k) = A(pre, i, j, k) * (aRK[s] + bRK[s]) * dtime;

        /* explicit */
        for (int q = 0; q < 3; ++q) range_op(c, q, vel + q * n, rhs + q * n, 0.0, 0);
        orc_divergence2(c, vel, res);
        if (p->jor_sweeps > 0) {
            if (p->jor_fp32)
                orc_jor_mixed(c, p->jor_sweeps, p->jor_omega, ss, nn, pre, res);
            else if (p->pressure_bc && p->ct_state && p->ct_lab && s == 2) {
                /* device_jacobi_solver's substep-3 pressure before every sweep
                 * and once more after the last (the exit trip applies it too) */
                for (int r = 0; r < p->jor_sweeps; ++r) {
                    jacobi_pressure_bc(c, p->ct_lab, p->ct_state->k_start, 3, dtime, p->ct_state->pres, pre);
                    orc_jor(c, 1, p->jor_omega, ss, nn, pre, res, comp);
                }
                jacobi_pressure_bc(c, p->ct_lab, p->ct_state->k_start, 3, dtime, p->ct_state->pres, pre);
                orc_exchange_relax(c, pre);
            } else
                orc_jor(c, p->jor_sweeps, p->jor_omega, ss, nn, pre, res, comp);
        }
        orc_bicgstab_stats bst;
        const double eps = p->bicg_rel_tol * absmax(c, ss, nn, res);
        if (p->precond_sweeps > 0) {
            /* the device keeps the initial guess and restarts plain FP64 BiCGstab
             * from it when the preconditioned iteration does not converge */
            memcpy(comp, pre, sizeof(double) * (size_t)n);
            bicgstab_run(c, eps, p->bicg_max_it, ss, nn, res, pre, bw, p->precond_sweeps, p->jor_omega, NULL, 0,
                         p->precond_fp64, &bst);
            if (bst.exit_reason != 0) {
                const int mixed_its = bst.iterations;
                memcpy(pre, comp, sizeof(double) * (size_t)n);
                bicgstab_run(c, eps, p->bicg_max_it, ss, nn, res, pre, bw, 0, 0.0, NULL, 0, 0, &bst);
                bst.iterations += mixed_its;
            }
        } else if (mg && p->mg_fp32) {
            /* the FP32 cycle: as with the FP32 sweeps, a solve that does not converge
             * is redone in FP64 (here with the FP64 cycle) from the saved guess */
            memcpy(comp, pre, sizeof(double) * (size_t)n);
            bicgstab_run(c, eps, p->bicg_max_it, ss, nn, res, pre, bw, 0, 0.0, mg, 1, 0, &bst);
            if (bst.exit_reason != 0) {
                const int mixed_its = bst.iterations;
                memcpy(pre, comp, sizeof(double) * (size_t)n);
                bicgstab_run(c, eps, p->bicg_max_it, ss, nn, res, pre, bw, 0, 0.0, mg, 0, 0, &bst);
                bst.iterations += mixed_its;
            }
        } else {
            bicgstab_run(c, eps, p->bicg_max_it, ss, nn, res, pre, bw, 0, 0.0, mg, 0, 0, &bst);
        }
        if (st) {
            st->bicg_iterations[s] = bst.iterations;
            st->bicg_final[s] = bst.norm_inf_final;
        }
        for (int m = 1; m <= 3; ++m) {
            orc_gradient(c, m, pre, gpre);
            range_op(c, m - 1, vel + (m - 1) * n, gpre, 1.0, 1);
        }

        for (int k = ss[2]; k <= nn[2]; ++k)
            for (int j = ss[1]; j <= nn[1]; ++j)
                for (int i = ss[0]; i <= nn[0]; ++i)
                    A(pre, i, j, k) = A(pre, i, j, k) / (aRK[s] + bRK[s]) / dtime;
        orc_fill_corners(c, pre);
    }
    if (p->mask3)
        orc_apply_mask(c, p->mask3, vel);
    /* apply_CT_forcing follows the time update (src/mod_timeint.f90:405-416) */
    if (p->ct_lab && p->ct_state)
        orc_ct_apply_forcing(c, p->ct_lab, p->ct_state, time + p->dtime, vel);
    else if (p->ct_lab)
        orc_ct_forcing(c, p->ct_lab, p->ct_npts, p->ct_t, p->ct_q, time + p->dtime, vel);
    if (p->level_pressure)
        orc_level_pressure(c, pre);
    if (st) {
        orc_divergence2(c, vel, res);
        st->div_max_after = absmax(c, ss, nn, res);
    }
    free(work);
    free(rhs);
    free(res);
    free(gpre);
    free(pp);
    free(comp);
    orc_mg_destroy(mg);
    free(bw);
}
