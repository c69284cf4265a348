/* Oracle internals (test infrastructure only). */
#ifndef ORC_INTERNAL_H
#define ORC_INTERNAL_H

#include "impact_oracle.h"

#include <stdint.h>

struct orc_case {
    int N[3];
    int bc[3][2];
    double L[3];
    double stretch[3];
    /* index ranges (Fortran values) */
    int Sp[3], Np[3];
    int SB[3][3], NB[3][3]; /* [component m][direction d] */
    int SI[3][3], NI[3][3];
    int SNF[3][2];
    /* array geometry */
    int nx, ny, nz;
    int64_t nxy, ntot;
    /* coordinates (-3:N+3), element i at [i+3] */
    double *xp[3], *xu[3];
    /* coefficients c(lo:hi, 0:N) */
    double *cNp[3][2]; /* [d][0] = D (upwind -1), [1] = U (upwind +1), n = -3..3 */
    double *cNu[3][2]; /* velocity grid of component d */
    double *cpp[3];    /* cp11, cp22, cp33 */
    double *cuu[3];    /* cu11, cv22, cw33 */
    double *cG[3];     /* cGp1..3, -2..3 */
    double *cIpu[3];   /* cIpu, cIpv, cIpw, -2..3 */
    double *cD[3];     /* cDu1, cDv2, cDw3, -3..2 */
    double *cIup[3];   /* cIup, cIvp, cIwp, -3..2 */
    double *cdg[3];    /* (-1:1, 1:N) */
    int red_mode;
};

/* ORC_OMP builds (bench baseline only) run the point loops in parallel over k;
 * every point is written by one thread and no sum is split, so results are
 * identical to the serial build. */
#ifdef ORC_OMP
#define ORC_PAR _Pragma("omp parallel for schedule(static)")
#define ORC_PAR_MAX_NORM _Pragma("omp parallel for schedule(static) reduction(max : norm_inf)")
#else
#define ORC_PAR
#define ORC_PAR_MAX_NORM
#endif

#define ORC_IDX(c, i, j, k) ((int64_t)((i) + 3) + (int64_t)((j) + 3) * (c)->nx + (int64_t)((k) + 3) * (c)->nxy)
#define C7(a, ii, i) ((a)[((ii) + 3) + 7 * (i)])
#define CG6(a, ii, i) ((a)[((ii) + 2) + 6 * (i)])
#define CD6(a, ii, i) ((a)[((ii) + 3) + 6 * (i)])
#define CDG(a, ii, i) ((a)[((ii) + 1) + 3 * ((i) - 1)])

/* canonical reduction tree shared with the device kernels */
#define ORC_RED_TX 16
#define ORC_RED_TY 16
#define ORC_RED_KC 32
#define ORC_RED_THREADS 256

typedef double (*orc_point_fn)(const void *ctx, int64_t idx);
double orc_reduce_sum(const orc_case *c, const int ss[3], const int nn[3], orc_point_fn f, const void *ctx);
double orc_reduce_max(const orc_case *c, const int ss[3], const int nn[3], orc_point_fn f, const void *ctx);

#endif
