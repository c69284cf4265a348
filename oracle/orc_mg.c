/*
 * Oracle (test infrastructure only): the multigrid V cycle that preconditions
 * BiCGstab (preconditioner 1), single block (NB = 1, n_gather = 1, every level
 * on this rank), Jacobi smoothing.  Build with -ffp-contract=off.
 *
 *   grid hierarchy        src/mod_setup.f90:96-143
 *   level coordinates     src/mod_geometry.f90:104-166
 *   cdg per level         src/mod_coeffs.f90:1048-1248 (get_stencil)
 *   interpolation coeffs  src/mod_coeffs.f90:1683-1746 (interp_coeffs)
 *   restriction coeffs    src/mod_coeffs.f90:1879-2035 (restr_coeffs)
 *   V cycle               src/mod_solvers.f90:353-585 (relax_restrict, interpolate_relax,
 *                         relax_bottom, multigridV with init_yes = .TRUE., gstart = 1)
 *   restriction           src/mod_solvers.f90:588-808 (restrict, add_yes = .TRUE.)
 *   interpolation         src/mod_solvers.f90:955-1141 (interpolate, add_yes = .TRUE.)
 *   smoother              src/mod_laplace.f90:466-597 and 712-790 (relaxation_div_grad and
 *                         relaxation_div_grad_inv, Jacobi_yes, omega = 0.98: identical)
 *   residual product      src/mod_laplace.f90:200-300 (product_div_grad_relax)
 *   exchange              src/mod_exchange.f90:1273-1480 (exchange_relax)
 *
 * Relaxation range.  The reference relaxes and forms the residual product on
 * SNF (src/mod_laplace.f90:472-518, 270-280), which with walls is 2..N-1, and
 * then copies its local work array back over all of 1..N (:595), so the wall
 * planes receive values that were never computed, while BiCGstab's unknowns
 * (S?p..N?p) include them.  Here every level relaxes, and forms the product
 * on, its pressure range S?p..N?p, as the JOR sweeps do: with walls 1..N, with
 * periodic directions 1..N-1.  The only points of 1..N outside it are periodic
 * image planes; sweeps and the first relaxation set them to 0, and every
 * read of them goes through exchange_relax first.  With this a V cycle is a
 * fixed linear map of bb (with only the reference's SNF, BiCGstab breaks down
 * on wall cases because M then ignores the wall residuals).  The edge averaging
 * of restrict and interpolate (src/mod_solvers.f90:632-661, 1028-1043) is
 * kept: it changes the caller's bb on the edges of wall pairs, which BiCGstab
 * reads back.
 */
#include "orc_internal.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define MG_MAX_LEVELS 15 /* n_grids_max */

typedef struct {
    int N[3];
    int bc[3][2];
    int per[3];
    int SNF[3][2]; /* relaxation range */
    int SB[3][2];  /* SNB: the range exchange_relax copies from (S?p..N?p) */
    int nx, ny;
    int64_t nxy, ntot;
    double *cdg[3]; /* (-1:1, 1:N) at (ii+1) + 3(i-1) */
    double *cR[3];  /* (-1:1, 1:N) at (ii+1) + 3(i-1); levels >= 2 */
    double *cI[3];  /* (1:2, 1:N) at (t-1) + 2(i-1); levels <= G-1 */
    double *xp[3], *xu[3]; /* level coordinates (-3:N+3) at [i+3] */
    double *A, *B, *C, *D; /* bb, phi, work (residual product / interpolation), relaxation copy */
} mg_level;

struct orc_mg {
    const orc_case *c;
    int G;
    int n_down, n_up, n_bottom;
    double omega;
    mg_level L[MG_MAX_LEVELS + 1]; /* 1-based */
};

#define LX(L, i, j, k) ((int64_t)((i) + 3) + (int64_t)((j) + 3) * (L)->nx + (int64_t)((k) + 3) * (L)->nxy)
#define F(L, p, i, j, k) (p)[LX(L, i, j, k)]
#define CDG3(a, ii, i) ((a)[((ii) + 1) + 3 * ((i) - 1)])
#define CI2(a, t, i) ((a)[((t) - 1) + 2 * ((i) - 1)])

/* ------------------------------------------------------------------------ */
/* hierarchy: src/mod_setup.f90:96-143 with NB = 1                           */
/* ------------------------------------------------------------------------ */
static int hierarchy(const int N[3], int limit, int NN[MG_MAX_LEVELS + 1][3])
{
    int n_grids = 1;
    for (int m = 0; m < 3; ++m) NN[1][m] = N[m];
    for (int g = 2; g <= MG_MAX_LEVELS; ++g)
        for (int m = 0; m < 3; ++m) {
            if ((NN[g - 1][m] - 1) % 2 == 0 && NN[g - 1][m] >= 5) {
                NN[g][m] = (NN[g - 1][m] - 1) / 2 + 1;
                n_grids = g;
            } else {
                NN[g][m] = NN[g - 1][m];
            }
        }
    if (n_grids > limit) n_grids = limit;
    return n_grids;
}

/* level coordinates, src/mod_geometry.f90:104-166 (g >= 2; g = 1 is the case's) */
static void level_coords(const orc_case *c, int d, int Nl, double *xp, double *xu)
{
    const int M = c->N[d];
    const double *Y = c->xp[d];
    const double Lx = c->L[d];
    const int di = (M - 1) / (Nl - 1);
    double tp[4], tu[4];
    for (int i = 0; i < Nl + 7; ++i) xp[i] = xu[i] = 0.0;
    for (int i = 1; i <= Nl; ++i) xp[i + 3] = Y[(1 + di * i - di) + 3];
    for (int i = 1; i <= Nl - 1; ++i) xu[i + 3] = Y[(1 + di * i - di / 2) + 3];
    if (c->bc[d][0] == -1) {
        /* array assignments: the right-hand sides are read before any store */
        for (int t = 0; t < 4; ++t) {
            tp[t] = xp[(Nl - 1 - 3 + t) + 3] - Lx;
            tu[t] = xu[(Nl - 1 - 3 + t) + 3] - Lx;
        }
        for (int t = 0; t < 4; ++t) {
            xp[(-3 + t) + 3] = tp[t];
            xu[(-3 + t) + 3] = tu[t];
        }
        for (int t = 0; t < 4; ++t) {
            tp[t] = xp[(1 + t) + 3] + Lx;
            tu[t] = xu[(1 + t) + 3] + Lx;
        }
        for (int t = 0; t < 4; ++t) {
            xp[(Nl + t) + 3] = tp[t];
            xu[(Nl + t) + 3] = tu[t];
        }
    }
    if (c->bc[d][0] > 0) {
        for (int b = -3; b <= 0; ++b) xp[b + 3] = 2.0 * xp[1 + 3] - xp[(2 - b) + 3];
        for (int b = -3; b <= 0; ++b) xu[b + 3] = 2.0 * xp[1 + 3] - xu[(1 - b) + 3];
    }
    if (c->bc[d][1] > 0) {
        for (int t = 1; t <= 3; ++t) xp[(Nl + t) + 3] = 2.0 * xp[Nl + 3] - xp[(Nl - t) + 3];
        for (int t = 0; t <= 3; ++t) xu[(Nl + t) + 3] = 2.0 * xp[Nl + 3] - xu[(Nl - 1 - t) + 3];
    }
}

/* get_stencil for one level and direction (src/mod_coeffs.f90:1077-1128) */
static void level_cdg(const orc_case *c, int d, const mg_level *L, double *cdg)
{
    const int N = L->N[d];
    const double *xp = L->xp[d], *xu = L->xu[d];
    const double *YU = c->xu[d]; /* the finest grid's y?u: used on every level at walls */
    const int M = c->N[d];
    double *cD = calloc(2 * (size_t)(N + 1), sizeof(double)); /* (-1:0, 1:N) */
    double *cG = calloc(2 * (size_t)(N + 1), sizeof(double)); /* (0:1, 0:N) */
#define CDv(ii, i) cD[((ii) + 1) + 2 * (i)]
#define CGv(ii, i) cG[(ii) + 2 * (i)]
#define XP(i) xp[(i) + 3]
#define XU(i) xu[(i) + 3]
    for (int i = 1; i <= N; ++i) {
        CDv(-1, i) = -1.0 / (XU(i) - XU(i - 1));
        CDv(0, i) = 1.0 / (XU(i) - XU(i - 1));
    }
    for (int i = 1; i <= N - 1; ++i) {
        CGv(0, i) = -1.0 / (XP(i + 1) - XP(i));
        CGv(1, i) = 1.0 / (XP(i + 1) - XP(i));
    }
    if (c->bc[d][0] > 0) {
        CGv(0, 0) = CGv(1, 0) = 0.0;
        CDv(-1, 1) = CDv(0, 1) = 0.0;
        CDv(0, 1) = 1.0 / (YU[1 + 3] - YU[0 + 3]);
    } else {
        CGv(0, 0) = -1.0 / (XP(1) - XP(0));
        CGv(1, 0) = 1.0 / (XP(1) - XP(0));
    }
    if (c->bc[d][1] > 0) {
        CGv(0, N) = CGv(1, N) = 0.0;
        CDv(-1, N) = CDv(0, N) = 0.0;
        CDv(-1, N) = -1.0 / (YU[M + 3] - YU[(M - 1) + 3]);
    } else {
        CGv(0, N) = -1.0 / (XP(N + 1) - XP(N));
        CGv(1, N) = 1.0 / (XP(N + 1) - XP(N));
    }
    for (int i = 1; i <= N; ++i) {
        CDG3(cdg, -1, i) = CDv(-1, i) * CGv(0, i - 1);
        CDG3(cdg, 0, i) = CDv(-1, i) * CGv(1, i - 1) + CDv(0, i) * CGv(0, i);
        CDG3(cdg, 1, i) = CDv(0, i) * CGv(1, i);
    }
    if (c->bc[d][0] > 0)
        for (int q = -1; q <= 1; ++q) CDG3(cdg, q, 1) = 2.0 * CDG3(cdg, q, 1);
    if (c->bc[d][1] > 0)
        for (int q = -1; q <= 1; ++q) CDG3(cdg, q, N) = 2.0 * CDG3(cdg, q, N);
#undef CDv
#undef CGv
#undef XP
#undef XU
    free(cD);
    free(cG);
}

/* restr_coeffs for level g >= 2 (src/mod_coeffs.f90:1879-2035) */
static void level_cR(const orc_case *c, int d, const mg_level *L, double *cR)
{
    const int n = L->N[d];
    for (int ii = 1; ii <= n; ++ii) {
        CDG3(cR, -1, ii) = 1.0 / 4.0;
        CDG3(cR, 0, ii) = 2.0 / 4.0;
        CDG3(cR, 1, ii) = 1.0 / 4.0;
    }
    if (c->bc[d][0] > 0)
        for (int ii = 1; ii <= 2; ++ii) {
            CDG3(cR, -1, ii) = 0.0;
            CDG3(cR, 0, ii) = 1.0;
            CDG3(cR, 1, ii) = 0.0;
        }
    if (c->bc[d][1] > 0)
        for (int ii = n; ii >= n - 1; --ii) {
            CDG3(cR, -1, ii) = 0.0;
            CDG3(cR, 0, ii) = 1.0;
            CDG3(cR, 1, ii) = 0.0;
        }
}

/* interp_coeffs for level g <= G-1 (src/mod_coeffs.f90:1683-1746) */
static void level_cI(const orc_case *c, int d, const mg_level *L, double *cI)
{
    const int n = L->N[d];
    const int di = (c->N[d] - 1) / (n - 1);
    const double *Y = c->xp[d];
    for (int i = 2; i <= n - 1; i += 2) {
        const int i0 = 1 + (i - 1) * di;
        const double Dx10 = Y[i0 + 3] - Y[(i0 - di) + 3];
        const double Dx12 = Y[(i0 + di) + 3] - Y[(i0 - di) + 3];
        CI2(cI, 1, i) = 1.0 - Dx10 / Dx12;
        CI2(cI, 2, i) = Dx10 / Dx12;
    }
}

orc_mg *orc_mg_create(const orc_case *c, int n_grids_limit, int n_down, int n_up, int n_bottom)
{
    if (n_grids_limit < 1 || n_grids_limit > MG_MAX_LEVELS) return NULL;
    for (int d = 0; d < 3; ++d)
        if (c->bc[d][0] != -1 && !(c->bc[d][0] > 0 && c->bc[d][1] > 0)) return NULL;
    orc_mg *m = calloc(1, sizeof *m);
    m->c = c;
    m->n_down = n_down;
    m->n_up = n_up;
    m->n_bottom = n_bottom;
    m->omega = 0.98; /* src/mod_laplace.f90:470 */
    int NN[MG_MAX_LEVELS + 1][3];
    m->G = hierarchy(c->N, n_grids_limit, NN);
    for (int g = 1; g <= m->G; ++g) {
        mg_level *L = &m->L[g];
        for (int d = 0; d < 3; ++d) {
            const int N = NN[g][d];
            L->N[d] = N;
            L->bc[d][0] = c->bc[d][0];
            L->bc[d][1] = c->bc[d][1];
            L->per[d] = c->bc[d][0] == -1;
            /* SNB / SNF with ls = -1 (src/mod_setup.f90:828-867) */
            L->SB[d][0] = 1;
            L->SB[d][1] = L->per[d] ? N - 1 : N;
            /* relaxation range: the level's pressure range (see the header) */
            L->SNF[d][0] = L->SB[d][0];
            L->SNF[d][1] = L->SB[d][1];
            L->xp[d] = calloc((size_t)N + 7, sizeof(double));
            L->xu[d] = calloc((size_t)N + 7, sizeof(double));
            if (g == 1) {
                memcpy(L->xp[d], c->xp[d], sizeof(double) * ((size_t)N + 7));
                memcpy(L->xu[d], c->xu[d], sizeof(double) * ((size_t)N + 7));
            } else {
                level_coords(c, d, N, L->xp[d], L->xu[d]);
            }
            L->cdg[d] = calloc(3 * (size_t)N, sizeof(double));
            level_cdg(c, d, L, L->cdg[d]);
            L->cR[d] = calloc(3 * (size_t)N, sizeof(double));
            if (g >= 2) level_cR(c, d, L, L->cR[d]);
            L->cI[d] = calloc(2 * (size_t)N, sizeof(double));
            if (g <= m->G - 1) level_cI(c, d, L, L->cI[d]);
        }
        L->nx = L->N[0] + 7;
        L->ny = L->N[1] + 7;
        L->nxy = (int64_t)L->nx * L->ny;
        L->ntot = L->nxy * (L->N[2] + 7);
        if (g >= 2) { /* level 1 works on the caller's bb and phi */
            L->A = calloc((size_t)L->ntot, sizeof(double));
            L->B = calloc((size_t)L->ntot, sizeof(double));
        }
        L->C = calloc((size_t)L->ntot, sizeof(double));
        L->D = calloc((size_t)L->ntot, sizeof(double));
    }
    return m;
}

void orc_mg_destroy(orc_mg *m)
{
    if (!m) return;
    for (int g = 1; g <= m->G; ++g) {
        mg_level *L = &m->L[g];
        for (int d = 0; d < 3; ++d) {
            free(L->cdg[d]);
            free(L->cR[d]);
            free(L->cI[d]);
            free(L->xp[d]);
            free(L->xu[d]);
        }
        free(L->A);
        free(L->B);
        free(L->C);
        free(L->D);
    }
    free(m);
}

int orc_mg_levels(const orc_mg *m, int *NN)
{
    for (int g = 1; g <= m->G; ++g)
        for (int d = 0; d < 3; ++d) NN[3 * (g - 1) + d] = m->L[g].N[d];
    return m->G;
}

int orc_mg_coeff(const orc_mg *m, int g, const char *name, double *out)
{
    if (g < 1 || g > m->G || strlen(name) < 3) return -1;
    const mg_level *L = &m->L[g];
    const int d = name[strlen(name) - 1] - '1';
    if (d < 0 || d > 2) return -1;
    const int n = L->N[d];
    if (!strncmp(name, "cdg", 3)) {
        memcpy(out, L->cdg[d], sizeof(double) * 3 * (size_t)n);
        return 3 * n;
    }
    if (!strncmp(name, "cR", 2)) {
        memcpy(out, L->cR[d], sizeof(double) * 3 * (size_t)n);
        return 3 * n;
    }
    if (!strncmp(name, "cI", 2)) {
        memcpy(out, L->cI[d], sizeof(double) * 2 * (size_t)n);
        return 2 * n;
    }
    if (!strncmp(name, "xp", 2)) {
        memcpy(out, L->xp[d], sizeof(double) * ((size_t)n + 7));
        return n + 7;
    }
    if (!strncmp(name, "xu", 2)) {
        memcpy(out, L->xu[d], sizeof(double) * ((size_t)n + 7));
        return n + 7;
    }
    return -1;
}

int64_t orc_mg_field_size(const orc_mg *m, int g)
{
    return (g >= 1 && g <= m->G) ? m->L[g].ntot : -1;
}

/* exchange_relax (src/mod_exchange.f90:1273-1480), single block: periodic
 * ghosts from the opposite side; lower (fb >= 0) and/or upper (fb <= 0) */
static void lvl_exchange(const mg_level *L, double *phi, int lower, int upper)
{
    const int N1 = L->N[0], N2 = L->N[1], N3 = L->N[2];
    if (L->per[0]) {
        const int S = L->SB[0][0], Nn = L->SB[0][1];
        for (int k = 1; k <= N3; ++k)
            for (int j = 1; j <= N2; ++j) {
                if (lower) F(L, phi, S - 1, j, k) = F(L, phi, Nn, j, k);
                if (upper) F(L, phi, Nn + 1, j, k) = F(L, phi, S, j, k);
            }
    }
    if (L->per[1]) {
        const int S = L->SB[1][0], Nn = L->SB[1][1];
        for (int k = 1; k <= N3; ++k)
            for (int i = 1; i <= N1; ++i) {
                if (lower) F(L, phi, i, S - 1, k) = F(L, phi, i, Nn, k);
                if (upper) F(L, phi, i, Nn + 1, k) = F(L, phi, i, S, k);
            }
    }
    if (L->per[2]) {
        const int S = L->SB[2][0], Nn = L->SB[2][1];
        for (int j = 1; j <= N2; ++j)
            for (int i = 1; i <= N1; ++i) {
                if (lower) F(L, phi, i, j, S - 1) = F(L, phi, i, j, Nn);
                if (upper) F(L, phi, i, j, Nn + 1) = F(L, phi, i, j, S);
            }
    }
}

static int in_snf(const mg_level *L, int i, int j, int k)
{
    return i >= L->SNF[0][0] && i <= L->SNF[0][1] && j >= L->SNF[1][0] && j <= L->SNF[1][1] &&
           k >= L->SNF[2][0] && k <= L->SNF[2][1];
}

/* relaxation_div_grad / relaxation_div_grad_inv, Jacobi branch */
static void mg_relax(const orc_mg *m, const mg_level *L, int init, int n_relax, const double *bb, double *rel)
{
    const double *g1 = L->cdg[0], *g2 = L->cdg[1], *g3 = L->cdg[2];
    const double omega = m->omega;
    double *comp = L->D;
    for (int r = 1; r <= n_relax; ++r) {
        if (r == 1 && init) {
            for (int k = 1; k <= L->N[2]; ++k)
                for (int j = 1; j <= L->N[1]; ++j)
                    for (int i = 1; i <= L->N[0]; ++i) {
                        if (in_snf(L, i, j, k))
                            F(L, rel, i, j, k) =
                                omega * F(L, bb, i, j, k) / ((CDG3(g1, 0, i) + CDG3(g2, 0, j)) + CDG3(g3, 0, k));
                        else
                            F(L, rel, i, j, k) = 0.0;
                    }
        } else {
            lvl_exchange(L, rel, 1, 1);
            for (int k = L->SNF[2][0]; k <= L->SNF[2][1]; ++k)
                for (int j = L->SNF[1][0]; j <= L->SNF[1][1]; ++j)
                    for (int i = L->SNF[0][0]; i <= L->SNF[0][1]; ++i) {
                        double t = F(L, bb, i, j, k);
                        t = t - CDG3(g1, -1, i) * F(L, rel, i - 1, j, k);
                        t = t - CDG3(g1, 1, i) * F(L, rel, i + 1, j, k);
                        t = t - CDG3(g2, -1, j) * F(L, rel, i, j - 1, k);
                        t = t - CDG3(g2, 1, j) * F(L, rel, i, j + 1, k);
                        t = t - CDG3(g3, -1, k) * F(L, rel, i, j, k - 1);
                        t = t - CDG3(g3, 1, k) * F(L, rel, i, j, k + 1);
                        const double diag = (CDG3(g1, 0, i) + CDG3(g2, 0, j)) + CDG3(g3, 0, k);
                        F(L, comp, i, j, k) = omega * t / diag + (1.0 - omega) * F(L, rel, i, j, k);
                    }
            for (int k = 1; k <= L->N[2]; ++k)
                for (int j = 1; j <= L->N[1]; ++j)
                    for (int i = 1; i <= L->N[0]; ++i)
                        F(L, rel, i, j, k) = in_snf(L, i, j, k) ? F(L, comp, i, j, k) : 0.0;
        }
    }
}

/* product_div_grad_relax (src/mod_laplace.f90:200-300) */
static void mg_lap(const mg_level *L, double *phi, double *lap)
{
    const double *g1 = L->cdg[0], *g2 = L->cdg[1], *g3 = L->cdg[2];
    lvl_exchange(L, phi, 1, 1);
    for (int k = 1; k <= L->N[2]; ++k)
        for (int j = 1; j <= L->N[1]; ++j)
            for (int i = 1; i <= L->N[0]; ++i) {
                if (!in_snf(L, i, j, k)) {
                    F(L, lap, i, j, k) = 0.0;
                    continue;
                }
                double t = CDG3(g1, -1, i) * F(L, phi, i - 1, j, k);
                t = t + CDG3(g1, 1, i) * F(L, phi, i + 1, j, k);
                t = t + CDG3(g2, -1, j) * F(L, phi, i, j - 1, k);
                t = t + CDG3(g2, 1, j) * F(L, phi, i, j + 1, k);
                t = t + CDG3(g3, -1, k) * F(L, phi, i, j, k - 1);
                t = t + CDG3(g3, 1, k) * F(L, phi, i, j, k + 1);
                t = t + ((CDG3(g1, 0, i) + CDG3(g2, 0, j)) + CDG3(g3, 0, k)) * F(L, phi, i, j, k);
                F(L, lap, i, j, k) = t;
            }
}

/* Edge averaging on the edges of wall pairs: restrict (three = 0, two
 * neighbours / 2, src/mod_solvers.f90:632-661) and interpolate (three = 1,
 * three neighbours / 3, src/mod_solvers.f90:1028-1043).  Statements in the
 * reference's order: direction pairs (1,2), (1,3), (2,3); per pair the sides
 * (L,L), (L,U), (U,L), (U,U); each a line over the whole range 1..N of the
 * third direction. */
static void mg_edges(const mg_level *L, double *f, int three)
{
    static const int pairs[3][3] = {{0, 1, 2}, {0, 2, 1}, {1, 2, 0}}; /* p, q, line t */
    for (int pr = 0; pr < 3; ++pr) {
        const int p = pairs[pr][0], q = pairs[pr][1], t = pairs[pr][2];
        for (int sp = 0; sp < 2; ++sp)
            for (int sq = 0; sq < 2; ++sq) {
                if (!(L->bc[p][sp] > 0 && L->bc[q][sq] > 0)) continue;
                const int ep = sp ? L->N[p] : 1, ip = sp ? L->N[p] - 1 : 2;
                const int eq = sq ? L->N[q] : 1, iq = sq ? L->N[q] - 1 : 2;
                for (int s = 1; s <= L->N[t]; ++s) {
                    int e[3], a[3], b[3], cc[3];
                    e[p] = ep; e[q] = eq; e[t] = s;
                    a[p] = ip; a[q] = eq; a[t] = s;
                    b[p] = ep; b[q] = iq; b[t] = s;
                    cc[p] = ip; cc[q] = iq; cc[t] = s;
                    if (three)
                        F(L, f, e[0], e[1], e[2]) =
                            (F(L, f, a[0], a[1], a[2]) + F(L, f, b[0], b[1], b[2]) + F(L, f, cc[0], cc[1], cc[2])) /
                            3.0;
                    else
                        F(L, f, e[0], e[1], e[2]) = (F(L, f, a[0], a[1], a[2]) + F(L, f, b[0], b[1], b[2])) / 2.0;
                }
            }
    }
}

/* restrict(.TRUE., g, coarse, fine1, fine2): coarse = R (fine1 - fine2) */
static void mg_restrict(const mg_level *L, const mg_level *Lc, double *coarse, double *fine1, double *fine2)
{
    mg_edges(L, fine1, 0);
    mg_edges(L, fine2, 0);
    int nmax[3], dd[3];
    for (int d = 0; d < 3; ++d) {
        nmax[d] = Lc->N[d];
        dd[d] = (L->N[d] - 1) / (nmax[d] - 1);
        if (L->bc[d][1] == -1 || L->bc[d][1] == 0) nmax[d] = nmax[d] - 1; /* ls = -1 */
    }
    lvl_exchange(L, fine1, 1, 1);
    lvl_exchange(L, fine2, 1, 1);
    const double *r1 = Lc->cR[0], *r2 = Lc->cR[1], *r3 = Lc->cR[2];
#define DF(i, j, k) (F(L, fine1, i, j, k) - F(L, fine2, i, j, k))
    for (int kk = 1; kk <= nmax[2]; ++kk) {
        const int k = dd[2] * (kk - 1) + 1;
        for (int jj = 1; jj <= nmax[1]; ++jj) {
            const int j = dd[1] * (jj - 1) + 1;
            for (int ii = 1; ii <= nmax[0]; ++ii) {
                const int i = dd[0] * (ii - 1) + 1;
                double t = ((CDG3(r1, 0, ii) + CDG3(r2, 0, jj)) + CDG3(r3, 0, kk)) * DF(i, j, k);
                t = t + CDG3(r1, -1, ii) * DF(i - 1, j, k);
                t = t + CDG3(r1, 1, ii) * DF(i + 1, j, k);
                t = t + CDG3(r2, -1, jj) * DF(i, j - 1, k);
                t = t + CDG3(r2, 1, jj) * DF(i, j + 1, k);
                t = t + CDG3(r3, -1, kk) * DF(i, j, k - 1);
                t = t + CDG3(r3, 1, kk) * DF(i, j, k + 1);
                F(Lc, coarse, ii, jj, kk) = t / 3.0;
            }
        }
    }
#undef DF
}

/* interpolate(.TRUE., g, coarse, fine, work): fine += P coarse */
static void mg_interpolate(const mg_level *L, const mg_level *Lc, double *coarse, double *fine, double *work)
{
    mg_edges(Lc, coarse, 1);
    lvl_exchange(Lc, coarse, 0, 1); /* ex = -1: upper ghosts only */
    const int imax = L->N[0], jmax = L->N[1], kmax = L->N[2];
    const int di = (imax - 1) / (Lc->N[0] - 1), dj = (jmax - 1) / (Lc->N[1] - 1), dk = (kmax - 1) / (Lc->N[2] - 1);
    for (int kk = 1; kk <= Lc->N[2]; ++kk)
        for (int jj = 1; jj <= Lc->N[1]; ++jj)
            for (int ii = 1; ii <= Lc->N[0]; ++ii)
                F(L, work, di * (ii - 1) + 1, dj * (jj - 1) + 1, dk * (kk - 1) + 1) = F(Lc, coarse, ii, jj, kk);
    const double *c1 = L->cI[0], *c2 = L->cI[1], *c3 = L->cI[2];
    if (dk != 1)
        for (int k = 2; k <= kmax - 1; k += dk)
            for (int j = 1; j <= jmax; j += dj)
                for (int i = 1; i <= imax; i += di)
                    F(L, work, i, j, k) = CI2(c3, 1, k) * F(L, work, i, j, k - 1) + CI2(c3, 2, k) * F(L, work, i, j, k + 1);
    if (dj != 1)
        for (int k = 1; k <= kmax; ++k)
            for (int j = 2; j <= jmax - 1; j += dj)
                for (int i = 1; i <= imax; i += di)
                    F(L, work, i, j, k) = CI2(c2, 1, j) * F(L, work, i, j - 1, k) + CI2(c2, 2, j) * F(L, work, i, j + 1, k);
    if (di != 1)
        for (int k = 1; k <= kmax; ++k)
            for (int j = 1; j <= jmax; ++j)
                for (int i = 2; i <= imax - 1; i += di)
                    F(L, work, i, j, k) = CI2(c1, 1, i) * F(L, work, i - 1, j, k) + CI2(c1, 2, i) * F(L, work, i + 1, j, k);
    for (int k = 1; k <= kmax; ++k)
        for (int j = 1; j <= jmax; ++j)
            for (int i = 1; i <= imax; ++i) F(L, fine, i, j, k) = F(L, fine, i, j, k) + F(L, work, i, j, k);
}

/* multigridV(.TRUE., 1, bb, phi) (src/mod_solvers.f90:456-585) */
void orc_mg_vcycle(orc_mg *m, double *bb, double *phi)
{
    const int G = m->G;
    mg_level *L = m->L;
    if (G == 1) {
        mg_relax(m, &L[1], 1, m->n_bottom, bb, phi);
        mg_relax(m, &L[1], 0, m->n_bottom, bb, phi);
        return;
    }
    for (int g = 1; g <= G - 1; ++g) {
        double *b = g == 1 ? bb : L[g].A, *x = g == 1 ? phi : L[g].B;
        mg_relax(m, &L[g], 1, m->n_down, b, x);
        mg_lap(&L[g], x, L[g].C);
        mg_restrict(&L[g], &L[g + 1], L[g + 1].A, b, L[g].C);
    }
    mg_relax(m, &L[G], 1, m->n_bottom, L[G].A, L[G].B);
    mg_relax(m, &L[G], 0, m->n_bottom, L[G].A, L[G].B);
    for (int g = G - 1; g >= 1; --g) {
        double *b = g == 1 ? bb : L[g].A, *x = g == 1 ? phi : L[g].B;
        mg_interpolate(&L[g], &L[g + 1], L[g + 1].B, x, L[g].C);
        mg_relax(m, &L[g], 0, m->n_up, b, x);
    }
}

/* test hooks: the single operations of the cycle on level g */
void orc_mg_restrict_residual(orc_mg *m, int g, double *bb, double *x, double *coarse)
{
    mg_lap(&m->L[g], x, m->L[g].C);
    mg_restrict(&m->L[g], &m->L[g + 1], coarse, bb, m->L[g].C);
}

void orc_mg_interpolate_add(orc_mg *m, int g, double *coarse, double *fine)
{
    mg_interpolate(&m->L[g], &m->L[g + 1], coarse, fine, m->L[g].C);
}

void orc_mg_relax(orc_mg *m, int g, int init, int n_relax, const double *bb, double *rel)
{
    mg_relax(m, &m->L[g], init, n_relax, bb, rel);
}

int orc_mg_grids(const orc_mg *m) { return m->G; }
