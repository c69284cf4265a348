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
    /* the FP32 cycle: its tables (rounded once) and arrays; level 1's Af holds (float) bb */
    float *cdgf[3], *cRf[3], *cIf[3];
    float *Af, *Bf, *Cf, *Df;
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
        for (int d = 0; d < 3; ++d) {
            const int N = L->N[d];
            L->cdgf[d] = calloc(3 * (size_t)N, sizeof(float));
            L->cRf[d] = calloc(3 * (size_t)N, sizeof(float));
            L->cIf[d] = calloc(2 * (size_t)N, sizeof(float));
            for (int t = 0; t < 3 * N; ++t) {
                L->cdgf[d][t] = (float)L->cdg[d][t];
                L->cRf[d][t] = (float)L->cR[d][t];
            }
            for (int t = 0; t < 2 * N; ++t) L->cIf[d][t] = (float)L->cI[d][t];
        }
        L->Af = calloc((size_t)L->ntot, sizeof(float));
        L->Bf = calloc((size_t)L->ntot, sizeof(float));
        L->Cf = calloc((size_t)L->ntot, sizeof(float));
        L->Df = calloc((size_t)L->ntot, sizeof(float));
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
            free(L->cdgf[d]);
            free(L->cRf[d]);
            free(L->cIf[d]);
        }
        free(L->Af);
        free(L->Bf);
        free(L->Cf);
        free(L->Df);
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

static int in_snf(const mg_level *L, int i, int j, int k)
{
    return i >= L->SNF[0][0] && i <= L->SNF[0][1] && j >= L->SNF[1][0] && j <= L->SNF[1][1] &&
           k >= L->SNF[2][0] && k <= L->SNF[2][1];
}

#define REAL double
#define FN(name) name##_d
#define LCDG(L) ((L)->cdg)
#define LCR(L) ((L)->cR)
#define LCI(L) ((L)->cI)
#define LD(L) ((L)->D)
#include "orc_mg_cycle.h"
#undef REAL
#undef FN
#undef LCDG
#undef LCR
#undef LCI
#undef LD
#define REAL float
#define FN(name) name##_f
#define LCDG(L) ((L)->cdgf)
#define LCR(L) ((L)->cRf)
#define LCI(L) ((L)->cIf)
#define LD(L) ((L)->Df)
#include "orc_mg_cycle.h"
#undef REAL
#undef FN
#undef LCDG
#undef LCR
#undef LCI
#undef LD

/* multigridV(.TRUE., 1, bb, phi) (src/mod_solvers.f90:456-585) */
void orc_mg_vcycle(orc_mg *m, double *bb, double *phi)
{
    const int G = m->G;
    mg_level *L = m->L;
    if (G == 1) {
        mg_relax_d(m, &L[1], 1, m->n_bottom, bb, phi);
        mg_relax_d(m, &L[1], 0, m->n_bottom, bb, phi);
        return;
    }
    for (int g = 1; g <= G - 1; ++g) {
        double *b = g == 1 ? bb : L[g].A, *x = g == 1 ? phi : L[g].B;
        mg_relax_d(m, &L[g], 1, m->n_down, b, x);
        mg_lap_d(&L[g], x, L[g].C);
        mg_restrict_d(&L[g], &L[g + 1], L[g + 1].A, b, L[g].C);
    }
    mg_relax_d(m, &L[G], 1, m->n_bottom, L[G].A, L[G].B);
    mg_relax_d(m, &L[G], 0, m->n_bottom, L[G].A, L[G].B);
    for (int g = G - 1; g >= 1; --g) {
        double *b = g == 1 ? bb : L[g].A, *x = g == 1 ? phi : L[g].B;
        mg_interpolate_d(&L[g], &L[g + 1], L[g + 1].B, x, L[g].C);
        mg_relax_d(m, &L[g], 0, m->n_up, b, x);
    }
}

/* The FP32 cycle of the mixed-precision solve (no counterpart in the
 * reference): the same steps in binary32 on float copies of the tables, from
 * b32 = (float) bb over the pressure range; phi = (double) of the level-1
 * result.  Where the FP64 cycle averages bb's edges (restrict on level 1) the
 * double bb is averaged as well, so BiCGstab sees the same change to pp and rr. */
void orc_mg_vcycle_f32(orc_mg *m, double *bb, double *phi)
{
    const int G = m->G;
    mg_level *L = m->L;
    mg_level *L1 = &L[1];
    float *b1 = L1->Af, *x1 = L1->Bf;
    memset(b1, 0, sizeof(float) * (size_t)L1->ntot);
    for (int k = L1->SNF[2][0]; k <= L1->SNF[2][1]; ++k)
        for (int j = L1->SNF[1][0]; j <= L1->SNF[1][1]; ++j)
            for (int i = L1->SNF[0][0]; i <= L1->SNF[0][1]; ++i) F(L1, b1, i, j, k) = (float)F(L1, bb, i, j, k);
    if (G == 1) {
        mg_relax_f(m, L1, 1, m->n_bottom, b1, x1);
        mg_relax_f(m, L1, 0, m->n_bottom, b1, x1);
    } else {
        for (int g = 1; g <= G - 1; ++g) {
            float *b = L[g].Af, *x = L[g].Bf;
            mg_relax_f(m, &L[g], 1, m->n_down, b, x);
            mg_lap_f(&L[g], x, L[g].Cf);
            if (g == 1) mg_edges_d(L1, bb, 0);
            mg_restrict_f(&L[g], &L[g + 1], L[g + 1].Af, b, L[g].Cf);
        }
        mg_relax_f(m, &L[G], 1, m->n_bottom, L[G].Af, L[G].Bf);
        mg_relax_f(m, &L[G], 0, m->n_bottom, L[G].Af, L[G].Bf);
        for (int g = G - 1; g >= 1; --g) {
            mg_interpolate_f(&L[g], &L[g + 1], L[g + 1].Bf, L[g].Bf, L[g].Cf);
            mg_relax_f(m, &L[g], 0, m->n_up, L[g].Af, L[g].Bf);
        }
    }
    for (int k = L1->SNF[2][0]; k <= L1->SNF[2][1]; ++k)
        for (int j = L1->SNF[1][0]; j <= L1->SNF[1][1]; ++j)
            for (int i = L1->SNF[0][0]; i <= L1->SNF[0][1]; ++i) F(L1, phi, i, j, k) = (double)F(L1, x1, i, j, k);
}

/* test hooks: the single operations of the cycle on level g */
void orc_mg_restrict_residual(orc_mg *m, int g, double *bb, double *x, double *coarse)
{
    mg_lap_d(&m->L[g], x, m->L[g].C);
    mg_restrict_d(&m->L[g], &m->L[g + 1], coarse, bb, m->L[g].C);
}

void orc_mg_interpolate_add(orc_mg *m, int g, double *coarse, double *fine)
{
    mg_interpolate_d(&m->L[g], &m->L[g + 1], coarse, fine, m->L[g].C);
}

void orc_mg_relax(orc_mg *m, int g, int init, int n_relax, const double *bb, double *rel)
{
    mg_relax_d(m, &m->L[g], init, n_relax, bb, rel);
}

int orc_mg_grids(const orc_mg *m) { return m->G; }
