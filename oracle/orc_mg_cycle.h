/*
 * Oracle (test infrastructure only): the steps of the multigrid V cycle, written
 * once for the arithmetic type REAL and included by orc_mg.c twice (double:
 * the reference's cycle; float: the FP32 cycle of the mixed-precision solve).
 * The including file defines REAL, FN(name) and the level accessors LCDG(L),
 * LCR(L), LCI(L), LD(L).  Sources are cited in orc_mg.c.
 */
/* exchange_relax (src/mod_exchange.f90:1273-1480), single block: periodic
 * ghosts from the opposite side; lower (fb >= 0) and/or upper (fb <= 0) */
static void FN(lvl_exchange)(const mg_level *L, REAL *phi, int lower, int upper)
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

/* relaxation_div_grad / relaxation_div_grad_inv, Jacobi branch */
static void FN(mg_relax)(const orc_mg *m, const mg_level *L, int init, int n_relax, const REAL *bb, REAL *rel)
{
    const REAL *g1 = LCDG(L)[0], *g2 = LCDG(L)[1], *g3 = LCDG(L)[2];
    const REAL omega = (REAL)m->omega;
    REAL *comp = LD(L);
    for (int r = 1; r <= n_relax; ++r) {
        if (r == 1 && init) {
            for (int k = 1; k <= L->N[2]; ++k)
                for (int j = 1; j <= L->N[1]; ++j)
                    for (int i = 1; i <= L->N[0]; ++i) {
                        if (in_snf(L, i, j, k))
                            F(L, rel, i, j, k) =
                                omega * F(L, bb, i, j, k) / ((CDG3(g1, 0, i) + CDG3(g2, 0, j)) + CDG3(g3, 0, k));
                        else
                            F(L, rel, i, j, k) = (REAL)0;
                    }
        } else {
            FN(lvl_exchange)(L, rel, 1, 1);
            for (int k = L->SNF[2][0]; k <= L->SNF[2][1]; ++k)
                for (int j = L->SNF[1][0]; j <= L->SNF[1][1]; ++j)
                    for (int i = L->SNF[0][0]; i <= L->SNF[0][1]; ++i) {
                        REAL t = F(L, bb, i, j, k);
                        t = t - CDG3(g1, -1, i) * F(L, rel, i - 1, j, k);
                        t = t - CDG3(g1, 1, i) * F(L, rel, i + 1, j, k);
                        t = t - CDG3(g2, -1, j) * F(L, rel, i, j - 1, k);
                        t = t - CDG3(g2, 1, j) * F(L, rel, i, j + 1, k);
                        t = t - CDG3(g3, -1, k) * F(L, rel, i, j, k - 1);
                        t = t - CDG3(g3, 1, k) * F(L, rel, i, j, k + 1);
                        const REAL diag = (CDG3(g1, 0, i) + CDG3(g2, 0, j)) + CDG3(g3, 0, k);
                        F(L, comp, i, j, k) = omega * t / diag + ((REAL)1 - omega) * F(L, rel, i, j, k);
                    }
            for (int k = 1; k <= L->N[2]; ++k)
                for (int j = 1; j <= L->N[1]; ++j)
                    for (int i = 1; i <= L->N[0]; ++i)
                        F(L, rel, i, j, k) = in_snf(L, i, j, k) ? F(L, comp, i, j, k) : (REAL)0;
        }
    }
}

/* product_div_grad_relax (src/mod_laplace.f90:200-300) */
static void FN(mg_lap)(const mg_level *L, REAL *phi, REAL *lap)
{
    const REAL *g1 = LCDG(L)[0], *g2 = LCDG(L)[1], *g3 = LCDG(L)[2];
    FN(lvl_exchange)(L, phi, 1, 1);
    for (int k = 1; k <= L->N[2]; ++k)
        for (int j = 1; j <= L->N[1]; ++j)
            for (int i = 1; i <= L->N[0]; ++i) {
                if (!in_snf(L, i, j, k)) {
                    F(L, lap, i, j, k) = (REAL)0;
                    continue;
                }
                REAL t = CDG3(g1, -1, i) * F(L, phi, i - 1, j, k);
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
static void FN(mg_edges)(const mg_level *L, REAL *f, int three)
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
                        F(L, f, e[0], e[1], e[2]) = (F(L, f, a[0], a[1], a[2]) + F(L, f, b[0], b[1], b[2])) / (REAL)2;
                }
            }
    }
}

/* restrict(.TRUE., g, coarse, fine1, fine2): coarse = R (fine1 - fine2) */
static void FN(mg_restrict)(const mg_level *L, const mg_level *Lc, REAL *coarse, REAL *fine1, REAL *fine2)
{
    FN(mg_edges)(L, fine1, 0);
    FN(mg_edges)(L, fine2, 0);
    int nmax[3], dd[3];
    for (int d = 0; d < 3; ++d) {
        nmax[d] = Lc->N[d];
        dd[d] = (L->N[d] - 1) / (nmax[d] - 1);
        if (L->bc[d][1] == -1 || L->bc[d][1] == 0) nmax[d] = nmax[d] - 1; /* ls = -1 */
    }
    FN(lvl_exchange)(L, fine1, 1, 1);
    FN(lvl_exchange)(L, fine2, 1, 1);
    const REAL *r1 = LCR(Lc)[0], *r2 = LCR(Lc)[1], *r3 = LCR(Lc)[2];
#define DF(i, j, k) (F(L, fine1, i, j, k) - F(L, fine2, i, j, k))
    for (int kk = 1; kk <= nmax[2]; ++kk) {
        const int k = dd[2] * (kk - 1) + 1;
        for (int jj = 1; jj <= nmax[1]; ++jj) {
            const int j = dd[1] * (jj - 1) + 1;
            for (int ii = 1; ii <= nmax[0]; ++ii) {
                const int i = dd[0] * (ii - 1) + 1;
                REAL t = ((CDG3(r1, 0, ii) + CDG3(r2, 0, jj)) + CDG3(r3, 0, kk)) * DF(i, j, k);
                t = t + CDG3(r1, -1, ii) * DF(i - 1, j, k);
                t = t + CDG3(r1, 1, ii) * DF(i + 1, j, k);
                t = t + CDG3(r2, -1, jj) * DF(i, j - 1, k);
                t = t + CDG3(r2, 1, jj) * DF(i, j + 1, k);
                t = t + CDG3(r3, -1, kk) * DF(i, j, k - 1);
                t = t + CDG3(r3, 1, kk) * DF(i, j, k + 1);
                F(Lc, coarse, ii, jj, kk) = t / (REAL)3;
            }
        }
    }
#undef DF
}

/* interpolate(.TRUE., g, coarse, fine, work): fine += P coarse */
static void FN(mg_interpolate)(const mg_level *L, const mg_level *Lc, REAL *coarse, REAL *fine, REAL *work)
{
    FN(mg_edges)(Lc, coarse, 1);
    FN(lvl_exchange)(Lc, coarse, 0, 1); /* ex = -1: upper ghosts only */
    const int imax = L->N[0], jmax = L->N[1], kmax = L->N[2];
    const int di = (imax - 1) / (Lc->N[0] - 1), dj = (jmax - 1) / (Lc->N[1] - 1), dk = (kmax - 1) / (Lc->N[2] - 1);
    for (int kk = 1; kk <= Lc->N[2]; ++kk)
        for (int jj = 1; jj <= Lc->N[1]; ++jj)
            for (int ii = 1; ii <= Lc->N[0]; ++ii)
                F(L, work, di * (ii - 1) + 1, dj * (jj - 1) + 1, dk * (kk - 1) + 1) = F(Lc, coarse, ii, jj, kk);
    const REAL *c1 = LCI(L)[0], *c2 = LCI(L)[1], *c3 = LCI(L)[2];
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

