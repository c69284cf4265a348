/*
 * Oracle (test infrastructure only): grid, index ranges, coordinates and FD
 * coefficients.  Restates the reference Fortran; each function cites the
 * lines it follows.  Build with -ffp-contract=off.
 */
#include "orc_internal.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* index ranges: src/mod_setup.f90:465-700 (ls = -1, single block)           */
/* ------------------------------------------------------------------------ */
static void init_limits(orc_case *c)
{
    for (int d = 0; d < 3; ++d) {
        const int N = c->N[d];
        /* defaults (periodic / neighbour block): 2+ls .. N+ls */
        c->Sp[d] = 1;
        c->Np[d] = N - 1;
        c->SNF[d][0] = 1;
        c->SNF[d][1] = N - 1;
        for (int m = 0; m < 3; ++m) {
            c->SB[m][d] = 1;
            c->NB[m][d] = N - 1;
            c->SI[m][d] = 1;
            c->NI[m][d] = N - 1;
        }
        if (c->bc[d][0] > 0) { /* mod_setup.f90:608-620 */
            c->Sp[d] = 1;
            c->SNF[d][0] = 2;
            for (int m = 0; m < 3; ++m) {
                c->SB[m][d] = (m == d) ? 0 : 1;
                c->SI[m][d] = (m == d) ? 1 : 2;
            }
        }
        if (c->bc[d][1] > 0) { /* mod_setup.f90:634-646 */
            c->Np[d] = N;
            c->SNF[d][1] = N - 1;
            for (int m = 0; m < 3; ++m) {
                c->NB[m][d] = N;
                c->NI[m][d] = N - 1;
            }
        }
    }
}

/* ------------------------------------------------------------------------ */
/* coord_tanh: src/usr_func.f90:1964-2000                                    */
/* ------------------------------------------------------------------------ */
/* The reference takes its equidistant branch whenever ii0U == ii0L, so a
 * symmetric tanh grid cannot be requested through it; here factor 0 selects
 * the equidistant branch and any other factor evaluates the reference's tanh
 * mapping with ii0L = ii0U = factor (same expressions, same order). */
static void coord_tanh(double Lmax, double iimax, double ii0L, double ii0U, double ii, double *xx)
{
    if (ii0L == 0.0 && ii0U == 0.0) {
        *xx = (ii - 1.0) * Lmax / (iimax - 1.0);
    } else {
        double cmax = tanh(ii0U);
        double cmin = -tanh(ii0L);
        double yy = (ii - 1.0) / (iimax - 1.0) * (cmax - cmin) + cmin;
        *xx = Lmax * (atanh(yy) - atanh(cmin)) / (ii0U + ii0L);
    }
}

/* get_coords (src/usr_geometry.f90:93-157) + coordinates() for g = 1
 * (src/mod_geometry.f90:67-200): global y?p(1:M), y?u(0:M), extended into the
 * halo by periodic shift or by mirroring about the wall, then copied to x?p, x?u. */
static void init_coords(orc_case *c)
{
    for (int d = 0; d < 3; ++d) {
        const int M = c->N[d];
        const double L = c->L[d];
        const double s = c->stretch[d];
        double *yp = c->xp[d]; /* index i at yp[i+3], i = -3..M+3 */
        double *yu = c->xu[d];
        for (int i = -3; i <= M + 3; ++i) {
            yp[i + 3] = 0.0;
            yu[i + 3] = 0.0;
        }
        for (int i = 1; i <= M; ++i)
            coord_tanh(L, (double)M, s, s, (double)i, &yp[i + 3]);
        for (int i = 0; i <= M; ++i)
            coord_tanh(L, (double)M, s, s, (double)i + 0.5, &yu[i + 3]);

        if (c->bc[d][0] == -1) { /* mod_geometry.f90:129-135 */
            for (int i = -3; i <= 0; ++i) {
                yp[i + 3] = yp[(M - 1 + i) + 3] - L;
                yu[i + 3] = yu[(M - 1 + i) + 3] - L;
            }
            for (int i = M; i <= M + 3; ++i) {
                yp[i + 3] = yp[(1 + i - M) + 3] + L;
                yu[i + 3] = yu[(1 + i - M) + 3] + L;
            }
        }
        if (c->bc[d][0] > 0) { /* mod_geometry.f90:139-147, g == 1 */
            /* y?pR(-3:0) = 2*y?pR(1) - y?pR(5:2:-1) */
            for (int i = -3; i <= 0; ++i)
                yp[i + 3] = 2.0 * yp[1 + 3] - yp[(2 - i) + 3];
            /* y?uR(-3:-1) = 2*y?pR(1) - y?uR(4:2:-1) */
            for (int i = -3; i <= -1; ++i)
                yu[i + 3] = 2.0 * yp[1 + 3] - yu[(1 - i) + 3];
        }
        if (c->bc[d][1] > 0) { /* mod_geometry.f90:149-157, g == 1 */
            for (int i = M + 1; i <= M + 3; ++i)
                yp[i + 3] = 2.0 * yp[M + 3] - yp[(2 * M - i) + 3];
            for (int i = M + 1; i <= M + 3; ++i)
                yu[i + 3] = 2.0 * yp[M + 3] - yu[(2 * M - 1 - i) + 3];
        }
    }
}

/* ------------------------------------------------------------------------ */
/* diff_coeffs_exact: src/mod_coeffs.f90:598-772 (the reference's rational   */
/* tables, written out here as numerators over a common denominator).       */
/* ------------------------------------------------------------------------ */
typedef struct {
    int n, abl;
    double dx;
    double den;
    double num[7];
} exact_row;

static const exact_row EXACT[] = {
    /* n = 2 */
    {2, 0, 0.5, 2., {3., -1.}},
    {2, 0, -0.5, 2., {1., 1.}},
    {2, 0, -1.5, 2., {-1., 3.}},
    {2, 1, 0.5, 1., {-1., 1.}},
    {2, 1, -0.5, 1., {-1., 1.}},
    {2, 1, -1.5, 1., {-1., 1.}},
    {2, 1, 0.0, 1., {-1., 1.}},
    {2, 1, -1.0, 1., {-1., 1.}},
    /* n = 3 */
    {3, 0, 0.5, 8., {15., -10., 3.}},
    {3, 0, -0.5, 8., {3., 6., -1.}},
    {3, 0, -1.5, 8., {-1., 6., 3.}},
    {3, 0, -2.5, 8., {3., -10., 15.}},
    {3, 1, 0.5, 1., {-2., 3., -1.}},
    {3, 1, -0.5, 1., {-1., 1., 0.}},
    {3, 1, -1.5, 1., {0., -1., 1.}},
    {3, 1, -2.5, 1., {1., -3., 2.}},
    {3, 1, 0.0, 2., {-3., 4., -1.}},
    {3, 1, -1.0, 2., {-1., 0., 1.}},
    {3, 1, -2.0, 2., {1., -4., 3.}},
    {3, 2, 0.0, 1., {1., -2., 1.}},
    {3, 2, -1.0, 1., {1., -2., 1.}},
    {3, 2, -2.0, 1., {1., -2., 1.}},
    /* n = 4 */
    {4, 0, 0.5, 16., {35., -35., 21., -5.}},
    {4, 0, -0.5, 16., {5., 15., -5., 1.}},
    {4, 0, -1.5, 16., {-1., 9., 9., -1.}},
    {4, 0, -2.5, 16., {1., -5., 15., 5.}},
    {4, 0, -3.5, 16., {-5., 21., -35., 35.}},
    {4, 1, 0.5, 24., {-71., 141., -93., 23.}},
    {4, 1, -0.5, 24., {-23., 21., 3., -1.}},
    {4, 1, -1.5, 24., {1., -27., 27., -1.}},
    {4, 1, -2.5, 24., {1., -3., -21., 23.}},
    {4, 1, -3.5, 24., {-23., 93., -141., 71.}},
    {4, 1, 0.0, 6., {-11., 18., -9., 2.}},
    {4, 1, -1.0, 6., {-2., -3., 6., -1.}},
    {4, 1, -2.0, 6., {1., -6., 3., 2.}},
    {4, 1, -3.0, 6., {-2., 9., -18., 11.}},
    {4, 2, 0.0, 1., {2., -5., 4., -1.}},
    {4, 2, -1.0, 1., {1., -2., 1., 0.}},
    {4, 2, -2.0, 1., {0., 1., -2., 1.}},
    {4, 2, -3.0, 1., {-1., 4., -5., 2.}},
    /* n = 5 */
    {5, 0, 0.5, 128., {315., -420., 378., -180., 35.}},
    {5, 0, -0.5, 128., {35., 140., -70., 28., -5.}},
    {5, 0, -1.5, 128., {-5., 60., 90., -20., 3.}},
    {5, 0, -2.5, 128., {3., -20., 90., 60., -5.}},
    {5, 0, -3.5, 128., {-5., 28., -70., 140., 35.}},
    {5, 0, -4.5, 128., {35., -180., 378., -420., 315.}},
    {5, 1, 0.5, 24., {-93., 229., -225., 111., -22.}},
    {5, 1, -0.5, 24., {-22., 17., 9., -5., 1.}},
    {5, 1, -1.5, 24., {1., -27., 27., -1., 0.}},
    {5, 1, -2.5, 24., {0., 1., -27., 27., -1.}},
    {5, 1, -3.5, 24., {-1., 5., -9., -17., 22.}},
    {5, 1, -4.5, 24., {22., -111., 225., -229., 93.}},
    {5, 1, 0.0, 12., {-25., 48., -36., 16., -3.}},
    {5, 1, -1.0, 12., {-3., -10., 18., -6., 1.}},
    {5, 1, -2.0, 12., {1., -8., 0., 8., -1.}},
    {5, 1, -3.0, 12., {-1., 6., -18., 10., 3.}},
    {5, 1, -4.0, 12., {3., -16., 36., -48., 25.}},
    {5, 2, 0.0, 12., {35., -104., 114., -56., 11.}},
    {5, 2, -1.0, 12., {11., -20., 6., 4., -1.}},
    {5, 2, -2.0, 12., {-1., 16., -30., 16., -1.}},
    {5, 2, -3.0, 12., {-1., 4., 6., -20., 11.}},
    {5, 2, -4.0, 12., {11., -56., 114., -104., 35.}},
    /* n = 6 */
    {6, 0, 0.5, 256., {693., -1155., 1386., -990., 385., -63.}},
    {6, 0, -0.5, 256., {63., 315., -210., 126., -45., 7.}},
    {6, 0, -1.5, 256., {-7., 105., 210., -70., 21., -3.}},
    {6, 0, -2.5, 256., {3., -25., 150., 150., -25., 3.}},
    {6, 0, -3.5, 256., {-3., 21., -70., 210., 105., -7.}},
    {6, 0, -4.5, 256., {7., -45., 126., -210., 315., 63.}},
    {6, 0, -5.5, 256., {-63., 385., -990., 1386., -1155., 693.}},
    {6, 1, 0.5, 1920., {-9129., 26765., -34890., 25770., -10205., 1689.}},
    {6, 1, -0.5, 1920., {-1689., 1005., 1430., -1110., 435., -71.}},
    {6, 1, -1.5, 1920., {71., -2115., 2070., 10., -45., 9.}},
    {6, 1, -2.5, 1920., {-9., 125., -2250., 2250., -125., 9.}},
    {6, 1, -3.5, 1920., {-9., 45., -10., -2070., 2115., -71.}},
    {6, 1, -4.5, 1920., {71., -435., 1110., -1430., -1005., 1689.}},
    {6, 1, -5.5, 1920., {-1689., 10205., -25770., 34890., -26765., 9129.}},
    {6, 1, 0.0, 60., {-137., 300., -300., 200., -75., 12.}},
    {6, 1, -1.0, 60., {-12., -65., 120., -60., 20., -3.}},
    {6, 1, -2.0, 60., {3., -30., -20., 60., -15., 2.}},
    {6, 1, -3.0, 60., {-2., 15., -60., 20., 30., -3.}},
    {6, 1, -4.0, 60., {3., -20., 60., -120., 65., 12.}},
    {6, 1, -5.0, 60., {-12., 75., -200., 300., -300., 137.}},
    {6, 2, 0.0, 12., {45., -154., 214., -156., 61., -10.}},
    {6, 2, -1.0, 12., {10., -15., -4., 14., -6., 1.}},
    {6, 2, -2.0, 12., {-1., 16., -30., 16., -1., 0.}},
    {6, 2, -3.0, 12., {0., -1., 16., -30., 16., -1.}},
    {6, 2, -4.0, 12., {1., -6., 14., -4., -15., 10.}},
    {6, 2, -5.0, 12., {-10., 61., -156., 214., -154., 45.}},
    /* n = 7 */
    {7, 0, 0.5, 1024., {3003., -6006., 9009., -8580., 5005., -1638., 231.}},
    {7, 0, -0.5, 1024., {231., 1386., -1155., 924., -495., 154., -21.}},
    {7, 0, -1.5, 1024., {-21., 378., 945., -420., 189., -54., 7.}},
    {7, 0, -2.5, 1024., {7., -70., 525., 700., -175., 42., -5.}},
    {7, 0, -3.5, 1024., {-5., 42., -175., 700., 525., -70., 7.}},
    {7, 0, -4.5, 1024., {7., -54., 189., -420., 945., 378., -21.}},
    {7, 0, -5.5, 1024., {-21., 154., -495., 924., -1155., 1386., 231.}},
    {7, 0, -6.5, 1024., {231., -1638., 5005., -8580., 9009., -6006., 3003.}},
    {7, 1, 0.5, 1920., {-10756., 36527., -59295., 58310., -34610., 11451., -1627.}},
    {7, 1, -0.5, 1920., {-1627., 633., 2360., -2350., 1365., -443., 62.}},
    {7, 1, -1.5, 1920., {62., -2061., 1935., 190., -180., 63., -9.}},
    {7, 1, -2.5, 1920., {-9., 125., -2250., 2250., -125., 9., 0.}},
    {7, 1, -3.5, 1920., {0., -9., 125., -2250., 2250., -125., 9.}},
    {7, 1, -4.5, 1920., {9., -63., 180., -190., -1935., 2061., -62.}},
    {7, 1, -5.5, 1920., {-62., 443., -1365., 2350., -2360., -633., 1627.}},
    {7, 1, -6.5, 1920., {1627., -11451., 34610., -58310., 59295., -36527., 10756.}},
    {7, 1, 0.0, 60., {-147., 360., -450., 400., -225., 72., -10.}},
    {7, 1, -1.0, 60., {-10., -77., 150., -100., 50., -15., 2.}},
    {7, 1, -2.0, 60., {2., -24., -35., 80., -30., 8., -1.}},
    {7, 1, -3.0, 60., {-1., 9., -45., 0., 45., -9., 1.}},
    {7, 1, -4.0, 60., {1., -8., 30., -80., 35., 24., -2.}},
    {7, 1, -5.0, 60., {-2., 15., -50., 100., -150., 77., 10.}},
    {7, 1, -6.0, 60., {10., -72., 225., -400., 450., -360., 147.}},
    {7, 2, 0.0, 180., {812., -3132., 5265., -5080., 2970., -972., 137.}},
    {7, 2, -1.0, 180., {137., -147., -225., 470., -285., 93., -13.}},
    {7, 2, -2.0, 180., {-13., 228., -420., 200., 15., -12., 2.}},
    {7, 2, -3.0, 180., {2., -27., 270., -490., 270., -27., 2.}},
    {7, 2, -4.0, 180., {2., -12., 15., 200., -420., 228., -13.}},
    {7, 2, -5.0, 180., {-13., 93., -285., 470., -225., -147., 137.}},
    {7, 2, -6.0, 180., {137., -972., 2970., -5080., 5265., -3132., 812.}},
};

int orc_exact_table(int abl, int n_coeff, double deltaX, double *cc)
{
    const int nrows = (int)(sizeof(EXACT) / sizeof(EXACT[0]));
    for (int r = 0; r < nrows; ++r) {
        if (EXACT[r].n == n_coeff && EXACT[r].abl == abl && EXACT[r].dx == deltaX) {
            for (int q = 0; q < n_coeff; ++q)
                cc[q] = EXACT[r].num[q] / EXACT[r].den;
            return 1;
        }
    }
    return 0;
}

/* diff_coeffs: src/mod_coeffs.f90:200-596, restricted to mapping_yes = .TRUE.
 * with abl in {0 (grid types 2,3), 1, 2} and at most 7 coefficients, i.e.
 * every call the hot path makes (FD_coeffs, src/mod_coeffs.f90:37-120).
 * xC, xE are indexed (bL:Nmax+bU) -> xC[i - bL].  cc(bL:bU, 0:Nmax). */
static int diff_coeffs(int abl, int upwind, const int *ncb, int dim_ncb, const double *xC, int BCL, int BCU,
                       int grid_type, int Nmax, int bL, int bU, double *cc)
{
    const int w = bU - bL + 1;
    const int iStart = (grid_type == 0 || grid_type == 3) ? 1 : 0;
    for (int q = 0; q < w * (Nmax + 1); ++q)
        cc[q] = 0.0;

    for (int i = iStart; i <= Nmax; ++i) {
        int n_coeff;
        if (BCL > 0 && i <= iStart - 1 + dim_ncb)
            n_coeff = ncb[i + 1 - iStart - 1];
        else if (BCU > 0 && i >= Nmax + 1 - dim_ncb)
            n_coeff = ncb[Nmax + 1 - i - 1];
        else
            n_coeff = ncb[dim_ncb - 1];

        if (!(n_coeff > 0 && n_coeff > abl))
            continue;

        int right;
        if (grid_type == 0 || grid_type == 1) {
            if (BCL > 0 && i <= iStart - 1 + n_coeff / 2)
                right = n_coeff - i + iStart - 1;
            else if (BCU > 0 && i >= Nmax + 1 - n_coeff / 2)
                right = Nmax - i;
            else
                right = (n_coeff - 1) / 2;
        } else if (grid_type == 2) {
            if (BCL > 0 && i <= n_coeff / 2)
                right = n_coeff - i - iStart;
            else if (BCU > 0 && i >= Nmax + 0 - n_coeff / 2)
                right = Nmax - i;
            else
                right = n_coeff / 2;
        } else {
            if (BCL > 0 && i <= n_coeff / 2)
                right = n_coeff - i - iStart;
            else if (BCU > 0 && i >= Nmax + 1 - n_coeff / 2)
                right = Nmax - i;
            else
                right = n_coeff / 2 - 1;
        }
        int left = right - n_coeff + 1;

        if (upwind == -1) {
            if (!(BCL > 0 && i < iStart + n_coeff / 2)) {
                n_coeff = n_coeff - 1;
                left = left + 1;
            }
        }
        if (upwind == 1) {
            if (!(BCU > 0 && i > Nmax - n_coeff / 2)) {
                n_coeff = n_coeff - 1;
                right = right - 1;
            }
        }
        if (right > bU || left < bL) {
            fprintf(stderr, "oracle diff_coeffs: stencil does not fit (i=%d)\n", i);
            return -1;
        }

        /* index-space offsets (mapping_yes) */
        double deltaX0;
        if (grid_type == 2)
            deltaX0 = (double)left - 0.5;
        else if (grid_type == 3)
            deltaX0 = (double)left + 0.5;
        else
            deltaX0 = (double)left;

        if (n_coeff > 7) {
            fprintf(stderr, "oracle diff_coeffs: n_coeff > 7 not restated\n");
            return -1;
        }
        if (abl != 0) {
            /* mapped derivative, mod_coeffs.f90:497-537 */
            double cc_xi[2][7];
            double dxi[2] = {0.0, 0.0};
            for (int k = 1; k <= abl; ++k) {
                if (!orc_exact_table(k, n_coeff, deltaX0, cc_xi[k - 1])) {
                    fprintf(stderr, "oracle: no table entry abl=%d n=%d dx=%g\n", k, n_coeff, deltaX0);
                    return -1;
                }
                for (int ii = left; ii <= right; ++ii)
                    dxi[k - 1] = dxi[k - 1] + cc_xi[k - 1][ii - left] * xC[(i + ii) - bL];
            }
            for (int ii = left; ii <= right; ++ii) {
                double v;
                if (abl == 1) {
                    v = cc_xi[0][ii - left] / dxi[0];
                } else {
                    double d1sq = dxi[0] * dxi[0];
                    double d1cu = dxi[0] * dxi[0] * dxi[0];
                    v = cc_xi[1][ii - left] / d1sq - cc_xi[0][ii - left] * dxi[1] / d1cu;
                }
                cc[(ii - bL) + w * i] = v;
            }
        } else {
            /* mapped interpolation (not a filter), mod_coeffs.f90:539-556 */
            double tmp[7];
            if (!orc_exact_table(0, n_coeff, deltaX0, tmp)) {
                fprintf(stderr, "oracle: no table entry abl=0 n=%d dx=%g\n", n_coeff, deltaX0);
                return -1;
            }
            for (int ii = left; ii <= right; ++ii)
                cc[(ii - bL) + w * i] = tmp[ii - left];
        }
    }
    return 0;
}

/* get_stencil for the finest grid: src/mod_coeffs.f90:1048-1258 */
static void get_stencil(orc_case *c)
{
    for (int d = 0; d < 3; ++d) {
        const int imax = c->N[d];
        const double *xu = c->xu[d];
        const double *xp = c->xp[d];
        double *cdg = c->cdg[d];
        /* cDuR(-1:0, 1:N), cGpR(0:1, 0:N) */
        double *cDR = calloc((size_t)2 * (imax + 1), sizeof(double));
        double *cGR = calloc((size_t)2 * (imax + 1), sizeof(double));
#define CD(ii, i) cDR[((ii) + 1) + 2 * (i)]
#define CG(ii, i) cGR[(ii) + 2 * (i)]
#define XU(i) xu[(i) + 3]
#define XP(i) xp[(i) + 3]
        for (int i = 1; i <= imax; ++i) {
            CD(-1, i) = -1.0 / (XU(i) - XU(i - 1));
            CD(0, i) = 1.0 / (XU(i) - XU(i - 1));
        }
        for (int i = 1; i <= imax - 1; ++i) {
            CG(0, i) = -1.0 / (XP(i + 1) - XP(i));
            CG(1, i) = 1.0 / (XP(i + 1) - XP(i));
        }
        if (c->bc[d][0] > 0) {
            CG(0, 0) = 0.0;
            CG(1, 0) = 0.0;
            CD(-1, 1) = 0.0;
            CD(0, 1) = 0.0;
            /* global y?u(1) - y?u(0); single block: the same as the block coordinates */
            CD(0, 1) = 1.0 / (XU(1) - XU(0));
        } else {
            CG(0, 0) = -1.0 / (XP(1) - XP(0));
            CG(1, 0) = 1.0 / (XP(1) - XP(0));
        }
        if (c->bc[d][1] > 0) {
            CG(0, imax) = 0.0;
            CG(1, imax) = 0.0;
            CD(-1, imax) = 0.0;
            CD(0, imax) = 0.0;
            CD(-1, imax) = -1.0 / (XU(imax) - XU(imax - 1));
        } else {
            CG(0, imax) = -1.0 / (XP(imax + 1) - XP(imax));
            CG(1, imax) = 1.0 / (XP(imax + 1) - XP(imax));
        }
        for (int i = 1; i <= imax; ++i) {
            cdg[0 + 3 * (i - 1)] = CD(-1, i) * CG(0, i - 1);
            cdg[1 + 3 * (i - 1)] = CD(-1, i) * CG(1, i - 1) + CD(0, i) * CG(0, i);
            cdg[2 + 3 * (i - 1)] = CD(0, i) * CG(1, i);
        }
        if (c->bc[d][0] > 0)
            for (int q = 0; q < 3; ++q)
                cdg[q + 3 * 0] = 2.0 * cdg[q + 3 * 0];
        if (c->bc[d][1] > 0)
            for (int q = 0; q < 3; ++q)
                cdg[q + 3 * (imax - 1)] = 2.0 * cdg[q + 3 * (imax - 1)];
#undef CD
#undef CG
#undef XU
#undef XP
        free(cDR);
        free(cGR);
    }
}

/* FD_coeffs: src/mod_coeffs.f90:37-120 (only the arrays the hot path reads).
 * ncb?c = (4,5,5,7), ncb?g = (3,4,4,6), ncb?d = (4,4,6,6) (src/mod_dims.f90:64-84). */
static int fd_coeffs(orc_case *c)
{
    static const int ncbc[4] = {4, 5, 5, 7};
    static const int ncbg[4] = {3, 4, 4, 6};
    static const int ncbd[4] = {4, 4, 6, 6};
    int rc = 0;
    for (int d = 0; d < 3; ++d) {
        const int N = c->N[d];
        const int BL = c->bc[d][0], BU = c->bc[d][1];
        const double *xp = c->xp[d]; /* (-3:N+3) */
        const double *xu = c->xu[d];
        /* upwind: cNp?D/U (pressure grid), cN??D/U (velocity grid), n = -3..3 */
        rc |= diff_coeffs(1, -1, ncbc, 4, xp, BL, BU, 0, N, -3, 3, c->cNp[d][0]);
        rc |= diff_coeffs(1, 1, ncbc, 4, xp, BL, BU, 0, N, -3, 3, c->cNp[d][1]);
        rc |= diff_coeffs(1, -1, ncbc, 4, xu, BL, BU, 1, N, -3, 3, c->cNu[d][0]);
        rc |= diff_coeffs(1, 1, ncbc, 4, xu, BL, BU, 1, N, -3, 3, c->cNu[d][1]);
        /* second derivatives cp11/cu11, b = -3..3 */
        rc |= diff_coeffs(2, 0, ncbc, 4, xp, BL, BU, 0, N, -3, 3, c->cpp[d]);
        rc |= diff_coeffs(2, 0, ncbc, 4, xu, BL, BU, 1, N, -3, 3, c->cuu[d]);
        /* gradient and p->u interpolation: xC = x?p(g1L:N+g1U), g = -2..3 */
        rc |= diff_coeffs(1, 0, ncbg, 4, xp + 1, BL, BU, 2, N, -2, 3, c->cG[d]);
        rc |= diff_coeffs(0, 0, ncbg, 4, xp + 1, BL, BU, 2, N, -2, 3, c->cIpu[d]);
        /* divergence and u->p interpolation: xC = x?u(d1L:N+d1U), d = -3..2 */
        rc |= diff_coeffs(1, 0, ncbd, 4, xu, BL, BU, 3, N, -3, 2, c->cD[d]);
        rc |= diff_coeffs(0, 0, ncbd, 4, xu, BL, BU, 3, N, -3, 2, c->cIup[d]);
    }
    return rc;
}

orc_case *orc_case_create(const int N[3], const int bc[6], const double L[3], const double stretch[3])
{
    for (int d = 0; d < 3; ++d) {
        if (N[d] < 8)
            return NULL;
        for (int s = 0; s < 2; ++s)
            if (!(bc[2 * d + s] == -1 || bc[2 * d + s] == 1))
                return NULL;
        if ((bc[2 * d] == -1) != (bc[2 * d + 1] == -1))
            return NULL;
    }
    orc_case *c = calloc(1, sizeof(orc_case));
    for (int d = 0; d < 3; ++d) {
        c->N[d] = N[d];
        c->bc[d][0] = bc[2 * d];
        c->bc[d][1] = bc[2 * d + 1];
        c->L[d] = L[d];
        c->stretch[d] = stretch[d];
    }
    c->nx = N[0] + 7;
    c->ny = N[1] + 7;
    c->nz = N[2] + 7;
    c->nxy = (int64_t)c->nx * c->ny;
    c->ntot = c->nxy * c->nz;
    c->red_mode = 1;
    init_limits(c);
    for (int d = 0; d < 3; ++d) {
        const int n = N[d];
        c->xp[d] = calloc((size_t)n + 7, sizeof(double));
        c->xu[d] = calloc((size_t)n + 7, sizeof(double));
        for (int s = 0; s < 2; ++s) {
            c->cNp[d][s] = calloc((size_t)7 * (n + 1), sizeof(double));
            c->cNu[d][s] = calloc((size_t)7 * (n + 1), sizeof(double));
        }
        c->cpp[d] = calloc((size_t)7 * (n + 1), sizeof(double));
        c->cuu[d] = calloc((size_t)7 * (n + 1), sizeof(double));
        c->cG[d] = calloc((size_t)6 * (n + 1), sizeof(double));
        c->cIpu[d] = calloc((size_t)6 * (n + 1), sizeof(double));
        c->cD[d] = calloc((size_t)6 * (n + 1), sizeof(double));
        c->cIup[d] = calloc((size_t)6 * (n + 1), sizeof(double));
        c->cdg[d] = calloc((size_t)3 * n, sizeof(double));
    }
    init_coords(c);
    if (fd_coeffs(c) != 0) {
        orc_case_destroy(c);
        return NULL;
    }
    get_stencil(c);
    return c;
}

void orc_case_destroy(orc_case *c)
{
    if (!c)
        return;
    for (int d = 0; d < 3; ++d) {
        free(c->xp[d]);
        free(c->xu[d]);
        for (int s = 0; s < 2; ++s) {
            free(c->cNp[d][s]);
            free(c->cNu[d][s]);
        }
        free(c->cpp[d]);
        free(c->cuu[d]);
        free(c->cG[d]);
        free(c->cIpu[d]);
        free(c->cD[d]);
        free(c->cIup[d]);
        free(c->cdg[d]);
    }
    free(c);
}

int64_t orc_field_size(const orc_case *c) { return c->ntot; }

void orc_ranges(const orc_case *c, int out[48])
{
    int q = 0;
    for (int d = 0; d < 3; ++d) out[q++] = c->Sp[d];
    for (int d = 0; d < 3; ++d) out[q++] = c->Np[d];
    for (int m = 0; m < 3; ++m) for (int d = 0; d < 3; ++d) out[q++] = c->SB[m][d];
    for (int m = 0; m < 3; ++m) for (int d = 0; d < 3; ++d) out[q++] = c->NB[m][d];
    for (int m = 0; m < 3; ++m) for (int d = 0; d < 3; ++d) out[q++] = c->SI[m][d];
    for (int m = 0; m < 3; ++m) for (int d = 0; d < 3; ++d) out[q++] = c->NI[m][d];
    for (int d = 0; d < 3; ++d) out[q++] = c->SNF[d][0];
    for (int d = 0; d < 3; ++d) out[q++] = c->SNF[d][1];
}

void orc_coords(const orc_case *c, int d, int which, double *out)
{
    const double *src = which ? c->xu[d] : c->xp[d];
    memcpy(out, src, sizeof(double) * ((size_t)c->N[d] + 7));
}

int orc_coeff(const orc_case *c, const char *name, double *out)
{
    static const char dch[3] = {'1', '2', '3'};
    static const char *uname[3] = {"u1", "v2", "w3"};
    static const char *u2name[3] = {"u11", "v22", "w33"};
    for (int d = 0; d < 3; ++d) {
        const int n = c->N[d];
        char buf[16];
        const double *src = NULL;
        int len = 0;
        snprintf(buf, sizeof buf, "cGp%c", dch[d]);
        if (!strcmp(name, buf)) { src = c->cG[d]; len = 6 * (n + 1); }
        snprintf(buf, sizeof buf, "cD%s", uname[d]);
        if (!strcmp(name, buf)) { src = c->cD[d]; len = 6 * (n + 1); }
        snprintf(buf, sizeof buf, "cIp%c", "uvw"[d]);
        if (!strcmp(name, buf)) { src = c->cIpu[d]; len = 6 * (n + 1); }
        snprintf(buf, sizeof buf, "cI%cp", "uvw"[d]);
        if (!strcmp(name, buf)) { src = c->cIup[d]; len = 6 * (n + 1); }
        snprintf(buf, sizeof buf, "cNp%cD", dch[d]);
        if (!strcmp(name, buf)) { src = c->cNp[d][0]; len = 7 * (n + 1); }
        snprintf(buf, sizeof buf, "cNp%cU", dch[d]);
        if (!strcmp(name, buf)) { src = c->cNp[d][1]; len = 7 * (n + 1); }
        snprintf(buf, sizeof buf, "cN%sD", uname[d]);
        if (!strcmp(name, buf)) { src = c->cNu[d][0]; len = 7 * (n + 1); }
        snprintf(buf, sizeof buf, "cN%sU", uname[d]);
        if (!strcmp(name, buf)) { src = c->cNu[d][1]; len = 7 * (n + 1); }
        snprintf(buf, sizeof buf, "cp%c%c", dch[d], dch[d]);
        if (!strcmp(name, buf)) { src = c->cpp[d]; len = 7 * (n + 1); }
        snprintf(buf, sizeof buf, "c%s", u2name[d]);
        if (!strcmp(name, buf)) { src = c->cuu[d]; len = 7 * (n + 1); }
        snprintf(buf, sizeof buf, "cdg%c", dch[d]);
        if (!strcmp(name, buf)) { src = c->cdg[d]; len = 3 * n; }
        if (src) {
            memcpy(out, src, sizeof(double) * (size_t)len);
            return len;
        }
    }
    return -1;
}

void orc_set_reduction_mode(orc_case *c, int mode) { c->red_mode = mode; }

#ifdef ORC_OMP
#include <omp.h>
#endif
/* threads of the OpenMP build's point loops (the serial build ignores it) */
int orc_set_threads(int n)
{
#ifdef ORC_OMP
    if (n > 0) omp_set_num_threads(n);
    return omp_get_max_threads();
#else
    (void)n;
    return 1;
#endif
}
