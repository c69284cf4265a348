/*
 * Oracle (test infrastructure only): ghost fill, stencil operators and the
 * reductions.  Loop order and operation order follow the reference Fortran
 * cited per function.  Build with -ffp-contract=off.
 */
#include "orc_internal.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define A(p, i, j, k) (p)[ORC_IDX(c, i, j, k)]

/* ------------------------------------------------------------------------ */
/* exchange: src/mod_exchange.f90:47-294 (single block: BC -1 or > 0 only)   */
/* ------------------------------------------------------------------------ */
void orc_exchange(const orc_case *c, int dir, int vel_dir, double *phi)
{
    const int d = dir - 1;
    const int N1 = c->N[0], N2 = c->N[1], N3 = c->N[2];
    const int S = c->Sp[d], Nn = c->Np[d], N = c->N[d];
    if (c->bc[d][0] == -1) {
        /* lower ghosts from the upper interior, then upper ghosts from the lower interior */
        if (d == 0) {
            for (int k = 1; k <= N3; ++k)
                for (int j = 1; j <= N2; ++j) {
                    for (int s = -3; s <= -1; ++s) A(phi, S + s, j, k) = A(phi, Nn + 1 + s, j, k);
                    for (int s = 1; s <= 3; ++s) A(phi, Nn + s, j, k) = A(phi, S - 1 + s, j, k);
                }
        } else if (d == 1) {
            for (int k = 1; k <= N3; ++k)
                for (int i = 1; i <= N1; ++i) {
                    for (int s = -3; s <= -1; ++s) A(phi, i, S + s, k) = A(phi, i, Nn + 1 + s, k);
                    for (int s = 1; s <= 3; ++s) A(phi, i, Nn + s, k) = A(phi, i, S - 1 + s, k);
                }
        } else {
            for (int j = 1; j <= N2; ++j)
                for (int i = 1; i <= N1; ++i) {
                    for (int s = -3; s <= -1; ++s) A(phi, i, j, S + s) = A(phi, i, j, Nn + 1 + s);
                    for (int s = 1; s <= 3; ++s) A(phi, i, j, Nn + s) = A(phi, i, j, S - 1 + s);
                }
        }
    }
    /* wall ghosts: normal velocity keeps index 0, everything else zeroes 0 too */
    const int lo_end = (vel_dir == dir) ? -1 : 0;
    for (int side = 0; side < 2; ++side) {
        if (!(c->bc[d][side] > 0))
            continue;
        const int a = side == 0 ? -3 : N + 1;
        const int b = side == 0 ? lo_end : N + 3;
        for (int s = a; s <= b; ++s) {
            if (d == 0) {
                for (int k = 1; k <= N3; ++k)
                    for (int j = 1; j <= N2; ++j) A(phi, s, j, k) = 0.0;
            } else if (d == 1) {
                for (int k = 1; k <= N3; ++k)
                    for (int i = 1; i <= N1; ++i) A(phi, i, s, k) = 0.0;
            } else {
                for (int j = 1; j <= N2; ++j)
                    for (int i = 1; i <= N1; ++i) A(phi, i, j, s) = 0.0;
            }
        }
    }
}

/* exchange_all_all(.TRUE., vel): src/mod_exchange.f90:565-817 -- per direction,
 * every component is copied; the component normal to the wall keeps index 0. */
void orc_exchange_all_all(const orc_case *c, double *vel)
{
    for (int dir = 1; dir <= 3; ++dir)
        for (int m = 1; m <= 3; ++m)
            orc_exchange(c, dir, m, vel + (int64_t)(m - 1) * c->ntot);
}

/* exchange_relax(1,0,0,0,0,...): src/mod_exchange.f90:1273-1480, periodic copy
 * of one layer, S?R-1 <- N?R and N?R+1 <- S?R (walls untouched). */
void orc_exchange_relax(const orc_case *c, double *phi)
{
    const int N1 = c->N[0], N2 = c->N[1], N3 = c->N[2];
    if (c->bc[0][0] == -1) {
        const int S = c->Sp[0], Nn = c->Np[0];
        for (int k = 1; k <= N3; ++k)
            for (int j = 1; j <= N2; ++j) {
                A(phi, S - 1, j, k) = A(phi, Nn, j, k);
                A(phi, Nn + 1, j, k) = A(phi, S, j, k);
            }
    }
    if (c->bc[1][0] == -1) {
        const int S = c->Sp[1], Nn = c->Np[1];
        for (int k = 1; k <= N3; ++k)
            for (int i = 1; i <= N1; ++i) {
                A(phi, i, S - 1, k) = A(phi, i, Nn, k);
                A(phi, i, Nn + 1, k) = A(phi, i, S, k);
            }
    }
    if (c->bc[2][0] == -1) {
        const int S = c->Sp[2], Nn = c->Np[2];
        for (int j = 1; j <= N2; ++j)
            for (int i = 1; i <= N1; ++i) {
                A(phi, i, j, S - 1) = A(phi, i, j, Nn);
                A(phi, i, j, Nn + 1) = A(phi, i, j, S);
            }
    }
}

/* ------------------------------------------------------------------------ */
/* divergence2: src/mod_diff.f90:130-215                                     */
/* ------------------------------------------------------------------------ */
void orc_divergence2(const orc_case *c, double *vel, double *div)
{
    double *u = vel, *v = vel + c->ntot, *w = vel + 2 * c->ntot;
    orc_exchange(c, 1, 1, u);
    orc_exchange(c, 2, 2, v);
    orc_exchange(c, 3, 3, w);
    const double *c1 = c->cD[0], *c2 = c->cD[1], *c3 = c->cD[2];
    ORC_PAR
    for (int k = c->Sp[2]; k <= c->Np[2]; ++k)
        for (int j = c->Sp[1]; j <= c->Np[1]; ++j)
            for (int i = c->Sp[0]; i <= c->Np[0]; ++i) {
                double s = CD6(c1, -3, i) * A(u, i - 3, j, k);
                for (int ii = -2; ii <= 2; ++ii) s = s + CD6(c1, ii, i) * A(u, i + ii, j, k);
                for (int jj = -3; jj <= 2; ++jj) s = s + CD6(c2, jj, j) * A(v, i, j + jj, k);
                for (int kk = -3; kk <= 2; ++kk) s = s + CD6(c3, kk, k) * A(w, i, j, k + kk);
                A(div, i, j, k) = s;
            }
}

/* gradient: src/mod_diff.f90:309-414 */
void orc_gradient(const orc_case *c, int m, double *phi, double *grad)
{
    const int q = m - 1;
    orc_exchange(c, m, 0, phi);
    const double *cg = c->cG[q];
    ORC_PAR
    for (int k = c->SI[q][2]; k <= c->NI[q][2]; ++k)
        for (int j = c->SI[q][1]; j <= c->NI[q][1]; ++j)
            for (int i = c->SI[q][0]; i <= c->NI[q][0]; ++i) {
                double s;
                if (q == 0) {
                    s = CG6(cg, -2, i) * A(phi, i - 2, j, k);
                    for (int ii = -1; ii <= 3; ++ii) s = s + A(phi, i + ii, j, k) * CG6(cg, ii, i);
                } else if (q == 1) {
                    s = CG6(cg, -2, j) * A(phi, i, j - 2, k);
                    for (int jj = -1; jj <= 3; ++jj) s = s + CG6(cg, jj, j) * A(phi, i, j + jj, k);
                } else {
                    s = CG6(cg, -2, k) * A(phi, i, j, k - 2);
                    for (int kk = -1; kk <= 3; ++kk) s = s + CG6(cg, kk, k) * A(phi, i, j, k + kk);
                }
                A(grad, i, j, k) = s;
            }
    /* wall planes of the B range: index 0 for the normal direction, 1 otherwise */
    for (int e = 0; e < 3; ++e) {
        for (int side = 0; side < 2; ++side) {
            if (!(c->bc[e][side] > 0))
                continue;
            const int plane = side == 0 ? (e == q ? 0 : 1) : c->N[e];
            int lo[3], hi[3];
            for (int f = 0; f < 3; ++f) {
                lo[f] = c->SB[q][f];
                hi[f] = c->NB[q][f];
            }
            lo[e] = hi[e] = plane;
            for (int k = lo[2]; k <= hi[2]; ++k)
                for (int j = lo[1]; j <= hi[1]; ++j)
                    for (int i = lo[0]; i <= hi[0]; ++i) A(grad, i, j, k) = 0.0;
        }
    }
}

/* interpolate2_vel_pre (exch_yes = .FALSE.): src/mod_diff.f90:1740-1825 */
void orc_interpolate2_vel_pre(const orc_case *c, int m, double *phi, double *inter)
{
    const int q = m - 1;
    const double *ci = c->cIup[q];
    ORC_PAR
    for (int k = c->Sp[2]; k <= c->Np[2]; ++k)
        for (int j = c->Sp[1]; j <= c->Np[1]; ++j)
            for (int i = c->Sp[0]; i <= c->Np[0]; ++i) {
                double s;
                if (q == 0) {
                    s = CD6(ci, -3, i) * A(phi, i - 3, j, k);
                    for (int ii = -2; ii <= 2; ++ii) s = s + CD6(ci, ii, i) * A(phi, i + ii, j, k);
                } else if (q == 1) {
                    s = CD6(ci, -3, j) * A(phi, i, j - 3, k);
                    for (int jj = -2; jj <= 2; ++jj) s = s + CD6(ci, jj, j) * A(phi, i, j + jj, k);
                } else {
                    s = CD6(ci, -3, k) * A(phi, i, j, k - 3);
                    for (int kk = -2; kk <= 2; ++kk) s = s + CD6(ci, kk, k) * A(phi, i, j, k + kk);
                }
                A(inter, i, j, k) = s;
            }
}

/* interpolate2_pre_vel (exch_yes = .FALSE.): src/mod_diff.f90:1567-1648 */
void orc_interpolate2_pre_vel(const orc_case *c, int m, double *phi, double *inter)
{
    const int q = m - 1;
    const double *ci = c->cIpu[q];
    ORC_PAR
    for (int k = c->SI[q][2]; k <= c->NI[q][2]; ++k)
        for (int j = c->SI[q][1]; j <= c->NI[q][1]; ++j)
            for (int i = c->SI[q][0]; i <= c->NI[q][0]; ++i) {
                double s;
                if (q == 0) {
                    s = CG6(ci, -2, i) * A(phi, i - 2, j, k);
                    for (int ii = -1; ii <= 3; ++ii) s = s + CG6(ci, ii, i) * A(phi, i + ii, j, k);
                } else if (q == 1) {
                    s = CG6(ci, -2, j) * A(phi, i, j - 2, k);
                    for (int jj = -1; jj <= 3; ++jj) s = s + CG6(ci, jj, j) * A(phi, i, j + jj, k);
                } else {
                    s = CG6(ci, -2, k) * A(phi, i, j, k - 2);
                    for (int kk = -1; kk <= 3; ++kk) s = s + CG6(ci, kk, k) * A(phi, i, j, k + kk);
                }
                A(inter, i, j, k) = s;
            }
}

/* interpolate_vel(.FALSE.): src/mod_diff.f90:1436-1472 */
void orc_interpolate_vel(const orc_case *c, double *vel, double *work)
{
    for (int m = 1; m <= 3; ++m)
        orc_interpolate2_vel_pre(c, m, vel + (int64_t)(m - 1) * c->ntot, work + (int64_t)(m - 1) * c->ntot);
    for (int m = 1; m <= 3; ++m)
        for (int dir = 1; dir <= 3; ++dir)
            orc_exchange(c, dir, 0, work + (int64_t)(m - 1) * c->ntot);
}

/* nonlinear(.FALSE.) with upwind_yes: src/mod_diff.f90:784-1424.
 * For component m and direction d: pp = I_m(work_d); the upwind-biased stencil
 * (cN..U if pp >= 0, cN..D otherwise) of vel_m along d is scaled by pp and
 * accumulated into nl_m (assigned for d = 1). */
void orc_nonlinear(const orc_case *c, double *vel, double *work, double *nl, double *pp)
{
    for (int q = 0; q < 3; ++q) {
        double *vm = vel + (int64_t)q * c->ntot;
        double *nm = nl + (int64_t)q * c->ntot;
        for (int e = 0; e < 3; ++e) {
            orc_interpolate2_pre_vel(c, q + 1, work + (int64_t)e * c->ntot, pp);
            const double *cU = (e == q) ? c->cNu[e][1] : c->cNp[e][1];
            const double *cDn = (e == q) ? c->cNu[e][0] : c->cNp[e][0];
            ORC_PAR
            for (int k = c->SI[q][2]; k <= c->NI[q][2]; ++k)
                for (int j = c->SI[q][1]; j <= c->NI[q][1]; ++j)
                    for (int i = c->SI[q][0]; i <= c->NI[q][0]; ++i) {
                        const double p = A(pp, i, j, k);
                        const double *cc = (p >= 0.0) ? cU : cDn;
                        const int idx = e == 0 ? i : (e == 1 ? j : k);
                        double dd1;
                        if (e == 0) {
                            dd1 = C7(cc, -3, idx) * A(vm, i - 3, j, k);
                            for (int s = -2; s <= 3; ++s) dd1 = dd1 + C7(cc, s, idx) * A(vm, i + s, j, k);
                        } else if (e == 1) {
                            dd1 = C7(cc, -3, idx) * A(vm, i, j - 3, k);
                            for (int s = -2; s <= 3; ++s) dd1 = dd1 + C7(cc, s, idx) * A(vm, i, j + s, k);
                        } else {
                            dd1 = C7(cc, -3, idx) * A(vm, i, j, k - 3);
                            for (int s = -2; s <= 3; ++s) dd1 = dd1 + C7(cc, s, idx) * A(vm, i, j, k + s);
                        }
                        if (e == 0)
                            A(nm, i, j, k) = dd1 * p;
                        else
                            A(nm, i, j, k) = A(nm, i, j, k) + dd1 * p;
                    }
        }
    }
}

/* Helmholtz_explicit: src/mod_diff.f90:549-762 */
void orc_helmholtz_explicit(const orc_case *c, double multL, double *vel, double *nl)
{
    for (int q = 0; q < 3; ++q) {
        const double *vm = vel + (int64_t)q * c->ntot;
        double *nm = nl + (int64_t)q * c->ntot;
        const double *c1 = (q == 0) ? c->cuu[0] : c->cpp[0];
        const double *c2 = (q == 1) ? c->cuu[1] : c->cpp[1];
        const double *c3 = (q == 2) ? c->cuu[2] : c->cpp[2];
        ORC_PAR
        for (int k = c->SI[q][2]; k <= c->NI[q][2]; ++k)
            for (int j = c->SI[q][1]; j <= c->NI[q][1]; ++j)
                for (int i = c->SI[q][0]; i <= c->NI[q][0]; ++i) {
                    double dd1 = C7(c1, -3, i) * A(vm, i - 3, j, k);
                    for (int s = -2; s <= 3; ++s) dd1 = dd1 + C7(c1, s, i) * A(vm, i + s, j, k);
                    for (int s = -3; s <= 3; ++s) dd1 = dd1 + C7(c2, s, j) * A(vm, i, j + s, k);
                    for (int s = -3; s <= 3; ++s) dd1 = dd1 + C7(c3, s, k) * A(vm, i, j, k + s);
                    A(nm, i, j, k) = A(nm, i, j, k) - multL * dd1;
                }
    }
}

/* product_div_grad: src/mod_laplace.f90:38-197 = gradient(m) then divergence(m)
 * (src/mod_diff.f90:52-118) for m = 1..3, accumulating into lap. */
void orc_product_div_grad(const orc_case *c, double *phi, double *lap, double *dig)
{
    for (int m = 1; m <= 3; ++m) {
        orc_gradient(c, m, phi, dig);
        orc_exchange(c, m, m, dig);
        const double *cd = c->cD[m - 1];
        ORC_PAR
        for (int k = c->Sp[2]; k <= c->Np[2]; ++k)
            for (int j = c->Sp[1]; j <= c->Np[1]; ++j)
                for (int i = c->Sp[0]; i <= c->Np[0]; ++i) {
                    if (m == 1) {
                        double s = CD6(cd, -3, i) * A(dig, i - 3, j, k);
                        for (int ii = -2; ii <= 2; ++ii) s = s + CD6(cd, ii, i) * A(dig, i + ii, j, k);
                        A(lap, i, j, k) = s;
                    } else if (m == 2) {
                        double s = A(lap, i, j, k);
                        for (int jj = -3; jj <= 2; ++jj) s = s + CD6(cd, jj, j) * A(dig, i, j + jj, k);
                        A(lap, i, j, k) = s;
                    } else {
                        double s = A(lap, i, j, k);
                        for (int kk = -3; kk <= 2; ++kk) s = s + CD6(cd, kk, k) * A(dig, i, j, k + kk);
                        A(lap, i, j, k) = s;
                    }
                }
    }
}

/* ------------------------------------------------------------------------ */
/* reductions                                                                */
/* ------------------------------------------------------------------------ */
/* Fixed tree of the device kernels: threads of a 16x16 tile (lane tx + 16 ty)
 * each add the planes of their z chunk sequentially, the k range being cut
 * into ceil(nK / ORC_RED_KC) chunks of balanced length (chunk b holds planes
 * ss + b*nK/GZ .. ss + (b+1)*nK/GZ - 1); each wavefront of 64 lanes combines by
 * an xor butterfly (offsets 32..1), the four wave sums as (w0+w1)+(w2+w3); the
 * block partials are then added by 256 threads striding over them and the same
 * block tree. */
static double block_tree(double *a)
{
    double w[4];
    for (int q = 0; q < 4; ++q) {
        double t[64], u[64];
        memcpy(t, a + 64 * q, sizeof t);
        for (int off = 32; off >= 1; off >>= 1) {
            for (int l = 0; l < 64; ++l) u[l] = t[l] + t[l ^ off];
            memcpy(t, u, sizeof t);
        }
        w[q] = t[0];
    }
    return (w[0] + w[1]) + (w[2] + w[3]);
}

double orc_reduce_sum(const orc_case *c, const int ss[3], const int nn[3], orc_point_fn f, const void *ctx)
{
    if (c->red_mode == 0) {
        double s = 0.0;
        for (int k = ss[2]; k <= nn[2]; ++k)
            for (int j = ss[1]; j <= nn[1]; ++j)
                for (int i = ss[0]; i <= nn[0]; ++i) s = s + f(ctx, ORC_IDX(c, i, j, k));
        return s;
    }
    const int nI = nn[0] - ss[0] + 1, nJ = nn[1] - ss[1] + 1, nK = nn[2] - ss[2] + 1;
    if (nI <= 0 || nJ <= 0 || nK <= 0)
        return 0.0;
    const int GX = (nI + ORC_RED_TX - 1) / ORC_RED_TX;
    const int GY = (nJ + ORC_RED_TY - 1) / ORC_RED_TY;
    const int GZ = (nK + ORC_RED_KC - 1) / ORC_RED_KC;
    const int64_t nb = (int64_t)GX * GY * GZ;
    double *part = malloc(sizeof(double) * (size_t)nb);
    double tile[ORC_RED_THREADS];
    for (int bz = 0; bz < GZ; ++bz)
        for (int by = 0; by < GY; ++by)
            for (int bx = 0; bx < GX; ++bx) {
                for (int ty = 0; ty < ORC_RED_TY; ++ty)
                    for (int tx = 0; tx < ORC_RED_TX; ++tx) {
                        const int i = ss[0] + bx * ORC_RED_TX + tx;
                        const int j = ss[1] + by * ORC_RED_TY + ty;
                        double acc = 0.0;
                        if (i <= nn[0] && j <= nn[1]) {
                            const int k0 = ss[2] + (int)(((int64_t)bz * nK) / GZ);
                            const int k1 = ss[2] + (int)(((int64_t)(bz + 1) * nK) / GZ) - 1;
                            for (int k = k0; k <= k1; ++k) acc = acc + f(ctx, ORC_IDX(c, i, j, k));
                        }
                        tile[tx + ORC_RED_TX * ty] = acc;
                    }
                part[bx + (int64_t)GX * (by + (int64_t)GY * bz)] = block_tree(tile);
            }
    for (int t = 0; t < ORC_RED_THREADS; ++t) {
        double acc = 0.0;
        for (int64_t q = t; q < nb; q += ORC_RED_THREADS) acc = acc + part[q];
        tile[t] = acc;
    }
    free(part);
    return block_tree(tile);
}

double orc_reduce_max(const orc_case *c, const int ss[3], const int nn[3], orc_point_fn f, const void *ctx)
{
    double m = 0.0;
    for (int k = ss[2]; k <= nn[2]; ++k)
        for (int j = ss[1]; j <= nn[1]; ++j)
            for (int i = ss[0]; i <= nn[0]; ++i) {
                const double v = f(ctx, ORC_IDX(c, i, j, k));
                m = (v > m || v != v) ? v : m; /* a NaN wins */
            }
    return m;
}

typedef struct {
    const double *a, *b;
} dot_ctx;
static double dot_pt(const void *ctx, int64_t idx)
{
    const dot_ctx *d = ctx;
    return d->a[idx] * d->b[idx];
}

double orc_dot(const orc_case *c, const int ss[3], const int nn[3], const double *a, const double *b)
{
    dot_ctx d = {a, b};
    return orc_reduce_sum(c, ss, nn, dot_pt, &d);
}
