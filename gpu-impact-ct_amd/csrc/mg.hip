// Multigrid V cycle (BiCGstab preconditioner 1) on the device (gfx950).
//
// multigridV(.TRUE., 1, bb, phi) (src/mod_solvers.f90:456-585) with its parts:
// relax_restrict (relaxation, product_div_grad_relax, restrict), relax_bottom
// and interpolate_relax (interpolate, relaxation).  The relaxation is the
// Jacobi branch of relaxation_div_grad / relaxation_div_grad_inv with omega
// 0.98 (src/mod_laplace.f90:466-597, 712-790) and runs as the out-of-place JOR
// sweep of kernels.hip on each level's pressure range; the reference's
// exchange_relax (src/mod_exchange.f90:1273-1480) becomes wrapped reads across
// periodic directions.  Every value is formed with the reference's operation
// order, so the cycle is bit-identical to the loop restatement.
//
// Level arrays follow the case layout; restriction and interpolation are one
// thread per output point (coarse point, fine point) reading through the
// caches: on the finest level they touch each fine value about once, and the
// coarse levels are 8x, 64x, ... smaller.
#include "ih_device.h"
#include "ih_mg.h"

#include <utility>

namespace ih {

using namespace dev;

namespace {

// index along d after exchange_relax: periodic directions wrap onto S?p..N?p
__device__ inline int wrapd(const Geo &g, int d, int idx) {
    if (!g.per[d]) return idx;
    const int P = g.Np[d] - g.Sp[d] + 1;
    if (idx < g.Sp[d]) return idx + P;
    if (idx > g.Np[d]) return idx - P;
    return idx;
}

__device__ inline int64_t wix(const Geo &g, int i, int j, int k) {
    return gidx(g, wrapd(g, 0, i), wrapd(g, 1, j), wrapd(g, 2, k));
}

// first relaxation of a level (r == 1 with init_yes): rel = omega bb / diag
__global__ void k_mg_init(Geo g, Coef c, Range3 r, double omega, const double *__restrict__ bb,
                          double *__restrict__ x, const int *active) {
    if (active && !*active) return;
    IH_POINT(r);
    const int64_t p = gidx(g, i, j, k);
    const double d = (c.dg[0][g.cst[0] + i] + c.dg[1][g.cst[1] + j]) + c.dg[2][g.cst[2] + k];
    x[p] = omega * bb[p] / d;
}

// product_div_grad_relax (src/mod_laplace.f90:270-280) on the range
__global__ void k_mg_lap(Geo g, Coef c, Range3 r, const double *__restrict__ x, double *__restrict__ lap,
                         const int *active) {
    if (active && !*active) return;
    IH_POINT(r);
    const int c0 = g.cst[0], c1 = g.cst[1], c2 = g.cst[2];
    const double *g1 = c.dg[0], *g2 = c.dg[1], *g3 = c.dg[2];
    double t = g1[i] * x[wix(g, i - 1, j, k)];
    t = t + g1[2 * c0 + i] * x[wix(g, i + 1, j, k)];
    t = t + g2[j] * x[wix(g, i, j - 1, k)];
    t = t + g2[2 * c1 + j] * x[wix(g, i, j + 1, k)];
    t = t + g3[k] * x[wix(g, i, j, k - 1)];
    t = t + g3[2 * c2 + k] * x[wix(g, i, j, k + 1)];
    const int64_t p = gidx(g, i, j, k);
    t = t + ((g1[c0 + i] + g2[c1 + j]) + g3[c2 + k]) * x[p];
    lap[p] = t;
}

// Averaging on the edges of wall pairs, the reference's statements in order:
// direction pairs (1,2), (1,3), (2,3), per pair the sides (L,L), (L,U), (U,L),
// (U,U), each a line over 1..N of the third direction.  three = 0: two
// neighbours / 2 (restrict, src/mod_solvers.f90:632-661); three = 1: three
// neighbours / 3 (interpolate, :1028-1043).  One block; a barrier between
// statements, whose lines share their corner points.
__global__ __launch_bounds__(256) void k_mg_edges(Geo g, double *f1, double *f2, int three, const int *active) {
    if (active && !*active) return;
    const int pairs[3][3] = {{0, 1, 2}, {0, 2, 1}, {1, 2, 0}};
    for (int pr = 0; pr < 3; ++pr) {
        const int p = pairs[pr][0], q = pairs[pr][1], t = pairs[pr][2];
        if (g.per[p] || g.per[q]) continue;
        for (int sp = 0; sp < 2; ++sp)
            for (int sq = 0; sq < 2; ++sq) {
                const int ep = sp ? g.N[p] : 1, ip = sp ? g.N[p] - 1 : 2;
                const int eq = sq ? g.N[q] : 1, iq = sq ? g.N[q] - 1 : 2;
                for (int s = 1 + (int)threadIdx.x; s <= g.N[t]; s += (int)blockDim.x) {
                    int e[3], a[3], b[3], cc[3];
                    e[p] = ep, e[q] = eq, e[t] = s;
                    a[p] = ip, a[q] = eq, a[t] = s;
                    b[p] = ep, b[q] = iq, b[t] = s;
                    cc[p] = ip, cc[q] = iq, cc[t] = s;
                    const int64_t pe = gidx(g, e[0], e[1], e[2]), pa = gidx(g, a[0], a[1], a[2]),
                                  pb = gidx(g, b[0], b[1], b[2]), pc = gidx(g, cc[0], cc[1], cc[2]);
                    for (int u = 0; u < 2; ++u) {
                        double *f = u ? f2 : f1;
                        if (!f) continue;
                        if (three)
                            f[pe] = (f[pa] + f[pb] + f[pc]) / 3.0;
                        else
                            f[pe] = (f[pa] + f[pb]) / 2.0;
                    }
                }
                __syncthreads();
            }
    }
}

struct RestrictW {
    const double *w[3];  // weights of the coarse level, (ii+1)*cst + ii
    int cst[3];
};

// restrict(.TRUE., g, coarse, bb, lap) (src/mod_solvers.f90:699-714): one
// thread per coarse point of [1, nmax]
__global__ void k_mg_restrict(Geo gf, Geo gc, RestrictW R, Range3 r, int dd0, int dd1, int dd2,
                              const double *__restrict__ bb, const double *__restrict__ lap,
                              double *__restrict__ out, const int *active) {
    if (active && !*active) return;
    const int ii = r.lo0 + (int)(blockIdx.x * BX + threadIdx.x);
    const int jj = r.lo1 + (int)(blockIdx.y * BY + threadIdx.y);
    const int kk = r.lo2 + (int)blockIdx.z;
    if (ii >= r.lo0 + r.n0 || jj >= r.lo1 + r.n1) return;
    const int i = dd0 * (ii - 1) + 1, j = dd1 * (jj - 1) + 1, k = dd2 * (kk - 1) + 1;
    auto df = [&](int a, int b, int c) {
        const int64_t p = wix(gf, a, b, c);
        return bb[p] - lap[p];
    };
    const double *w1 = R.w[0], *w2 = R.w[1], *w3 = R.w[2];
    const int s1 = R.cst[0], s2 = R.cst[1], s3 = R.cst[2];
    double t = ((w1[s1 + ii] + w2[s2 + jj]) + w3[s3 + kk]) * df(i, j, k);
    t = t + w1[ii] * df(i - 1, j, k);
    t = t + w1[2 * s1 + ii] * df(i + 1, j, k);
    t = t + w2[jj] * df(i, j - 1, k);
    t = t + w2[2 * s2 + jj] * df(i, j + 1, k);
    t = t + w3[kk] * df(i, j, k - 1);
    t = t + w3[2 * s3 + kk] * df(i, j, k + 1);
    out[gidx(gc, ii, jj, kk)] = t / 3.0;
}

struct InterpW {
    const double *w[3];  // weights of the fine level, (t-1)*cst + i
    int cst[3];
    int dn[3];
};

// interpolate(.TRUE., g, coarse, fine, work) (src/mod_solvers.f90:1046-1141):
// the injected coarse values, then the z, y and x passes, evaluated per fine
// point with the operands each pass reads, and added to fine over 1..N
__global__ void k_mg_interp(Geo gf, Geo gc, InterpW W, Range3 r, const double *__restrict__ coarse,
                            double *__restrict__ x, const int *active) {
    if (active && !*active) return;
    IH_POINT(r);
    auto lat = [&](int d, int t) { return W.dn[d] == 1 || ((t - 1) & 1) == 0; };
    auto cix = [&](int d, int t) { return (t - 1) / W.dn[d] + 1; };
    auto C = [&](int I, int J, int K) { return coarse[wix(gc, I, J, K)]; };
    // z pass (i, j on the coarse lattice)
    auto wz = [&](int a, int b, int c) {
        const int I = cix(0, a), J = cix(1, b);
        if (lat(2, c)) return C(I, J, cix(2, c));
        const int K = c / 2;
        return W.w[2][c] * C(I, J, K) + W.w[2][W.cst[2] + c] * C(I, J, K + 1);
    };
    // y pass (i on the lattice)
    auto wy = [&](int a, int b, int c) {
        if (lat(1, b)) return wz(a, b, c);
        return W.w[1][b] * wz(a, b - 1, c) + W.w[1][W.cst[1] + b] * wz(a, b + 1, c);
    };
    double v;
    if (lat(0, i))
        v = wy(i, j, k);
    else
        v = W.w[0][i] * wy(i - 1, j, k) + W.w[0][W.cst[0] + i] * wy(i + 1, j, k);
    const int64_t p = gidx(gf, i, j, k);
    x[p] = x[p] + v;
}

inline Range3 full_range(const Geo &g) { return Range3{1, 1, 1, g.N[0], g.N[1], g.N[2]}; }

}  // namespace

void launch_mg_vcycle(hipStream_t s, const MgDev &m, double *v, double *z, const int *act) {
    const int G = m.G;
    double *b[MG_LEVELS_MAX + 1], *x[MG_LEVELS_MAX + 1], *y[MG_LEVELS_MAX + 1];
    // level 1 alternates between z and D over the cycle's sweeps; starting on the
    // buffer that makes the last sweep write z leaves the result in place
    const int flips1 = G > 1 ? (m.n_down - 1) + m.n_up : (m.n_bottom - 1) + m.n_bottom;
    b[1] = v;
    x[1] = (flips1 % 2 == 0) ? z : m.L[1].D;
    y[1] = (x[1] == z) ? m.L[1].D : z;
    for (int g = 2; g <= G; ++g) {
        b[g] = m.L[g].A;
        x[g] = m.L[g].B;
        y[g] = m.L[g].D;
    }
    auto relax = [&](int g, bool init, int n) {
        const MgLevel &L = m.L[g];
        const Range3 r = make_range(L.pr);
        int q = 1;
        if (init && n >= 1) {
            hipLaunchKernelGGL(k_mg_init, grid_for(r), dim3(BX, BY, 1), 0, s, L.g, L.c, r, m.omega, b[g], x[g], act);
            q = 2;
        }
        for (; q <= n; ++q) {
            launch_jor(s, L.g, L.c, L.pr, m.omega, x[g], b[g], y[g], act, L.shape);
            std::swap(x[g], y[g]);
        }
    };
    if (G == 1) {
        relax(1, true, m.n_bottom);
        relax(1, false, m.n_bottom);
        return;
    }
    for (int g = 1; g <= G - 1; ++g) {
        const MgLevel &L = m.L[g], &Lc = m.L[g + 1];
        relax(g, true, m.n_down);
        const Range3 pr = make_range(L.pr);
        hipLaunchKernelGGL(k_mg_lap, grid_for(pr), dim3(BX, BY, 1), 0, s, L.g, L.c, pr, x[g], L.C, act);
        if (!(L.g.per[0] && L.g.per[1] && L.g.per[2]))
            hipLaunchKernelGGL(k_mg_edges, dim3(1), dim3(256), 0, s, L.g, b[g], L.C, 0, act);
        RestrictW R;
        int nmax[3];
        for (int d = 0; d < 3; ++d) {
            R.w[d] = Lc.cR[d];
            R.cst[d] = Lc.g.cst[d];
            nmax[d] = Lc.g.N[d] - (L.g.per[d] ? 1 : 0);
        }
        const Range3 cr{1, 1, 1, nmax[0], nmax[1], nmax[2]};
        hipLaunchKernelGGL(k_mg_restrict, grid_for(cr), dim3(BX, BY, 1), 0, s, L.g, Lc.g, R, cr, L.dn[0], L.dn[1],
                           L.dn[2], b[g], L.C, Lc.A, act);
    }
    relax(G, true, m.n_bottom);
    relax(G, false, m.n_bottom);
    for (int g = G - 1; g >= 1; --g) {
        const MgLevel &L = m.L[g], &Lc = m.L[g + 1];
        if (!(Lc.g.per[0] && Lc.g.per[1] && Lc.g.per[2]))
            hipLaunchKernelGGL(k_mg_edges, dim3(1), dim3(256), 0, s, Lc.g, x[g + 1], (double *)nullptr, 1, act);
        InterpW W;
        for (int d = 0; d < 3; ++d) {
            W.w[d] = L.cI[d];
            W.cst[d] = L.g.cst[d];
            W.dn[d] = L.dn[d];
        }
        const Range3 fr = full_range(L.g);
        hipLaunchKernelGGL(k_mg_interp, grid_for(fr), dim3(BX, BY, 1), 0, s, L.g, Lc.g, W, fr, x[g + 1], x[g], act);
        relax(g, false, m.n_up);
    }
}

}  // namespace ih
