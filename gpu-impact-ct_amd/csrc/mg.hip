// Multigrid V cycle (BiCGstab preconditioner 1) on the device (gfx950).
//
// multigridV(.TRUE., 1, bb, phi) (src/mod_solvers.f90:456-585) with its parts:
// relax_restrict (relaxation, product_div_grad_relax, restrict), relax_bottom
// and interpolate_relax (interpolate, relaxation).  The relaxation is the
// Jacobi branch of relaxation_div_grad / relaxation_div_grad_inv with omega
// 0.98 (src/mod_laplace.f90:466-597, 712-790) on each level's pressure range;
// the reference's exchange_relax (src/mod_exchange.f90:1273-1480) becomes
// wrapped reads across periodic directions.  Every value is formed with the
// reference's operation order, so the cycle is bit-identical to the loop
// restatement.
//
// Large levels run one launch per step (the JOR sweep of kernels.hip, and one
// thread per output point for the rest).  A launch of a few-microsecond
// kernel costs about as much as the work of a small level, so every level of
// at most MG_FUSED_POINTS points runs inside one single-workgroup kernel that
// walks the rest of the cycle with workgroup barriers between its steps.
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

// ---- per-point steps (T: the arithmetic type of the cycle) -----------------

// first relaxation of a level (r == 1 with init_yes): rel = omega bb / diag
// The per-point steps take their field pointers and coefficient tables as
// template types: plain pointers from the launched kernels, global-qualified
// ones from the fused kernel, which reads them from its level table (as plain
// pointers the compiler could not tell they address global memory and would
// emit flat accesses, which wait on both the vector-memory and LDS counters).
template <typename T>
using gptr = __attribute__((address_space(1))) T *;
template <typename T>
using gcptr = const __attribute__((address_space(1))) T *;

template <typename T, typename TAB, typename PC, typename PM>
__device__ inline void pt_init(const Geo &g, TAB dg, int i, int j, int k, T omega, PC bb, PM x) {
    const int64_t p = gidx(g, i, j, k);
    const T d = (dg[0][g.cst[0] + i] + dg[1][g.cst[1] + j]) + dg[2][g.cst[2] + k];
    x[p] = omega * bb[p] / d;
}

// one Jacobi sweep point (src/mod_laplace.f90:514-518), out of place
template <typename T, typename TAB, typename PC, typename PM>
__device__ inline void pt_jor(const Geo &g, TAB dg, int i, int j, int k, T omega, PC x, PC bb, PM y) {
    const int c0 = g.cst[0], c1 = g.cst[1], c2 = g.cst[2];
    const int64_t p = gidx(g, i, j, k);
    T t = bb[p];
    t = t - dg[0][i] * x[wix(g, i - 1, j, k)];
    t = t - dg[0][2 * c0 + i] * x[wix(g, i + 1, j, k)];
    t = t - dg[1][j] * x[wix(g, i, j - 1, k)];
    t = t - dg[1][2 * c1 + j] * x[wix(g, i, j + 1, k)];
    t = t - dg[2][k] * x[wix(g, i, j, k - 1)];
    t = t - dg[2][2 * c2 + k] * x[wix(g, i, j, k + 1)];
    const T diag = (dg[0][c0 + i] + dg[1][c1 + j]) + dg[2][c2 + k];
    y[p] = omega * t / diag + (T(1) - omega) * x[p];
}

// product_div_grad_relax (src/mod_laplace.f90:270-280)
template <typename T, typename TAB, typename PC, typename PM>
__device__ inline void pt_lap(const Geo &g, TAB dg, int i, int j, int k, PC x, PM lap) {
    const int c0 = g.cst[0], c1 = g.cst[1], c2 = g.cst[2];
    T t = dg[0][i] * x[wix(g, i - 1, j, k)];
    t = t + dg[0][2 * c0 + i] * x[wix(g, i + 1, j, k)];
    t = t + dg[1][j] * x[wix(g, i, j - 1, k)];
    t = t + dg[1][2 * c1 + j] * x[wix(g, i, j + 1, k)];
    t = t + dg[2][k] * x[wix(g, i, j, k - 1)];
    t = t + dg[2][2 * c2 + k] * x[wix(g, i, j, k + 1)];
    const int64_t p = gidx(g, i, j, k);
    t = t + ((dg[0][c0 + i] + dg[1][c1 + j]) + dg[2][c2 + k]) * x[p];
    lap[p] = t;
}

// restrict(.TRUE., g, coarse, bb, lap) at coarse point (ii, jj, kk)
// (src/mod_solvers.f90:699-714); w: the coarse level's weights
template <typename T, typename TAB, typename PC, typename PM>
__device__ inline T pt_restrict(const Geo &gf, const Geo &gc, TAB w, const int *dd, int ii, int jj, int kk, PC bb,
                                PC lap, PM out) {
    const int i = dd[0] * (ii - 1) + 1, j = dd[1] * (jj - 1) + 1, k = dd[2] * (kk - 1) + 1;
    auto df = [&](int a, int b, int c) {
        const int64_t p = wix(gf, a, b, c);
        return bb[p] - lap[p];
    };
    const int s1 = gc.cst[0], s2 = gc.cst[1], s3 = gc.cst[2];
    T t = ((w[0][s1 + ii] + w[1][s2 + jj]) + w[2][s3 + kk]) * df(i, j, k);
    t = t + w[0][ii] * df(i - 1, j, k);
    t = t + w[0][2 * s1 + ii] * df(i + 1, j, k);
    t = t + w[1][jj] * df(i, j - 1, k);
    t = t + w[1][2 * s2 + jj] * df(i, j + 1, k);
    t = t + w[2][kk] * df(i, j, k - 1);
    t = t + w[2][2 * s3 + kk] * df(i, j, k + 1);
    const T a = t / T(3);
    out[gidx(gc, ii, jj, kk)] = a;
    return a;
}

// interpolate(.TRUE., g, coarse, fine, work) at fine point (i, j, k)
// (src/mod_solvers.f90:1046-1141): the injected coarse values, then the z, y
// and x passes, evaluated with the operands each pass reads, added to fine;
// w: the fine level's weights, dn: coarsening ratios
template <typename T, typename TAB, typename PC, typename PM>
__device__ inline void pt_interp(const Geo &gf, const Geo &gc, TAB w, const int *dn, int i, int j, int k, PC coarse,
                                 PM x) {
    auto lat = [&](int d, int t) { return dn[d] == 1 || ((t - 1) & 1) == 0; };
    auto cix = [&](int d, int t) { return (t - 1) / dn[d] + 1; };
    auto C = [&](int I, int J, int K) { return coarse[wix(gc, I, J, K)]; };
    auto wz = [&](int a, int b, int c) {  // z pass (i, j on the coarse lattice)
        const int I = cix(0, a), J = cix(1, b);
        if (lat(2, c)) return C(I, J, cix(2, c));
        const int K = c / 2;
        return w[2][c] * C(I, J, K) + w[2][gf.cst[2] + c] * C(I, J, K + 1);
    };
    auto wy = [&](int a, int b, int c) {  // y pass (i on the lattice)
        if (lat(1, b)) return wz(a, b, c);
        return w[1][b] * wz(a, b - 1, c) + w[1][gf.cst[1] + b] * wz(a, b + 1, c);
    };
    T v;
    if (lat(0, i))
        v = wy(i, j, k);
    else
        v = w[0][i] * wy(i - 1, j, k) + w[0][gf.cst[0] + i] * wy(i + 1, j, k);
    const int64_t p = gidx(gf, i, j, k);
    x[p] = x[p] + v;
}

// Averaging on the edges of wall pairs, the reference's statements in order:
// direction pairs (1,2), (1,3), (2,3), per pair the sides (L,L), (L,U), (U,L),
// (U,U), each a line over 1..N of the third direction.  three = 0: two
// neighbours / 2 (restrict, src/mod_solvers.f90:632-661); three = 1: three
// neighbours / 3 (interpolate, :1028-1043).  Lines of one pair share no point
// when every N >= 3, so a pair is one parallel step; pairs are ordered by
// barriers (a later pair reads corners an earlier one wrote).  Runs on one
// workgroup of nthr threads (tid); f2 may be null.
template <typename T, typename PM = T *>
__device__ inline void wg_edges(const Geo &g, PM f1, PM f2, int three, int tid, int nthr) {
    const int pairs[3][3] = {{0, 1, 2}, {0, 2, 1}, {1, 2, 0}};
    for (int pr = 0; pr < 3; ++pr) {
        const int p = pairs[pr][0], q = pairs[pr][1], t = pairs[pr][2];
        if (g.per[p] || g.per[q]) continue;
        const int nt = g.N[t];
        for (int u = tid; u < 4 * nt; u += nthr) {
            const int side = u / nt, s = 1 + u - side * nt;
            const int sp = side >> 1, sq = side & 1;
            const int ep = sp ? g.N[p] : 1, ip = sp ? g.N[p] - 1 : 2;
            const int eq = sq ? g.N[q] : 1, iq = sq ? g.N[q] - 1 : 2;
            int e[3], a[3], b[3], cc[3];
            e[p] = ep, e[q] = eq, e[t] = s;
            a[p] = ip, a[q] = eq, a[t] = s;
            b[p] = ep, b[q] = iq, b[t] = s;
            cc[p] = ip, cc[q] = iq, cc[t] = s;
            const int64_t pe = gidx(g, e[0], e[1], e[2]), pa = gidx(g, a[0], a[1], a[2]),
                          pb = gidx(g, b[0], b[1], b[2]), pc = gidx(g, cc[0], cc[1], cc[2]);
            // both fields' loads before either store: the loads of a pair never read a
            // point that the pair writes, and f1 and f2 are different fields
            T s1 = T(0), s2 = T(0);
            if (three) {
                if (f1) s1 = (f1[pa] + f1[pb] + f1[pc]) / T(3);
                if (f2) s2 = (f2[pa] + f2[pb] + f2[pc]) / T(3);
            } else {
                if (f1) s1 = (f1[pa] + f1[pb]) / T(2);
                if (f2) s2 = (f2[pa] + f2[pb]) / T(2);
            }
            if (f1) f1[pe] = s1;
            if (f2) f2[pe] = s2;
        }
        __syncthreads();
    }
}

template <typename T>
__device__ inline const T *const *dg_of(const Coef &c);
template <>
__device__ inline const double *const *dg_of<double>(const Coef &c) {
    return c.dg;
}
template <>
__device__ inline const float *const *dg_of<float>(const Coef &c) {
    return c.dgf;
}

// ---- one launch per step (large levels) ------------------------------------

template <typename T>
__global__ void k_mg_init(Geo g, Coef c, Range3 r, T omega, const T *__restrict__ bb, T *__restrict__ x,
                          const int *active) {
    if (active && !*active) return;
    IH_POINT(r);
    pt_init<T>(g, dg_of<T>(c), i, j, k, omega, bb, x);
}

// the FP32 cycle's first relaxation on level 1, from the FP64 right-hand side:
// b32 = (float) v is stored for the later steps
__global__ void k_mg_init_f32_first(Geo g, Coef c, Range3 r, float omega, const double *__restrict__ v,
                                    float *__restrict__ b32, float *__restrict__ x, const int *active) {
    if (active && !*active) return;
    IH_POINT(r);
    const int64_t p = gidx(g, i, j, k);
    b32[p] = (float)v[p];
    pt_init<float>(g, c.dgf, i, j, k, omega, b32, x);
}


// one workgroup of MG_EDGE_THREADS: the longest line set of the bench grid
// (4 x 257 points) takes two passes
constexpr int MG_EDGE_THREADS = 1024;

template <typename T>
__global__ __launch_bounds__(MG_EDGE_THREADS) void k_mg_edges(Geo g, T *f1, T *f2, int three, const int *active) {
    if (active && !*active) return;
    wg_edges<T>(g, f1, f2, three, (int)threadIdx.x, (int)blockDim.x);
}

template <typename T>
struct LevelW {
    const T *w[3];
    int d[3];
};

// restriction; with xinit the coarse level's first relaxation as well
// (omega A / diag at the same point, the value pt_init forms)
template <typename T>
__global__ void k_mg_restrict(Geo gf, Geo gc, LevelW<T> R, Coef cc, Range3 r, const T *__restrict__ bb,
                              const T *__restrict__ lap, T *__restrict__ out, T *__restrict__ xinit, T omega,
                              const int *active) {
    if (active && !*active) return;
    const int ii = r.lo0 + (int)(blockIdx.x * BX + threadIdx.x);
    const int jj = r.lo1 + (int)(blockIdx.y * BY + threadIdx.y);
    const int kk = r.lo2 + (int)blockIdx.z;
    if (ii >= r.lo0 + r.n0 || jj >= r.lo1 + r.n1) return;
    const T a = pt_restrict<T>(gf, gc, R.w, R.d, ii, jj, kk, bb, lap, out);
    if (xinit) {
        const T *const *dg = dg_of<T>(cc);
        const T d = (dg[0][gc.cst[0] + ii] + dg[1][gc.cst[1] + jj]) + dg[2][gc.cst[2] + kk];
        xinit[gidx(gc, ii, jj, kk)] = omega * a / d;
    }
}

// interpolate-and-add by coarse cells: the thread of cell (I, J, K) writes the
// fine points dn*(T-1)+1+s, s < dn (s = 0 only at T = Nc), from the eight
// coarse values around it; every fine value is formed with pt_interp's
// operands and order
template <typename T>
__global__ void k_mg_interp_cells(Geo gf, Geo gc, LevelW<T> W, const T *__restrict__ coarse, T *__restrict__ x,
                                  const int *active) {
    if (active && !*active) return;
    const int I = 1 + (int)(blockIdx.x * BX + threadIdx.x);
    const int J = 1 + (int)(blockIdx.y * BY + threadIdx.y);
    const int K = 1 + (int)blockIdx.z;
    if (I > gc.N[0] || J > gc.N[1]) return;
    const int Tc[3] = {I, J, K};
    int ns[3];
    for (int d = 0; d < 3; ++d) ns[d] = (W.d[d] == 2 && Tc[d] < gc.N[d]) ? 2 : 1;
    // coarse values at (I + a, J + b, K + e); a direction with one fine point
    // per cell reads index 0 only
    T cv[2][2][2];
    for (int e = 0; e < 2; ++e)
        for (int b = 0; b < 2; ++b)
            for (int a = 0; a < 2; ++a)
                cv[a][b][e] = coarse[wix(gc, I + (a < ns[0] ? a : 0), J + (b < ns[1] ? b : 0), K + (e < ns[2] ? e : 0))];
    const int i0 = W.d[0] * (I - 1) + 1, j0 = W.d[1] * (J - 1) + 1, k0 = W.d[2] * (K - 1) + 1;
    const T *w1 = W.w[0], *w2 = W.w[1], *w3 = W.w[2];
    const int s1 = gf.cst[0], s2 = gf.cst[1], s3 = gf.cst[2];
    for (int se = 0; se < ns[2]; ++se) {
        const int k = k0 + se;
        // z pass at the four lattice columns (a, b)
        T wz[2][2];
        for (int b = 0; b < 2; ++b)
            for (int a = 0; a < 2; ++a)
                wz[a][b] = se == 0 ? cv[a][b][0] : w3[k] * cv[a][b][0] + w3[s3 + k] * cv[a][b][1];
        for (int sb = 0; sb < ns[1]; ++sb) {
            const int j = j0 + sb;
            // y pass at the two lattice x positions
            T wy[2];
            for (int a = 0; a < 2; ++a) wy[a] = sb == 0 ? wz[a][0] : w2[j] * wz[a][0] + w2[s2 + j] * wz[a][1];
            for (int sa = 0; sa < ns[0]; ++sa) {
                const int i = i0 + sa;
                const T v = sa == 0 ? wy[0] : w1[i] * wy[0] + w1[s1 + i] * wy[1];
                const int64_t p = gidx(gf, i, j, k);
                x[p] = x[p] + v;
            }
        }
    }
}

// ---- levels g0..G in one workgroup ------------------------------------------

constexpr int MG_FUSED_THREADS = 1024;

// every point of [lo, lo + n) (three directions) once, over the workgroup
template <typename F>
__device__ inline void wg_points(const int *lo, const int *n, F f) {
    const int tot = n[0] * n[1] * n[2];
    for (int q = (int)threadIdx.x; q < tot; q += MG_FUSED_THREADS) {
        const int a = q % n[0], r = q / n[0];
        const int b = r % n[1], c = r / n[1];
        f(lo[0] + a, lo[1] + b, lo[2] + c);
    }
    __syncthreads();
}

// the rest of the cycle from level g0: its right-hand side A is in place; the
// iterate of g0 ends in B or D as the host's account of the same swaps says
template <typename T>
__global__ __launch_bounds__(MG_FUSED_THREADS) void k_mg_fused(const MgLevelDev<T> *__restrict__ L, int g0, int G,
                                                               int n_down, int n_up, int n_bottom, T omega,
                                                               const int *active) {
    if (active && !*active) return;
    const int tid = (int)threadIdx.x;
    gptr<T> x[MG_LEVELS_MAX + 1], y[MG_LEVELS_MAX + 1];
    for (int g = g0; g <= G; ++g) {
        x[g] = (gptr<T>)L[g].B;
        y[g] = (gptr<T>)L[g].D;
    }
    // a level's coefficient tables as global pointers
    struct Tab {
        gcptr<T> t[3];
    };
    auto tab = [](const T *const *w) {
        Tab r;
        for (int d = 0; d < 3; ++d) r.t[d] = (gcptr<T>)w[d];
        return r;
    };
    auto range = [&](int g, int *lo, int *n) {
        for (int d = 0; d < 3; ++d) {
            lo[d] = L[g].g.Sp[d];
            n[d] = L[g].g.Np[d] - L[g].g.Sp[d] + 1;
        }
    };
    auto relax = [&](int g, bool init, int nr) {
        const MgLevelDev<T> &l = L[g];
        int lo[3], n[3];
        range(g, lo, n);
        const Tab dg = tab(l.dg);
        const gcptr<T> A = (gcptr<T>)l.A;
        int q = 1;
        if (init && nr >= 1) {
            gptr<T> xx = x[g];
            wg_points(lo, n, [&](int i, int j, int k) { pt_init<T>(l.g, dg.t, i, j, k, omega, A, xx); });
            q = 2;
        }
        for (; q <= nr; ++q) {
            const gcptr<T> xx = x[g];
            gptr<T> yy = y[g];
            wg_points(lo, n, [&](int i, int j, int k) { pt_jor<T>(l.g, dg.t, i, j, k, omega, xx, A, yy); });
            gptr<T> tmp = x[g];
            x[g] = y[g];
            y[g] = tmp;
        }
    };
    for (int g = g0; g <= G - 1; ++g) {
        const MgLevelDev<T> &l = L[g], &lc = L[g + 1];
        relax(g, true, n_down);
        int lo[3], n[3];
        range(g, lo, n);
        const gcptr<T> xx = x[g];
        const Tab dg = tab(l.dg), cR = tab(lc.cR);
        const gptr<T> A = (gptr<T>)l.A, Cl = (gptr<T>)l.C, Ac = (gptr<T>)lc.A;
        wg_points(lo, n, [&](int i, int j, int k) { pt_lap<T>(l.g, dg.t, i, j, k, xx, Cl); });
        wg_edges<T, gptr<T>>(l.g, A, Cl, 0, tid, MG_FUSED_THREADS);
        int clo[3] = {1, 1, 1}, cn[3];
        for (int d = 0; d < 3; ++d) cn[d] = lc.g.N[d] - (l.g.per[d] ? 1 : 0);
        wg_points(clo, cn, [&](int ii, int jj, int kk) {
            pt_restrict<T>(l.g, lc.g, cR.t, l.dn, ii, jj, kk, (gcptr<T>)A, (gcptr<T>)Cl, Ac);
        });
    }
    relax(G, true, n_bottom);
    relax(G, false, n_bottom);
    for (int g = G - 1; g >= g0; --g) {
        const MgLevelDev<T> &l = L[g], &lc = L[g + 1];
        wg_edges<T, gptr<T>>(lc.g, x[g + 1], (gptr<T>)nullptr, 1, tid, MG_FUSED_THREADS);
        int flo[3] = {1, 1, 1}, fn[3] = {l.g.N[0], l.g.N[1], l.g.N[2]};
        const gcptr<T> co = x[g + 1];
        gptr<T> xx = x[g];
        const Tab cI = tab(l.cI);
        wg_points(flo, fn, [&](int i, int j, int k) { pt_interp<T>(l.g, lc.g, cI.t, l.dn, i, j, k, co, xx); });
        relax(g, false, n_up);
    }
}

inline Range3 full_range(const Geo &g) { return Range3{1, 1, 1, g.N[0], g.N[1], g.N[2]}; }


// The cycle for arithmetic type T.  Level 1: right-hand side b1 (v itself for
// FP64; the float copy the first relaxation stores for FP32), iterate buffers
// x/y; the last sweep of level 1 writes z (for FP32 widened to double).
template <typename T>
struct CycleBufs;
template <>
struct CycleBufs<double> {
    static double *A(const MgLevel &L) { return L.A; }
    static double *B(const MgLevel &L) { return L.B; }
    static double *C(const MgLevel &L) { return L.C; }
    static double *D(const MgLevel &L) { return L.D; }
    static const double *const *cR(const MgLevel &L) { return L.cR; }
    static const double *const *cI(const MgLevel &L) { return L.cI; }
    static JorLaunch shape(const MgLevel &L) { return L.shape; }
    static const MgLevelDev<double> *dev(const MgDev &m) { return m.Ldev; }
};
template <>
struct CycleBufs<float> {
    static float *A(const MgLevel &L) { return L.Af; }
    static float *B(const MgLevel &L) { return L.Bf; }
    static float *C(const MgLevel &L) { return L.Cf; }
    static float *D(const MgLevel &L) { return L.Df; }
    static const float *const *cR(const MgLevel &L) { return L.cRf; }
    static const float *const *cI(const MgLevel &L) { return L.cIf; }
    static JorLaunch shape(const MgLevel &L) { return L.shape32; }
    static const MgLevelDev<float> *dev(const MgDev &m) { return m.Ldev32; }
};

inline void jor_any(hipStream_t s, const MgLevel &L, double omega, const double *x, const double *b, double *y,
                    const int *act, JorLaunch sh) {
    launch_jor(s, L.g, L.c, L.pr, omega, x, b, y, act, sh);
}
inline void jor_any(hipStream_t s, const MgLevel &L, float omega, const float *x, const float *b, float *y,
                    const int *act, JorLaunch sh) {
    launch_jor_f32(s, L.g, L.c, L.pr, omega, x, b, y, act, sh);
}

template <typename T>
void mg_cycle(hipStream_t s, const MgDev &m, double *v, double *z, const int *act) {
    using B_ = CycleBufs<T>;
    const int G = m.G;
    const int g0 = m.g0;  // levels g0..G run fused (g0 >= 2)
    const T omega = (T)m.omega;
    const bool fp64 = sizeof(T) == 8;
    T *b[MG_LEVELS_MAX + 1], *x[MG_LEVELS_MAX + 1], *y[MG_LEVELS_MAX + 1];
    if (fp64) {
        // level 1 alternates between z and D over the cycle's sweeps; starting on
        // the buffer that makes the last sweep write z leaves the result in place
        const int flips1 = G > 1 ? (m.n_down - 1) + m.n_up : (m.n_bottom - 1) + m.n_bottom;
        b[1] = (T *)v;
        x[1] = (flips1 % 2 == 0) ? (T *)z : B_::D(m.L[1]);
        y[1] = (x[1] == (T *)z) ? B_::D(m.L[1]) : (T *)z;
    } else {
        b[1] = B_::A(m.L[1]);
        x[1] = B_::B(m.L[1]);
        y[1] = B_::D(m.L[1]);
    }
    for (int g = 2; g <= G; ++g) {
        b[g] = B_::A(m.L[g]);
        x[g] = B_::B(m.L[g]);
        y[g] = B_::D(m.L[g]);
    }
    // relaxation of level g; launch = false only follows the swaps (the fused kernel
    // runs them); init_done: the restriction into g already wrote the first
    // relaxation; last_to_z: the FP32 cycle's final sweep, written to z as double
    auto relax = [&](int g, bool init, int n, bool launch, bool init_done = false, bool last_to_z = false) {
        const MgLevel &L = m.L[g];
        const Range3 r = make_range(L.pr);
        int q = 1;
        if (init && n >= 1) {
            if (launch && !init_done) {
                if (!fp64 && g == 1)
                    hipLaunchKernelGGL(k_mg_init_f32_first, grid_for(r), dim3(BX, BY, 1), 0, s, L.g, L.c, r,
                                       (float)omega, (const double *)v, (float *)b[1], (float *)x[1], act);
                else
                    hipLaunchKernelGGL(k_mg_init<T>, grid_for(r), dim3(BX, BY, 1), 0, s, L.g, L.c, r, omega, b[g],
                                       x[g], act);
            }
            q = 2;
        }
        for (; q <= n; ++q) {
            if (launch) {
                if (last_to_z && q == n)
                    launch_jor_f32_to64(s, L.g, L.c, L.pr, (float)omega, (const float *)x[g], (const float *)b[g], z,
                                        act, B_::shape(L));
                else
                    jor_any(s, L, omega, x[g], b[g], y[g], act, B_::shape(L));
            }
            std::swap(x[g], y[g]);
        }
    };
    if (G == 1) {
        relax(1, true, m.n_bottom, true);
        relax(1, false, m.n_bottom, true, false, !fp64);
        return;
    }
    const int gl = std::min(g0, G) - 1;  // last level with launches of its own on the way down
    for (int g = 1; g <= gl; ++g) {
        const MgLevel &L = m.L[g], &Lc = m.L[g + 1];
        relax(g, true, m.n_down, true, g >= 2);
        // the residual product with the JOR sweep's 16-byte lanes and launch shape
        // (one point a lane with seven 8-byte loads is bound by address issue)
        launch_lap(s, L.g, L.c, L.pr, (const T *)x[g], B_::C(L), act, B_::shape(L));
        const bool walls = !(L.g.per[0] && L.g.per[1] && L.g.per[2]);
        if (walls) {
            hipLaunchKernelGGL(k_mg_edges<T>, dim3(1), dim3(MG_EDGE_THREADS), 0, s, L.g, b[g], B_::C(L), 0, act);
            // the FP32 cycle works on a copy of v; v's edges change as the reference's do
            if (!fp64 && g == 1)
                hipLaunchKernelGGL(k_mg_edges<double>, dim3(1), dim3(MG_EDGE_THREADS), 0, s, L.g, v, (double *)nullptr,
                                   0, act);
        }
        LevelW<T> R;
        int nmax[3];
        for (int d = 0; d < 3; ++d) {
            R.w[d] = B_::cR(Lc)[d];
            R.d[d] = L.dn[d];
            nmax[d] = Lc.g.N[d] - (L.g.per[d] ? 1 : 0);
        }
        const Range3 cr{1, 1, 1, nmax[0], nmax[1], nmax[2]};
        // a launched coarse level takes its first relaxation from here (the fused
        // levels run their own); nmax spans that level's pressure range
        const bool coarse_launched = g + 1 < g0;
        hipLaunchKernelGGL(k_mg_restrict<T>, grid_for(cr), dim3(BX, BY, 1), 0, s, L.g, Lc.g, R, Lc.c, cr, b[g],
                           B_::C(L), B_::A(Lc), coarse_launched ? x[g + 1] : (T *)nullptr, omega, act);
    }
    if (g0 <= G) {
        hipLaunchKernelGGL(k_mg_fused<T>, dim3(1), dim3(MG_FUSED_THREADS), 0, s, B_::dev(m), g0, G, m.n_down, m.n_up,
                           m.n_bottom, omega, act);
        // the fused levels' swaps, for the buffer that holds level g0's iterate
        for (int g = g0; g <= G - 1; ++g) relax(g, true, m.n_down, false);
        relax(G, true, m.n_bottom, false);
        relax(G, false, m.n_bottom, false);
        for (int g = G - 1; g >= g0; --g) relax(g, false, m.n_up, false);
    } else {
        relax(G, true, m.n_bottom, true, true);
        relax(G, false, m.n_bottom, true);
    }
    for (int g = gl; g >= 1; --g) {
        const MgLevel &L = m.L[g], &Lc = m.L[g + 1];
        if (!(Lc.g.per[0] && Lc.g.per[1] && Lc.g.per[2]))
            hipLaunchKernelGGL(k_mg_edges<T>, dim3(1), dim3(MG_EDGE_THREADS), 0, s, Lc.g, x[g + 1], (T *)nullptr, 1, act);
        LevelW<T> W;
        for (int d = 0; d < 3; ++d) {
            W.w[d] = B_::cI(L)[d];
            W.d[d] = L.dn[d];
        }
        const Range3 cr = full_range(Lc.g);
        hipLaunchKernelGGL(k_mg_interp_cells<T>, grid_for(cr), dim3(BX, BY, 1), 0, s, L.g, Lc.g, W, x[g + 1], x[g],
                           act);
        relax(g, false, m.n_up, true, false, !fp64 && g == 1);
    }
}

}  // namespace

void launch_mg_vcycle(hipStream_t s, const MgDev &m, double *v, double *z, const int *act) {
    mg_cycle<double>(s, m, v, z, act);
}

void launch_mg_vcycle_f32(hipStream_t s, const MgDev &m, double *v, double *z, const int *act) {
    mg_cycle<float>(s, m, v, z, act);
}

}  // namespace ih
