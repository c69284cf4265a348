// Fused Poisson operator and device-controlled BiCGstab (gfx950).
//
// k_apply_A evaluates product_div_grad (src/mod_laplace.f90:38-197): for each
// direction m the gradient G_m p (src/mod_diff.f90:309-414, zero on the wall
// planes of its component, exchanged with the vel_dir = m rules) followed by
// the divergence D_m (src/mod_diff.f90:52-118), accumulated in the reference's
// order.  Instead of writing the three gradient fields to HBM it keeps them in
// LDS: a 64 x 4 column tile marches in z; per plane the pressure cross
// (x and y halos of 5) is staged in LDS, G_1 and G_2 are formed there once per
// position, and G_3 lives in a per-thread register queue along z.  Every
// gradient value is computed with the reference's operation order, so Ap is
// bit-identical to the three-pass evaluation, while HBM traffic drops to one
// read of p and one write of Ap.
//
// The BiCGstab recurrences (src/mod_solvers.f90:1509-1935, preconditioner 0)
// run without host round trips: one-block control kernels reduce the partial
// dot products with the fixed tree and update the scalars in device memory; an
// active flag turns the remaining launches of a batch into no-ops once the
// iteration has stopped, so the host checks convergence once per batch.
#include "ih_device.h"
#include "ih_krylov.h"

namespace ih {

using namespace dev;

namespace {

constexpr int AX = 64, AY = 4, AZ = 8;   // operator tile (x, y) and z chunk (= the reduction chunk)
constexpr int HX = AX + 10, HY = AY + 10;  // pressure halo of 5 each side
constexpr int G1W = AX + 5;               // G_1 at x0-3 .. x0+AX+1
constexpr int G2H = AY + 5;               // G_2 at y0-3 .. y0+AY+1

__device__ inline int wrap_p(const Geo &g, int d, int idx) {
    const int P = g.Np[d] - g.Sp[d] + 1;
    if (idx < g.Sp[d]) return idx + P;
    if (idx > g.Np[d]) return idx - P;
    return idx;
}

// Source index of the pressure-like field along d after exchange(d, 0, .):
// periodic ghosts are copies, wall ghosts (<= 0, > N) are zero.
__device__ inline bool p_src(const Geo &g, int d, int &idx) {
    if (g.per[d]) {
        idx = wrap_p(g, d, idx);
        return true;
    }
    return idx >= 1 && idx <= g.N[d];
}

// Coefficient index of G_d at position idx along d; false where the exchanged
// gradient field is zero (wall planes 0 and N, wall ghosts).
__device__ inline bool g_src(const Geo &g, int d, int &idx) {
    if (g.per[d]) {
        idx = wrap_p(g, d, idx);
        return true;
    }
    return idx >= 1 && idx <= g.N[d] - 1;
}

__device__ inline bool wall_plane(const Geo &g, int e, int idx) { return !g.per[e] && (idx == 1 || idx == g.N[e]); }

__device__ inline double grad6(const double *cg, int cst, int ci, const double *p, int stride) {
    // G p = c(-2)*p(-2) + c(-1)*p(-1) + ... + c(3)*p(3), p[0] is offset -2
    double v = cg[0 * cst + ci] * p[0];
#pragma unroll
    for (int s = 1; s < 6; ++s) v = v + cg[s * cst + ci] * p[s * stride];
    return v;
}

// dot modes of k_apply_A
enum { DOT_NONE = 0, DOT_W = 1, DOT_SELF = 2 };

// One thread per (x, y) column of a 64 x 4 tile, AZ planes.  Per plane the
// thread reads p(x-5..x+5, y) and p(x, y-5..y+5) into registers and forms its
// six G_1 and six G_2 values itself (a value shared by neighbouring threads is
// evaluated identically by each), G_3 comes from the z register queue.  The
// wrap/zero rules and coefficients of the tile's x and y positions are resolved
// once per block into LDS; there is no barrier inside the plane loop.
__global__ __launch_bounds__(256) void k_apply_A(Geo g, Coef c, Range3 r, int nbx, int nby, int nbz,
                                                 const double *__restrict__ src, double *__restrict__ dst,
                                                 const double *__restrict__ w, int dot_mode, double *part1,
                                                 double *part2, const int *active) {
    if (active && !*active) return;
    int bx, by, bz;
    if (!xcd_tile(nbx, nby, nbz, bx, by, bz)) return;
    __shared__ double sG1[6][G1W];  // cG_1 at the tile's G_1 positions, 0 where G_1 is zero
    __shared__ double sD1[6][AX];
    __shared__ double sG2[6][G2H];
    __shared__ double sD2[6][AY];
    __shared__ int sPx[HX];         // source x of p at x0-5+u, -1 = zero
    __shared__ int sPy[HY];
    __shared__ int sV1[G1W];        // 1 where G_1 (G_2) at the position is a computed value
    __shared__ int sV2[G2H];
    __shared__ double sm[8];

    const int tx = threadIdx.x, ty = threadIdx.y;
    const int tid = tx + AX * ty;
    const int x0 = r.lo0 + bx * AX;
    const int y0 = r.lo1 + by * AY;
    const int x = x0 + tx, y = y0 + ty;
    const int xe = r.lo0 + r.n0 - 1, ye = r.lo1 + r.n1 - 1, ze = r.lo2 + r.n2 - 1;
    const bool own = (x <= xe) && (y <= ye);
    const int kc0 = r.lo2 + bz * AZ;
    const int kc1 = min(kc0 + AZ - 1, ze);
    const int c0 = g.cst[0], c1 = g.cst[1], c2 = g.cst[2];
    const double *cg3 = c.G[2], *cd3 = c.D[2];

    // ---- per-block tables (positions past the box halo are never read) ----
    for (int e = tid; e < 6 * G1W; e += 256) {
        const int s = e / G1W, u = e % G1W;
        int xs = x0 - 3 + u;
        const bool v = xs <= xe + 2 && g_src(g, 0, xs);
        sG1[s][u] = v ? c.G[0][s * c0 + xs] : 0.0;
        if (s == 0) sV1[u] = v ? 1 : 0;
    }
    for (int e = tid; e < 6 * G2H; e += 256) {
        const int s = e / G2H, u = e % G2H;
        int ys = y0 - 3 + u;
        const bool v = ys <= ye + 2 && g_src(g, 1, ys);
        sG2[s][u] = v ? c.G[1][s * c1 + ys] : 0.0;
        if (s == 0) sV2[u] = v ? 1 : 0;
    }
    for (int e = tid; e < 6 * AX; e += 256) {
        const int t = e / AX, u = e % AX;
        sD1[t][u] = (x0 + u <= xe) ? c.D[0][t * c0 + x0 + u] : 0.0;
    }
    if (tid < 6 * AY) {
        const int t = tid / AY, u = tid % AY;
        sD2[t][u] = (y0 + u <= ye) ? c.D[1][t * c1 + y0 + u] : 0.0;
    }
    if (tid < HX) {
        int xx = x0 - 5 + tid;
        sPx[tid] = (xx <= xe + 5 && p_src(g, 0, xx)) ? xx : -1;
    }
    if (tid >= 128 && tid < 128 + HY) {
        int yy = y0 - 5 + (tid - 128);
        sPy[tid - 128] = (yy <= ye + 5 && p_src(g, 1, yy)) ? yy : -1;
    }
    __syncthreads();

    // G_1 positions x-3..x+2 and G_2 positions y-3..y+2 that are identically zero
    // because of the wall planes of their component in the transverse directions
    const bool zy1 = wall_plane(g, 1, y);   // G_1 on a y wall plane
    const bool zx2 = wall_plane(g, 0, x);   // G_2 on an x wall plane
    const bool z3 = zx2 || zy1;             // G_3 on an x or y wall plane

    double q[6], g3[6];
    {
        double pz[11];
#pragma unroll
        for (int s = 0; s < 11; ++s) {
            int kk = kc0 - 5 + s;
            pz[s] = (own && p_src(g, 2, kk)) ? src[gidx(g, x, y, kk)] : 0.0;
        }
#pragma unroll
        for (int t = 0; t < 6; ++t) {
            int kk = kc0 - 3 + t;
            g3[t] = (!z3 && g_src(g, 2, kk)) ? grad6(cg3, c2, kk, &pz[t], 1) : 0.0;
        }
#pragma unroll
        for (int s = 0; s < 6; ++s) q[s] = pz[5 + s];
    }

    double acc1 = 0.0, acc2 = 0.0;
    for (int k = kc0; k <= kc1; ++k) {
        double lap = 0.0;
        if (own) {
            const bool zk = wall_plane(g, 2, k);
            const int64_t xk = gidx(g, 0, y, k);  // row base of (., y, k) at i = 0
            double px[11], py[11];
#pragma unroll
            for (int u = 0; u < 11; ++u) {
                const int xs = sPx[tx + u];
                px[u] = (xs >= 0) ? src[xk + xs] : 0.0;
            }
            const int64_t yk = gidx(g, x, 0, k);
#pragma unroll
            for (int u = 0; u < 11; ++u) {
                const int ys = sPy[ty + u];
                py[u] = (ys >= 0) ? src[yk + (int64_t)ys * g.px] : 0.0;
            }
            // direction 1: lap = D_1 G_1
#pragma unroll
            for (int t = 0; t < 6; ++t) {
                double g1 = 0.0;
                if (!zy1 && !zk && sV1[tx + t]) {
                    g1 = sG1[0][tx + t] * px[t];
#pragma unroll
                    for (int s = 1; s < 6; ++s) g1 = g1 + sG1[s][tx + t] * px[t + s];
                }
                lap = (t == 0) ? sD1[0][tx] * g1 : lap + sD1[t][tx] * g1;
            }
#pragma unroll
            for (int t = 0; t < 6; ++t) {
                double g2 = 0.0;
                if (!zx2 && !zk && sV2[ty + t]) {
                    g2 = sG2[0][ty + t] * py[t];
#pragma unroll
                    for (int s = 1; s < 6; ++s) g2 = g2 + sG2[s][ty + t] * py[t + s];
                }
                lap = lap + sD2[t][ty] * g2;
            }
#pragma unroll
            for (int t = 0; t < 6; ++t) lap = lap + cd3[t * c2 + k] * g3[t];
            const int64_t xi = gidx(g, x, y, k);
            dst[xi] = lap;
            if (dot_mode == DOT_W) {
                acc1 = acc1 + lap * w[xi];
            } else if (dot_mode == DOT_SELF) {
                acc1 = acc1 + src[xi] * lap;
                acc2 = acc2 + lap * lap;
            }
        }
        // ---- advance the z queues: G_3(k+3) from p(k+1 .. k+6) ----
        {
            int kn = k + 6;
            const double pn = (own && p_src(g, 2, kn)) ? src[gidx(g, x, y, kn)] : 0.0;
#pragma unroll
            for (int s = 0; s < 5; ++s) q[s] = q[s + 1];
            q[5] = pn;
            int kk = k + 3;
#pragma unroll
            for (int t = 0; t < 5; ++t) g3[t] = g3[t + 1];
            g3[5] = (!z3 && g_src(g, 2, kk)) ? grad6(cg3, c2, kk, q, 1) : 0.0;
        }
    }
    // ---- fixed-tree partials (the z chunk is exactly one reduction chunk) ----
    if (dot_mode != DOT_NONE) {
        const int64_t pi = (int64_t)bx + (int64_t)nbx * ((int64_t)by + (int64_t)nby * bz);
        const double s1 = block_sum(acc1, sm);
        __syncthreads();
        if (tid == 0) part1[pi] = s1;
        if (dot_mode == DOT_SELF) {
            const double s2 = block_sum(acc2, sm + 4);
            if (tid == 0) part2[pi] = s2;
        }
    }
}

// ---------------------------------------------------------------------------
// BiCGstab vector kernels (tile 64 x 4 x RED_KC so that fused partials follow
// the fixed tree) and the one-block control kernels.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_init_rr(Geo g, Range3 r, const double *bb, double *rr, double *rh,
                                                 double *pmax, double *pdot) {
    __shared__ double sm[8];
    const int i = r.lo0 + (int)(blockIdx.x * RED_TX + threadIdx.x);
    const int j = r.lo1 + (int)(blockIdx.y * RED_TY + threadIdx.y);
    double m = 0.0, d = 0.0;
    if (i < r.lo0 + r.n0 && j < r.lo1 + r.n1) {
        for (int t = 0; t < RED_KC; ++t) {
            const int k = r.lo2 + (int)blockIdx.z * RED_KC + t;
            if (k >= r.lo2 + r.n2) break;
            const int64_t x = gidx(g, i, j, k);
            const double v = bb[x] - rr[x];
            rr[x] = v;
            rh[x] = v;
            const double a = fabs(v);
            m = (a > m) ? a : m;
            d = d + v * v;
        }
    }
    const double bm = block_max(m, sm);
    __syncthreads();
    const double bs = block_sum(d, sm + 4);
    if (threadIdx.x == 0 && threadIdx.y == 0) {
        pmax[part_index()] = bm;
        pdot[part_index()] = bs;
    }
}

__global__ void k_p_update(Geo g, Range3 r, const double *rr, double *pk, const double *Ap, const double *sc,
                           const int *fl) {
    if (!fl[KF_ACTIVE]) return;
    IH_POINT(r);
    const int64_t x = gidx(g, i, j, k);
    if (fl[KF_MODE])
        pk[x] = rr[x] + sc[KS_BETA] * pk[x] + sc[KS_OM2] * Ap[x];
    else
        pk[x] = rr[x];
}

__global__ void k_s_update(Geo g, Range3 r, double *rr, const double *Ap, const double *sc, const int *fl) {
    if (!fl[KF_ACTIVE]) return;
    IH_POINT(r);
    const int64_t x = gidx(g, i, j, k);
    rr[x] = rr[x] - sc[KS_ALPHA] * Ap[x];
}

__global__ __launch_bounds__(256) void k_x_update(Geo g, Range3 r, double *phi, double *rr, const double *pk,
                                                  const double *Ar, const double *rh, const double *sc, const int *fl,
                                                  double *pmax, double *pdot) {
    if (!fl[KF_ACTIVE]) return;
    __shared__ double sm[8];
    const double om = sc[KS_OMEGA], alpha = sc[KS_ALPHA];
    const int i = r.lo0 + (int)(blockIdx.x * RED_TX + threadIdx.x);
    const int j = r.lo1 + (int)(blockIdx.y * RED_TY + threadIdx.y);
    double m = 0.0, d = 0.0;
    if (i < r.lo0 + r.n0 && j < r.lo1 + r.n1) {
        for (int t = 0; t < RED_KC; ++t) {
            const int k = r.lo2 + (int)blockIdx.z * RED_KC + t;
            if (k >= r.lo2 + r.n2) break;
            const int64_t x = gidx(g, i, j, k);
            phi[x] = phi[x] + om * rr[x] + alpha * pk[x];
            const double v = rr[x] - om * Ar[x];
            rr[x] = v;
            const double a = fabs(v);
            m = (a > m) ? a : m;
            d = d + v * rh[x];
        }
    }
    const double bm = block_max(m, sm);
    __syncthreads();
    const double bs = block_sum(d, sm + 4);
    if (threadIdx.x == 0 && threadIdx.y == 0) {
        pmax[part_index()] = bm;
        pdot[part_index()] = bs;
    }
}

__device__ inline double reduce_sum_parts(const double *part, int64_t nb, double *sm) {
    const int tid = threadIdx.x + threadIdx.y * blockDim.x;
    double acc = 0.0;
    for (int64_t q = tid; q < nb; q += RED_THREADS) acc = acc + part[q];
    return block_sum(acc, sm);
}

__device__ inline double reduce_max_parts(const double *part, int64_t nb, double *sm) {
    const int tid = threadIdx.x + threadIdx.y * blockDim.x;
    double m = 0.0;
    for (int64_t q = tid; q < nb; q += RED_THREADS) m = (part[q] > m) ? part[q] : m;
    return block_max(m, sm);
}

// after the initial residual: norm, rhr_next; reset the iteration state
__global__ __launch_bounds__(256) void k_ctl_setup(const double *pmax, const double *pdot, int64_t nb, double *sc,
                                                   int *fl, double eps, int nmax) {
    __shared__ double sm[8];
    const double nm = reduce_max_parts(pmax, nb, sm);
    __syncthreads();
    const double dt = reduce_sum_parts(pdot, nb, sm + 4);
    if (threadIdx.x == 0 && threadIdx.y == 0) {
        sc[KS_NORM] = nm;
        sc[KS_NORM0] = nm;
        sc[KS_RHR_NEXT] = dt;
        sc[KS_RHR] = 0.0;
        sc[KS_ALPHA] = 0.0;
        sc[KS_OMEGA] = 0.0;
        sc[KS_EPS] = eps;
        fl[KF_ACTIVE] = 1;
        fl[KF_COUNTER] = 0;
        fl[KF_NMAX] = nmax;
        fl[KF_EXIT] = 0;
        fl[KF_RESET] = 0;
        fl[KF_MODE] = 0;
        fl[KF_DONE] = 0;
    }
}

// status_iteration + the scalar part of the iteration head (mod_solvers.f90:1589-1666)
__device__ inline void ctl_head_tail(double *sc, int *fl, int counter) {
    if (counter >= 2) {
        if (sc[KS_OMEGA] == 0.0 || sc[KS_RHR_PREV] == 0.0) {
            fl[KF_EXIT] = 2;
            fl[KF_ACTIVE] = 0;
            fl[KF_DONE] = 1;
            return;
        }
        const double beta = (sc[KS_ALPHA] / sc[KS_OMEGA]) * (sc[KS_RHR] / sc[KS_RHR_PREV]);
        const double om2 = -beta * sc[KS_OMEGA];
        sc[KS_BETA] = beta;
        sc[KS_OM2] = om2;
        sc[KS_OMEGA] = om2;
        fl[KF_MODE] = 1;
    } else {
        fl[KF_MODE] = 0;
    }
}

__global__ void k_ctl_begin(double *sc, int *fl) {
    if (!fl[KF_ACTIVE]) return;
    const double norm = sc[KS_NORM];
    int counter = fl[KF_COUNTER];
    bool ex = norm < sc[KS_EPS];
    if (!ex && counter == fl[KF_NMAX]) {
        ex = true;
        fl[KF_EXIT] = 1;
    }
    if (counter < 1 && norm != 0.0) ex = false;
    if (ex) {
        fl[KF_ACTIVE] = 0;
        fl[KF_DONE] = 1;
        return;
    }
    counter = counter + 1;
    fl[KF_COUNTER] = counter;
    sc[KS_RHR_PREV] = sc[KS_RHR];
    sc[KS_RHR] = sc[KS_RHR_NEXT];
    if (fabs(sc[KS_RHR]) == 0.0) {
        // new reference vector needed: the host resets rh and resumes
        fl[KF_RESET] = 1;
        fl[KF_ACTIVE] = 0;
        return;
    }
    ctl_head_tail(sc, fl, counter);
}

__global__ void k_ctl_resume(double *sc, int *fl) {
    fl[KF_RESET] = 0;
    fl[KF_ACTIVE] = 1;
    ctl_head_tail(sc, fl, fl[KF_COUNTER]);
}

__global__ __launch_bounds__(256) void k_ctl_alpha(const double *part, int64_t nb, double *sc, int *fl) {
    if (!fl[KF_ACTIVE]) return;
    __shared__ double sm[4];
    const double rhAp = reduce_sum_parts(part, nb, sm);
    if (threadIdx.x == 0 && threadIdx.y == 0) {
        if (fabs(rhAp) == 0.0) {
            fl[KF_EXIT] = 2;
            fl[KF_ACTIVE] = 0;
            fl[KF_DONE] = 1;
        } else {
            sc[KS_ALPHA] = sc[KS_RHR] / rhAp;
        }
    }
}

__global__ __launch_bounds__(256) void k_ctl_omega(const double *p1, const double *p2, int64_t nb, double *sc,
                                                   int *fl) {
    if (!fl[KF_ACTIVE]) return;
    __shared__ double sm[8];
    const double rAr = reduce_sum_parts(p1, nb, sm);
    __syncthreads();
    const double ArAr = reduce_sum_parts(p2, nb, sm + 4);
    if (threadIdx.x == 0 && threadIdx.y == 0) {
        if (fabs(ArAr) == 0.0) {
            fl[KF_EXIT] = 2;
            fl[KF_ACTIVE] = 0;
            fl[KF_DONE] = 1;
        } else {
            sc[KS_OMEGA] = rAr / ArAr;
        }
    }
}

__global__ __launch_bounds__(256) void k_ctl_end(const double *pmax, const double *pdot, int64_t nb, double *sc,
                                                 int *fl) {
    if (!fl[KF_ACTIVE]) return;
    __shared__ double sm[8];
    const double nm = reduce_max_parts(pmax, nb, sm);
    __syncthreads();
    const double dt = reduce_sum_parts(pdot, nb, sm + 4);
    if (threadIdx.x == 0 && threadIdx.y == 0) {
        sc[KS_NORM] = nm;
        sc[KS_RHR_NEXT] = dt;
    }
}

__global__ __launch_bounds__(256) void k_set_rhr(const double *part, int64_t nb, double *sc) {
    __shared__ double sm[4];
    const double v = reduce_sum_parts(part, nb, sm);
    if (threadIdx.x == 0 && threadIdx.y == 0) sc[KS_RHR] = v;
}

}  // namespace

// ===========================================================================
static inline Range3 Rk(const Box &b) { return make_range(b); }

void launch_apply_A(hipStream_t s, const Geo &g, const Coef &c, const Box &pb, const double *src, double *dst,
                    const double *w, int dot_mode, double *part1, double *part2, const int *active) {
    Range3 r = Rk(pb);
    if (empty(r)) return;
    const int nbx = (r.n0 + AX - 1) / AX, nby = (r.n1 + AY - 1) / AY, nbz = (r.n2 + AZ - 1) / AZ;
    hipLaunchKernelGGL(k_apply_A, dim3(xcd_grid((int64_t)nbx * nby * nbz)), dim3(AX, AY, 1), 0, s, g, c, r, nbx, nby,
                       nbz, src, dst, w, dot_mode, part1, part2, active);
}

void launch_krylov_init(hipStream_t s, const Geo &g, const Box &b, const double *bb, double *rr, double *rh,
                        double *pmax, double *pdot, double *sc, int *fl, double eps, int nmax) {
    Range3 r = Rk(b);
    hipLaunchKernelGGL(k_init_rr, red_grid(r), dim3(RED_TX, RED_TY, 1), 0, s, g, r, bb, rr, rh, pmax, pdot);
    hipLaunchKernelGGL(k_ctl_setup, dim3(1), dim3(RED_TX, RED_TY, 1), 0, s, pmax, pdot, red_blocks(b), sc, fl, eps,
                       nmax);
}

void launch_krylov_iteration(hipStream_t s, const Geo &g, const Coef &c, const Box &b, const KrylovPtrs &k,
                             int from_stage) {
    Range3 r = Rk(b);
    const int64_t nb = red_blocks(b);
    if (from_stage <= 0) hipLaunchKernelGGL(k_ctl_begin, dim3(1), dim3(1), 0, s, k.sc, k.fl);
    hipLaunchKernelGGL(k_p_update, grid_for(r), dim3(BX, BY, 1), 0, s, g, r, k.rr, k.pk, k.Ap, k.sc, k.fl);
    launch_apply_A(s, g, c, b, k.pk, k.Ap, k.rh, DOT_W, k.part1, nullptr, k.fl + KF_ACTIVE);
    hipLaunchKernelGGL(k_ctl_alpha, dim3(1), dim3(RED_TX, RED_TY, 1), 0, s, k.part1, nb, k.sc, k.fl);
    hipLaunchKernelGGL(k_s_update, grid_for(r), dim3(BX, BY, 1), 0, s, g, r, k.rr, k.Ap, k.sc, k.fl);
    launch_apply_A(s, g, c, b, k.rr, k.Ar, nullptr, DOT_SELF, k.part1, k.part2, k.fl + KF_ACTIVE);
    hipLaunchKernelGGL(k_ctl_omega, dim3(1), dim3(RED_TX, RED_TY, 1), 0, s, k.part1, k.part2, nb, k.sc, k.fl);
    hipLaunchKernelGGL(k_x_update, red_grid(r), dim3(RED_TX, RED_TY, 1), 0, s, g, r, k.phi, k.rr, k.pk, k.Ar, k.rh,
                       k.sc, k.fl, k.part1, k.part2);
    hipLaunchKernelGGL(k_ctl_end, dim3(1), dim3(RED_TX, RED_TY, 1), 0, s, k.part1, k.part2, nb, k.sc, k.fl);
}

void launch_krylov_resume(hipStream_t s, const Geo &g, const Box &b, const KrylovPtrs &k) {
    // rh = rr, rhr = <rr, rh>, then the tail of the iteration head
    launch_copy_box(s, g, b, k.rr, k.rh);
    launch_dot_partial(s, g, b, k.rr, k.rh, k.part1);
    hipLaunchKernelGGL(k_set_rhr, dim3(1), dim3(RED_TX, RED_TY, 1), 0, s, k.part1, red_blocks(b), k.sc);
    hipLaunchKernelGGL(k_ctl_resume, dim3(1), dim3(1), 0, s, k.sc, k.fl);
}

}  // namespace ih
