// Fused Poisson operator and device-controlled BiCGstab (gfx950).
//
// k_apply evaluates product_div_grad (src/mod_laplace.f90:38-197): for each
// direction m the gradient G_m p (src/mod_diff.f90:309-414, zero on the wall
// planes of its component, exchanged with the vel_dir = m rules) followed by
// the divergence D_m (src/mod_diff.f90:52-118), accumulated in the reference's
// order.  Instead of writing the three gradient fields to HBM it keeps them on
// chip: a 16 x 16 column tile marches through a z chunk of up to 32 planes;
// per plane the pressure cross (the tile plus x and y halos of 5) is staged in
// LDS, every G_1 and G_2 value of the tile and its halo is formed once, and G_3
// lives in a per-thread register queue along z.  Every gradient value is
// computed with the reference's operation order, so Ap is bit-identical to the
// three-pass evaluation, while HBM traffic drops to one read of p and one
// write of Ap.
//
// The BiCGstab recurrences (src/mod_solvers.f90:1509-1935; preconditioner 0,
// or the right-preconditioned form with FP32 JOR sweeps or the multigrid V
// cycle as M) run without host round trips: one-block control kernels reduce the partial
// dot products with the fixed tree and update the scalars in device memory; an
// active flag turns the remaining launches of a batch into no-ops once the
// iteration has stopped, so the host checks convergence once per batch.
#include "ih_device.h"
#include "ih_krylov.h"
#include "ih_mg.h"

#include <utility>

namespace ih {

using namespace dev;

namespace {


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

// p_src, g_src and wall_plane along z with the periodicity of z a compile-time
// constant: the plane loop of k_apply then has no branches on it (they cost
// scalar instructions and branch joins every plane)
template <bool PZ>
__device__ inline bool p_src_z(const Geo &g, int &idx) {
    if (PZ) {
        idx = wrap_p(g, 2, idx);
        return true;
    }
    return idx >= 1 && idx <= g.N[2];
}
template <bool PZ>
__device__ inline bool g_src_z(const Geo &g, int &idx) {
    if (PZ) {
        idx = wrap_p(g, 2, idx);
        return true;
    }
    return idx >= 1 && idx <= g.N[2] - 1;
}
template <bool PZ>
__device__ inline bool wall_z(const Geo &g, int idx) {
    return !PZ && (idx == 1 || idx == g.N[2]);
}

// dot modes of k_apply
enum { DOT_NONE = 0, DOT_W = 1, DOT_SELF = 2, DOT_PAIR = 3 };

// One thread per (x, y) column of a TX x TY tile, the planes of one z chunk
// (balanced chunks of at most AZ planes, the reduction chunks).  Pressure cross
// and gradient tiles are double buffered, and the divergence of plane k-1 is
// formed after the gradients of plane k, so each plane costs one barrier.  The
// halo of plane k+2, p(k+7) of the column queue and w(k+1) are loaded while
// planes k and k+1 are computed (the plane loop is unrolled twice so that the
// two register sets alternate).  The gradient coefficients of a thread's
// positions stay in registers for the whole chunk; every thread forms one G_1,
// one G_2 and at most one halo gradient per plane without divergent branches.
// DM = DOT_W accumulates <Ap, w>, DOT_SELF <p, Ap> and <Ap, Ap>, DOT_PAIR
// <w, Ap> and <Ap, Ap>, per
// thread in plane order, then over the block with the fixed tree.
// The bound asks for 3 wavefronts per SIMD (at most 168 VGPRs): the pair-dot
// form otherwise takes 170, which the 8-register allocation rounds to 176, and
// runs at 2 wavefronts per SIMD, 40% slower.
// The block function of tile (bx, by, bz); k_apply and the batched operator
// (k_apply_batch) both run it, so a batched case is bit-identical to its own launch.
template <int TX, int TY, int AZ, int DM, bool PZ>
__device__ inline void apply_block(const Geo &g, const Coef &c, const Range3 &r, int nbx, int nby, int nbz, int bx,
                                   int by, int bz, const double *__restrict__ src, double *__restrict__ dst,
                                   const double *__restrict__ w, double *part1, double *part2) {
    constexpr int NT = TX * TY;
    constexpr int PW = TX + 10, PH = TY + 10, PS = PW * PH;
    constexpr int NHX = 10 * TY;
    constexpr int NH = NHX + 10 * TX;
    constexpr int HL = (NH + NT - 1) / NT;
    constexpr int NG1X = 5 * TY, NG2X = 5 * TX;
    static_assert(NG1X + NG2X <= NT, "one extra gradient job per thread at most");
    constexpr int G1S = TX + 5;                 // row stride of the G_1 tile
    constexpr int OG2 = TY * G1S;               // offset of the G_2 tile
    constexpr int GS = OG2 + (TY + 5) * TX + 1; // + one dummy slot
    constexpr int ZW = AZ + 7;                  // G_3 positions kc0-4 .. kc0+AZ+2
    __shared__ double sP[2][PS];
    __shared__ double sG[2][GS];
    __shared__ double sCG3[6][ZW];
    __shared__ double sCD3[6][AZ];
    __shared__ int sVZ[ZW];
    __shared__ double sm[8];

    const int tx = threadIdx.x, ty = threadIdx.y;
    const int tid = tx + TX * ty;
    const int x0 = r.lo0 + bx * TX, y0 = r.lo1 + by * TY;
    const int x = x0 + tx, y = y0 + ty;
    const int xe = r.lo0 + r.n0 - 1, ye = r.lo1 + r.n1 - 1;
    const bool own = (x <= xe) && (y <= ye);
    int kc0, kc1;
    red_chunk(r, bz, nbz, kc0, kc1);
    const int c0 = g.cst[0], c1 = g.cst[1], c2 = g.cst[2];
    const int64_t pxy = g.pxy;
    // loads outside the field read c.zero, so no value is masked after its load
    // (a select on the loaded value would make the wave wait for it at once)
    const double *zero = c.zero;

    for (int e = tid; e < 6 * ZW; e += NT) {
        const int s = e / ZW, u = e % ZW;
        int kk = kc0 - 4 + u;
        const bool v = kk <= kc1 + 2 && g_src_z<PZ>(g, kk);
        sCG3[s][u] = v ? c.G[2][s * c2 + kk] : 0.0;
        if (s == 0) sVZ[u] = v ? 1 : 0;
    }
    for (int e = tid; e < 6 * AZ; e += NT) {
        const int t = e / AZ, u = e % AZ;
        sCD3[t][u] = (kc0 + u <= kc1) ? c.D[2][t * c2 + kc0 + u] : 0.0;
    }

    const int64_t hbase = gidx(g, -3, -3, 0);  // a valid element; offsets below are >= 0 from it
    int xs = x, ys = y;
    const bool vcol = x <= xe + 5 && y <= ye + 5 && p_src(g, 0, xs) && p_src(g, 1, ys);
    const int64_t colb = vcol ? gidx(g, xs, ys, 0) : hbase;
    const int64_t selfb = own ? gidx(g, x, y, 0) : hbase;

    int hoff[HL], hslot[HL];
#pragma unroll
    for (int h = 0; h < HL; ++h) {
        const int e = tid + h * NT;
        hoff[h] = -1;
        hslot[h] = PS;  // past the tile: not written
        if (e < NHX) {
            const int row = e / 10, cc = e % 10;
            const int col = cc < 5 ? cc : TX + cc;
            int hx = x0 - 5 + col, hy = y0 + row;
            hslot[h] = (row + 5) * PW + col;
            if (hx <= xe + 5 && hy <= ye + 5 && p_src(g, 0, hx) && p_src(g, 1, hy))
                hoff[h] = (int)(gidx(g, hx, hy, 0) - hbase);
        } else if (e < NH) {
            const int e2 = e - NHX;
            const int rr = e2 / TX, col = e2 % TX;
            const int row = rr < 5 ? rr : TY + rr;
            int hx = x0 + col, hy = y0 - 5 + row;
            hslot[h] = row * PW + col + 5;
            if (hx <= xe + 5 && hy <= ye + 5 && p_src(g, 0, hx) && p_src(g, 1, hy))
                hoff[h] = (int)(gidx(g, hx, hy, 0) - hbase);
        }
    }

    // gradient jobs (LDS base, stride, destination) and coefficients
    double cg1[6], cg2[6], cgx[6], cd1[6], cd2[6];
    bool v1, v2, vx;
    const int b1 = (ty + 5) * PW + tx, w1 = ty * G1S + tx;           // G_1 at x0-3+tx, row ty
    const int b2 = ty * PW + tx + 5, w2 = OG2 + ty * TX + tx;         // G_2 at y0-3+ty, column tx
    int bxj = 0, sxj = 1, wxj = GS - 1;
    {
        int p1 = x0 - 3 + tx;
        v1 = p1 <= xe + 2 && g_src(g, 0, p1) && !wall_plane(g, 1, y);
        int p2 = y0 - 3 + ty;
        v2 = p2 <= ye + 2 && g_src(g, 1, p2) && !wall_plane(g, 0, x);
        const double *cx = c.G[0];
        int cix = 0, cstx = c0;
        vx = false;
        const int e1 = tid - NG2X;
        if (e1 >= 0 && e1 < NG1X) {
            const int r1x = e1 / 5, u1x = TX + e1 % 5;
            int px_ = x0 - 3 + u1x;
            vx = px_ <= xe + 2 && g_src(g, 0, px_) && !wall_plane(g, 1, y0 + r1x);
            cix = px_;
            bxj = (r1x + 5) * PW + u1x;
            sxj = 1;
            wxj = r1x * G1S + u1x;
        } else if (tid < NG2X) {
            const int c2x = tid % TX, v2r = TY + tid / TX;
            int py_ = y0 - 3 + v2r;
            vx = py_ <= ye + 2 && g_src(g, 1, py_) && !wall_plane(g, 0, x0 + c2x);
            cx = c.G[1];
            cix = py_;
            cstx = c1;
            bxj = v2r * PW + c2x + 5;
            sxj = PW;
            wxj = OG2 + v2r * TX + c2x;
        }
#pragma unroll
        for (int s = 0; s < 6; ++s) {
            cg1[s] = v1 ? c.G[0][s * c0 + p1] : 0.0;
            cg2[s] = v2 ? c.G[1][s * c1 + p2] : 0.0;
            cgx[s] = vx ? cx[s * cstx + cix] : 0.0;
            cd1[s] = own ? c.D[0][s * c0 + x] : 0.0;
            cd2[s] = own ? c.D[1][s * c1 + y] : 0.0;
        }
    }
    const bool g3on = own && !(wall_plane(g, 0, x) || wall_plane(g, 1, y));
    __syncthreads();

    // queue at iteration k: q = p(k-1 .. k+4), g3 = G_3(k-4 .. k+1)
    double q[6], g3[6];
    {
        double pz[11];
#pragma unroll
        for (int s = 0; s < 11; ++s) {
            int kk = kc0 - 6 + s;
            const bool v = vcol && p_src_z<PZ>(g, kk);
            pz[s] = *(v ? src + colb + (int64_t)kk * pxy : zero);
        }
#pragma unroll
        for (int t = 0; t < 6; ++t) {
            double v = sCG3[0][t] * pz[t];
#pragma unroll
            for (int s = 1; s < 6; ++s) v = v + sCG3[s][t] * pz[t + s];
            g3[t] = (g3on && sVZ[t]) ? v : 0.0;
        }
#pragma unroll
        for (int s = 0; s < 6; ++s) q[s] = pz[5 + s];
    }
    auto load_halo = [&](int kk, double (&hv)[HL]) {
        const bool vk = kk <= kc1 && p_src_z<PZ>(g, kk);
#pragma unroll
        for (int h = 0; h < HL; ++h) {
            const bool v = vk && hoff[h] >= 0;
            hv[h] = *(v ? src + hbase + hoff[h] + (int64_t)kk * pxy : zero);
        }
    };
    auto load_col = [&](int kk) -> double {
        const bool v = vcol && kk <= kc1 + 5 && p_src_z<PZ>(g, kk);
        return *(v ? src + colb + (int64_t)kk * pxy : zero);
    };
    auto load_w = [&](int kk) -> double {
        if (DM != DOT_W && DM != DOT_PAIR) return 0.0;  // w is only a valid pointer in these modes
        const bool v = own && kk <= kc1;
        return *(v ? w + selfb + (int64_t)kk * pxy : zero);
    };
    double hvA[HL], hvB[HL];
    load_halo(kc0, hvA);
    load_halo(kc0 + 1, hvB);
    double pA = load_col(kc0 + 5), pB = load_col(kc0 + 6);
    double wA = load_w(kc0), wB = load_w(kc0 + 1);
    double acc1 = 0.0, acc2 = 0.0;

    auto divergence = [&](int k, const double *G, double wv) {
        // plane k from the gradient tile G, g3 = G_3(k-3 .. k+2), q[0] = p(k)
        double lap = cd1[0] * G[ty * G1S + tx];
#pragma unroll
        for (int t = 1; t < 6; ++t) lap = lap + cd1[t] * G[ty * G1S + tx + t];
#pragma unroll
        for (int t = 0; t < 6; ++t) lap = lap + cd2[t] * G[OG2 + (ty + t) * TX + tx];
        const int kz = k - kc0;
#pragma unroll
        for (int t = 0; t < 6; ++t) lap = lap + sCD3[t][kz] * g3[t];
        if (own) dst[selfb + (int64_t)k * pxy] = lap;
        if (DM == DOT_W) {
            acc1 = acc1 + (own ? lap * wv : 0.0);
        } else if (DM == DOT_SELF) {
            acc1 = acc1 + (own ? q[0] * lap : 0.0);
            acc2 = acc2 + (own ? lap * lap : 0.0);
        } else if (DM == DOT_PAIR) {
            acc1 = acc1 + (own ? wv * lap : 0.0);
            acc2 = acc2 + (own ? lap * lap : 0.0);
        }
    };
    auto body = [&](int k, int buf, double (&hv)[HL], double &pn, double &wprev, double &wcur) {
        // stage plane k
        double *P = sP[buf];
        P[(ty + 5) * PW + tx + 5] = q[1];
#pragma unroll
        for (int h = 0; h < HL; ++h)
            if (hslot[h] < PS) P[hslot[h]] = hv[h];
        __syncthreads();
        load_halo(k + 2, hv);
        // gradients of plane k
        const bool zk = wall_z<PZ>(g, k);
        double *G = sG[buf];
        {
            double a = cg1[0] * P[b1];
#pragma unroll
            for (int s = 1; s < 6; ++s) a = a + cg1[s] * P[b1 + s];
            G[w1] = (v1 && !zk) ? a : 0.0;
            double b = cg2[0] * P[b2];
#pragma unroll
            for (int s = 1; s < 6; ++s) b = b + cg2[s] * P[b2 + s * PW];
            G[w2] = (v2 && !zk) ? b : 0.0;
            if (tid < NG1X + NG2X) {  // halo gradient jobs; the last wavefront has none
                double e = cgx[0] * P[bxj];
#pragma unroll
                for (int s = 1; s < 6; ++s) e = e + cgx[s] * P[bxj + s * sxj];
                G[wxj] = (vx && !zk) ? e : 0.0;
            }
        }
        // divergence of plane k-1 (its gradient tile is the other buffer)
        if (k > kc0) {
            divergence(k - 1, sG[buf ^ 1], wprev);
            wprev = load_w(k + 1);
        }
        // advance: q = p(k .. k+5), g3 = G_3(k-3 .. k+2)
#pragma unroll
        for (int s = 0; s < 5; ++s) q[s] = q[s + 1];
        q[5] = pn;
        pn = load_col(k + 7);
#pragma unroll
        for (int t = 0; t < 5; ++t) g3[t] = g3[t + 1];
        const int u = k - kc0 + 6;  // G_3 position k+2
        double v = sCG3[0][u] * q[0];
#pragma unroll
        for (int s = 1; s < 6; ++s) v = v + sCG3[s][u] * q[s];
        g3[5] = (g3on && sVZ[u]) ? v : 0.0;
        (void)wcur;
    };
    // w(k-1) is consumed in iteration k: the A set serves even offsets k-kc0, the B set odd ones
    for (int k = kc0; k <= kc1; k += 2) {
        body(k, 0, hvA, pA, wB, wA);
        if (k + 1 <= kc1) body(k + 1, 1, hvB, pB, wA, wB);
    }
    __syncthreads();
    {
        const int n = kc1 - kc0 + 1;
        const int buf = (n - 1) & 1;
        divergence(kc1, sG[buf], buf == 0 ? wA : wB);
    }
    if (DM != DOT_NONE) {
        const int64_t pi = (int64_t)bx + (int64_t)nbx * ((int64_t)by + (int64_t)nby * bz);
        const double s1 = block_sum(acc1, sm);
        __syncthreads();
        if (tid == 0) part1[pi] = s1;
        if (DM == DOT_SELF || DM == DOT_PAIR) {
            const double s2 = block_sum(acc2, sm + 4);
            if (tid == 0) part2[pi] = s2;
        }
    }
}

template <int TX, int TY, int AZ, int DM, bool PZ>
__global__ __launch_bounds__(TX *TY, 3) void k_apply(
    Geo g, Coef c, Range3 r, int nbx, int nby, int nbz, const double *__restrict__ src, double *__restrict__ dst,
    const double *__restrict__ w, double *part1, double *part2, const int *active) {
    if (active && !*active) return;
    int bx, by, bz;
    if (!xcd_tile(nbx, nby, nbz, bx, by, bz)) return;
    apply_block<TX, TY, AZ, DM, PZ>(g, c, r, nbx, nby, nbz, bx, by, bz, src, dst, w, part1, part2);
}

// ---------------------------------------------------------------------------
// BiCGstab vector kernels (reduction tiles, so that fused partials follow
// the fixed tree) and the one-block control kernels.
// ---------------------------------------------------------------------------
// The reduction-tile kernels below are block functions of tile (bx, by, bz)
// out of nbz z chunks, partial index pi = bx + nbx*(by + nby*bz), run by the
// single-case kernels (tile = block index) and by the batched ones.
__device__ inline void init_rr_block(const Geo &g, const Range3 &r, int bx, int by, int bz, int nbz, int64_t pi,
                                     const double *bb, double *rr, double *rh, double *pmax, double *pdot) {
    __shared__ double sm[8];
    const int i = r.lo0 + bx * RED_TX + (int)threadIdx.x;
    const int j = r.lo1 + by * RED_TY + (int)threadIdx.y;
    double m = 0.0, d = 0.0;
    if (i < r.lo0 + r.n0 && j < r.lo1 + r.n1) {
        int kc0, kc1;
        red_chunk(r, bz, nbz, kc0, kc1);
        for (int k = kc0; k <= kc1; ++k) {
            const int64_t x = gidx(g, i, j, k);
            const double v = bb[x] - rr[x];
            rr[x] = v;
            rh[x] = v;
            const double a = fabs(v);
            m = max_nan(a, m);
            d = d + v * v;
        }
    }
    const double bm = block_max(m, sm);
    __syncthreads();
    const double bs = block_sum(d, sm + 4);
    if (threadIdx.x == 0 && threadIdx.y == 0) {
        pmax[pi] = bm;
        pdot[pi] = bs;
    }
}

__global__ __launch_bounds__(256) void k_init_rr(Geo g, Range3 r, const double *bb, double *rr, double *rh,
                                                 double *pmax, double *pdot) {
    init_rr_block(g, r, (int)blockIdx.x, (int)blockIdx.y, (int)blockIdx.z, (int)gridDim.z, part_index(), bb, rr, rh,
                  pmax, pdot);
}

// The first FP32 preconditioner sweep at one point, from the zero start: b32 =
// (float) v and e = the JOR update with every neighbour 0, in the operands and
// order of the sweep kernels (jor_f32_first_block), so the values are the same
// bits.  The update kernels that produce v call it when the preconditioner is
// FP32 sweeps (FIRST), which saves the first sweep's own pass over v.
struct FirstSweep {
    const float *g1, *g2, *g3;
    float omega;
    float *b32, *e;
};

__device__ inline void first_sweep_point(const Geo &g, const FirstSweep &f, int i, int j, int k, int64_t x,
                                         double v) {
    const int c0 = g.cst[0], c1 = g.cst[1], c2 = g.cst[2];
    const float z = 0.0f;
    const float bv = (float)v;
    float s = bv;
    s = s - f.g1[0 * c0 + i] * z;
    s = s - f.g1[2 * c0 + i] * z;
    s = s - f.g2[0 * c1 + j] * z;
    s = s - f.g2[2 * c1 + j] * z;
    s = s - f.g3[0 * c2 + k] * z;
    s = s - f.g3[2 * c2 + k] * z;
    const float diag = (f.g1[1 * c0 + i] + f.g2[1 * c1 + j]) + f.g3[1 * c2 + k];
    f.b32[x] = bv;
    f.e[x] = f.omega * s / diag + (1.0f - f.omega) * z;
}

template <bool FIRST>
__global__ void k_p_update(Geo g, Range3 r, const double *rr, double *pk, const double *Ap, const double *sc,
                           const int *fl, FirstSweep fs) {
    if (!fl[KF_ACTIVE]) return;
    IH_POINT(r);
    const int64_t x = gidx(g, i, j, k);
    double p;
    if (fl[KF_MODE])
        p = rr[x] + sc[KS_BETA] * pk[x] + sc[KS_OM2] * Ap[x];
    else
        p = rr[x];
    pk[x] = p;
    if (FIRST) first_sweep_point(g, fs, i, j, k, x, p);
}

__device__ inline double reduce_sum_parts(const double *part, int64_t nb, double *sm) {
    const int tid = threadIdx.x + threadIdx.y * blockDim.x;
    return block_sum(thread_sum_parts(part, nb, tid), sm);
}

__device__ inline double reduce_max_parts(const double *part, int64_t nb, double *sm) {
    const int tid = threadIdx.x + threadIdx.y * blockDim.x;
    return block_max(thread_max_parts(part, nb, tid), sm);
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

// The head of an iteration: stop test, counter, rho rotation and beta.  It runs
// in the same thread right after the scalars it reads were written, at the end
// of k_ctl_setup and of k_ctl_end (the head of the next iteration), so no
// launch of its own is needed.
__device__ inline void ctl_begin(double *sc, int *fl) {
    if (!fl[KF_ACTIVE]) return;
    const double norm = sc[KS_NORM];
    int counter = fl[KF_COUNTER];
    bool ex = norm < sc[KS_EPS];
    if (!ex && counter == fl[KF_NMAX]) {
        ex = true;
        fl[KF_EXIT] = 1;
    }
    if (counter < 1 && norm != 0.0) ex = false;
    if (!(norm == norm) || norm > 1.0e300) {  // a non-finite residual: stop as a breakdown
        ex = true;
        fl[KF_EXIT] = 2;
    }
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
        ctl_begin(sc, fl);  // the head of the first iteration
    }
}

__global__ void k_ctl_resume(double *sc, int *fl) {
    fl[KF_RESET] = 0;
    fl[KF_ACTIVE] = 1;
    ctl_head_tail(sc, fl, fl[KF_COUNTER]);
}

// Every workgroup of the two update kernels below first reduces the partials
// of the operator kernel before it (the same fixed tree as the single-workgroup
// control kernels), so the scalar step of BiCGstab needs no launch of its own.
// Workgroup 0 publishes the scalar, or the breakdown flags, in which case no
// workgroup updates.  The active flag is read once per workgroup: workgroup 0
// may clear it while later workgroups start, and either value they see leads
// to the same decision.
__device__ inline bool wg_active(const int *fl, int *s_go) {
    if (threadIdx.x == 0 && threadIdx.y == 0) *s_go = fl[KF_ACTIVE];
    __syncthreads();
    return *s_go != 0;
}

__device__ inline void breakdown(int *fl) {
    fl[KF_EXIT] = 2;
    fl[KF_ACTIVE] = 0;
    fl[KF_DONE] = 1;
}

// alpha = rhr/<rh, Ap> (a vanishing <rh, Ap> is a breakdown), then rr = rr - alpha*Ap
// (src/mod_solvers.f90:1709-1740)
template <bool FIRST = false>
__device__ inline void s_update_block(const Geo &g, const Range3 &r, int bx, int by, int bz, int nbz, int64_t pi,
                                      double *rr, const double *Ap, const double *part, int64_t nb, double *sc,
                                      int *fl, const FirstSweep &fs = FirstSweep{}) {
    __shared__ double sm[4];
    __shared__ double s_alpha;
    __shared__ int s_go;
    if (!wg_active(fl, &s_go)) return;
    const double rhAp = reduce_sum_parts(part, nb, sm);
    if (threadIdx.x == 0 && threadIdx.y == 0) {
        const bool ok = !(fabs(rhAp) == 0.0);
        const double alpha = ok ? sc[KS_RHR] / rhAp : 0.0;
        s_alpha = alpha;
        s_go = ok;
        if (pi == 0) {
            if (ok)
                sc[KS_ALPHA] = alpha;
            else
                breakdown(fl);
        }
    }
    __syncthreads();
    if (!s_go) return;
    const double alpha = s_alpha;
    const int i = r.lo0 + bx * RED_TX + (int)threadIdx.x;
    const int j = r.lo1 + by * RED_TY + (int)threadIdx.y;
    if (i >= r.lo0 + r.n0 || j >= r.lo1 + r.n1) return;
    int kc0, kc1;
    red_chunk(r, bz, nbz, kc0, kc1);
    for (int k = kc0; k <= kc1; ++k) {
        const int64_t x = gidx(g, i, j, k);
        const double v = rr[x] - alpha * Ap[x];
        rr[x] = v;
        if (FIRST) first_sweep_point(g, fs, i, j, k, x, v);
    }
}

template <bool FIRST>
__global__ __launch_bounds__(256) void k_s_update_alpha(Geo g, Range3 r, double *rr, const double *Ap,
                                                        const double *part, int64_t nb, double *sc, int *fl,
                                                        FirstSweep fs) {
    s_update_block<FIRST>(g, r, (int)blockIdx.x, (int)blockIdx.y, (int)blockIdx.z, (int)gridDim.z, part_index(), rr,
                          Ap, part, nb, sc, fl, fs);
}

// omega = <r, Ar>/<Ar, Ar> (a vanishing <Ar, Ar> is a breakdown), then
// phi = phi + omega*ra + alpha*pa; rr = rr - omega*Ar, with the max-norm and
// <rr, rh> partials of the new residual in pmax/pdot (ra = rr, pa = pk without
// a preconditioner; ra = z2, pa = z1 with one, src/mod_solvers.f90:1782-1797, 1855-1899).
// pmax/pdot must not alias p1/p2, which other workgroups may still be reducing.
__device__ inline void x_update_block(const Geo &g, const Range3 &r, int bx, int by, int bz, int nbz, int64_t pi,
                                      double *phi, double *rr, const double *ra, const double *pa, const double *Ar,
                                      const double *rh, const double *p1, const double *p2, int64_t nb, double *sc,
                                      int *fl, double *pmax, double *pdot) {
    __shared__ double sm[8];
    __shared__ double s_om;
    __shared__ int s_go;
    if (!wg_active(fl, &s_go)) return;
    double t1, t2;
    thread_sum_parts2(p1, p2, nb, (int)(threadIdx.x + threadIdx.y * blockDim.x), t1, t2);
    const double rAr = block_sum(t1, sm);
    __syncthreads();
    const double ArAr = block_sum(t2, sm + 4);
    if (threadIdx.x == 0 && threadIdx.y == 0) {
        const bool ok = !(fabs(ArAr) == 0.0);
        const double om = ok ? rAr / ArAr : 0.0;
        s_om = om;
        s_go = ok;
        if (pi == 0) {
            if (ok)
                sc[KS_OMEGA] = om;
            else
                breakdown(fl);
        }
    }
    __syncthreads();
    if (!s_go) return;
    const double om = s_om, alpha = sc[KS_ALPHA];
    const int i = r.lo0 + bx * RED_TX + (int)threadIdx.x;
    const int j = r.lo1 + by * RED_TY + (int)threadIdx.y;
    double m = 0.0, d = 0.0;
    if (i < r.lo0 + r.n0 && j < r.lo1 + r.n1) {
        int kc0, kc1;
        red_chunk(r, bz, nbz, kc0, kc1);
        // two register sets: the loads of plane k+1 are in flight while plane k is
        // updated (a thread's planes are distinct elements, so no store changes a
        // value loaded for a later plane; ra is rr itself without a preconditioner,
        // and both then read the old r).  No memory access sits on a branch, so the
        // waits can count exactly: past the chunk, the loads read its last plane
        // again, and the one update past it (odd chunk lengths) rewrites that plane
        // from the values loaded before its first update, the same bits again,
        // while the norm and the dot skip it.
        const int64_t x0 = gidx(g, i, j, kc0);
        const int64_t pxy = g.pxy;
        struct In {
            double ph, ra, pa, rr, Ar, rh;
        };
        In A, B;
        auto ld = [&](int k, In &v) {
            const int64_t x = x0 + (int64_t)((k < kc1 ? k : kc1) - kc0) * pxy;
            v.ph = phi[x];
            v.ra = ra[x];
            v.pa = pa[x];
            v.rr = rr[x];
            v.Ar = Ar[x];
            v.rh = rh[x];
        };
        auto up = [&](int k, const In &v) {
            const int64_t x = x0 + (int64_t)((k < kc1 ? k : kc1) - kc0) * pxy;
            phi[x] = v.ph + om * v.ra + alpha * v.pa;
            const double w = v.rr - om * v.Ar;
            rr[x] = w;
            const double a = fabs(w);
            const bool in = k <= kc1;
            m = in ? max_nan(a, m) : m;
            d = in ? d + w * v.rh : d;
        };
        if (kc0 <= kc1) ld(kc0, A);
        for (int k = kc0; k <= kc1; k += 2) {
            ld(k + 1, B);
            up(k, A);
            ld(k + 2, A);
            up(k + 1, B);
        }
    }
    const double bm = block_max(m, sm);
    __syncthreads();
    const double bs = block_sum(d, sm + 4);
    if (threadIdx.x == 0 && threadIdx.y == 0) {
        pmax[pi] = bm;
        pdot[pi] = bs;
    }
}

__global__ __launch_bounds__(256) void k_x_update_omega(Geo g, Range3 r, double *phi, double *rr, const double *ra,
                                                        const double *pa, const double *Ar, const double *rh,
                                                        const double *p1, const double *p2, int64_t nb, double *sc,
                                                        int *fl, double *pmax, double *pdot) {
    x_update_block(g, r, (int)blockIdx.x, (int)blockIdx.y, (int)blockIdx.z, (int)gridDim.z, part_index(), phi, rr,
                   ra, pa, Ar, rh, p1, p2, nb, sc, fl, pmax, pdot);
}

__global__ __launch_bounds__(256) void k_ctl_end(const double *pmax, const double *pdot, int64_t nb, double *sc,
                                                 int *fl) {
    if (!fl[KF_ACTIVE]) return;
    __shared__ double sm[8];
    double tm, ts;
    thread_max_sum_parts(pmax, pdot, nb, (int)(threadIdx.x + threadIdx.y * blockDim.x), tm, ts);
    const double nm = block_max(tm, sm);
    __syncthreads();
    const double dt = block_sum(ts, sm + 4);
    if (threadIdx.x == 0 && threadIdx.y == 0) {
        sc[KS_NORM] = nm;
        sc[KS_RHR_NEXT] = dt;
        ctl_begin(sc, fl);  // the head of the next iteration
    }
}

__global__ __launch_bounds__(256) void k_set_rhr(const double *part, int64_t nb, double *sc) {
    __shared__ double sm[4];
    const double v = reduce_sum_parts(part, nb, sm);
    if (threadIdx.x == 0 && threadIdx.y == 0) sc[KS_RHR] = v;
}

// ---------------------------------------------------------------------------
// Batched BiCGstab.  One launch covers the tiles of many cases (any grid
// sizes): tile_case maps a global tile to its case, and a case's tiles are its
// operator/reduction tiles in the single-case numbering from its tile0.  Every
// block runs the block function of the single-case kernel for that tile and
// writes the partial of the same index into the case's own buffers, and every
// scalar step reads the case's own scalars, so each case's iteration is
// bit-identical to its own solve.  A case whose solve has stopped (its active
// flag cleared) makes its blocks return at once.
// ---------------------------------------------------------------------------
__device__ inline int batch_tile(const KBatchCase *cases, const int *tile_case, int ntiles, int &bx, int &by,
                                 int &bz, int64_t &pi) {
    int t, d1, d2;
    if (!xcd_tile(ntiles, 1, 1, t, d1, d2)) return -1;
    const int q = tile_case[t];
    const int local = t - cases[q].tile0;
    const int nbx = cases[q].nbx, nby = cases[q].nby;
    bx = local % nbx;
    const int rem = local / nbx;
    by = rem % nby;
    bz = rem / nby;
    pi = local;
    return q;
}

template <int DM, bool PZ>
__global__ __launch_bounds__(RED_TX *RED_TY, 3) void k_apply_batch(const KBatchCase *__restrict__ cases,
                                                                   const int *__restrict__ tile_case, int ntiles,
                                                                   int form, int pre) {
    int bx, by, bz;
    int64_t pi;
    const int q = batch_tile(cases, tile_case, ntiles, bx, by, bz, pi);
    if (q < 0) return;
    const KBatchCase &K = cases[q];
    if ((K.g.per[2] != 0) != PZ) return;  // the other launch of this form covers it
    if (form != KB_APPLY_PHI && !K.fl[KF_ACTIVE]) return;
    const double *src = K.phi, *w = nullptr;
    double *dst = K.rr;
    if (form == KB_APPLY_W) {
        src = pre ? K.z1 : K.pk;
        dst = K.Ap;
        w = K.rh;
    } else if (form == KB_APPLY_SELF) {
        src = K.rr;
        dst = K.Ar;
    } else if (form == KB_APPLY_PAIR) {
        src = K.z2;
        dst = K.Ar;
        w = K.rr;
    }
    apply_block<RED_TX, RED_TY, RED_KC, DM, PZ>(K.g, K.c, K.r, K.nbx, K.nby, K.nbz, bx, by, bz, src, dst, w, K.part1,
                                                K.part2);
}

// the first-sweep operands of a batched case
__device__ inline FirstSweep batch_first(const KBatchCase &K, float omega) {
    return FirstSweep{K.c.dgf[0], K.c.dgf[1], K.c.dgf[2], omega, K.b32, K.e32};
}

// pk = rr + beta*pk + om2*Ap (rr in the first iteration) over a reduction tile;
// FIRST: with M's first FP32 sweep of pk
template <bool FIRST>
__global__ __launch_bounds__(256) void k_p_update_batch(const KBatchCase *__restrict__ cases,
                                                        const int *__restrict__ tile_case, int ntiles, float omega) {
    int bx, by, bz;
    int64_t pi;
    const int q = batch_tile(cases, tile_case, ntiles, bx, by, bz, pi);
    if (q < 0) return;
    const KBatchCase &K = cases[q];
    if (!K.fl[KF_ACTIVE]) return;
    const Range3 &r = K.r;
    const int i = r.lo0 + bx * RED_TX + (int)threadIdx.x;
    const int j = r.lo1 + by * RED_TY + (int)threadIdx.y;
    if (i >= r.lo0 + r.n0 || j >= r.lo1 + r.n1) return;
    int kc0, kc1;
    red_chunk(r, bz, K.nbz, kc0, kc1);
    const bool mode = K.fl[KF_MODE] != 0;
    const double beta = K.sc[KS_BETA], om2 = K.sc[KS_OM2];
    const FirstSweep fs = batch_first(K, omega);
    for (int k = kc0; k <= kc1; ++k) {
        const int64_t x = gidx(K.g, i, j, k);
        const double p = mode ? K.rr[x] + beta * K.pk[x] + om2 * K.Ap[x] : K.rr[x];
        K.pk[x] = p;
        if (FIRST) first_sweep_point(K.g, fs, i, j, k, x, p);
    }
}

template <bool FIRST>
__global__ __launch_bounds__(256) void k_s_update_batch(const KBatchCase *__restrict__ cases,
                                                        const int *__restrict__ tile_case, int ntiles, float omega) {
    int bx, by, bz;
    int64_t pi;
    const int q = batch_tile(cases, tile_case, ntiles, bx, by, bz, pi);
    if (q < 0) return;
    const KBatchCase &K = cases[q];
    s_update_block<FIRST>(K.g, K.r, bx, by, bz, K.nbz, pi, K.rr, K.Ap, K.part1, K.nb, K.sc, K.fl,
                          batch_first(K, omega));
}

__global__ __launch_bounds__(256) void k_x_update_batch(const KBatchCase *__restrict__ cases,
                                                        const int *__restrict__ tile_case, int ntiles, int pre) {
    int bx, by, bz;
    int64_t pi;
    const int q = batch_tile(cases, tile_case, ntiles, bx, by, bz, pi);
    if (q < 0) return;
    const KBatchCase &K = cases[q];
    x_update_block(K.g, K.r, bx, by, bz, K.nbz, pi, K.phi, K.rr, pre ? K.z2 : K.rr, pre ? K.z1 : K.pk, K.Ar, K.rh,
                   K.part1, K.part2, K.nb, K.sc, K.fl, K.part3, K.part4);
}

// one workgroup per case
__global__ __launch_bounds__(256) void k_ctl_end_batch(const KBatchCase *__restrict__ cases) {
    const KBatchCase &K = cases[blockIdx.x];
    if (!K.fl[KF_ACTIVE]) return;
    __shared__ double sm[8];
    double tm, ts;
    thread_max_sum_parts(K.part3, K.part4, K.nb, (int)(threadIdx.x + threadIdx.y * blockDim.x), tm, ts);
    const double nm = block_max(tm, sm);
    __syncthreads();
    const double dt = block_sum(ts, sm + 4);
    if (threadIdx.x == 0 && threadIdx.y == 0) {
        K.sc[KS_NORM] = nm;
        K.sc[KS_RHR_NEXT] = dt;
        ctl_begin(K.sc, K.fl);
    }
}

// rr = bb - rr, rh = rr with their partials, and max|bb| partials (part3)
__global__ __launch_bounds__(256) void k_init_rr_batch(const KBatchCase *__restrict__ cases,
                                                       const int *__restrict__ tile_case, int ntiles) {
    int bx, by, bz;
    int64_t pi;
    const int q = batch_tile(cases, tile_case, ntiles, bx, by, bz, pi);
    if (q < 0) return;
    const KBatchCase &K = cases[q];
    init_rr_block(K.g, K.r, bx, by, bz, K.nbz, pi, K.bb, K.rr, K.rh, K.part1, K.part2);
    __syncthreads();
    __shared__ double sm[4];
    const Range3 &r = K.r;
    const int i = r.lo0 + bx * RED_TX + (int)threadIdx.x;
    const int j = r.lo1 + by * RED_TY + (int)threadIdx.y;
    double m = 0.0;
    if (i < r.lo0 + r.n0 && j < r.lo1 + r.n1) {
        int kc0, kc1;
        red_chunk(r, bz, K.nbz, kc0, kc1);
        for (int k = kc0; k <= kc1; ++k) {
            const double a = fabs(K.bb[gidx(K.g, i, j, k)]);
            m = max_nan(a, m);
        }
    }
    const double bm = block_max(m, sm);
    if (threadIdx.x == 0 && threadIdx.y == 0) K.part3[pi] = bm;
}

// per case: eps = tol * max|bb| (the host's bnorm product, here on the device),
// then k_ctl_setup's work on the initial residual's partials
__global__ __launch_bounds__(256) void k_ctl_setup_batch(const KBatchCase *__restrict__ cases, double tol, int nmax) {
    const KBatchCase &K = cases[blockIdx.x];
    __shared__ double sm[8];
    const double bmax = reduce_max_parts(K.part3, K.nb, sm);
    __syncthreads();
    const double nm = reduce_max_parts(K.part1, K.nb, sm + 4);
    __syncthreads();
    const double dt = reduce_sum_parts(K.part2, K.nb, sm);
    if (threadIdx.x == 0 && threadIdx.y == 0) {
        double *sc = K.sc;
        int *fl = K.fl;
        sc[KS_NORM] = nm;
        sc[KS_NORM0] = nm;
        sc[KS_RHR_NEXT] = dt;
        sc[KS_RHR] = 0.0;
        sc[KS_ALPHA] = 0.0;
        sc[KS_OMEGA] = 0.0;
        sc[KS_EPS] = tol * bmax;
        fl[KF_ACTIVE] = 1;
        fl[KF_COUNTER] = 0;
        fl[KF_NMAX] = nmax;
        fl[KF_EXIT] = 0;
        fl[KF_RESET] = 0;
        fl[KF_MODE] = 0;
        fl[KF_DONE] = 0;
        ctl_begin(sc, fl);
    }
}

// every case's flags and scalars into one array each, for a single copy to the host
__global__ void k_gather_batch(const KBatchCase *__restrict__ cases, int ncases, int *fl_out, double *sc_out) {
    const int q = blockIdx.x;
    const int t = threadIdx.x;
    if (q >= ncases) return;
    if (t < KF_NSLOTS) fl_out[q * KF_NSLOTS + t] = cases[q].fl[t];
    if (t < KS_NSLOTS) sc_out[q * KS_NSLOTS + t] = cases[q].sc[t];
}

}  // namespace

// ===========================================================================
static inline Range3 Rk(const Box &b) { return make_range(b); }

void launch_apply_A(hipStream_t s, const Geo &g, const Coef &c, const Box &pb, const double *src, double *dst,
                    const double *w, int dot_mode, double *part1, double *part2, const int *active) {
    Range3 r = Rk(pb);
    if (empty(r)) return;
    // the operator tile is the reduction tile, so the fused dot partials follow the fixed tree
    const int nbx = (r.n0 + RED_TX - 1) / RED_TX, nby = (r.n1 + RED_TY - 1) / RED_TY, nbz = (r.n2 + RED_KC - 1) / RED_KC;
    const dim3 grid(xcd_grid((int64_t)nbx * nby * nbz)), block(RED_TX, RED_TY, 1);
#define IH_APPLY(DMV)                                                                                           \
    do {                                                                                                        \
        if (g.per[2])                                                                                           \
            hipLaunchKernelGGL((k_apply<RED_TX, RED_TY, RED_KC, DMV, true>), grid, block, 0, s, g, c, r, nbx, nby, \
                               nbz, src, dst, w, part1, part2, active);                                         \
        else                                                                                                    \
            hipLaunchKernelGGL((k_apply<RED_TX, RED_TY, RED_KC, DMV, false>), grid, block, 0, s, g, c, r, nbx,     \
                               nby, nbz, src, dst, w, part1, part2, active);                                    \
    } while (0)
    if (dot_mode == DOT_W)
        IH_APPLY(DOT_W);
    else if (dot_mode == DOT_SELF)
        IH_APPLY(DOT_SELF);
    else if (dot_mode == DOT_PAIR)
        IH_APPLY(DOT_PAIR);
    else
        IH_APPLY(DOT_NONE);
#undef IH_APPLY
}

void launch_krylov_init(hipStream_t s, const Geo &g, const Box &b, const double *bb, double *rr, double *rh,
                        double *pmax, double *pdot, double *sc, int *fl, double eps, int nmax) {
    Range3 r = Rk(b);
    hipLaunchKernelGGL(k_init_rr, red_grid(r), dim3(RED_TX, RED_TY, 1), 0, s, g, r, bb, rr, rh, pmax, pdot);
    hipLaunchKernelGGL(k_ctl_setup, dim3(1), dim3(RED_TX, RED_TY, 1), 0, s, pmax, pdot, red_blocks(b), sc, fl, eps,
                       nmax);
}

// z = M v: `sweeps` FP32 JOR sweeps on A_cdg e = (float)v from e = 0, z = (double)e.
// The first sweep reads v itself (and stores b = (float)v), the last writes z
// widened to double, so M costs sweeps passes over the box and no more.  With
// first_done the kernel that wrote v has formed the first sweep already.
static void launch_precond(hipStream_t s, const Geo &g, const Coef &c, const Box &b, double *v, double *z,
                           const KrylovPtrs &k, bool first_done = false) {
    const int *act = k.fl + KF_ACTIVE;
    if (k.mg) {
        // multigridV(.TRUE.,1,v,z) (src/mod_solvers.f90:1688, 1751)
        if (k.mg_fp32)
            launch_mg_vcycle_f32(s, *k.mg, v, z, act);
        else
            launch_mg_vcycle(s, *k.mg, v, z, act);
        return;
    }
    if (k.sweeps64 > 0) {
        // FP64 sweeps on A e = v from e = 0 (the like-for-like FP64 form of the FP32 M):
        // the first reads the zero field, the last writes z
        const double *x = k.zero64;
        for (int q = 0; q < k.sweeps64; ++q) {
            double *y = q == k.sweeps64 - 1 ? z : (q % 2 == 0 ? k.e64a : k.e64b);
            launch_jor(s, g, c, b, k.omega64, x, v, y, act, k.shape64);
            x = y;
        }
        return;
    }
    if (k.sweeps == 1) {
        launch_jor_f32_first(s, g, c, b, k.omega, v, k.b32, nullptr, z, act);
        return;
    }
    if (!first_done) launch_jor_f32_first(s, g, c, b, k.omega, v, k.b32, k.e0, nullptr, act);
    float *x = k.e0, *y = k.e1;
    for (int q = 1; q < k.sweeps - 1; ++q) {
        launch_jor_f32(s, g, c, b, k.omega, x, k.b32, y, act, k.shape);
        std::swap(x, y);
    }
    launch_jor_f32_to64(s, g, c, b, k.omega, x, k.b32, z, act, k.shape);
}

void launch_krylov_iteration(hipStream_t s, const Geo &g, const Coef &c, const Box &b, const KrylovPtrs &k) {
    Range3 r = Rk(b);
    const int64_t nb = red_blocks(b);
    const bool pre = k.sweeps > 0 || k.sweeps64 > 0 || k.mg;
    // FP32 sweeps as M (two or more): the p and s updates also form M's first sweep
    const bool fuse = !k.mg && k.sweeps64 == 0 && k.sweeps >= 2;
    const FirstSweep fs{c.dgf[0], c.dgf[1], c.dgf[2], k.omega, k.b32, k.e0};
    if (fuse)
        hipLaunchKernelGGL(k_p_update<true>, grid_for(r), dim3(BX, BY, 1), 0, s, g, r, k.rr, k.pk, k.Ap, k.sc, k.fl,
                           fs);
    else
        hipLaunchKernelGGL(k_p_update<false>, grid_for(r), dim3(BX, BY, 1), 0, s, g, r, k.rr, k.pk, k.Ap, k.sc, k.fl,
                           FirstSweep{});
    if (pre) launch_precond(s, g, c, b, k.pk, k.z1, k, fuse);
    launch_apply_A(s, g, c, b, pre ? k.z1 : k.pk, k.Ap, k.rh, DOT_W, k.part1, nullptr, k.fl + KF_ACTIVE);
    if (fuse)
        hipLaunchKernelGGL(k_s_update_alpha<true>, red_grid(r), dim3(RED_TX, RED_TY, 1), 0, s, g, r, k.rr, k.Ap,
                           k.part1, nb, k.sc, k.fl, fs);
    else
        hipLaunchKernelGGL(k_s_update_alpha<false>, red_grid(r), dim3(RED_TX, RED_TY, 1), 0, s, g, r, k.rr, k.Ap,
                           k.part1, nb, k.sc, k.fl, FirstSweep{});
    if (pre) {
        launch_precond(s, g, c, b, k.rr, k.z2, k, fuse);
        launch_apply_A(s, g, c, b, k.z2, k.Ar, k.rr, DOT_PAIR, k.part1, k.part2, k.fl + KF_ACTIVE);
    } else {
        launch_apply_A(s, g, c, b, k.rr, k.Ar, nullptr, DOT_SELF, k.part1, k.part2, k.fl + KF_ACTIVE);
    }
    hipLaunchKernelGGL(k_x_update_omega, red_grid(r), dim3(RED_TX, RED_TY, 1), 0, s, g, r, k.phi, k.rr,
                       pre ? k.z2 : k.rr, pre ? k.z1 : k.pk, k.Ar, k.rh, k.part1, k.part2, nb, k.sc, k.fl, k.part3,
                       k.part4);
    hipLaunchKernelGGL(k_ctl_end, dim3(1), dim3(RED_TX, RED_TY, 1), 0, s, k.part3, k.part4, nb, k.sc, k.fl);
}

void launch_krylov_part(hipStream_t s, const Geo &g, const Coef &c, const Box &b, const KrylovPtrs &k, int which) {
    Range3 r = Rk(b);
    const int64_t nb = red_blocks(b);
    switch (which) {
    case KPART_P_UPDATE:
        hipLaunchKernelGGL(k_p_update<false>, grid_for(r), dim3(BX, BY, 1), 0, s, g, r, k.rr, k.pk, k.Ap, k.sc, k.fl,
                           FirstSweep{});
        break;
    case KPART_APPLY_W:
        launch_apply_A(s, g, c, b, k.pk, k.Ap, k.rh, DOT_W, k.part1, nullptr, k.fl + KF_ACTIVE);
        break;
    case KPART_S_UPDATE:
        hipLaunchKernelGGL(k_s_update_alpha<false>, red_grid(r), dim3(RED_TX, RED_TY, 1), 0, s, g, r, k.rr, k.Ap,
                           k.part1, nb, k.sc, k.fl, FirstSweep{});
        break;
    case KPART_APPLY_SELF:
        launch_apply_A(s, g, c, b, k.rr, k.Ar, nullptr, DOT_SELF, k.part1, k.part2, k.fl + KF_ACTIVE);
        break;
    case KPART_X_UPDATE:
        hipLaunchKernelGGL(k_x_update_omega, red_grid(r), dim3(RED_TX, RED_TY, 1), 0, s, g, r, k.phi, k.rr, k.rr,
                           k.pk, k.Ar, k.rh, k.part1, k.part2, nb, k.sc, k.fl, k.part3, k.part4);
        break;
    default:
        break;
    }
}

void launch_apply_batch(hipStream_t s, const KBatchCase *cases, const int *tile_case, int ntiles, int form, bool pre,
                        bool any_pz, bool any_nonpz) {
    if (ntiles <= 0) return;
    const dim3 grid(xcd_grid(ntiles)), block(RED_TX, RED_TY, 1);
#define IH_APB(DMV)                                                                                               \
    do {                                                                                                          \
        if (any_nonpz)                                                                                            \
            hipLaunchKernelGGL((k_apply_batch<DMV, false>), grid, block, 0, s, cases, tile_case, ntiles, form,      \
                               (int)pre);                                                                         \
        if (any_pz)                                                                                               \
            hipLaunchKernelGGL((k_apply_batch<DMV, true>), grid, block, 0, s, cases, tile_case, ntiles, form,       \
                               (int)pre);                                                                         \
    } while (0)
    if (form == KB_APPLY_W)
        IH_APB(DOT_W);
    else if (form == KB_APPLY_SELF)
        IH_APB(DOT_SELF);
    else if (form == KB_APPLY_PAIR)
        IH_APB(DOT_PAIR);
    else
        IH_APB(DOT_NONE);
#undef IH_APB
}

void launch_krylov_init_batch(hipStream_t s, const KBatchCase *cases, int ncases, const int *tile_case, int ntiles,
                              double tol, int nmax) {
    if (ntiles <= 0) return;
    hipLaunchKernelGGL(k_init_rr_batch, dim3(xcd_grid(ntiles)), dim3(RED_TX, RED_TY, 1), 0, s, cases, tile_case,
                       ntiles);
    hipLaunchKernelGGL(k_ctl_setup_batch, dim3(ncases), dim3(RED_TX, RED_TY, 1), 0, s, cases, tol, nmax);
}

void launch_p_update_batch(hipStream_t s, const KBatchCase *cases, const int *tile_case, int ntiles, bool first,
                           float omega) {
    if (ntiles <= 0) return;
    if (first)
        hipLaunchKernelGGL(k_p_update_batch<true>, dim3(xcd_grid(ntiles)), dim3(RED_TX, RED_TY, 1), 0, s, cases,
                           tile_case, ntiles, omega);
    else
        hipLaunchKernelGGL(k_p_update_batch<false>, dim3(xcd_grid(ntiles)), dim3(RED_TX, RED_TY, 1), 0, s, cases,
                           tile_case, ntiles, omega);
}

void launch_s_update_batch(hipStream_t s, const KBatchCase *cases, const int *tile_case, int ntiles, bool first,
                           float omega) {
    if (ntiles <= 0) return;
    if (first)
        hipLaunchKernelGGL(k_s_update_batch<true>, dim3(xcd_grid(ntiles)), dim3(RED_TX, RED_TY, 1), 0, s, cases,
                           tile_case, ntiles, omega);
    else
        hipLaunchKernelGGL(k_s_update_batch<false>, dim3(xcd_grid(ntiles)), dim3(RED_TX, RED_TY, 1), 0, s, cases,
                           tile_case, ntiles, omega);
}

void launch_x_update_batch(hipStream_t s, const KBatchCase *cases, const int *tile_case, int ntiles, bool pre) {
    if (ntiles > 0)
        hipLaunchKernelGGL(k_x_update_batch, dim3(xcd_grid(ntiles)), dim3(RED_TX, RED_TY, 1), 0, s, cases, tile_case,
                           ntiles, (int)pre);
}

void launch_gather_batch(hipStream_t s, const KBatchCase *cases, int ncases, int *fl_out, double *sc_out) {
    if (ncases > 0) hipLaunchKernelGGL(k_gather_batch, dim3(ncases), dim3(64), 0, s, cases, ncases, fl_out, sc_out);
}

void launch_ctl_end_batch(hipStream_t s, const KBatchCase *cases, int ncases) {
    if (ncases > 0) hipLaunchKernelGGL(k_ctl_end_batch, dim3(ncases), dim3(RED_TX, RED_TY, 1), 0, s, cases);
}

void launch_krylov_resume(hipStream_t s, const Geo &g, const Box &b, const KrylovPtrs &k) {
    // rh = rr, rhr = <rr, rh>, then the tail of the iteration head
    launch_copy_box(s, g, b, k.rr, k.rh);
    launch_dot_partial(s, g, b, k.rr, k.rh, k.part1);
    hipLaunchKernelGGL(k_set_rhr, dim3(1), dim3(RED_TX, RED_TY, 1), 0, s, k.part1, red_blocks(b), k.sc);
    hipLaunchKernelGGL(k_ctl_resume, dim3(1), dim3(1), 0, s, k.sc, k.fl);
}

}  // namespace ih
