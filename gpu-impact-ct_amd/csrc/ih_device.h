// Device-side helpers shared by the kernel files of libimpact_hip.
#pragma once

#include "ih_internal.h"

namespace ih {
namespace dev {

constexpr int BX = 64;
constexpr int BY = 4;

using Range3 = ih::Range3;

inline Range3 make_range(const Box &b) {
    return Range3{b.lo[0], b.lo[1], b.lo[2], b.hi[0] - b.lo[0] + 1, b.hi[1] - b.lo[1] + 1, b.hi[2] - b.lo[2] + 1};
}

inline dim3 grid_for(const Range3 &r) {
    return dim3((unsigned)((r.n0 + BX - 1) / BX), (unsigned)((r.n1 + BY - 1) / BY), (unsigned)r.n2);
}

inline bool empty(const Range3 &r) { return r.n0 <= 0 || r.n1 <= 0 || r.n2 <= 0; }

inline dim3 red_grid(const Range3 &r) {
    return dim3((unsigned)((r.n0 + RED_TX - 1) / RED_TX), (unsigned)((r.n1 + RED_TY - 1) / RED_TY),
                (unsigned)((r.n2 + RED_KC - 1) / RED_KC));
}

// Planes kc0..kc1 of z chunk bz out of nbz over the range r (balanced chunks).
// 32-bit unsigned arithmetic: bz * n2 stays far below 2^32 for any field that
// fits a device, and a 64-bit division costs over a hundred scalar instructions.
__device__ inline void red_chunk(const Range3 &r, int bz, int nbz, int &kc0, int &kc1) {
    const unsigned n = (unsigned)r.n2, nb = (unsigned)nbz, b = (unsigned)bz;
    kc0 = r.lo2 + (int)((b * n) / nb);
    kc1 = r.lo2 + (int)(((b + 1u) * n) / nb) - 1;
}

#define IH_POINT(r)                                                  \
    const int i = (r).lo0 + (int)(blockIdx.x * BX + threadIdx.x);    \
    const int j = (r).lo1 + (int)(blockIdx.y * BY + threadIdx.y);    \
    const int k = (r).lo2 + (int)blockIdx.z;                         \
    if (i >= (r).lo0 + (r).n0 || j >= (r).lo1 + (r).n1) return;

__device__ inline int64_t sh(const Geo &g, int d) { return d == 0 ? 1 : (d == 1 ? (int64_t)g.px : g.pxy); }

__device__ inline int pick(int d, int i, int j, int k) { return d == 0 ? i : (d == 1 ? j : k); }

// Reduction tree: xor butterfly within the 64-lane wavefront (offsets 32..1),
// then the four wave sums of a 256-thread block as (w0+w1)+(w2+w3); lanes are
// numbered threadIdx.x + blockDim.x * threadIdx.y.
__device__ inline double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v = v + __shfl_xor(v, off, 64);
    return v;
}

// The larger of a and m, where a NaN in either wins: max-norms built from it are
// NaN as soon as one entry is, so a non-finite residual cannot pass a stop test
// as a small one.  For finite values it is the plain (a > m) ? a : m.
__device__ inline double max_nan(double a, double m) { return (a > m || a != a) ? a : m; }

__device__ inline double wave_max(double v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v = max_nan(__shfl_xor(v, off, 64), v);
    return v;
}

__device__ inline double block_sum(double v, double *sm) {
    v = wave_sum(v);
    const int tid = threadIdx.x + threadIdx.y * blockDim.x;
    if ((tid & 63) == 0) sm[tid >> 6] = v;
    __syncthreads();
    double r = 0.0;
    if (tid == 0) r = (sm[0] + sm[1]) + (sm[2] + sm[3]);
    return r;
}

__device__ inline double block_max(double v, double *sm) {
    v = wave_max(v);
    const int tid = threadIdx.x + threadIdx.y * blockDim.x;
    if ((tid & 63) == 0) sm[tid >> 6] = v;
    __syncthreads();
    double r = 0.0;
    if (tid == 0) {
        r = sm[0];
        for (int q = 1; q < 4; ++q) r = max_nan(sm[q], r);
    }
    return r;
}

// One thread's share of a reduction over nb partials: part[tid], part[tid + RED_THREADS],
// ... in that order.  The loads run RED_AHEAD strides ahead of the adds (one
// after another they would cost a memory latency per stride), and a slot past
// nb is skipped rather than added as 0.0, so the result is the plain loop's bit
// for bit.  thread_sum_parts2 forms two such sums with all loads first.
constexpr int RED_AHEAD = 4;

__device__ inline double thread_sum_parts(const double *part, int64_t nb, int tid) {
    double acc = 0.0;
    for (int64_t q0 = tid; q0 < nb; q0 += (int64_t)RED_AHEAD * RED_THREADS) {
        double v[RED_AHEAD];
#pragma unroll
        for (int u = 0; u < RED_AHEAD; ++u) {
            const int64_t q = q0 + (int64_t)u * RED_THREADS;
            v[u] = q < nb ? part[q] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < RED_AHEAD; ++u)
            if (q0 + (int64_t)u * RED_THREADS < nb) acc = acc + v[u];
    }
    return acc;
}

__device__ inline void thread_sum_parts2(const double *p1, const double *p2, int64_t nb, int tid, double &s1,
                                         double &s2) {
    double a1 = 0.0, a2 = 0.0;
    for (int64_t q0 = tid; q0 < nb; q0 += (int64_t)RED_AHEAD * RED_THREADS) {
        double v1[RED_AHEAD], v2[RED_AHEAD];
#pragma unroll
        for (int u = 0; u < RED_AHEAD; ++u) {
            const int64_t q = q0 + (int64_t)u * RED_THREADS;
            v1[u] = q < nb ? p1[q] : 0.0;
            v2[u] = q < nb ? p2[q] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < RED_AHEAD; ++u)
            if (q0 + (int64_t)u * RED_THREADS < nb) {
                a1 = a1 + v1[u];
                a2 = a2 + v2[u];
            }
    }
    s1 = a1;
    s2 = a2;
}

// thread_max_parts of pm and thread_sum_parts of ps, with all loads first
__device__ inline void thread_max_sum_parts(const double *pm, const double *ps, int64_t nb, int tid, double &m_out,
                                            double &s_out) {
    double m = 0.0, a = 0.0;
    for (int64_t q0 = tid; q0 < nb; q0 += (int64_t)RED_AHEAD * RED_THREADS) {
        double vm[RED_AHEAD], vs[RED_AHEAD];
#pragma unroll
        for (int u = 0; u < RED_AHEAD; ++u) {
            const int64_t q = q0 + (int64_t)u * RED_THREADS;
            vm[u] = q < nb ? pm[q] : 0.0;
            vs[u] = q < nb ? ps[q] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < RED_AHEAD; ++u)
            if (q0 + (int64_t)u * RED_THREADS < nb) {
                m = max_nan(vm[u], m);
                a = a + vs[u];
            }
    }
    m_out = m;
    s_out = a;
}

__device__ inline double thread_max_parts(const double *part, int64_t nb, int tid) {
    double m = 0.0;
    for (int64_t q0 = tid; q0 < nb; q0 += (int64_t)RED_AHEAD * RED_THREADS) {
        double v[RED_AHEAD];
#pragma unroll
        for (int u = 0; u < RED_AHEAD; ++u) {
            const int64_t q = q0 + (int64_t)u * RED_THREADS;
            v[u] = q < nb ? part[q] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < RED_AHEAD; ++u)
            if (q0 + (int64_t)u * RED_THREADS < nb) m = max_nan(v[u], m);
    }
    return m;
}

__device__ inline int64_t part_index() {
    return (int64_t)blockIdx.x + (int64_t)gridDim.x * ((int64_t)blockIdx.y + (int64_t)gridDim.y * blockIdx.z);
}

// XCD-aware tile order.  The command processor deals consecutive workgroups of
// a launch round robin to the 8 XCDs (each with its own L2).  Launched as a
// flat grid padded to a multiple of 8, workgroup id maps to a logical tile so
// that every XCD walks one contiguous range of tiles, and tiles that share
// halo rows and planes hit the same L2.  Returns false for the padding.
constexpr int NXCD = 8;

inline unsigned xcd_grid(int64_t tiles) { return (unsigned)((tiles + NXCD - 1) / NXCD * NXCD); }

// Tile counts stay below 2^31, so the index arithmetic is 32-bit unsigned.
__device__ inline bool xcd_tile(int nbx, int nby, int nbz, int &bx, int &by, int &bz) {
    const unsigned total = (unsigned)nbx * (unsigned)nby * (unsigned)nbz;
    const unsigned per = gridDim.x / NXCD;
    const unsigned id = blockIdx.x;
    const unsigned logical = (id % NXCD) * per + id / NXCD;
    if (logical >= total) return false;
    const unsigned row = logical / (unsigned)nbx;
    bx = (int)(logical - row * (unsigned)nbx);
    by = (int)(row % (unsigned)nby);
    bz = (int)(row / (unsigned)nby);
    return true;
}

}  // namespace dev

}  // namespace ih
