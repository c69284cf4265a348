// Internal declarations shared by the HIP kernels and the host side of
// libimpact_hip.  Not part of the public ABI (include/impact_hip.h).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

namespace ih {

// Geometry of one block as the kernels see it.  Field element (i,j,k) with the
// reference's Fortran indices (-3..N+3) lives at lead + (i+3) + (j+3)*px + (k+3)*pxy.
struct Geo {
    int N[3];
    int px;
    int64_t pxy;
    int lead;
    int64_t fe;          // elements allocated per field (padded)
    int per[3];          // 1 if the direction is periodic
    int Sp[3], Np[3];    // pressure range
    int SB[3][3], NB[3][3];  // [component][direction] velocity ranges incl. boundary
    int SI[3][3], NI[3][3];  // velocity ranges without boundary
    int cst[3];          // coefficient stride per direction (>= N+1)
};

__host__ __device__ inline int64_t gidx(const Geo &g, int i, int j, int k) {
    return (int64_t)g.lead + (i + 3) + (int64_t)(j + 3) * g.px + (int64_t)(k + 3) * g.pxy;
}

// Device coefficient tables, structure-of-arrays: c(ii, i) at base[(ii - lo) * stride + i].
struct Coef {
    const double *G[3];    // cGp    lo -2, 6 taps
    const double *Ip[3];   // cIpu.. lo -2, 6 taps
    const double *D[3];    // cDu1.. lo -3, 6 taps
    const double *Iu[3];   // cIup.. lo -3, 6 taps
    const double *NpD[3];  // lo -3, 7 taps
    const double *NpU[3];
    const double *NuD[3];
    const double *NuU[3];
    const double *p2[3];   // cp11..
    const double *u2[3];   // cu11..
    const double *dg[3];   // cdg: 3 taps (lo -1), index i = 1..N
    const float *dgf[3];   // cdg in FP32 for the mixed-precision sweep
    const double *xp[3];   // coordinates (-3:N+3) at [i+3]
    const double *xu[3];
    const double *zero;    // 64 doubles of 0.0, never written: the load source for positions outside the field
};

// Reduction tree constants, shared with the oracle's restatement: blocks of
// RED_TX x RED_TY threads over (i, j); along k the range is cut into
// ceil(n / RED_KC) chunks of balanced length (chunk b holds the planes
// [lo + b*n/nb, lo + (b+1)*n/nb) in integer arithmetic).
constexpr int RED_TX = 16;
constexpr int RED_TY = 16;
constexpr int RED_KC = 32;
constexpr int RED_THREADS = 256;

// Launch-side helpers implemented in kernels.hip ------------------------------
struct Box {
    int lo[3], hi[3];
};

// An index range as the kernels take it: lo and extent per direction.
struct Range3 {
    int lo0, lo1, lo2, n0, n1, n2;
};

void launch_exchange(hipStream_t s, const Geo &g, int dir, int nf, double *const *f, const int *vel_dir);
void launch_exchange_relax(hipStream_t s, const Geo &g, double *f);
void launch_interp_vel_pre3(hipStream_t s, const Geo &g, const Coef &c, const double *u, const double *v,
                            const double *w, double *w1, double *w2, double *w3);
void launch_interp_vel_pre(hipStream_t s, const Geo &g, const Coef &c, int q, const double *phi, double *inter);
void launch_interp_pre_vel(hipStream_t s, const Geo &g, const Coef &c, int q, const double *phi, double *inter);
void launch_nl_helm_rhs(hipStream_t s, const Geo &g, const Coef &c, int q, const double *w1, const double *w2,
                        const double *w3, const double *vel, double *nl, double *rhs, double multL, double mult_a,
                        double mult_b, int use_b);
void launch_nonlinear_one(hipStream_t s, const Geo &g, const Coef &c, int q, int d, const double *pp,
                          const double *vel, double *nl, int assign);
void launch_helmholtz_one(hipStream_t s, const Geo &g, const Coef &c, int q, double multL, const double *vel,
                          double *nl);
void launch_bc_planes(hipStream_t s, const Geo &g, const Coef &c, double *rhs1, double *rhs2, double *rhs3,
                      int inflow_kind, double amp, const uint8_t *mask_u);
void launch_scale_box(hipStream_t s, const Geo &g, const Box &b, double *f, double a, double bb, int divide);
void launch_divergence2(hipStream_t s, const Geo &g, const Coef &c, const double *u, const double *v,
                        const double *w, double *div);
void launch_gradient(hipStream_t s, const Geo &g, const Coef &c, int q, const double *phi, double *grad);
void launch_grad_walls(hipStream_t s, const Geo &g, int q, double *grad);
void launch_grad_correct(hipStream_t s, const Geo &g, const Coef &c, int q, const double *phi, double *vel);
void launch_div_dir(hipStream_t s, const Geo &g, const Coef &c, int q, const double *dig, double *lap, int assign);
// Launch shape of a JOR sweep: threads per tile (128, 256 or 512) and planes
// per block along z; 0 selects the default of the type.
struct JorLaunch {
    int tpb = 0;
    int kc = 0;
};
// JOR sweeps
void launch_jor(hipStream_t s, const Geo &g, const Coef &c, const Box &b, double omega, const double *rel,
                const double *bb, double *out, JorLaunch shape = JorLaunch());
// the same, a no-op once *active is 0
void launch_jor(hipStream_t s, const Geo &g, const Coef &c, const Box &b, double omega, const double *rel,
                const double *bb, double *out, const int *active, JorLaunch shape);
void launch_jor_f32(hipStream_t s, const Geo &g, const Coef &c, const Box &b, float omega, const float *rel,
                    const float *bb, float *out, const int *active = nullptr, JorLaunch shape = JorLaunch());
void launch_residual_cdg(hipStream_t s, const Geo &g, const Coef &c, const Box &b, const double *rel,
                         const double *bb, double *out_res /* may be null */, double *partial_max);
void launch_copy_box(hipStream_t s, const Geo &g, const Box &b, const double *src, double *dst);
// strided box lo + st*(0..n-1) of a field <-> a dense array of n[0]*n[1]*n[2] (x fastest)
void launch_pack_box(hipStream_t s, const Geo &g, const int lo[3], const int n[3], const int st[3],
                     const double *field, double *dense);
void launch_unpack_box(hipStream_t s, const Geo &g, const int lo[3], const int n[3], const int st[3],
                       const double *dense, double *field);
void launch_axpby(hipStream_t s, const Geo &g, double a, double b, const double *x, const double *y, double *z,
                  int64_t n);
void launch_fill(hipStream_t s, double *f, double v, int64_t n);
void launch_fill_f32(hipStream_t s, float *f, float v, int64_t n, const int *active = nullptr);
void launch_to_f32(hipStream_t s, const Geo &g, const Box &b, const double *src, float *dst,
                   const int *active = nullptr);
void launch_from_f32(hipStream_t s, const Geo &g, const Box &b, const float *e, double *z, const int *active = nullptr);
void launch_correct_f32(hipStream_t s, const Geo &g, const Box &b, const float *e, double *x);

// canonical reductions: returns number of partial blocks written
int64_t red_blocks(const Box &b);

// Vector numbering of a relaxation range for the JOR sweep: every row is nv
// aligned V-point vectors plus a halo vector at either end (nent entries), the
// rows one flat sequence of nent_all entries, covered by `waves` wavefronts
// that advance by 62 entries.
struct JorVecShape {
    int i0;        // first point of the first vector of each row (aligned)
    int nv;        // vectors per row covering the range
    int nent;      // entries per row: nv + 2
    int nent_all;  // nent * rows
    int waves;     // wavefronts over one plane
};

// One case of a batched relaxation (k_jor_batch): geometry, relaxation box, its
// vector numbering, cdg in FP64 and FP32, the arrays of the sweep direction,
// tiles per plane chunk (ntxy), chunks (nbz) and the case's first global tile.
struct JorBatchItem {
    Geo g;
    Range3 r;
    JorVecShape sh;
    const double *dg[3];
    const float *dgf[3];
    const void *rel, *bb;
    void *out;
    int ntxy, nbz, tile0, pad;
    const int *active;  // the case's BiCGstab active flag (nullptr: always sweep)
};
// One case of the batched first FP32 preconditioner sweep (k_jor_f32_first):
// 4-point vectors from i0, nv a row, ntile tiles a chunk, nbz chunks.
struct JorFirstItem {
    Geo g;
    Range3 r;
    const float *dgf[3];
    int i0, nv, ntile, nbz, tile0;
    const double *v;
    float *b32;
    void *e;  // float, or double when the first sweep is the last one
    const int *active;
};
// tiles of box b for the batch kernels: V points per lane (2 FP64, 4 FP32),
// tpb threads per tile, z chunks of at most zchunk planes
int jor_batch_tiles(const Box &b, const Geo &g, int V, int tpb, int zchunk, JorVecShape &sh, int &ntxy, int &nbz);
// layout 0: one flat grid over the global tiles tile_base .. tile_base + ntiles - 1
// with a tile -> item table; 1: a row of row_tiles blocks for each of nitems items
int launch_jor_batch(hipStream_t s, const JorBatchItem *items, const int *tile_item, int tile_base, int ntiles,
                     int nitems, int row_tiles, int layout, int tpb, bool fp32, double omega, bool any_periodic_x);
// the last FP32 sweep written widened to double, and the first one from e = 0
// on A e = (float)v (stores b32; writes e, or e64 widened when e64 != nullptr)
// the residual product of a multigrid level (product_div_grad_relax) over box b
// with the JOR sweep's vector lanes, in FP64 or FP32 (cdg or its FP32 copy)
void launch_lap(hipStream_t s, const Geo &g, const Coef &c, const Box &b, const double *x, double *lap,
                const int *active, JorLaunch shape);
void launch_lap(hipStream_t s, const Geo &g, const Coef &c, const Box &b, const float *x, float *lap,
                const int *active, JorLaunch shape);
void launch_jor_f32_to64(hipStream_t s, const Geo &g, const Coef &c, const Box &b, float omega, const float *rel,
                         const float *bb, double *out, const int *active, JorLaunch shape = JorLaunch());
void launch_jor_f32_first(hipStream_t s, const Geo &g, const Coef &c, const Box &b, float omega, const double *v,
                          float *b32, float *e, double *e64, const int *active);
// batched FP32 preconditioner sweeps (flat tile -> item table over ntiles tiles)
void launch_jor_batch_f32(hipStream_t s, const JorBatchItem *items, const int *tile_item, int ntiles, float omega,
                          bool to64, bool any_periodic_x);
int jor_first_tiles(const Box &b, const Geo &g, JorFirstItem &it);  // sets r, i0, nv, ntile, nbz; returns tiles
void launch_jor_f32_first_batch(hipStream_t s, const JorFirstItem *items, const int *tile_item, int ntiles,
                                float omega, bool to64);
void launch_dot_partial(hipStream_t s, const Geo &g, const Box &b, const double *a, const double *c, double *part);
void launch_sum_partial(hipStream_t s, const Geo &g, const Box &b, const double *a, double *part);
void launch_final_sum(hipStream_t s, const double *part, int64_t nb, double *out);
void launch_max_partial(hipStream_t s, const Geo &g, const Box &b, const double *a, double *part);
void launch_absdiff_partial(hipStream_t s, const Geo &g, const Box &b, const double *a, const double *c,
                            double *part);
void launch_final_max(hipStream_t s, const double *part, int64_t nb, double *out);

// BiCGstab vector updates
void launch_copy_field_box(hipStream_t s, const Geo &g, const Box &b, const double *src, double *dst);

void launch_fill_corners(hipStream_t s, const Geo &g, double *f);
void launch_sub_box(hipStream_t s, const Geo &g, const Box &b, double *f, const double *val);
void launch_apply_mask(hipStream_t s, const Geo &g, const uint8_t *mask, double *u, double *v, double *w);
void launch_ct_forcing(hipStream_t s, const Geo &g, const uint8_t *lab, double q, int found, int ideal,
                       double w_ideal, double *u, double *v, double *w);
void launch_set_pressure_bc(hipStream_t s, const Geo &g, const int cover[3], int kplane, const uint8_t *lab,
                            int type1, int type2, int type3, double p_in, double p_out, double p_out2, double *rel);

}  // namespace ih
