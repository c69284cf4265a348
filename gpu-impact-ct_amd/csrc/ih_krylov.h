// Device-controlled BiCGstab: scalar and flag slots in device memory and the
// launch entry points (krylov.hip).
#pragma once

#include "ih_internal.h"

namespace ih {

struct MgDev;

// double slots
enum {
    KS_RHR = 0,
    KS_RHR_PREV,
    KS_RHR_NEXT,
    KS_ALPHA,
    KS_BETA,
    KS_OMEGA,
    KS_OM2,
    KS_NORM,
    KS_NORM0,
    KS_EPS,
    KS_NSLOTS = 16
};
// int slots
enum { KF_ACTIVE = 0, KF_COUNTER, KF_NMAX, KF_EXIT, KF_RESET, KF_MODE, KF_DONE, KF_NSLOTS = 8 };

struct KrylovPtrs {
    double *rr, *rh, *pk, *Ap, *Ar, *phi;
    double *part1, *part2;
    double *sc;
    int *fl;
    // right preconditioner M = `sweeps` FP32 JOR sweeps on cdg from zero (0: none):
    // z1 = M pk, z2 = M rr; e0/e1 ping-pong, b32 the FP32 right-hand side
    double *z1 = nullptr, *z2 = nullptr;
    float *e0 = nullptr, *e1 = nullptr, *b32 = nullptr;
    int sweeps = 0;
    float omega = 0.8f;
    JorLaunch shape;  // launch shape of the FP32 sweeps
    // the same M with the sweeps in FP64 (sweeps64 > 0 instead of sweeps): e64a/e64b
    // ping-pong, zero64 a field of zeros (the first sweep's start), omega64 the factor
    int sweeps64 = 0;
    double omega64 = 0.8;
    double *e64a = nullptr, *e64b = nullptr;
    const double *zero64 = nullptr;
    JorLaunch shape64;
    // right preconditioner M = the multigrid V cycle (preconditioner 1) instead
    const MgDev *mg = nullptr;
    bool mg_fp32 = false;  // the cycle in FP32 (z = (double) of its result)
    // max-norm and <r, rh> partials of the solution update: their own buffers,
    // because that kernel reduces the operator's partials in part1/part2 itself
    double *part3 = nullptr, *part4 = nullptr;
};

// dot modes of the operator kernel
enum { KDOT_NONE = 0, KDOT_W = 1, KDOT_SELF = 2, KDOT_PAIR = 3 };

// One case of a batched BiCGstab: its geometry, coefficients and pressure
// box, its operator/reduction tiles (16 x 16 columns, z chunks of <= 32
// planes; nb = nbx*nby*nbz) numbered from tile0 in the launch's global tile
// range, its work vectors, partial buffers and device scalars/flags.  bb is
// the right-hand side; z1/z2 the preconditioned directions (or null).
struct KBatchCase {
    Geo g;
    Coef c;
    Range3 r;
    int nbx, nby, nbz, tile0;
    int64_t nb;
    double *rr, *rh, *pk, *Ap, *Ar, *phi, *z1, *z2, *bb;
    double *part1, *part2, *part3, *part4;
    double *sc;
    int *fl;
    float *b32, *e32;  // FP32 preconditioner scratch: the sweeps' right side and first iterate
};
// batched operator forms: A phi -> rr (initial residual, no dot), A (z1 or pk) -> Ap
// with <Ap, rh>, A rr -> Ar with <rr, Ar>, <Ar, Ar>, A z2 -> Ar with <rr, Ar>, <Ar, Ar>
enum { KB_APPLY_PHI = 0, KB_APPLY_W = 1, KB_APPLY_SELF = 2, KB_APPLY_PAIR = 3 };
void launch_apply_batch(hipStream_t s, const KBatchCase *cases, const int *tile_case, int ntiles, int form, bool pre,
                        bool any_pz, bool any_nonpz);
// rr = bb - rr (rr holding A phi), rh = rr, the max-norm/<rr, rh> partials, and
// max|bb| partials in part3; then per case eps = tol * max|bb| and the iteration
// head (k_ctl_setup's work)
void launch_krylov_init_batch(hipStream_t s, const KBatchCase *cases, int ncases, const int *tile_case, int ntiles,
                              double tol, int nmax);
// one batched iteration: p update, [M pk], A, alpha/s update, [M rr], A, omega/x update,
// head of the next iteration.  M (pre) is applied by the caller-provided functor
// first: the update also forms the first FP32 preconditioner sweep of its result
// into b32/e32 (M = two or more FP32 sweeps), as the single-case iteration does
void launch_p_update_batch(hipStream_t s, const KBatchCase *cases, const int *tile_case, int ntiles, bool first,
                           float omega);
void launch_s_update_batch(hipStream_t s, const KBatchCase *cases, const int *tile_case, int ntiles, bool first,
                           float omega);
void launch_x_update_batch(hipStream_t s, const KBatchCase *cases, const int *tile_case, int ntiles, bool pre);
void launch_ctl_end_batch(hipStream_t s, const KBatchCase *cases, int ncases);
// fl_out[q*KF_NSLOTS + s] = case q's flag s, sc_out[q*KS_NSLOTS + s] its scalar s
void launch_gather_batch(hipStream_t s, const KBatchCase *cases, int ncases, int *fl_out, double *sc_out);

void launch_apply_A(hipStream_t s, const Geo &g, const Coef &c, const Box &pb, const double *src, double *dst,
                    const double *w, int dot_mode, double *part1, double *part2, const int *active);
void launch_krylov_init(hipStream_t s, const Geo &g, const Box &b, const double *bb, double *rr, double *rh,
                        double *pmax, double *pdot, double *sc, int *fl, double eps, int nmax);
// one iteration from p_update to the next iteration's head (k_ctl_end); the head
// of the first one runs in launch_krylov_init's last kernel
void launch_krylov_iteration(hipStream_t s, const Geo &g, const Coef &c, const Box &b, const KrylovPtrs &k);
void launch_krylov_resume(hipStream_t s, const Geo &g, const Box &b, const KrylovPtrs &k);
// one kernel of the unpreconditioned iteration, for timing it alone
enum { KPART_P_UPDATE = 0, KPART_APPLY_W, KPART_S_UPDATE, KPART_APPLY_SELF, KPART_X_UPDATE };
void launch_krylov_part(hipStream_t s, const Geo &g, const Coef &c, const Box &b, const KrylovPtrs &k, int which);

}  // namespace ih
