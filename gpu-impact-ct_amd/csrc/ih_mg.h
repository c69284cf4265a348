// Multigrid V cycle on the device: the BiCGstab preconditioner 1
// (multigridV, src/mod_solvers.f90:456-585) for one case.
#pragma once

#include "ih_internal.h"

namespace ih {

constexpr int MG_LEVELS_MAX = 15;

// One level.  Arrays follow the case layout rules (halo 3, rows padded to 16,
// lead 12), so the JOR sweep runs on them unchanged.
struct MgLevel {
    Geo g;
    Coef c;            // dg of the level (other tables unused)
    Box pr;            // relaxation range: the level's pressure range S?p..N?p
    double *A = nullptr;  // right-hand side (levels >= 2)
    double *B = nullptr;  // iterate (levels >= 2)
    double *D = nullptr;  // second iterate of the out-of-place sweeps
    double *C = nullptr;  // residual product, then the interpolation work of the level below
    const double *cR[3] = {nullptr, nullptr, nullptr};  // restriction weights, (ii+1)*cst + i (levels >= 2)
    const double *cI[3] = {nullptr, nullptr, nullptr};  // interpolation weights, (t-1)*cst + i (levels < G)
    int dn[3] = {1, 1, 1};  // coarsening ratio to the next level
    JorLaunch shape;
    // the FP32 cycle: float arrays (level 1: Af holds (float) of the right-hand
    // side) and tables rounded once (c.dgf holds cdg)
    float *Af = nullptr, *Bf = nullptr, *Cf = nullptr, *Df = nullptr;
    const float *cRf[3] = {nullptr, nullptr, nullptr};
    const float *cIf[3] = {nullptr, nullptr, nullptr};
    JorLaunch shape32;
};

// A level as the fused coarse-level kernel reads it from device memory, for
// the arithmetic type T of the cycle.
template <typename T>
struct MgLevelDev {
    Geo g;
    const T *dg[3];
    const T *cR[3];
    const T *cI[3];
    int dn[3];
    T *A, *B, *C, *D;
};

struct MgDev {
    int G = 0;
    int n_down = 4, n_up = 4, n_bottom = 4;
    double omega = 0.98;  // src/mod_laplace.f90:470
    MgLevel L[MG_LEVELS_MAX + 1];  // 1-based; level 1 uses the case's geometry and cdg
    // levels g0..G run inside one single-workgroup launch (g0 > G: none);
    // Ldev: device copy of levels 1..G (index 0 unused)
    int g0 = MG_LEVELS_MAX + 1;
    const MgLevelDev<double> *Ldev = nullptr;
    const MgLevelDev<float> *Ldev32 = nullptr;
};
// points up to which a level runs inside the fused coarse-level launch
// (one workgroup is latency-bound: a level of 2048 points takes two trips of
// every step per thread, 9.5K points measured slower than its own launches)
constexpr int64_t MG_FUSED_POINTS = 2048;

// z = M v.  v's edges of wall pairs are averaged in place (restrict,
// src/mod_solvers.f90:632-661), as the reference does to its pp and rr.
void launch_mg_vcycle(hipStream_t s, const MgDev &m, double *v, double *z, const int *active);
// the same cycle in FP32 on b32 = (float) v, z = (double) of the result; v's
// edges are averaged in FP64 as well (no counterpart in the reference)
void launch_mg_vcycle_f32(hipStream_t s, const MgDev &m, double *v, double *z, const int *active);

}  // namespace ih
