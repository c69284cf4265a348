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
};

struct MgDev {
    int G = 0;
    int n_down = 4, n_up = 4, n_bottom = 4;
    double omega = 0.98;  // src/mod_laplace.f90:470
    MgLevel L[MG_LEVELS_MAX + 1];  // 1-based; level 1 uses the case's geometry and cdg
};

// z = M v.  v's edges of wall pairs are averaged in place (restrict,
// src/mod_solvers.f90:632-661), as the reference does to its pp and rr.
void launch_mg_vcycle(hipStream_t s, const MgDev &m, double *v, double *z, const int *active);

}  // namespace ih
