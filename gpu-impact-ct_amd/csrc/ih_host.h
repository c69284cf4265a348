// Host-side structures of libimpact_hip (not part of the public ABI).
#pragma once

#include <cstdint>
#include <map>
#include <string>
#include <vector>

#include "../../include/impact_hip.h"
#include "ih_internal.h"
#include "ih_krylov.h"

namespace ih {

struct HostGrid {
    int N[3];
    int bc[3][2];
    double L[3];
    double stretch[3];
    std::vector<double> xp[3], xu[3];  // (-3:N+3) at [i+3]
};

// Coefficients in the reference layout c(lo:hi, 0:N), cdg (-1:1, 1:N).
struct HostCoeffs {
    std::vector<double> G[3], Ip[3], D[3], Iu[3], NpD[3], NpU[3], NuD[3], NuU[3], p2[3], u2[3], dg[3];
};

void host_coords(HostGrid &hg);
bool host_coeffs(const HostGrid &hg, HostCoeffs &hc, std::string &err);

// One level of the multigrid hierarchy (preconditioner 1): sizes, cdg (-1:1,
// 1:N) at (ii+1) + 3(i-1), restriction weights cR (same layout, levels >= 2),
// interpolation weights cI (1:2, 1:N) at (t-1) + 2(i-1) (levels <= G-1), and the
// coarsening ratio dn to the next level.
struct HostMgLevel {
    int N[3];
    int dn[3];
    std::vector<double> dg[3], cR[3], cI[3];
};
constexpr int MG_MAX_LEVELS = 15;
// levels 1..G of the grid (index 0 holds level 1), at most n_grids_limit
void host_mg_levels(const HostGrid &hg, int n_grids_limit, std::vector<HostMgLevel> &out);

// CT geometry (csrc/ct.cpp)
bool vtk_read_ascii(const char *path, int dims[3], std::vector<double> &values, std::string &err);
void waveform_convert(const double *t_ms, const double *q, int npts, int M1, bool idealized,
                      std::vector<double> &t_out, std::vector<double> &q_out, double ref[3]);
double waveform_inflow(const std::vector<double> &t, const std::vector<double> &q, double time, bool &found);
void wk_pressure_flow(int branch, double time, const double ref[3], double state[3], double &p, double &q);
void wk_echo(const std::vector<double> &t_ms, const std::vector<double> &q_raw, const std::vector<double> &t_conv,
             double time, double reference_velocity, double &pre_inlet, double &pre_outlet);

// The module state of the reference's CT setup that the forcing and the GPU
// Jacobi solver read: branch switches, the waveform as read and converted, the
// reference scales, the inlet jet and labels present in the forcing range, the
// Windkessel state, the boundary pressures, and the RK substep and time step.
struct CtState {
    int ct_based = 1, idealized = 0;  // CT_based_aorta, idealized_aorta
    std::vector<double> t_ms, q_raw, t_conv, q_conv;
    double ref[3] = {0.0, 0.0, 0.0};  // reference velocity, length, time
    int inlet_voxels = 0;             // inlet_voxels_count
    int k_start = 0;                  // k_start_index
    bool has5 = false, has9 = false, has11 = false;
    double wk_state[3] = {0.0, 0.0, 0.0};  // p_end_systole_in, _out, _out_2
    double pres[3] = {0.0, 0.0, 0.0};      // pre_inlet, pre_outlet, pre_outlet_2
    int substep = 0;
    double dtime = 0.0;
};

// the error of the calling thread (ih_last_error) and ih_dot's index check, for
// entry points outside case.cpp
int set_fail(int code, const char *msg);
int box_from(const ih_case *c, const int ss[3], const int nn[3], Box &b);
// states of a bound field: the host copy is newer (uploaded before the next
// device use), both agree, the device copy is newer (downloaded by ih_sync)
enum { BIND_NONE = 0, BIND_HOST_NEWER = 1, BIND_CLEAN = 2, BIND_DEVICE_NEWER = 3 };
// the slot of ih_case::hscal that ih_dot_noalloc reads its result through
constexpr int HSCAL_NOALLOC = 31;

struct KernelTiming {
    double total_ms = 0.0;
    int64_t launches = 0;
};

struct PendingTiming {
    hipEvent_t a = nullptr, b = nullptr;
    int64_t launches = 0;
};

}  // namespace ih

struct ih_case {
    int device = 0;
    ih::Geo g{};
    ih::Coef c{};
    hipStream_t stream = nullptr;
    double *fields[IH_NFIELDS] = {};
    float *f32[3] = {};        // mixed precision: relaxed correction, right-hand side, ping-pong
    uint8_t *mask = nullptr;   // 3 * fe bytes
    int has_mask = 0;
    double *dcoef = nullptr;   // coefficient tables (SoA) + coordinates
    float *dcoef32 = nullptr;  // cdg in FP32
    double *part = nullptr;    // reduction partials
    int64_t part_cap = 0;
    double *dscal = nullptr;   // device scalars
    double *hscal = nullptr;   // pinned mirror
    double *part2 = nullptr;   // second partial buffer (fused dots)
    double *part3 = nullptr;   // partials of the BiCGstab solution update (read while part/part2 are reduced)
    double *part4 = nullptr;
    double *ksc = nullptr;     // BiCGstab scalars (device)
    int *kfl = nullptr;        // BiCGstab flags (device)
    double *hksc = nullptr;    // pinned mirrors
    int *hkfl = nullptr;
    int64_t bytes = 0;
    ih::HostGrid hg{};
    ih::HostCoeffs hc{};
    bool has_coeffs = false;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    bool ktime = false;
    std::map<std::string, ih::KernelTiming> ktimes;
    std::map<std::string, std::vector<ih::PendingTiming>> kpending;
    std::map<std::vector<int>, ih::JorLaunch> jor_shape;  // (fp32, box, level) -> launch shape of the JOR sweep
    // CT label volume (uint8 labels on the field layout) and the converted waveform
    uint8_t *ctlab = nullptr;
    ih::CtState ct;
    // multigrid V cycle (preconditioner 1), built on first use
    ih::MgDev *mg = nullptr;
    double *mgmem = nullptr;          // level arrays and coefficient tables
    void *mgldev = nullptr;           // level descriptors of the fused coarse-level kernel
    float *mgmem32 = nullptr;         // the FP32 cycle's arrays and tables
    void *mgldev32 = nullptr;
    int mg_cfg[4] = {15, 4, 4, 4};    // n_grids_limit, n_relax_down, n_relax_up, n_relax_bottom
    int64_t mg_bytes = 0;             // device bytes of the multigrid levels (in bytes while built)
    std::vector<ih::HostMgLevel> mgh;
    // residency registry of the reference-signature entry points: the host array
    // each field mirrors (or null) and which copy is newer (ih::BIND_*)
    const double *bind_host[IH_NFIELDS] = {};
    uint8_t bind_state[IH_NFIELDS] = {};
    int64_t n_up = 0, n_down = 0, bytes_up = 0, bytes_down = 0;  // whole-field transfers
};
