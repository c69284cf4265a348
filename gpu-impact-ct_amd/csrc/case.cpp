// C ABI of libimpact_hip: case lifetime, field transfer, operators, Poisson
// solvers and the RK3 projection step, all ordered on the case's HIP stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <functional>
#include <string>

#include "ih_host.h"
#include "ih_mg.h"

using namespace ih;

namespace {

thread_local std::string g_err;

int fail(int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

#define HIPCHK(expr)                                                                                   \
    do {                                                                                               \
        hipError_t e_ = (expr);                                                                        \
        if (e_ != hipSuccess) return fail(IH_EHIP, "%s: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__, __LINE__); \
    } while (0)

#define LAUNCHCHK()                                                                                    \
    do {                                                                                               \
        hipError_t e_ = hipGetLastError();                                                             \
        if (e_ != hipSuccess) return fail(IH_EHIP, "kernel launch: %s (%s:%d)", hipGetErrorString(e_), __FILE__, __LINE__); \
    } while (0)

// RK3 coefficients (src/mod_vars.f90:774-775), evaluated in double precision.
const double aRK[3] = {8.0 / 15.0, 5.0 / 12.0, 3.0 / 4.0};
const double bRK[3] = {0.0, -17.0 / 60.0, -5.0 / 12.0};

// Scalar slots in ih_case::dscal / hscal (slot HSCAL_NOALLOC of hscal belongs to ih_dot_noalloc)
enum { S_DOT0 = 0, S_DOT1, S_DOT2, S_DOT3, S_MAX, S_BETA, S_OM2, S_ALPHA, S_OMEGA, S_ALPHA2, S_SUM, S_NSLOTS = 32 };
static_assert(HSCAL_NOALLOC > S_SUM + 1 && HSCAL_NOALLOC < S_NSLOTS, "a slot of its own");

bool valid_field(int f) { return f >= 0 && f < IH_NFIELDS; }

Box pbox(const Geo &g) {
    Box b;
    for (int d = 0; d < 3; ++d) {
        b.lo[d] = g.Sp[d];
        b.hi[d] = g.Np[d];
    }
    return b;
}

bool box_ok(const Geo &g, const Box &b) {
    for (int d = 0; d < 3; ++d) {
        if (b.lo[d] < 0 || b.hi[d] > g.N[d] || b.lo[d] > b.hi[d]) return false;
        // a stencil of the box reads one layer beyond it
        if (b.lo[d] - 1 < -3 || b.hi[d] + 1 > g.N[d] + 3) return false;
    }
    return true;
}

// init_limits, src/mod_setup.f90:465-700 (single block, ls = -1)
void init_limits(Geo &g, const int bc[6]) {
    for (int d = 0; d < 3; ++d) {
        const int N = g.N[d];
        g.per[d] = bc[2 * d] == -1 ? 1 : 0;
        g.Sp[d] = 1;
        g.Np[d] = N - 1;
        for (int m = 0; m < 3; ++m) {
            g.SB[m][d] = 1;
            g.NB[m][d] = N - 1;
            g.SI[m][d] = 1;
            g.NI[m][d] = N - 1;
        }
        if (bc[2 * d] > 0)
            for (int m = 0; m < 3; ++m) {
                g.SB[m][d] = (m == d) ? 0 : 1;
                g.SI[m][d] = (m == d) ? 1 : 2;
            }
        if (bc[2 * d + 1] > 0) {
            g.Np[d] = N;
            for (int m = 0; m < 3; ++m) {
                g.NB[m][d] = N;
                g.NI[m][d] = N - 1;
            }
        }
    }
}

// the four partial buffers hold at least nb block partials
int ensure_part(ih_case *c, int64_t nb) {
    if (nb <= c->part_cap && c->part4) return IH_OK;
    nb = std::max(nb, c->part_cap);
    HIPCHK(hipStreamSynchronize(c->stream));
    for (double **q : {&c->part, &c->part2, &c->part3, &c->part4}) {
        if (*q) (void)hipFree(*q);
        *q = nullptr;
    }
    for (double **q : {&c->part, &c->part2, &c->part3, &c->part4})
        HIPCHK(hipMalloc(q, sizeof(double) * (size_t)nb));
    c->part_cap = nb;
    return IH_OK;
}

// canonical reductions into dscal[slot] (device); host reads with sync_scalars
int dot_to(ih_case *c, const Box &b, const double *a, const double *bb, int slot) {
    const int64_t nb = red_blocks(b);
    int rc = ensure_part(c, nb);
    if (rc) return rc;
    launch_dot_partial(c->stream, c->g, b, a, bb, c->part);
    launch_final_sum(c->stream, c->part, nb, c->dscal + slot);
    LAUNCHCHK();
    return IH_OK;
}

int sum_to(ih_case *c, const Box &b, const double *a, int slot) {
    const int64_t nb = red_blocks(b);
    int rc = ensure_part(c, nb);
    if (rc) return rc;
    launch_sum_partial(c->stream, c->g, b, a, c->part);
    launch_final_sum(c->stream, c->part, nb, c->dscal + slot);
    LAUNCHCHK();
    return IH_OK;
}

int max_to(ih_case *c, const Box &b, const double *a, int slot) {
    const int64_t nb = red_blocks(b);
    int rc = ensure_part(c, nb);
    if (rc) return rc;
    launch_max_partial(c->stream, c->g, b, a, c->part);
    launch_final_max(c->stream, c->part, nb, c->dscal + slot);
    LAUNCHCHK();
    return IH_OK;
}

int read_scalars(ih_case *c, int first, int count) {
    HIPCHK(hipMemcpyAsync(c->hscal + first, c->dscal + first, sizeof(double) * (size_t)count, hipMemcpyDeviceToHost,
                          c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return IH_OK;
}

int write_scalars(ih_case *c, int first, int count) {
    HIPCHK(hipMemcpyAsync(c->dscal + first, c->hscal + first, sizeof(double) * (size_t)count, hipMemcpyHostToDevice,
                          c->stream));
    return IH_OK;
}

double *F(ih_case *c, int f) { return c->fields[f]; }

// Event pair around a phase of launches on the case's stream (when enabled).
struct PhaseTimer {
    ih_case *c;
    PendingTiming p;
    const char *family;
    PhaseTimer(ih_case *cc, const char *fam) : c(cc), family(fam) {
        if (!c->ktime) return;
        if (hipEventCreate(&p.a) != hipSuccess || hipEventCreate(&p.b) != hipSuccess) {
            p.a = p.b = nullptr;
            return;
        }
        (void)hipEventRecord(p.a, c->stream);
    }
    void done(int64_t launches) {
        if (!c->ktime || !p.a) return;
        (void)hipEventRecord(p.b, c->stream);
        p.launches = launches;
        c->kpending[family].push_back(p);
        p.a = p.b = nullptr;
    }
};

void resolve_timings(ih_case *c) {
    for (auto &kv : c->kpending) {
        KernelTiming &t = c->ktimes[kv.first];
        for (PendingTiming &p : kv.second) {
            (void)hipEventSynchronize(p.b);
            float ms = 0.f;
            if (hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
                t.total_ms += ms;
                t.launches += p.launches;
            }
            (void)hipEventDestroy(p.a);
            (void)hipEventDestroy(p.b);
        }
        kv.second.clear();
    }
}

// exchange of one field along one direction
int exch(ih_case *c, int dir, int vel_dir, double *f) {
    double *fs[1] = {f};
    int vd[1] = {vel_dir};
    launch_exchange(c->stream, c->g, dir, 1, fs, vd);
    LAUNCHCHK();
    return IH_OK;
}

// product_div_grad (src/mod_laplace.f90:38-197) with the gradients kept on chip.
// The reference also refreshes phi's ghost layers (exchange(m, 0, phi)); the
// fused kernel reads them through the same rules, and the exchange is kept so
// phi leaves the call in the same state.
int apply_div_grad(ih_case *c, double *phi, double *lap) {
    for (int m = 1; m <= 3; ++m) {
        int rc = exch(c, m, 0, phi);
        if (rc) return rc;
    }
    launch_apply_A(c->stream, c->g, c->c, pbox(c->g), phi, lap, nullptr, KDOT_NONE, nullptr, nullptr, nullptr);
    LAUNCHCHK();
    return IH_OK;
}

void drop_mg(ih_case *c);

int upload_coeffs(ih_case *c) {
    const Geo &g = c->g;
    drop_mg(c);  // level 1 of the multigrid cycle holds the old tables
    // SoA layout: per direction 10 arrays of up to 7 taps plus cdg (3 taps) and coordinates
    std::vector<double> h;
    std::vector<int64_t> off_G(3), off_Ip(3), off_D(3), off_Iu(3), off_NpD(3), off_NpU(3), off_NuD(3), off_NuU(3),
        off_p2(3), off_u2(3), off_dg(3), off_xp(3), off_xu(3);
    auto push = [&](const std::vector<double> &src, int taps, int cst, int N, bool is_dg) -> int64_t {
        const int64_t o = (int64_t)h.size();
        h.resize(h.size() + (size_t)taps * cst, 0.0);
        if (!is_dg) {
            for (int i = 0; i <= N; ++i)
                for (int t = 0; t < taps; ++t) h[(size_t)o + (size_t)t * cst + i] = src[(size_t)t + (size_t)taps * i];
        } else {
            for (int i = 1; i <= N; ++i)
                for (int t = 0; t < 3; ++t) h[(size_t)o + (size_t)t * cst + i] = src[(size_t)t + 3 * (size_t)(i - 1)];
        }
        return o;
    };
    const HostCoeffs &hc = c->hc;
    for (int d = 0; d < 3; ++d) {
        const int N = g.N[d], cst = g.cst[d];
        off_G[d] = push(hc.G[d], 6, cst, N, false);
        off_Ip[d] = push(hc.Ip[d], 6, cst, N, false);
        off_D[d] = push(hc.D[d], 6, cst, N, false);
        off_Iu[d] = push(hc.Iu[d], 6, cst, N, false);
        off_NpD[d] = push(hc.NpD[d], 7, cst, N, false);
        off_NpU[d] = push(hc.NpU[d], 7, cst, N, false);
        off_NuD[d] = push(hc.NuD[d], 7, cst, N, false);
        off_NuU[d] = push(hc.NuU[d], 7, cst, N, false);
        off_p2[d] = push(hc.p2[d], 7, cst, N, false);
        off_u2[d] = push(hc.u2[d], 7, cst, N, false);
        off_dg[d] = push(hc.dg[d], 3, cst, N, true);
        off_xp[d] = (int64_t)h.size();
        h.insert(h.end(), c->hg.xp[d].begin(), c->hg.xp[d].end());
        off_xu[d] = (int64_t)h.size();
        h.insert(h.end(), c->hg.xu[d].begin(), c->hg.xu[d].end());
    }
    const int64_t off_zero = (int64_t)h.size();
    h.resize(h.size() + 64, 0.0);
    if (c->dcoef) (void)hipFree(c->dcoef);
    if (c->dcoef32) (void)hipFree(c->dcoef32);
    c->dcoef = nullptr;
    c->dcoef32 = nullptr;
    HIPCHK(hipMalloc(&c->dcoef, sizeof(double) * h.size()));
    HIPCHK(hipMemcpy(c->dcoef, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice));
    std::vector<float> h32;
    std::vector<int64_t> off32(3);
    for (int d = 0; d < 3; ++d) {
        off32[d] = (int64_t)h32.size();
        for (int t = 0; t < 3 * g.cst[d]; ++t) h32.push_back((float)h[(size_t)off_dg[d] + t]);
    }
    HIPCHK(hipMalloc(&c->dcoef32, sizeof(float) * h32.size()));
    HIPCHK(hipMemcpy(c->dcoef32, h32.data(), sizeof(float) * h32.size(), hipMemcpyHostToDevice));
    c->c.zero = c->dcoef + off_zero;
    for (int d = 0; d < 3; ++d) {
        c->c.G[d] = c->dcoef + off_G[d];
        c->c.Ip[d] = c->dcoef + off_Ip[d];
        c->c.D[d] = c->dcoef + off_D[d];
        c->c.Iu[d] = c->dcoef + off_Iu[d];
        c->c.NpD[d] = c->dcoef + off_NpD[d];
        c->c.NpU[d] = c->dcoef + off_NpU[d];
        c->c.NuD[d] = c->dcoef + off_NuD[d];
        c->c.NuU[d] = c->dcoef + off_NuU[d];
        c->c.p2[d] = c->dcoef + off_p2[d];
        c->c.u2[d] = c->dcoef + off_u2[d];
        c->c.dg[d] = c->dcoef + off_dg[d];
        c->c.dgf[d] = c->dcoef32 + off32[d];
        c->c.xp[d] = c->dcoef + off_xp[d];
        c->c.xu[d] = c->dcoef + off_xu[d];
    }
    c->has_coeffs = true;
    return IH_OK;
}

int need_coeffs(ih_case *c) {
    if (!c) return fail(IH_EINVAL, "null case");
    if (!c->has_coeffs) return fail(IH_ESTATE, "coefficients not set (call ih_case_set_coeffs or ih_case_generate_coeffs)");
    return IH_OK;
}

int to_box(const ih_case *c, const int ss[3], const int nn[3], Box &b) {
    for (int d = 0; d < 3; ++d) {
        b.lo[d] = ss[d];
        b.hi[d] = nn[d];
    }
    if (!box_ok(c->g, b)) return fail(IH_EINVAL, "index range outside the block");
    return IH_OK;
}

}  // namespace

namespace ih {
int set_fail(int code, const char *msg) { return fail(code, "%s", msg); }
int box_from(const ih_case *c, const int ss[3], const int nn[3], Box &b) { return to_box(c, ss, nn, b); }
}  // namespace ih

namespace {

double absmax_host(ih_case *c, const Box &b, const double *a, int *rc) {
    *rc = max_to(c, b, a, S_MAX);
    if (*rc) return 0.0;
    *rc = read_scalars(c, S_MAX, 1);
    return c->hscal[S_MAX];
}

// set_pressure_BC_kernel over the points pressure_bc_launcher's grid reaches:
// b = (N-1)/8 blocks of t = (N-1)/b threads per direction cover 1..b*t; a
// direction without a block, or a block of more than 1024 threads, is a launch
// that fails in the reference and writes nothing.  The plane is Fortran
// k_start_index - 4 (storage offset k+4 compared with k_start_index - 1).
int pressure_bc(ih_case *c, double *rel, int type1, int type2, int type3, double p_in, double p_out, double p_out2) {
    int cover[3];
    int64_t threads = 1;
    for (int d = 0; d < 3; ++d) {
        const int m = c->g.N[d] - 1;
        const int blocks = m / 8;
        if (blocks < 1) return IH_OK;
        const int t = m / blocks;
        threads *= t;
        cover[d] = blocks * t;
    }
    if (threads > 1024) return IH_OK;
    launch_set_pressure_bc(c->stream, c->g, cover, c->ct.k_start - 4, c->ctlab, type1, type2, type3, p_in, p_out,
                           p_out2, rel);
    LAUNCHCHK();
    return IH_OK;
}

// device_jacobi_solver's substep-3 pressure (src/jacobi_solver_host_v4.f90:
// 302-305): types 1, 1, 0, the pressures times (aRK + bRK) dtime
int jacobi_pressure_bc(ih_case *c, double *rel, int substep, double dtime) {
    if (!c->ctlab || substep != 3) return IH_OK;
    const double f = aRK[substep - 1] + bRK[substep - 1];
    const double *pr = c->ct.pres;
    return pressure_bc(c, rel, 1, 1, 0, pr[0] * f * dtime, pr[1] * f * dtime, pr[2] * f * dtime);
}

void drop_mg(ih_case *c) {
    if (!c->mg && !c->mgmem && !c->mgldev && !c->mgmem32 && !c->mgldev32) return;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    delete c->mg;
    c->mg = nullptr;
    if (c->mgmem) (void)hipFree(c->mgmem);
    c->mgmem = nullptr;
    if (c->mgldev) (void)hipFree(c->mgldev);
    c->mgldev = nullptr;
    if (c->mgmem32) (void)hipFree(c->mgmem32);
    c->mgmem32 = nullptr;
    if (c->mgldev32) (void)hipFree(c->mgldev32);
    c->mgldev32 = nullptr;
    c->mgh.clear();
    c->bytes -= c->mg_bytes;
    c->mg_bytes = 0;
}

}  // namespace

static int grid_from_desc(const ih_grid_desc *gd, HostGrid &hg);
static int64_t mg_alloc_bytes(const std::vector<HostMgLevel> &mgh);

// ===========================================================================
extern "C" {

int ih_abi_version(void) { return IH_ABI_VERSION; }

int ih_last_error(char *buf, size_t n) {
    if (!buf || n == 0) return (int)g_err.size();
    std::snprintf(buf, n, "%s", g_err.c_str());
    return (int)g_err.size();
}

int ih_device_count(int *count) {
    if (!count) return fail(IH_EINVAL, "null count");
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess) n = 0;
    *count = n;
    return IH_OK;
}

int ih_case_create(const ih_grid_desc *gd, int device, ih_case **out) {
    if (!gd || !out) return fail(IH_EINVAL, "null argument");
    *out = nullptr;
    {
        HostGrid probe;
        int rc0 = grid_from_desc(gd, probe);
        if (rc0) return rc0;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return fail(IH_EHIP, "no HIP device available");
    if (device < 0 || device >= ndev) return fail(IH_EINVAL, "device %d out of range (%d devices)", device, ndev);
    HIPCHK(hipSetDevice(device));

    ih_case *c = new ih_case();
    c->device = device;
    Geo &g = c->g;
    for (int d = 0; d < 3; ++d) g.N[d] = gd->N[d];
    init_limits(g, gd->bc);
    g.lead = 12;  // puts i = 1 of every row on a 128-byte boundary
    g.px = ((g.N[0] + 7) + 15) / 16 * 16;
    g.pxy = (int64_t)g.px * (g.N[1] + 7);
    g.fe = ((g.lead + g.pxy * (g.N[2] + 7)) + 63) / 64 * 64;
    for (int d = 0; d < 3; ++d) g.cst[d] = ((g.N[d] + 1) + 7) / 8 * 8;
    for (int d = 0; d < 3; ++d) {
        c->hg.N[d] = gd->N[d];
        c->hg.bc[d][0] = gd->bc[2 * d];
        c->hg.bc[d][1] = gd->bc[2 * d + 1];
        c->hg.L[d] = gd->L[d];
        c->hg.stretch[d] = gd->stretch[d];
    }
    auto cleanup = [&](int rc) {
        ih_case_destroy(c);
        return rc;
    };
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess)
        return cleanup(fail(IH_EHIP, "hipStreamCreate failed"));
    const size_t fbytes = sizeof(double) * (size_t)g.fe;
    for (int f = 0; f < IH_NFIELDS; ++f) {
        if (hipMalloc(&c->fields[f], fbytes) != hipSuccess)
            return cleanup(fail(IH_ENOMEM, "out of device memory allocating field %d (%zu bytes)", f, fbytes));
        if (hipMemsetAsync(c->fields[f], 0, fbytes, c->stream) != hipSuccess) return cleanup(fail(IH_EHIP, "memset"));
        c->bytes += (int64_t)fbytes;
    }
    for (int q = 0; q < 3; ++q) {
        if (hipMalloc(&c->f32[q], sizeof(float) * (size_t)g.fe) != hipSuccess)
            return cleanup(fail(IH_ENOMEM, "out of device memory (fp32 field)"));
        if (hipMemsetAsync(c->f32[q], 0, sizeof(float) * (size_t)g.fe, c->stream) != hipSuccess)
            return cleanup(fail(IH_EHIP, "memset"));
        c->bytes += (int64_t)(sizeof(float) * (size_t)g.fe);
    }
    if (hipMalloc(&c->mask, 3 * (size_t)g.fe) != hipSuccess) return cleanup(fail(IH_ENOMEM, "mask"));
    if (hipMemsetAsync(c->mask, 1, 3 * (size_t)g.fe, c->stream) != hipSuccess) return cleanup(fail(IH_EHIP, "memset"));
    c->bytes += 3 * (int64_t)g.fe;
    if (hipMalloc(&c->dscal, sizeof(double) * S_NSLOTS) != hipSuccess) return cleanup(fail(IH_ENOMEM, "scalars"));
    if (hipHostMalloc((void **)&c->hscal, sizeof(double) * S_NSLOTS, hipHostMallocDefault) != hipSuccess)
        return cleanup(fail(IH_ENOMEM, "pinned scalars"));
    std::memset(c->hscal, 0, sizeof(double) * S_NSLOTS);
    if (hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess)
        return cleanup(fail(IH_EHIP, "event create"));
    Box pb = pbox(g);
    int rc = ensure_part(c, std::max<int64_t>(red_blocks(pb), 4096));
    if (rc) return cleanup(rc);
    if (hipMalloc(&c->ksc, sizeof(double) * KS_NSLOTS) != hipSuccess) return cleanup(fail(IH_ENOMEM, "ksc"));
    if (hipMalloc(&c->kfl, sizeof(int) * KF_NSLOTS) != hipSuccess) return cleanup(fail(IH_ENOMEM, "kfl"));
    if (hipHostMalloc((void **)&c->hksc, sizeof(double) * KS_NSLOTS, hipHostMallocDefault) != hipSuccess)
        return cleanup(fail(IH_ENOMEM, "pinned"));
    if (hipHostMalloc((void **)&c->hkfl, sizeof(int) * KF_NSLOTS, hipHostMallocDefault) != hipSuccess)
        return cleanup(fail(IH_ENOMEM, "pinned"));
    if (hipStreamSynchronize(c->stream) != hipSuccess) return cleanup(fail(IH_EHIP, "stream sync"));
    *out = c;
    return IH_OK;
}

void ih_case_destroy(ih_case *c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (int f = 0; f < IH_NFIELDS; ++f)
        if (c->fields[f]) (void)hipFree(c->fields[f]);
    for (int q = 0; q < 3; ++q)
        if (c->f32[q]) (void)hipFree(c->f32[q]);
    if (c->mask) (void)hipFree(c->mask);
    if (c->ctlab) (void)hipFree(c->ctlab);
    if (c->dcoef) (void)hipFree(c->dcoef);
    if (c->dcoef32) (void)hipFree(c->dcoef32);
    if (c->part) (void)hipFree(c->part);
    if (c->part2) (void)hipFree(c->part2);
    if (c->part3) (void)hipFree(c->part3);
    if (c->part4) (void)hipFree(c->part4);
    if (c->ksc) (void)hipFree(c->ksc);
    if (c->kfl) (void)hipFree(c->kfl);
    if (c->hksc) (void)hipHostFree(c->hksc);
    if (c->hkfl) (void)hipHostFree(c->hkfl);
    if (c->dscal) (void)hipFree(c->dscal);
    if (c->hscal) (void)hipHostFree(c->hscal);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    delete c->mg;
    if (c->mgmem) (void)hipFree(c->mgmem);
    if (c->mgldev) (void)hipFree(c->mgldev);
    if (c->mgmem32) (void)hipFree(c->mgmem32);
    if (c->mgldev32) (void)hipFree(c->mgldev32);
    resolve_timings(c);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

int ih_case_set_coeffs(ih_case *c, const ih_coeff_set *cs) {
    if (!c || !cs) return fail(IH_EINVAL, "null argument");
    HIPCHK(hipSetDevice(c->device));
    for (int d = 0; d < 3; ++d) {
        const int N = c->g.N[d];
        auto take = [&](std::vector<double> &dst, const double *src, size_t n) -> bool {
            if (!src) return false;
            dst.assign(src, src + n);
            return true;
        };
        const size_t n6 = 6 * (size_t)(N + 1), n7 = 7 * (size_t)(N + 1);
        bool ok = take(c->hc.G[d], cs->cGp[d], n6) && take(c->hc.D[d], cs->cD[d], n6) &&
                  take(c->hc.Ip[d], cs->cIp[d], n6) && take(c->hc.Iu[d], cs->cIu[d], n6) &&
                  take(c->hc.NpD[d], cs->cNpD[d], n7) && take(c->hc.NpU[d], cs->cNpU[d], n7) &&
                  take(c->hc.NuD[d], cs->cNuD[d], n7) && take(c->hc.NuU[d], cs->cNuU[d], n7) &&
                  take(c->hc.p2[d], cs->cp2[d], n7) && take(c->hc.u2[d], cs->cu2[d], n7) &&
                  take(c->hc.dg[d], cs->cdg[d], 3 * (size_t)N);
        if (!ok) return fail(IH_EINVAL, "coefficient set: missing array in direction %d", d);
    }
    // coordinates for the inflow profile: generated from L/stretch
    host_coords(c->hg);
    return upload_coeffs(c);
}

int ih_case_generate_coeffs(ih_case *c) {
    if (!c) return fail(IH_EINVAL, "null case");
    for (int d = 0; d < 3; ++d)
        if (!(c->hg.L[d] > 0.0)) return fail(IH_EINVAL, "L[%d] must be positive", d);
    HIPCHK(hipSetDevice(c->device));
    host_coords(c->hg);
    std::string err;
    if (!host_coeffs(c->hg, c->hc, err)) return fail(IH_EINVAL, "coefficients: %s", err.c_str());
    return upload_coeffs(c);
}

static const std::vector<double> *coeff_by_name(const HostCoeffs &hc, const char *name) {
    static const char *uv[3] = {"u1", "v2", "w3"};
    static const char *uvw2[3] = {"u11", "v22", "w33"};
    const std::vector<double> *src = nullptr;
    for (int d = 0; d < 3 && !src; ++d) {
        char b[16];
        const char dc = (char)('1' + d);
        auto is = [&](const char *s) { return std::strcmp(name, s) == 0; };
        std::snprintf(b, sizeof b, "cGp%c", dc);
        if (is(b)) src = &hc.G[d];
        std::snprintf(b, sizeof b, "cD%s", uv[d]);
        if (is(b)) src = &hc.D[d];
        std::snprintf(b, sizeof b, "cIp%c", "uvw"[d]);
        if (is(b)) src = &hc.Ip[d];
        std::snprintf(b, sizeof b, "cI%cp", "uvw"[d]);
        if (is(b)) src = &hc.Iu[d];
        std::snprintf(b, sizeof b, "cNp%cD", dc);
        if (is(b)) src = &hc.NpD[d];
        std::snprintf(b, sizeof b, "cNp%cU", dc);
        if (is(b)) src = &hc.NpU[d];
        std::snprintf(b, sizeof b, "cN%sD", uv[d]);
        if (is(b)) src = &hc.NuD[d];
        std::snprintf(b, sizeof b, "cN%sU", uv[d]);
        if (is(b)) src = &hc.NuU[d];
        std::snprintf(b, sizeof b, "cp%c%c", dc, dc);
        if (is(b)) src = &hc.p2[d];
        std::snprintf(b, sizeof b, "c%s", uvw2[d]);
        if (is(b)) src = &hc.u2[d];
        std::snprintf(b, sizeof b, "cdg%c", dc);
        if (is(b)) src = &hc.dg[d];
    }
    return src;
}

static int grid_from_desc(const ih_grid_desc *gd, HostGrid &hg) {
    if (!gd) return fail(IH_EINVAL, "null grid");
    // the descriptor's layout is the same in every ABI version so far
    if (gd->version < 1 || gd->version > IH_ABI_VERSION)
        return fail(IH_EINVAL, "grid desc version %d, library %d", gd->version, IH_ABI_VERSION);
    for (int d = 0; d < 3; ++d) {
        if (gd->N[d] < 8 || gd->N[d] > 8192) return fail(IH_EINVAL, "N[%d] = %d out of range [8, 8192]", d, gd->N[d]);
        for (int s = 0; s < 2; ++s)
            if (!(gd->bc[2 * d + s] == -1 || gd->bc[2 * d + s] == 1))
                return fail(IH_EINVAL, "bc[%d] = %d: only -1 (periodic) and 1 (wall) are supported", 2 * d + s,
                            gd->bc[2 * d + s]);
        if ((gd->bc[2 * d] == -1) != (gd->bc[2 * d + 1] == -1))
            return fail(IH_EINVAL, "direction %d: periodic must be set on both sides", d);
        hg.N[d] = gd->N[d];
        hg.bc[d][0] = gd->bc[2 * d];
        hg.bc[d][1] = gd->bc[2 * d + 1];
        hg.L[d] = gd->L[d];
        hg.stretch[d] = gd->stretch[d];
    }
    return IH_OK;
}

int ih_case_get_coeff(ih_case *c, const char *name, double *out, int64_t *n) {
    if (!c || !name || !n) return fail(IH_EINVAL, "null argument");
    if (!c->has_coeffs) return fail(IH_ESTATE, "no coefficients");
    const std::vector<double> *src = coeff_by_name(c->hc, name);
    if (!src) return fail(IH_EINVAL, "unknown coefficient '%s'", name);
    *n = (int64_t)src->size();
    if (out) std::memcpy(out, src->data(), sizeof(double) * src->size());
    return IH_OK;
}

int ih_grid_coeff(const ih_grid_desc *gd, const char *name, double *out, int64_t *n) {
    if (!name || !n) return fail(IH_EINVAL, "null argument");
    HostGrid hg;
    int rc = grid_from_desc(gd, hg);
    if (rc) return rc;
    host_coords(hg);
    HostCoeffs hc;
    std::string err;
    if (!host_coeffs(hg, hc, err)) return fail(IH_EINVAL, "coefficients: %s", err.c_str());
    const std::vector<double> *src = coeff_by_name(hc, name);
    if (!src) return fail(IH_EINVAL, "unknown coefficient '%s'", name);
    *n = (int64_t)src->size();
    if (out) std::memcpy(out, src->data(), sizeof(double) * src->size());
    return IH_OK;
}

int ih_grid_coords(const ih_grid_desc *gd, int d, int which, double *out) {
    if (!out || d < 0 || d > 2) return fail(IH_EINVAL, "bad argument");
    HostGrid hg;
    int rc = grid_from_desc(gd, hg);
    if (rc) return rc;
    host_coords(hg);
    const std::vector<double> &v = which ? hg.xu[d] : hg.xp[d];
    std::memcpy(out, v.data(), sizeof(double) * v.size());
    return IH_OK;
}

int ih_case_get_grid(const ih_case *c, ih_grid_desc *out) {
    if (!c || !out) return fail(IH_EINVAL, "null argument");
    out->version = IH_ABI_VERSION;
    for (int d = 0; d < 3; ++d) {
        out->N[d] = c->hg.N[d];
        out->bc[2 * d] = c->hg.bc[d][0];
        out->bc[2 * d + 1] = c->hg.bc[d][1];
        out->L[d] = c->hg.L[d];
        out->stretch[d] = c->hg.stretch[d];
    }
    return IH_OK;
}

int ih_case_get_coords(ih_case *c, int d, int which, double *out) {
    if (!c || !out || d < 0 || d > 2) return fail(IH_EINVAL, "bad argument");
    const std::vector<double> &v = which ? c->hg.xu[d] : c->hg.xp[d];
    if (v.empty()) return fail(IH_ESTATE, "coordinates not generated");
    std::memcpy(out, v.data(), sizeof(double) * v.size());
    return IH_OK;
}

int ih_case_get_ranges(ih_case *c, int out[48]) {
    if (!c || !out) return fail(IH_EINVAL, "null argument");
    const Geo &g = c->g;
    int q = 0;
    for (int d = 0; d < 3; ++d) out[q++] = g.Sp[d];
    for (int d = 0; d < 3; ++d) out[q++] = g.Np[d];
    for (int m = 0; m < 3; ++m) for (int d = 0; d < 3; ++d) out[q++] = g.SB[m][d];
    for (int m = 0; m < 3; ++m) for (int d = 0; d < 3; ++d) out[q++] = g.NB[m][d];
    for (int m = 0; m < 3; ++m) for (int d = 0; d < 3; ++d) out[q++] = g.SI[m][d];
    for (int m = 0; m < 3; ++m) for (int d = 0; d < 3; ++d) out[q++] = g.NI[m][d];
    for (int d = 0; d < 3; ++d) out[q++] = g.per[d] ? 1 : 2;  // relaxation range SNF (mod_setup.f90:850-867)
    for (int d = 0; d < 3; ++d) out[q++] = g.per[d] ? g.N[d] - 1 : g.N[d] - 1;
    return IH_OK;
}

int64_t ih_field_elems(const ih_case *c) {
    if (!c) return 0;
    return (int64_t)(c->g.N[0] + 7) * (c->g.N[1] + 7) * (c->g.N[2] + 7);
}

int64_t ih_case_bytes(const ih_case *c) { return c ? c->bytes : 0; }

int64_t ih_grid_bytes(const ih_grid_desc *gd) {
    if (!gd) return 0;
    for (int d = 0; d < 3; ++d)
        if (gd->N[d] < 1) return 0;
    const int64_t px = ((gd->N[0] + 7) + 15) / 16 * 16;
    const int64_t pxy = px * (gd->N[1] + 7);
    const int64_t fe = ((12 + pxy * (gd->N[2] + 7)) + 63) / 64 * 64;
    int64_t tables = 0;
    for (int d = 0; d < 3; ++d) tables += (int64_t)(((gd->N[d] + 1) + 7) / 8 * 8) * 80 * 12 + 2 * (gd->N[d] + 7) * 8;
    const int64_t parts = 2 * 8 * std::max<int64_t>(4096, ((gd->N[0] + 15) / 16) * ((gd->N[1] + 15) / 16) *
                                                             ((gd->N[2] + 31) / 32));
    return IH_NFIELDS * 8 * fe + 3 * 4 * fe + 3 * fe + tables + parts + (1 << 20);
}

int64_t ih_grid_bytes_step(const ih_grid_desc *gd, const ih_step_params *p) {
    const int64_t base = ih_grid_bytes(gd);
    if (!base || !p || p->precond_poisson != 1 || p->precond_sweeps > 0) return base;
    HostGrid hg;
    if (grid_from_desc(gd, hg) != IH_OK) return 0;
    host_coords(hg);
    std::vector<HostMgLevel> mgh;
    host_mg_levels(hg, 15, mgh);
    return base + mg_alloc_bytes(mgh);
}

int ih_field_upload(ih_case *c, ih_field f, const double *host) {
    if (!c || !host || !valid_field(f)) return fail(IH_EINVAL, "bad argument");
    HIPCHK(hipSetDevice(c->device));
    const Geo &g = c->g;
    const size_t w = sizeof(double) * (size_t)(g.N[0] + 7);
    HIPCHK(hipMemcpy2DAsync(c->fields[f] + g.lead, sizeof(double) * (size_t)g.px, host, w, w,
                            (size_t)(g.N[1] + 7) * (g.N[2] + 7), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    c->n_up += 1;
    c->bytes_up += (int64_t)w * (g.N[1] + 7) * (g.N[2] + 7);
    return IH_OK;
}

int ih_field_download(ih_case *c, ih_field f, double *host) {
    if (!c || !host || !valid_field(f)) return fail(IH_EINVAL, "bad argument");
    HIPCHK(hipSetDevice(c->device));
    const Geo &g = c->g;
    const size_t w = sizeof(double) * (size_t)(g.N[0] + 7);
    HIPCHK(hipMemcpy2DAsync(host, w, c->fields[f] + g.lead, sizeof(double) * (size_t)g.px, w,
                            (size_t)(g.N[1] + 7) * (g.N[2] + 7), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    c->n_down += 1;
    c->bytes_down += (int64_t)w * (g.N[1] + 7) * (g.N[2] + 7);
    return IH_OK;
}

// The strided box lo, lo+s, .. <= hi of a field as a dense host array, x fastest
// (the hyperslab write_hdf writes, src/mod_inout.f90:570-575, 657-667).  The box
// is gathered on the device, so only its bytes cross the host link.
static int box_transfer(ih_case *c, ih_field f, const int lo[3], const int hi[3], const int stride[3], double *host,
                        bool to_device) {
    if (!c || !host || !lo || !hi || !stride || !valid_field(f)) return fail(IH_EINVAL, "bad argument");
    int n[3];
    int64_t cnt = 1;
    for (int d = 0; d < 3; ++d) {
        if (stride[d] < 1 || lo[d] > hi[d] || lo[d] < -3 || hi[d] > c->g.N[d] + 3)
            return fail(IH_EINVAL, "box %d..%d stride %d outside the field's index range -3..%d along %d", lo[d], hi[d],
                        stride[d], c->g.N[d] + 3, d + 1);
        n[d] = (hi[d] - lo[d]) / stride[d] + 1;
        cnt *= n[d];
    }
    HIPCHK(hipSetDevice(c->device));
    double *buf = nullptr;
    HIPCHK(hipMalloc(&buf, sizeof(double) * (size_t)cnt));
    hipError_t e = hipSuccess;
    if (to_device) {
        e = hipMemcpyAsync(buf, host, sizeof(double) * (size_t)cnt, hipMemcpyHostToDevice, c->stream);
        if (e == hipSuccess) launch_unpack_box(c->stream, c->g, lo, n, stride, buf, c->fields[f]);
    } else {
        launch_pack_box(c->stream, c->g, lo, n, stride, c->fields[f], buf);
        e = hipGetLastError();
        if (e == hipSuccess)
            e = hipMemcpyAsync(host, buf, sizeof(double) * (size_t)cnt, hipMemcpyDeviceToHost, c->stream);
    }
    if (e == hipSuccess) e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    (void)hipFree(buf);
    if (e != hipSuccess) return fail(IH_EHIP, "box transfer: %s", hipGetErrorString(e));
    return IH_OK;
}

int ih_field_get_box(ih_case *c, ih_field f, const int lo[3], const int hi[3], const int stride[3], double *host) {
    return box_transfer(c, f, lo, hi, stride, host, false);
}

int ih_field_set_box(ih_case *c, ih_field f, const int lo[3], const int hi[3], const int stride[3],
                     const double *host) {
    return box_transfer(c, f, lo, hi, stride, const_cast<double *>(host), true);
}

int ih_field_fill(ih_case *c, ih_field f, double value) {
    if (!c || !valid_field(f)) return fail(IH_EINVAL, "bad argument");
    launch_fill(c->stream, c->fields[f], value, c->g.fe);
    LAUNCHCHK();
    return IH_OK;
}

int ih_field_copy(ih_case *c, ih_field dst, ih_field src) {
    if (!c || !valid_field(dst) || !valid_field(src)) return fail(IH_EINVAL, "bad argument");
    HIPCHK(hipMemcpyAsync(c->fields[dst], c->fields[src], sizeof(double) * (size_t)c->g.fe, hipMemcpyDeviceToDevice,
                          c->stream));
    return IH_OK;
}

int ih_synchronize(ih_case *c) {
    if (!c) return fail(IH_EINVAL, "null case");
    HIPCHK(hipStreamSynchronize(c->stream));
    return IH_OK;
}

// ---------------------------------------------------------------------------
int ih_exchange(ih_case *c, int dir, int vel_dir, ih_field f) {
    if (!c || !valid_field(f) || dir < 1 || dir > 3 || vel_dir < 0 || vel_dir > 3) return fail(IH_EINVAL, "bad argument");
    return exch(c, dir, vel_dir, F(c, f));
}

int ih_exchange_relax(ih_case *c, ih_field f) {
    if (!c || !valid_field(f)) return fail(IH_EINVAL, "bad argument");
    launch_exchange_relax(c->stream, c->g, F(c, f));
    LAUNCHCHK();
    return IH_OK;
}

int ih_divergence2(ih_case *c, ih_field u, ih_field v, ih_field w, ih_field div) {
    int rc = need_coeffs(c);
    if (rc) return rc;
    if (!valid_field(u) || !valid_field(v) || !valid_field(w) || !valid_field(div)) return fail(IH_EINVAL, "bad field");
    if ((rc = exch(c, 1, 1, F(c, u)))) return rc;
    if ((rc = exch(c, 2, 2, F(c, v)))) return rc;
    if ((rc = exch(c, 3, 3, F(c, w)))) return rc;
    launch_divergence2(c->stream, c->g, c->c, F(c, u), F(c, v), F(c, w), F(c, div));
    LAUNCHCHK();
    return IH_OK;
}

int ih_gradient(ih_case *c, int m, ih_field phi, ih_field grad) {
    int rc = need_coeffs(c);
    if (rc) return rc;
    if (m < 1 || m > 3 || !valid_field(phi) || !valid_field(grad)) return fail(IH_EINVAL, "bad argument");
    if ((rc = exch(c, m, 0, F(c, phi)))) return rc;
    launch_gradient(c->stream, c->g, c->c, m - 1, F(c, phi), F(c, grad));
    LAUNCHCHK();
    return IH_OK;
}

int ih_interpolate2_vel_pre(ih_case *c, int m, ih_field phi, ih_field inter) {
    int rc = need_coeffs(c);
    if (rc) return rc;
    if (m < 1 || m > 3 || !valid_field(phi) || !valid_field(inter)) return fail(IH_EINVAL, "bad argument");
    launch_interp_vel_pre(c->stream, c->g, c->c, m - 1, F(c, phi), F(c, inter));
    LAUNCHCHK();
    return IH_OK;
}

int ih_interpolate2_pre_vel(ih_case *c, int m, ih_field phi, ih_field inter) {
    int rc = need_coeffs(c);
    if (rc) return rc;
    if (m < 1 || m > 3 || !valid_field(phi) || !valid_field(inter)) return fail(IH_EINVAL, "bad argument");
    launch_interp_pre_vel(c->stream, c->g, c->c, m - 1, F(c, phi), F(c, inter));
    LAUNCHCHK();
    return IH_OK;
}

int ih_interpolate_vel(ih_case *c) {
    int rc = need_coeffs(c);
    if (rc) return rc;
    launch_interp_vel_pre3(c->stream, c->g, c->c, F(c, IH_VEL1), F(c, IH_VEL2), F(c, IH_VEL3), F(c, IH_WORK1),
                           F(c, IH_WORK2), F(c, IH_WORK3));
    LAUNCHCHK();
    double *ws[3] = {F(c, IH_WORK1), F(c, IH_WORK2), F(c, IH_WORK3)};
    int vd[3] = {0, 0, 0};
    for (int dir = 1; dir <= 3; ++dir) {
        launch_exchange(c->stream, c->g, dir, 3, ws, vd);
        LAUNCHCHK();
    }
    return IH_OK;
}

int ih_nonlinear_upwind(ih_case *c, int component_id, ih_field pp, ih_field vel_c, ih_field nl_c) {
    int rc = need_coeffs(c);
    if (rc) return rc;
    if (component_id < 1 || component_id > 9 || !valid_field(pp) || !valid_field(vel_c) || !valid_field(nl_c))
        return fail(IH_EINVAL, "bad argument");
    const int q = (component_id - 1) / 3, d = (component_id - 1) % 3;
    launch_nonlinear_one(c->stream, c->g, c->c, q, d, F(c, pp), F(c, vel_c), F(c, nl_c), d == 0 ? 1 : 0);
    LAUNCHCHK();
    return IH_OK;
}

int ih_helmholtz(ih_case *c, int component, double multL, ih_field vel_c, ih_field nl_c) {
    int rc = need_coeffs(c);
    if (rc) return rc;
    if (component < 1 || component > 3 || !valid_field(vel_c) || !valid_field(nl_c)) return fail(IH_EINVAL, "bad argument");
    launch_helmholtz_one(c->stream, c->g, c->c, component - 1, multL, F(c, vel_c), F(c, nl_c));
    LAUNCHCHK();
    return IH_OK;
}

int ih_axpby(ih_case *c, double a, double b, ih_field x, ih_field y, ih_field z) {
    if (!c || !valid_field(x) || !valid_field(y) || !valid_field(z)) return fail(IH_EINVAL, "bad argument");
    launch_axpby(c->stream, c->g, a, b, F(c, x), F(c, y), F(c, z), c->g.fe);
    LAUNCHCHK();
    return IH_OK;
}

int ih_product_div_grad(ih_case *c, ih_field phi, ih_field lap) {
    int rc = need_coeffs(c);
    if (rc) return rc;
    if (!valid_field(phi) || !valid_field(lap) || phi == IH_DIG || lap == IH_DIG) return fail(IH_EINVAL, "bad field");
    return apply_div_grad(c, F(c, phi), F(c, lap));
}

int ih_dot(ih_case *c, const int ss[3], const int nn[3], ih_field a, ih_field b, double *out) {
    if (!c || !ss || !nn || !out || !valid_field(a) || !valid_field(b)) return fail(IH_EINVAL, "bad argument");
    Box bx;
    int rc = to_box(c, ss, nn, bx);
    if (rc) return rc;
    if ((rc = dot_to(c, bx, F(c, a), F(c, b), S_DOT0))) return rc;
    if ((rc = read_scalars(c, S_DOT0, 1))) return rc;
    *out = c->hscal[S_DOT0];
    return IH_OK;
}

// Launch shape (threads per tile, planes per block along z) of the JOR sweep
// over box b, chosen once per case, precision and box by timing each
// candidate: which is fastest depends on how the blocks of a launch fill the
// compute units.  A sweep is a pure function of its inputs, so `sweep` writes
// the same output however often it runs and whatever shape it is given.
// While the stream is being captured no timing is possible and the default
// is returned (not remembered).
extern "C++" {
template <typename F>
static JorLaunch jor_shape(ih_case *c, const Box &b, bool fp32, F sweep, int level = 0) {
    const std::vector<int> key{fp32 ? 1 : 0, b.lo[0], b.lo[1], b.lo[2], b.hi[0], b.hi[1], b.hi[2], level};
    auto it = c->jor_shape.find(key);
    if (it != c->jor_shape.end()) return it->second;
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(c->stream, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone) return JorLaunch();
    const int n2 = b.hi[2] - b.lo[2] + 1;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
        if (e0) (void)hipEventDestroy(e0);
        return JorLaunch();
    }
    JorLaunch best;
    float best_ms = 0.0f;
    bool have = false, failed = false;
    for (int tpb : {128, 256, 512}) {
        int last_nbz = -1;
        for (int kc : {2, 4, 6, 8, 10, 12, 14, 16, 20, 24, 32}) {  // short chunks suit small boxes
            const int nbz = (n2 + kc - 1) / kc;
            if (nbz == last_nbz) continue;  // same launch as the previous candidate
            last_nbz = nbz;
            JorLaunch l;
            l.tpb = tpb;
            l.kc = kc;
            sweep(l);  // warm
            (void)hipEventRecord(e0, c->stream);
            for (int q = 0; q < 3; ++q) sweep(l);
            (void)hipEventRecord(e1, c->stream);
            float ms = 0.0f;
            if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
                failed = true;
                break;
            }
            if (!have || ms < best_ms) {
                best = l;
                best_ms = ms;
                have = true;
            }
        }
        if (failed) break;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (failed) best = JorLaunch();
    c->jor_shape[key] = best;
    return best;
}
}

// n_sweeps JOR sweeps of rel on bb over box b, out of place through COMP;
// pbc_substep 3 (with CT labels) applies device_jacobi_solver's Dirichlet
// pressure before every sweep and once after the last.
static int jor_run(ih_case *c, int n_sweeps, double omega, const Box &b, ih_field rel, ih_field bb, int pbc_substep,
                   double pbc_dtime) {
    double *x = F(c, rel), *y = F(c, IH_COMP);
    const double *B = F(c, bb);
    const JorLaunch sh = n_sweeps > 0 ? jor_shape(c, b, false, [&](JorLaunch l) {
        launch_jor(c->stream, c->g, c->c, b, omega, x, B, y, l);
    }) : JorLaunch();
    // Sweeps after the first read the ping-pong field outside the box (the
    // one-cell frame, and across a periodic direction the wrapped neighbour).
    // Over the whole pressure range those reads are ghosts with zero wall
    // coupling or wrapped points of the box itself; over a smaller box they are
    // points the reference loop reads from rel, unchanged, so COMP starts as a
    // copy of rel.
    const Box pb = pbox(c->g);
    bool whole = true;
    for (int d = 0; d < 3; ++d) whole = whole && b.lo[d] == pb.lo[d] && b.hi[d] == pb.hi[d];
    if (n_sweeps > 1 && !whole)
        HIPCHK(hipMemcpyAsync(y, x, sizeof(double) * (size_t)c->g.fe, hipMemcpyDeviceToDevice, c->stream));
    const bool with_bc = pbc_substep == 3 && c->ctlab;
    int rc;
    PhaseTimer pt(c, "jor");
    for (int s = 0; s < n_sweeps; ++s) {
        if (with_bc && (rc = jacobi_pressure_bc(c, x, pbc_substep, pbc_dtime))) return rc;
        launch_jor(c->stream, c->g, c->c, b, omega, x, B, y, sh);
        std::swap(x, y);
    }
    pt.done(n_sweeps);
    LAUNCHCHK();
    if (with_bc && (rc = jacobi_pressure_bc(c, x, pbc_substep, pbc_dtime))) return rc;
    if (x != F(c, rel)) {
        launch_copy_box(c->stream, c->g, b, x, F(c, rel));
        LAUNCHCHK();
    }
    launch_exchange_relax(c->stream, c->g, F(c, rel));
    LAUNCHCHK();
    return IH_OK;
}

int ih_jor(ih_case *c, int n_sweeps, double omega, const int ss[3], const int nn[3], ih_field rel, ih_field bb) {
    int rc = need_coeffs(c);
    if (rc) return rc;
    if (!ss || !nn || n_sweeps < 0 || !valid_field(rel) || !valid_field(bb) || rel == IH_COMP || bb == IH_COMP)
        return fail(IH_EINVAL, "bad argument");
    Box b;
    if ((rc = to_box(c, ss, nn, b))) return rc;
    return jor_run(c, n_sweeps, omega, b, rel, bb, 0, 0.0);
}

int ih_residual_cdg(ih_case *c, const int ss[3], const int nn[3], ih_field rel, ih_field bb, double *out) {
    int rc = need_coeffs(c);
    if (rc) return rc;
    if (!ss || !nn || !out || !valid_field(rel) || !valid_field(bb)) return fail(IH_EINVAL, "bad argument");
    Box b;
    if ((rc = to_box(c, ss, nn, b))) return rc;
    const int64_t nb = red_blocks(b);
    if ((rc = ensure_part(c, nb))) return rc;
    launch_exchange_relax(c->stream, c->g, F(c, rel));
    launch_residual_cdg(c->stream, c->g, c->c, b, F(c, rel), F(c, bb), nullptr, c->part);
    launch_final_max(c->stream, c->part, nb, c->dscal + S_MAX);
    LAUNCHCHK();
    if ((rc = read_scalars(c, S_MAX, 1))) return rc;
    *out = c->hscal[S_MAX];
    return IH_OK;
}

// BiCGstab_orig (src/mod_solvers.f90:1509-1935), preconditioner 0,
// init_yes = .FALSE.: the iteration runs on the device (scalars, stop test and
// breakdown exits in one-block control kernels); the host enqueues batches of
// iterations and reads the flags once per batch.
static int mg_setup(ih_case *c);

// mg: 0 no multigrid, 1 the FP64 cycle, 2 the FP32 cycle as M; sweeps > 0: M =
// that many JOR sweeps from zero, in FP32, or in FP64 when fp64_sweeps
static int bicgstab_run(ih_case *c, double eps, int n_it_max, const int ss[3], const int nn[3], ih_field bbf,
                        ih_field phif, int sweeps, double omega, int mg, ih_bicgstab_stats *st,
                        bool fp64_sweeps = false) {
    int rc = need_coeffs(c);
    if (rc) return rc;
    if (!ss || !nn || !valid_field(bbf) || !valid_field(phif) || n_it_max < 0 || sweeps < 0)
        return fail(IH_EINVAL, "bad argument");
    for (int f : {IH_RR, IH_RH, IH_PK, IH_AP, IH_AR, IH_DIG, IH_Z1, IH_Z2})
        if (bbf == f || phif == f) return fail(IH_EINVAL, "bb/phi may not be a BiCGstab work field");
    if (fp64_sweeps)
        for (int f : {IH_TMP3, IH_TMP4})
            if (bbf == f || phif == f) return fail(IH_EINVAL, "bb/phi may not be the FP64 preconditioner's work fields");
    Box b;
    if ((rc = to_box(c, ss, nn, b))) return rc;
    const Box pb = pbox(c->g);
    for (int d = 0; d < 3; ++d)
        if (b.lo[d] != pb.lo[d] || b.hi[d] != pb.hi[d])
            return fail(IH_EINVAL, "BiCGstab runs on the pressure range S?p..N?p");
    const int64_t nb = red_blocks(b);
    if ((rc = ensure_part(c, nb))) return rc;
    hipStream_t s = c->stream;
    KrylovPtrs k{F(c, IH_RR), F(c, IH_RH), F(c, IH_PK), F(c, IH_AP), F(c, IH_AR), F(c, phif), c->part, c->part2,
                 c->ksc, c->kfl};
    k.part3 = c->part3;
    k.part4 = c->part4;
    if (sweeps > 0 && fp64_sweeps) {
        k.z1 = F(c, IH_Z1);
        k.z2 = F(c, IH_Z2);
        k.sweeps64 = sweeps;
        k.omega64 = omega;
        k.e64a = F(c, IH_TMP3);
        k.e64b = F(c, IH_TMP4);
        k.zero64 = F(c, IH_DIG);
        k.shape64 = jor_shape(c, b, false, [&](JorLaunch l) {
            launch_jor(s, c->g, c->c, b, omega, F(c, IH_DIG), F(c, IH_TMP3), F(c, IH_TMP4), l);
        });
        // the zero start and finite ghost layers of the ping-pong fields
        launch_fill(s, F(c, IH_DIG), 0.0, c->g.fe);
        launch_fill(s, k.e64a, 0.0, c->g.fe);
        launch_fill(s, k.e64b, 0.0, c->g.fe);
    } else if (sweeps > 0) {
        k.z1 = F(c, IH_Z1);
        k.z2 = F(c, IH_Z2);
        k.e0 = c->f32[0];
        k.e1 = c->f32[2];
        k.b32 = c->f32[1];
        k.sweeps = sweeps;
        k.omega = (float)omega;
        // the timing sweeps (first solve only) write e1, cleared below
        k.shape = jor_shape(c, b, true, [&](JorLaunch l) {
            launch_jor_f32(s, c->g, c->c, b, k.omega, k.e0, k.b32, k.e1, nullptr, l);
        });
        // ghost layers of the ping-pong buffers must hold finite values (zero)
        launch_fill_f32(s, k.e0, 0.0f, c->g.fe);
        launch_fill_f32(s, k.e1, 0.0f, c->g.fe);
    }
    if (mg) {
        if ((rc = mg_setup(c))) return rc;
        k.z1 = F(c, IH_Z1);
        k.z2 = F(c, IH_Z2);
        k.mg = c->mg;
        k.mg_fp32 = mg == 2;
    }
    // rr = bb - A phi, rh = rr, norm and <rr, rh>
    if ((rc = apply_div_grad(c, F(c, phif), k.rr))) return rc;
    launch_krylov_init(s, c->g, b, F(c, bbf), k.rr, k.rh, k.part1, k.part2, k.sc, k.fl, eps, n_it_max);
    LAUNCHCHK();
    // iterations enqueued per convergence check: a preconditioned iteration is
    // many launches, and those enqueued past the stop run as no-ops
    const int batch = mg ? 2 : 8;
    for (;;) {
        for (int q = 0; q < batch; ++q) launch_krylov_iteration(s, c->g, c->c, b, k);
        LAUNCHCHK();
        HIPCHK(hipMemcpyAsync(c->hkfl, c->kfl, sizeof(int) * KF_NSLOTS, hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        if (c->hkfl[KF_DONE]) break;
        if (c->hkfl[KF_RESET]) {
            // rhr == 0: new reference vector rh = rr, then finish this iteration
            launch_krylov_resume(s, c->g, b, k);
            launch_krylov_iteration(s, c->g, c->c, b, k);
            LAUNCHCHK();
        }
    }
    HIPCHK(hipMemcpyAsync(c->hksc, c->ksc, sizeof(double) * KS_NSLOTS, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    if (st) {
        st->iterations = c->hkfl[KF_COUNTER];
        st->norm_inf_initial = c->hksc[KS_NORM0];
        st->norm_inf_final = c->hksc[KS_NORM];
        st->exit_reason = c->hkfl[KF_EXIT];
    }
    return IH_OK;
}

int ih_bicgstab(ih_case *c, double eps, int n_it_max, const int ss[3], const int nn[3], ih_field bbf, ih_field phif,
                ih_bicgstab_stats *st) {
    return bicgstab_run(c, eps, n_it_max, ss, nn, bbf, phif, 0, 0.0, 0, st);
}

int ih_bicgstab_precond(ih_case *c, double eps, int n_it_max, const int ss[3], const int nn[3], ih_field bbf,
                        ih_field phif, int preconditioner, ih_bicgstab_stats *st) {
    if (preconditioner == 0) return bicgstab_run(c, eps, n_it_max, ss, nn, bbf, phif, 0, 0.0, 0, st);
    if (preconditioner == 1) return bicgstab_run(c, eps, n_it_max, ss, nn, bbf, phif, 0, 0.0, 1, st);
    return fail(IH_EINVAL, "preconditioner %d (0: none, 1: multigrid V cycle)", preconditioner);
}

int ih_bicgstab_mg_f32(ih_case *c, double eps, int n_it_max, const int ss[3], const int nn[3], ih_field bbf,
                       ih_field phif, ih_bicgstab_stats *st) {
    return bicgstab_run(c, eps, n_it_max, ss, nn, bbf, phif, 0, 0.0, 2, st);
}

int ih_bicgstab_mixed(ih_case *c, double eps, int n_it_max, const int ss[3], const int nn[3], ih_field bbf,
                      ih_field phif, int sweeps, double omega, ih_bicgstab_stats *st) {
    if (sweeps < 1) return fail(IH_EINVAL, "the mixed-precision preconditioner needs at least one sweep");
    return bicgstab_run(c, eps, n_it_max, ss, nn, bbf, phif, sweeps, omega, 0, st);
}

int ih_bicgstab_jor64(ih_case *c, double eps, int n_it_max, const int ss[3], const int nn[3], ih_field bbf,
                      ih_field phif, int sweeps, double omega, ih_bicgstab_stats *st) {
    if (sweeps < 1) return fail(IH_EINVAL, "the JOR preconditioner needs at least one sweep");
    return bicgstab_run(c, eps, n_it_max, ss, nn, bbf, phif, sweeps, omega, 0, st, true);
}

// ---------------------------------------------------------------------------
// Multigrid V cycle (preconditioner 1): levels, arrays and tables on the device
// ---------------------------------------------------------------------------
static void mg_free(ih_case *c) { drop_mg(c); }

// Device bytes mg_setup allocates for the levels of mgh: the FP64 level arrays
// (2 fields on level 1, 4 below), the FP64 tables (cdg, cR: 3 taps, cI: 2 taps
// per direction) and 64 zeros, the FP32 arrays (4 per level) and tables, and
// the two descriptor arrays.
static int64_t mg_alloc_bytes(const std::vector<HostMgLevel> &mgh) {
    const int G = (int)mgh.size();
    int64_t arr = 0, tab = 0, arr32 = 0;
    for (int g = 1; g <= G; ++g) {
        const int *N = mgh[g - 1].N;
        const int64_t px = ((N[0] + 7) + 15) / 16 * 16;
        const int64_t fe = ((12 + px * (N[1] + 7) * (N[2] + 7)) + 63) / 64 * 64;
        arr += (g >= 2 ? 4 : 2) * fe;
        arr32 += 4 * fe;
        for (int d = 0; d < 3; ++d) tab += 8 * (int64_t)(((N[d] + 1) + 7) / 8 * 8);
    }
    return (int64_t)sizeof(double) * (arr + tab + 64) + (int64_t)sizeof(float) * (arr32 + tab) +
           (int64_t)(G + 1) * (int64_t)(sizeof(MgLevelDev<float>) + sizeof(MgLevelDev<double>));
}

static int mg_setup(ih_case *c) {
    if (c->mg) return IH_OK;
    if (c->hg.xp[0].empty())
        return fail(IH_ESTATE, "the multigrid levels need generated coordinates (ih_case_generate_coeffs)");
    int bc6[6];
    for (int d = 0; d < 3; ++d) {
        bc6[2 * d] = c->hg.bc[d][0];
        bc6[2 * d + 1] = c->hg.bc[d][1];
        if (!(c->g.per[d] || (bc6[2 * d] > 0 && bc6[2 * d + 1] > 0)))
            return fail(IH_EINVAL, "the multigrid cycle takes walls or periodic pairs in every direction");
    }
    host_mg_levels(c->hg, c->mg_cfg[0], c->mgh);
    const int G = (int)c->mgh.size();
    MgDev *m = new MgDev();
    m->G = G;
    m->n_down = c->mg_cfg[1];
    m->n_up = c->mg_cfg[2];
    m->n_bottom = c->mg_cfg[3];
    // level geometry, then one allocation: level arrays (each a multiple of 64
    // elements, so every array starts 512-byte aligned), then the tables
    int64_t arr = 0;
    int64_t offA[MG_LEVELS_MAX + 1], offB[MG_LEVELS_MAX + 1], offC[MG_LEVELS_MAX + 1], offD[MG_LEVELS_MAX + 1];
    for (int g = 1; g <= G; ++g) {
        MgLevel &L = m->L[g];
        if (g == 1) {
            L.g = c->g;
        } else {
            Geo &q = L.g;
            q = Geo{};
            for (int d = 0; d < 3; ++d) q.N[d] = c->mgh[g - 1].N[d];
            init_limits(q, bc6);
            q.lead = 12;
            q.px = ((q.N[0] + 7) + 15) / 16 * 16;
            q.pxy = (int64_t)q.px * (q.N[1] + 7);
            q.fe = ((q.lead + q.pxy * (q.N[2] + 7)) + 63) / 64 * 64;
            for (int d = 0; d < 3; ++d) q.cst[d] = ((q.N[d] + 1) + 7) / 8 * 8;
        }
        L.pr = pbox(L.g);
        for (int d = 0; d < 3; ++d) L.dn[d] = c->mgh[g - 1].dn[d];
        offA[g] = offB[g] = -1;
        if (g >= 2) {
            offA[g] = arr;
            arr += L.g.fe;
            offB[g] = arr;
            arr += L.g.fe;
        }
        offC[g] = arr;
        arr += L.g.fe;
        offD[g] = arr;
        arr += L.g.fe;
    }
    std::vector<double> tab;
    int64_t offdg[MG_LEVELS_MAX + 1][3], offR[MG_LEVELS_MAX + 1][3], offI[MG_LEVELS_MAX + 1][3];
    for (int g = 1; g <= G; ++g)
        for (int d = 0; d < 3; ++d) {
            const HostMgLevel &h = c->mgh[g - 1];
            const int N = h.N[d], cst = m->L[g].g.cst[d];
            auto put = [&](const std::vector<double> &src, int taps) -> int64_t {
                const int64_t o = (int64_t)tab.size();
                tab.resize(tab.size() + (size_t)taps * cst, 0.0);
                for (int i = 1; i <= N; ++i)
                    for (int t = 0; t < taps; ++t) tab[(size_t)o + (size_t)t * cst + i] = src[(size_t)t + (size_t)taps * (i - 1)];
                return o;
            };
            offdg[g][d] = put(h.dg[d], 3);
            offR[g][d] = put(h.cR[d], 3);
            offI[g][d] = put(h.cI[d], 2);
        }
    const int64_t offZ = (int64_t)tab.size();
    tab.resize(tab.size() + 64, 0.0);
    const int64_t total = arr + (int64_t)tab.size();
    if (hipMalloc(&c->mgmem, sizeof(double) * (size_t)total) != hipSuccess) {
        delete m;
        c->mgmem = nullptr;
        return fail(IH_ENOMEM, "out of device memory for the multigrid levels (%lld bytes)",
                    (long long)(sizeof(double) * total));
    }
    HIPCHK(hipMemsetAsync(c->mgmem, 0, sizeof(double) * (size_t)arr, c->stream));
    HIPCHK(hipMemcpyAsync(c->mgmem + arr, tab.data(), sizeof(double) * tab.size(), hipMemcpyHostToDevice, c->stream));
    double *base = c->mgmem, *tb = c->mgmem + arr;
    for (int g = 1; g <= G; ++g) {
        MgLevel &L = m->L[g];
        if (g == 1) {
            L.c = c->c;
        } else {
            L.c = Coef{};
            for (int d = 0; d < 3; ++d) L.c.dg[d] = tb + offdg[g][d];
            L.c.zero = tb + offZ;
            L.A = base + offA[g];
            L.B = base + offB[g];
        }
        L.C = base + offC[g];
        L.D = base + offD[g];
        for (int d = 0; d < 3; ++d) {
            L.cR[d] = tb + offR[g][d];
            L.cI[d] = tb + offI[g][d];
        }
    }
    // levels from g0 on run inside one workgroup; their descriptors in device memory
    for (int g = 2; g <= G; ++g) {
        const Geo &q = m->L[g].g;
        if ((int64_t)q.N[0] * q.N[1] * q.N[2] <= MG_FUSED_POINTS) {
            m->g0 = g;
            break;
        }
    }
    // the FP32 cycle: float arrays (level 1 included: its Af is (float) of the
    // right-hand side) and the tables rounded once
    int64_t arr32 = 0;
    int64_t offA32[MG_LEVELS_MAX + 1], offB32[MG_LEVELS_MAX + 1], offC32[MG_LEVELS_MAX + 1],
        offD32[MG_LEVELS_MAX + 1];
    for (int g = 1; g <= G; ++g) {
        const int64_t fe = m->L[g].g.fe;
        offA32[g] = arr32;
        offB32[g] = arr32 + fe;
        offC32[g] = arr32 + 2 * fe;
        offD32[g] = arr32 + 3 * fe;
        arr32 += 4 * fe;
    }
    std::vector<float> tab32;
    int64_t offdg32[MG_LEVELS_MAX + 1][3], offR32[MG_LEVELS_MAX + 1][3], offI32[MG_LEVELS_MAX + 1][3];
    for (int g = 1; g <= G; ++g)
        for (int d = 0; d < 3; ++d) {
            const int cst = m->L[g].g.cst[d];
            auto put32 = [&](int64_t off, int taps) -> int64_t {
                const int64_t o = (int64_t)tab32.size();
                for (int64_t q = 0; q < (int64_t)taps * cst; ++q) tab32.push_back((float)tab[(size_t)(off + q)]);
                return o;
            };
            offdg32[g][d] = put32(offdg[g][d], 3);
            offR32[g][d] = put32(offR[g][d], 3);
            offI32[g][d] = put32(offI[g][d], 2);
        }
    if (hipMalloc(&c->mgmem32, sizeof(float) * (size_t)(arr32 + (int64_t)tab32.size())) != hipSuccess) {
        delete m;
        return fail(IH_ENOMEM, "out of device memory for the FP32 multigrid levels");
    }
    HIPCHK(hipMemsetAsync(c->mgmem32, 0, sizeof(float) * (size_t)arr32, c->stream));
    HIPCHK(hipMemcpyAsync(c->mgmem32 + arr32, tab32.data(), sizeof(float) * tab32.size(), hipMemcpyHostToDevice,
                          c->stream));
    {
        float *fb = c->mgmem32, *ft = c->mgmem32 + arr32;
        for (int g = 1; g <= G; ++g) {
            MgLevel &L = m->L[g];
            L.Af = fb + offA32[g];
            L.Bf = fb + offB32[g];
            L.Cf = fb + offC32[g];
            L.Df = fb + offD32[g];
            for (int d = 0; d < 3; ++d) {
                if (g >= 2) L.c.dgf[d] = ft + offdg32[g][d];
                L.cRf[d] = ft + offR32[g][d];
                L.cIf[d] = ft + offI32[g][d];
            }
        }
    }
    std::vector<MgLevelDev<float>> ld32((size_t)G + 1);
    for (int g = 1; g <= G; ++g) {
        const MgLevel &L = m->L[g];
        MgLevelDev<float> &d = ld32[g];
        d.g = L.g;
        for (int e = 0; e < 3; ++e) {
            d.dg[e] = L.c.dgf[e];
            d.cR[e] = L.cRf[e];
            d.cI[e] = L.cIf[e];
            d.dn[e] = L.dn[e];
        }
        d.A = L.Af;
        d.B = L.Bf;
        d.C = L.Cf;
        d.D = L.Df;
    }
    if (hipMalloc(&c->mgldev32, sizeof(MgLevelDev<float>) * ld32.size()) != hipSuccess) {
        delete m;
        return fail(IH_ENOMEM, "out of device memory for the multigrid level descriptors");
    }
    HIPCHK(hipMemcpyAsync(c->mgldev32, ld32.data(), sizeof(MgLevelDev<float>) * ld32.size(), hipMemcpyHostToDevice,
                          c->stream));
    m->Ldev32 = (const MgLevelDev<float> *)c->mgldev32;
    std::vector<MgLevelDev<double>> ld((size_t)G + 1);
    for (int g = 1; g <= G; ++g) {
        const MgLevel &L = m->L[g];
        MgLevelDev<double> &d = ld[g];
        d.g = L.g;
        for (int e = 0; e < 3; ++e) {
            d.dg[e] = L.c.dg[e];
            d.cR[e] = L.cR[e];
            d.cI[e] = L.cI[e];
            d.dn[e] = L.dn[e];
        }
        d.A = L.A;
        d.B = L.B;
        d.C = L.C;
        d.D = L.D;
    }
    if (hipMalloc(&c->mgldev, sizeof(MgLevelDev<double>) * ld.size()) != hipSuccess) {
        delete m;
        return fail(IH_ENOMEM, "out of device memory for the multigrid level descriptors");
    }
    HIPCHK(hipMemcpyAsync(c->mgldev, ld.data(), sizeof(MgLevelDev<double>) * ld.size(), hipMemcpyHostToDevice,
                          c->stream));
    m->Ldev = (const MgLevelDev<double> *)c->mgldev;
    c->mg_bytes = mg_alloc_bytes(c->mgh);
    c->bytes += c->mg_bytes;
    // launch shape of each level's sweep (a sweep is a pure function of its inputs)
    for (int g = 1; g < std::min(m->g0, G + 1); ++g) {
        MgLevel &L = m->L[g];
        const double *bb = g == 1 ? L.C : L.A;
        const double *x0 = g == 1 ? L.C : L.B;
        L.shape = jor_shape(c, L.pr, false, [&](JorLaunch l) {
            launch_jor(c->stream, L.g, L.c, L.pr, m->omega, x0, bb, L.D, l);
        }, g);
        L.shape32 = jor_shape(c, L.pr, true, [&](JorLaunch l) {
            launch_jor_f32(c->stream, L.g, L.c, L.pr, (float)m->omega, L.Bf, L.Af, L.Df, nullptr, l);
        }, g);
    }
    LAUNCHCHK();
    HIPCHK(hipStreamSynchronize(c->stream));
    c->mg = m;
    return IH_OK;
}

int ih_mg_configure(ih_case *c, int n_grids_limit, int n_relax_down, int n_relax_up, int n_relax_bottom) {
    if (!c) return fail(IH_EINVAL, "null case");
    if (n_grids_limit < 1 || n_grids_limit > MG_LEVELS_MAX || n_relax_down < 1 || n_relax_up < 1 ||
        n_relax_bottom < 1)
        return fail(IH_EINVAL, "multigrid: 1 <= n_grids_limit <= %d and at least one relaxation per phase",
                    MG_LEVELS_MAX);
    HIPCHK(hipSetDevice(c->device));
    mg_free(c);
    c->mg_cfg[0] = n_grids_limit;
    c->mg_cfg[1] = n_relax_down;
    c->mg_cfg[2] = n_relax_up;
    c->mg_cfg[3] = n_relax_bottom;
    return IH_OK;
}

int ih_mg_levels(ih_case *c, int *n_grids, int *NN) {
    int rc = need_coeffs(c);
    if (rc) return rc;
    if (!n_grids || !NN) return fail(IH_EINVAL, "null argument");
    HIPCHK(hipSetDevice(c->device));
    if ((rc = mg_setup(c))) return rc;
    *n_grids = c->mg->G;
    for (int g = 1; g <= c->mg->G; ++g)
        for (int d = 0; d < 3; ++d) NN[3 * (g - 1) + d] = c->mgh[g - 1].N[d];
    return IH_OK;
}

int ih_mg_get_coeff(ih_case *c, int level, const char *name, double *out, int64_t *n) {
    int rc = need_coeffs(c);
    if (rc) return rc;
    if (!name || !n) return fail(IH_EINVAL, "null argument");
    HIPCHK(hipSetDevice(c->device));
    if ((rc = mg_setup(c))) return rc;
    if (level < 1 || level > c->mg->G) return fail(IH_EINVAL, "level %d out of 1..%d", level, c->mg->G);
    const HostMgLevel &h = c->mgh[level - 1];
    const size_t len = strlen(name);
    const int d = len >= 3 ? name[len - 1] - '1' : -1;
    if (d < 0 || d > 2) return fail(IH_EINVAL, "unknown multigrid table '%s'", name);
    const std::vector<double> *v = nullptr;
    if (!strncmp(name, "cdg", 3) && len == 4) v = &h.dg[d];
    else if (!strncmp(name, "cR", 2) && len == 3) v = &h.cR[d];
    else if (!strncmp(name, "cI", 2) && len == 3) v = &h.cI[d];
    if (!v) return fail(IH_EINVAL, "unknown multigrid table '%s'", name);
    if (out) std::copy(v->begin(), v->end(), out);
    *n = (int64_t)v->size();
    return IH_OK;
}

int ih_multigrid_v_f32(ih_case *c, ih_field bb, ih_field phi) {
    int rc = need_coeffs(c);
    if (rc) return rc;
    if (!valid_field(bb) || !valid_field(phi) || bb == phi) return fail(IH_EINVAL, "bad argument");
    HIPCHK(hipSetDevice(c->device));
    if ((rc = mg_setup(c))) return rc;
    launch_mg_vcycle_f32(c->stream, *c->mg, F(c, bb), F(c, phi), nullptr);
    LAUNCHCHK();
    return IH_OK;
}

int ih_multigrid_v(ih_case *c, ih_field bb, ih_field phi) {
    int rc = need_coeffs(c);
    if (rc) return rc;
    if (!valid_field(bb) || !valid_field(phi) || bb == phi) return fail(IH_EINVAL, "bad argument");
    HIPCHK(hipSetDevice(c->device));
    if ((rc = mg_setup(c))) return rc;
    launch_mg_vcycle(c->stream, *c->mg, F(c, bb), F(c, phi), nullptr);
    LAUNCHCHK();
    return IH_OK;
}

int ih_fill_corners(ih_case *c, ih_field f) {
    if (!c || !valid_field(f)) return fail(IH_EINVAL, "bad argument");
    launch_fill_corners(c->stream, c->g, F(c, f));
    LAUNCHCHK();
    return IH_OK;
}

// level_pressure, src/mod_lib.f90:507-530
int ih_level_pressure(ih_case *c, ih_field pre) {
    if (!c || !valid_field(pre)) return fail(IH_EINVAL, "bad argument");
    const Geo &g = c->g;
    Box b = pbox(g);
    int rc = sum_to(c, b, F(c, pre), S_SUM);
    if (rc) return rc;
    if ((rc = read_scalars(c, S_SUM, 1))) return rc;
    const int dim1 = g.per[0] ? g.N[0] - 1 : g.N[0];
    const int dim2 = g.per[1] ? g.N[1] - 1 : g.N[1];
    const int dim3 = g.per[2] ? g.N[2] - 1 : g.N[2];
    c->hscal[S_SUM + 1] = c->hscal[S_SUM] / (double)dim1 / (double)dim2 / (double)dim3;
    if ((rc = write_scalars(c, S_SUM + 1, 1))) return rc;
    launch_sub_box(c->stream, g, b, F(c, pre), c->dscal + S_SUM + 1);
    LAUNCHCHK();
    HIPCHK(hipStreamSynchronize(c->stream));
    return IH_OK;
}

int ih_set_tube_mask(ih_case *c, double yc, double zc, const double *r_u, const double *r_p) {
    if (!c || !r_u || !r_p) return fail(IH_EINVAL, "null argument");
    if (c->hg.xp[0].empty()) return fail(IH_ESTATE, "coordinates not generated");
    const Geo &g = c->g;
    const int64_t fe = g.fe;
    std::vector<uint8_t> m((size_t)(3 * fe), 1);
    for (int q = 0; q < 3; ++q)
        for (int k = g.SB[q][2]; k <= g.NB[q][2]; ++k)
            for (int j = g.SB[q][1]; j <= g.NB[q][1]; ++j)
                for (int i = g.SB[q][0]; i <= g.NB[q][0]; ++i) {
                    const double y = (q == 1) ? c->hg.xu[1][j + 3] : c->hg.xp[1][j + 3];
                    const double z = (q == 2) ? c->hg.xu[2][k + 3] : c->hg.xp[2][k + 3];
                    const double r = (q == 0) ? r_u[i + 3] : r_p[i + 3];
                    const double dy = y - yc, dz = z - zc;
                    m[(size_t)(q * fe + gidx(g, i, j, k))] = (std::sqrt(dy * dy + dz * dz) <= r) ? 1 : 0;
                }
    HIPCHK(hipMemcpy(c->mask, m.data(), m.size(), hipMemcpyHostToDevice));
    c->has_mask = 1;
    return IH_OK;
}

int ih_get_mask(ih_case *c, uint8_t *host3) {
    if (!c || !host3) return fail(IH_EINVAL, "null argument");
    const Geo &g = c->g;
    std::vector<uint8_t> m((size_t)(3 * g.fe));
    HIPCHK(hipMemcpy(m.data(), c->mask, m.size(), hipMemcpyDeviceToHost));
    const int64_t nx = g.N[0] + 7, ny = g.N[1] + 7, nz = g.N[2] + 7;
    for (int q = 0; q < 3; ++q)
        for (int64_t k = 0; k < nz; ++k)
            for (int64_t j = 0; j < ny; ++j)
                for (int64_t i = 0; i < nx; ++i)
                    host3[q * nx * ny * nz + i + nx * (j + ny * k)] =
                        m[(size_t)(q * g.fe + gidx(g, (int)i - 3, (int)j - 3, (int)k - 3))];
    return IH_OK;
}

// ---------------------------------------------------------------------------
// CT geometry: read_vtk_3D (src/usr_func.f90:187-555) and apply_CT_forcing's
// CT-based branch (src/usr_func.f90:909-945)
// ---------------------------------------------------------------------------
int ih_ct_load_vtk(ih_case *c, const char *path, int x_offset, int y_offset, int z_offset, const double *wave_t_ms,
                   const double *wave_q, int npts, ih_ct_info *info) {
    if (!c || !path || !wave_t_ms || !wave_q) return fail(IH_EINVAL, "null argument");
    if (npts < 2) return fail(IH_EINVAL, "the waveform needs at least two points");
    if (c->hg.xp[0].empty()) return fail(IH_ESTATE, "coordinates not generated (ih_case_generate_coeffs)");
    HIPCHK(hipSetDevice(c->device));
    const Geo &g = c->g;
    const int M1 = g.N[0], M2 = g.N[1], M3 = g.N[2];
    int dims[3];
    std::vector<double> img;
    std::string err;
    if (!vtk_read_ascii(path, dims, img, err)) return fail(IH_EINVAL, "%s", err.c_str());
    // the image's first axis is the grid's y, its second the grid's x
    if (x_offset < 0 || y_offset < 0 || dims[1] + y_offset > M1 || dims[0] + x_offset > M2 ||
        dims[2] + z_offset > M3 || z_offset < 10 || z_offset + dims[2] + 10 > M3)
        return fail(IH_EINVAL, "the image with its offsets and the 10-plane inlet and outlet buffers must fit the grid");
    double ref[3];
    std::vector<double> tconv, qconv;
    waveform_convert(wave_t_ms, wave_q, npts, M1, c->ct.idealized != 0, tconv, qconv, ref);
    std::vector<uint8_t> lab((size_t)g.fe, 0);
    auto L = [&](int i, int j, int k) -> uint8_t & { return lab[(size_t)gidx(g, i, j, k)]; };
    for (int k = 1; k <= M3; ++k)
        for (int j = 1; j <= M2; ++j)
            for (int i = 1; i <= M1; ++i) L(i, j, k) = 2;
    for (int i = 1; i <= dims[0]; ++i)
        for (int j = 1; j <= dims[1]; ++j)
            for (int k = 1; k <= dims[2]; ++k) {
                const double v = img[(size_t)(k - 1) * dims[1] * dims[0] + (size_t)(j - 1) * dims[0] + (i - 1)];
                if (!(v >= 0.0 && v <= 255.0 && v == std::floor(v)))
                    return fail(IH_EINVAL, "image value %g is not an integer label 0..255", v);
                L(j + y_offset, i + x_offset, k + z_offset) = (uint8_t)v;
            }
    // inlet (left half) and outlet (right half, from floor(M1/2)) on the image's
    // first plane, the second outlet on its last plane: label band [0.5, 2)
    const std::vector<double> &xp = c->hg.xp[0], &yp = c->hg.xp[1];
    const int half = M1 / 2, ks = z_offset + 1, ks2 = z_offset + dims[2];
    auto fluid = [&](int i, int j, int k) { return L(i, j, k) == 1; };
    std::vector<int> in, out, out2;
    double xi = 0.0, yi = 0.0, xo = 0.0, yo = 0.0;
    for (int j = 1; j <= M2; ++j)
        for (int i = 1; i <= half; ++i)
            if (fluid(i, j, ks)) {
                xi = xi + xp[i + 3];
                yi = yi + yp[j + 3];
                in.push_back(i);
                in.push_back(j);
            }
    for (int j = 1; j <= M2; ++j)
        for (int i = half; i <= M1; ++i)
            if (fluid(i, j, ks)) {
                xo = xo + xp[i + 3];
                yo = yo + yp[j + 3];
                out.push_back(i);
                out.push_back(j);
            }
    for (int j = 1; j <= M2; ++j)
        for (int i = 1; i <= M1; ++i)
            if (fluid(i, j, ks2)) {
                out2.push_back(i);
                out2.push_back(j);
            }
    const int ni = (int)in.size() / 2, no = (int)out.size() / 2, no2 = (int)out2.size() / 2;
    if (ni == 0 || no == 0) return fail(IH_EINVAL, "no fluid cells on the inlet or outlet side of the first plane");
    xi = xi / ni;
    yi = yi / ni;
    xo = xo / no;
    yo = yo / no;
    // the incoming jet: cells within R_inlet of the inlet centre, ten planes below
    const double R_inlet = std::sqrt(0.0004 / (4.0 * std::atan(1.0) * ref[0])) / ref[1];
    int jet = 0;
    for (int q = 0; q < ni; ++q)
        for (int k = -10; k <= -1; ++k) {
            const int i = in[2 * q], j = in[2 * q + 1];
            L(i, j, ks + k) = 7;
            const double dx = xp[i + 3] - xi, dy = yp[j + 3] - yi;
            if (std::sqrt(dx * dx + dy * dy) <= R_inlet) {
                L(i, j, ks + k) = 5;
                if (k == -1) ++jet;
            }
        }
    for (int q = 0; q < no; ++q)
        for (int k = -10; k <= -1; ++k) L(out[2 * q], out[2 * q + 1], ks + k) = 9;
    for (int q = 0; q < no2; ++q)
        for (int k = 1; k <= 10; ++k) L(out2[2 * q], out2[2 * q + 1], ks2 + k) = 11;
    if (!c->ctlab) {
        if (hipMalloc(&c->ctlab, (size_t)g.fe) != hipSuccess) {
            c->ctlab = nullptr;
            return fail(IH_ENOMEM, "out of device memory for the CT labels");
        }
        c->bytes += g.fe;
    }
    HIPCHK(hipMemcpy(c->ctlab, lab.data(), lab.size(), hipMemcpyHostToDevice));
    CtState &ct = c->ct;
    ct.t_ms.assign(wave_t_ms, wave_t_ms + npts);
    ct.q_raw.assign(wave_q, wave_q + npts);
    ct.t_conv = tconv;
    ct.q_conv = qconv;
    for (int d = 0; d < 3; ++d) ct.ref[d] = ref[d];
    ct.inlet_voxels = jet;
    ct.k_start = ks;
    // labels present in the forcing's range (the component-1 velocity range;
    // points outside 1..N read label 2): the idealized branch evaluates a
    // Windkessel model only where a point of its label lies in that range
    ct.has5 = ct.has9 = ct.has11 = false;
    for (int k = g.SB[0][2]; k <= g.NB[0][2]; ++k)
        for (int j = g.SB[0][1]; j <= g.NB[0][1]; ++j)
            for (int i = g.SB[0][0]; i <= g.NB[0][0]; ++i) {
                const bool inside = i >= 1 && i <= M1 && j >= 1 && j <= M2 && k >= 1 && k <= M3;
                const int Lb = inside ? L(i, j, k) : 2;
                ct.has5 = ct.has5 || Lb == 5;
                ct.has9 = ct.has9 || Lb == 9;
                ct.has11 = ct.has11 || Lb == 11;
            }
    for (int q = 0; q < 3; ++q) ct.wk_state[q] = ct.pres[q] = 0.0;
    if (info) {
        info->inlet_cells = ni;
        info->outlet_cells = no;
        info->outlet2_cells = no2;
        info->inlet_jet_voxels = jet;
        info->k_start = ks;
        info->k_start2 = ks2;
        info->inlet_centre[0] = xi;
        info->inlet_centre[1] = yi;
        info->outlet_centre[0] = xo;
        info->outlet_centre[1] = yo;
        info->reference_velocity = ref[0];
        info->reference_length = ref[1];
        info->reference_time = ref[2];
    }
    return IH_OK;
}

int ih_ct_get_labels(ih_case *c, uint8_t *host) {
    if (!c || !host) return fail(IH_EINVAL, "null argument");
    if (!c->ctlab) return fail(IH_ESTATE, "no CT labels loaded");
    const Geo &g = c->g;
    std::vector<uint8_t> m((size_t)g.fe);
    HIPCHK(hipMemcpy(m.data(), c->ctlab, m.size(), hipMemcpyDeviceToHost));
    const int64_t nx = g.N[0] + 7, ny = g.N[1] + 7, nz = g.N[2] + 7;
    for (int64_t k = 0; k < nz; ++k)
        for (int64_t j = 0; j < ny; ++j)
            for (int64_t i = 0; i < nx; ++i)
                host[i + nx * (j + ny * k)] = m[(size_t)gidx(g, (int)i - 3, (int)j - 3, (int)k - 3)];
    return IH_OK;
}

int ih_ct_inflow(ih_case *c, double time, double *q, int *found) {
    if (!c || !q || !found) return fail(IH_EINVAL, "null argument");
    if (c->ct.t_conv.size() < 2) return fail(IH_ESTATE, "no waveform loaded");
    bool f = false;
    *q = waveform_inflow(c->ct.t_conv, c->ct.q_conv, time, f);
    *found = f ? 1 : 0;
    return IH_OK;
}

int ih_ct_configure(ih_case *c, int ct_based, int idealized) {
    if (!c || (ct_based != 0 && ct_based != 1) || (idealized != 0 && idealized != 1))
        return fail(IH_EINVAL, "ct_based and idealized are 0 or 1");
    c->ct.ct_based = ct_based;
    c->ct.idealized = idealized;
    return IH_OK;
}

int ih_ct_forcing(ih_case *c, double time) {
    if (!c) return fail(IH_EINVAL, "null case");
    CtState &ct = c->ct;
    if (!c->ctlab || ct.t_conv.size() < 2) return fail(IH_ESTATE, "no CT labels loaded");
    bool found = false;
    double q = 0.0, w_ideal = 0.0;
    if (ct.ct_based) {
        wk_echo(ct.t_ms, ct.q_raw, ct.t_conv, time, ct.ref[0], ct.pres[0], ct.pres[1]);
        q = waveform_inflow(ct.t_conv, ct.q_conv, time, found);
    }
    if (ct.idealized) {
        double q_inlet = 0.0, q_other = 0.0;
        if (ct.has5) wk_pressure_flow(1, time, ct.ref, ct.wk_state, ct.pres[0], q_inlet);
        if (ct.has9) wk_pressure_flow(2, time, ct.ref, ct.wk_state, ct.pres[1], q_other);
        if (ct.has11) wk_pressure_flow(3, time, ct.ref, ct.wk_state, ct.pres[2], q_other);
        const double h = c->hg.L[0] / (double)(c->g.N[0] - 1);
        w_ideal = q_inlet / ((double)ct.inlet_voxels * (h * h));
    }
    if (!ct.ct_based && !ct.idealized) return IH_OK;
    launch_ct_forcing(c->stream, c->g, c->ctlab, q, found ? 1 : 0, ct.idealized, w_ideal, F(c, IH_VEL1),
                      F(c, IH_VEL2), F(c, IH_VEL3));
    LAUNCHCHK();
    return IH_OK;
}

int ih_ct_get_state(ih_case *c, double pres[3], double wk[3]) {
    if (!c || !pres || !wk) return fail(IH_EINVAL, "null argument");
    for (int q = 0; q < 3; ++q) {
        pres[q] = c->ct.pres[q];
        wk[q] = c->ct.wk_state[q];
    }
    return IH_OK;
}

int ih_ct_set_state(ih_case *c, const double pres[3], const double wk[3]) {
    if (!c || !pres || !wk) return fail(IH_EINVAL, "null argument");
    for (int q = 0; q < 3; ++q) {
        c->ct.pres[q] = pres[q];
        c->ct.wk_state[q] = wk[q];
    }
    return IH_OK;
}

int ih_set_pressure_bc(ih_case *c, ih_field rel, int type1, int type2, int type3, double p_in, double p_out,
                       double p_out2) {
    if (!c || !valid_field(rel)) return fail(IH_EINVAL, "bad argument");
    if (!c->ctlab) return fail(IH_ESTATE, "no CT labels loaded");
    return pressure_bc(c, F(c, rel), type1, type2, type3, p_in, p_out, p_out2);
}

int ih_case_set_substep(ih_case *c, int substep, double dtime) {
    if (!c || substep < 0 || substep > 3) return fail(IH_EINVAL, "substep is 0..3");
    c->ct.substep = substep;
    c->ct.dtime = dtime;
    return IH_OK;
}

int ih_apply_mask(ih_case *c) {
    if (!c) return fail(IH_EINVAL, "null case");
    launch_apply_mask(c->stream, c->g, c->mask, F(c, IH_VEL1), F(c, IH_VEL2), F(c, IH_VEL3));
    LAUNCHCHK();
    return IH_OK;
}

// ---------------------------------------------------------------------------
// One time step: 3 RK substeps of src/mod_timeint.f90:275-377 with rhs_vel
// (src/mod_rhs.f90:35-241) and explicit (src/mod_solvers.f90:64-346); after
// the step the mask (apply_CT_forcing pattern) and level_pressure.
// ---------------------------------------------------------------------------
static int poisson_solve(ih_case *c, const ih_step_params *p, const Box &b, int substep, ih_bicgstab_stats *bst) {
    int rc;
    const int ss[3] = {b.lo[0], b.lo[1], b.lo[2]};
    const int nn[3] = {b.hi[0], b.hi[1], b.hi[2]};
    if (p->jor_sweeps > 0) {
        if (p->jor_fp32) {
            // FP32 correction sweeps: A e = r solved approximately in FP32 from
            // e = 0, with r = b - A p and the update p += e kept in FP64.
            const Geo &g = c->g;
            const int64_t nb = red_blocks(b);
            if ((rc = ensure_part(c, nb))) return rc;
            launch_exchange_relax(c->stream, g, F(c, IH_PRE));
            launch_residual_cdg(c->stream, g, c->c, b, F(c, IH_PRE), F(c, IH_RES), F(c, IH_TMP1), c->part);
            launch_to_f32(c->stream, g, b, F(c, IH_TMP1), c->f32[1]);
            launch_fill_f32(c->stream, c->f32[0], 0.0f, g.fe);
            launch_fill_f32(c->stream, c->f32[2], 0.0f, g.fe);
            float *x = c->f32[0], *y = c->f32[2];
            const JorLaunch sh = jor_shape(c, b, true, [&](JorLaunch l) {
                launch_jor_f32(c->stream, g, c->c, b, (float)p->jor_omega, x, c->f32[1], y, nullptr, l);
            });
            PhaseTimer pt(c, "jor_f32");
            for (int s = 0; s < p->jor_sweeps; ++s) {
                launch_jor_f32(c->stream, g, c->c, b, (float)p->jor_omega, x, c->f32[1], y, nullptr, sh);
                std::swap(x, y);
            }
            pt.done(p->jor_sweeps);
            launch_correct_f32(c->stream, g, b, x, F(c, IH_PRE));
            launch_exchange_relax(c->stream, g, F(c, IH_PRE));
            LAUNCHCHK();
        } else {
            rc = jor_run(c, p->jor_sweeps, p->jor_omega, b, IH_PRE, IH_RES, p->pressure_bc ? substep : 0, p->dtime);
            if (rc) return rc;
        }
    }
    int rc2;
    const double bnorm = absmax_host(c, b, F(c, IH_RES), &rc2);
    if (rc2) return rc2;
    const double eps = p->bicg_rel_tol * bnorm;
    PhaseTimer pt(c, "bicgstab");
    if (p->precond_sweeps > 0) {
        // keep the initial guess: if the preconditioned iteration breaks down or
        // does not converge, the plain FP64 solver restarts from it
        HIPCHK(hipMemcpyAsync(F(c, IH_COMP), F(c, IH_PRE), sizeof(double) * (size_t)c->g.fe, hipMemcpyDeviceToDevice,
                              c->stream));
        ih_bicgstab_stats mst{};
        rc = p->precond_fp64
                 ? ih_bicgstab_jor64(c, eps, p->bicg_max_it, ss, nn, IH_RES, IH_PRE, p->precond_sweeps, p->jor_omega,
                                     &mst)
                 : ih_bicgstab_mixed(c, eps, p->bicg_max_it, ss, nn, IH_RES, IH_PRE, p->precond_sweeps, p->jor_omega,
                                     &mst);
        if (!rc && mst.exit_reason != 0) {
            HIPCHK(hipMemcpyAsync(F(c, IH_PRE), F(c, IH_COMP), sizeof(double) * (size_t)c->g.fe,
                                  hipMemcpyDeviceToDevice, c->stream));
            rc = ih_bicgstab(c, eps, p->bicg_max_it, ss, nn, IH_RES, IH_PRE, bst);
            if (bst) bst->iterations += mst.iterations;
        } else if (bst) {
            *bst = mst;
        }
    } else if (p->precond_poisson == 1 && p->mg_fp32) {
        // the FP32 cycle; a solve that does not converge is redone with the FP64 cycle from the saved guess
        HIPCHK(hipMemcpyAsync(F(c, IH_COMP), F(c, IH_PRE), sizeof(double) * (size_t)c->g.fe, hipMemcpyDeviceToDevice,
                              c->stream));
        ih_bicgstab_stats mst{};
        rc = ih_bicgstab_mg_f32(c, eps, p->bicg_max_it, ss, nn, IH_RES, IH_PRE, &mst);
        if (!rc && mst.exit_reason != 0) {
            HIPCHK(hipMemcpyAsync(F(c, IH_PRE), F(c, IH_COMP), sizeof(double) * (size_t)c->g.fe,
                                  hipMemcpyDeviceToDevice, c->stream));
            rc = ih_bicgstab_precond(c, eps, p->bicg_max_it, ss, nn, IH_RES, IH_PRE, 1, bst);
            if (bst) bst->iterations += mst.iterations;
        } else if (bst) {
            *bst = mst;
        }
    } else {
        rc = ih_bicgstab_precond(c, eps, p->bicg_max_it, ss, nn, IH_RES, IH_PRE, p->precond_poisson == 1 ? 1 : 0, bst);
    }
    pt.done(bst ? bst->iterations : 0);
    return rc;
}

}  // extern "C"

namespace ih {

// The phases of a time step (ih_timestep and the batched step run the same
// functions, so a batched case is bit-identical to its own step).

// the checks of ih_timestep
int step_check(ih_case *c, const ih_step_params *p) {
    int rc = need_coeffs(c);
    if (rc) return rc;
    if (!p || !(p->dtime > 0.0) || !(p->Re > 0.0)) return fail(IH_EINVAL, "bad step parameters");
    if (p->inflow_kind == 2 && !c->has_mask) return fail(IH_ESTATE, "pulsatile inflow needs a mask");
    if (p->pressure_bc && (!c->ctlab || p->jor_fp32))
        return fail(p->jor_fp32 ? IH_EINVAL : IH_ESTATE,
                    "the Dirichlet pressure of the Jacobi sweeps needs CT labels and FP64 sweeps");
    if (p->use_mask && !c->has_mask) return fail(IH_ESTATE, "no mask set");
    return IH_OK;
}

// substep sub (0..2) up to the Poisson right-hand side: ghost fill, velocity
// interpolation, rhs_vel (nonlinear + Helmholtz fused per component), the
// Dirichlet planes, the pressure scaled as the increment's guess, vel = rhs and
// div vel into RES
int step_pre(ih_case *c, const ih_step_params *p, double time, int sub) {
    const Geo &g = c->g;
    hipStream_t s = c->stream;
    const double dtime = p->dtime;
    const Box pb = pbox(g);
    int rc;
    double subtime;
    if (sub == 0)
        subtime = time + dtime * aRK[0];
    else if (sub == 1)
        subtime = time + dtime * (aRK[0] + aRK[1] + bRK[1]);
    else
        subtime = time + dtime;

    // exchange_all_all(.TRUE., vel)
    double *vel[3] = {F(c, IH_VEL1), F(c, IH_VEL2), F(c, IH_VEL3)};
    for (int dir = 1; dir <= 3; ++dir) {
        int vd[3] = {1, 2, 3};
        launch_exchange(s, g, dir, 3, vel, vd);
    }
    LAUNCHCHK();
    if ((rc = ih_interpolate_vel(c))) return rc;

    // rhs_vel: nonlinear + Helmholtz + rhs assembly fused per component
    const double mult_b = dtime * bRK[sub];
    const double mult_a = dtime * aRK[sub];
    for (int q = 0; q < 3; ++q)
        launch_nl_helm_rhs(s, g, c->c, q, F(c, IH_WORK1), F(c, IH_WORK2), F(c, IH_WORK3), F(c, IH_VEL1 + q),
                           F(c, IH_NL1 + q), F(c, IH_RHS1 + q), 1.0 / p->Re, mult_a, mult_b, sub != 0 ? 1 : 0);
    LAUNCHCHK();
    double amp = p->inflow_amp;
    if (p->inflow_kind == 2) {
        const double pi = 3.14159265358979323846;
        amp = p->inflow_amp *
              (1.0 + 0.5 * std::sin(2.0 * pi * subtime) + 0.2 * std::sin(4.0 * pi * subtime + 0.7) +
               0.1 * std::sin(6.0 * pi * subtime + 1.3));
    }
    launch_bc_planes(s, g, c->c, F(c, IH_RHS1), F(c, IH_RHS2), F(c, IH_RHS3), p->inflow_kind, amp,
                     p->inflow_kind == 2 ? c->mask : nullptr);
    LAUNCHCHK();

    // pressure scaled as the increment's initial guess
    launch_scale_box(s, g, pb, F(c, IH_PRE), aRK[sub] + bRK[sub], dtime, 0);

    // explicit: vel = rhs (on the B range: swap the buffers), then div vel
    for (int q = 0; q < 3; ++q) std::swap(c->fields[IH_VEL1 + q], c->fields[IH_RHS1 + q]);
    return ih_divergence2(c, IH_VEL1, IH_VEL2, IH_VEL3, IH_RES);
}

// substep sub after the Poisson solve: u -= G p, the pressure scaled back, corners
int step_post(ih_case *c, const ih_step_params *p, int sub) {
    const Geo &g = c->g;
    hipStream_t s = c->stream;
    int rc;
    for (int m = 1; m <= 3; ++m) {
        if ((rc = exch(c, m, 0, F(c, IH_PRE)))) return rc;
        launch_grad_correct(s, g, c->c, m - 1, F(c, IH_PRE), F(c, IH_VEL1 + m - 1));
    }
    LAUNCHCHK();
    launch_scale_box(s, g, pbox(g), F(c, IH_PRE), aRK[sub] + bRK[sub], p->dtime, 1);
    launch_fill_corners(s, g, F(c, IH_PRE));
    LAUNCHCHK();
    return IH_OK;
}

// after the last substep: the IBM mask, the CT forcing, pressure levelling and
// the step's divergence statistic
int step_end(ih_case *c, const ih_step_params *p, double time, ih_step_stats *st) {
    int rc;
    if (p->use_mask)
        if ((rc = ih_apply_mask(c))) return rc;
    // apply_CT_forcing follows the time update (src/mod_timeint.f90:405-416)
    if (p->use_ct)
        if ((rc = ih_ct_forcing(c, time + p->dtime))) return rc;
    if (p->level_pressure)
        if ((rc = ih_level_pressure(c, IH_PRE))) return rc;
    if (st) {
        if ((rc = ih_divergence2(c, IH_VEL1, IH_VEL2, IH_VEL3, IH_TMP2))) return rc;
        st->div_max_after = absmax_host(c, pbox(c->g), F(c, IH_TMP2), &rc);
        if (rc) return rc;
    }
    return IH_OK;
}

}  // namespace ih

extern "C" {

int ih_timestep(ih_case *c, const ih_step_params *p, double time, ih_step_stats *st) {
    int rc = step_check(c, p);
    if (rc) return rc;
    HIPCHK(hipSetDevice(c->device));
    const Box pb = pbox(c->g);
    for (int sub = 0; sub < 3; ++sub) {
        if ((rc = step_pre(c, p, time, sub))) return rc;
        ih_bicgstab_stats bst{};
        if ((rc = poisson_solve(c, p, pb, sub + 1, &bst))) return rc;
        if (st) {
            st->bicg_iterations[sub] = bst.iterations;
            st->bicg_final[sub] = bst.norm_inf_final;
        }
        if ((rc = step_post(c, p, sub))) return rc;
    }
    if ((rc = step_end(c, p, time, st))) return rc;
    HIPCHK(hipStreamSynchronize(c->stream));
    return IH_OK;
}

// ---------------------------------------------------------------------------
// Reference-signature entry points: host arrays in and out, as the reference's
// *_host.f90 wrappers take them.  A host array bound to a field (ih_bind_field)
// is that field's mirror: it is uploaded only when the host copy is newer, the
// call works on the field in place, and a result stays on the device until
// ih_sync.  Any other array is staged through TMP1..TMP4 and crosses PCIe both
// ways on every call, as the reference's wrappers do.
// ---------------------------------------------------------------------------
namespace {

// staging fields of the wrappers and the relaxation's ping-pong field: never bound
bool bindable(int f) { return valid_field(f) && !(f >= IH_TMP1 && f <= IH_TMP4) && f != IH_COMP; }

int bound_field(const ih_case *c, const double *host) {
    if (!host) return -1;
    for (int f = 0; f < IH_NFIELDS; ++f)
        if (c->bind_state[f] != BIND_NONE && c->bind_host[f] == host) return f;
    return -1;
}

// One host-array argument of a wrapper: the field the call works on.
struct Arg {
    double *host;
    int field;
    bool bound;
};

// resolves `host` to its bound field (uploading a newer host copy) or uploads it
// into `staging`; `upload` false: an output whose old values are not read
int arg_in(ih_case *c, double *host, ih_field staging, Arg &a, bool upload = true) {
    if (!host) return fail(IH_EINVAL, "null array argument");
    a.host = host;
    const int f = bound_field(c, host);
    a.bound = f >= 0;
    a.field = a.bound ? f : (int)staging;
    int rc;
    if (a.bound) {
        if (upload && c->bind_state[f] == BIND_HOST_NEWER) {
            if ((rc = ih_field_upload(c, (ih_field)f, host))) return rc;
            c->bind_state[f] = BIND_CLEAN;
        }
    } else if (upload) {
        if ((rc = ih_field_upload(c, staging, host))) return rc;
    }
    return IH_OK;
}

// after the call: a written bound field is now the newer copy; a written staged
// array goes back to the host
int arg_out(ih_case *c, const Arg &a, bool written) {
    if (!written) return IH_OK;
    if (a.bound) {
        c->bind_state[a.field] = BIND_DEVICE_NEWER;
        return IH_OK;
    }
    return ih_field_download(c, (ih_field)a.field, a.host);
}

}  // namespace

int ih_bind_field(ih_case *c, ih_field f, double *host) {
    if (!c || !valid_field(f)) return fail(IH_EINVAL, "bad argument");
    if (!bindable(f)) return fail(IH_EINVAL, "field %d is a staging field of the wrappers and cannot be bound", (int)f);
    int rc;
    if (host) {
        const int g = bound_field(c, host);
        if (g >= 0 && g != (int)f) return fail(IH_EINVAL, "that host array is already bound to field %d", g);
    }
    // an unbound or rebound mirror gives its newer device copy back first
    if (c->bind_state[f] == BIND_DEVICE_NEWER && c->bind_host[f] != host) {
        if ((rc = ih_field_download(c, f, const_cast<double *>(c->bind_host[f])))) return rc;
    }
    c->bind_host[f] = host;
    c->bind_state[f] = host ? BIND_HOST_NEWER : BIND_NONE;
    return IH_OK;
}

int ih_sync(ih_case *c, ih_field f, int dir) {
    if (!c || !valid_field(f) || (dir != IH_SYNC_TO_HOST && dir != IH_SYNC_TO_DEVICE))
        return fail(IH_EINVAL, "bad argument");
    if (c->bind_state[f] == BIND_NONE) return fail(IH_ESTATE, "field %d is not bound", (int)f);
    int rc;
    if (dir == IH_SYNC_TO_HOST) {
        if (c->bind_state[f] == BIND_DEVICE_NEWER) {
            if ((rc = ih_field_download(c, f, const_cast<double *>(c->bind_host[f])))) return rc;
            c->bind_state[f] = BIND_CLEAN;
        }
    } else {
        // the host changed the array: its copy becomes the newer one
        if ((rc = ih_field_upload(c, f, c->bind_host[f]))) return rc;
        c->bind_state[f] = BIND_CLEAN;
    }
    return IH_OK;
}

int ih_sync_all(ih_case *c, int dir) {
    if (!c) return fail(IH_EINVAL, "null case");
    for (int f = 0; f < IH_NFIELDS; ++f)
        if (c->bind_state[f] != BIND_NONE) {
            const int rc = ih_sync(c, (ih_field)f, dir);
            if (rc) return rc;
        }
    return IH_OK;
}

int ih_transfer_stats(ih_case *c, int64_t out[4]) {
    if (!c || !out) return fail(IH_EINVAL, "bad argument");
    out[0] = c->n_up;
    out[1] = c->n_down;
    out[2] = c->bytes_up;
    out[3] = c->bytes_down;
    return IH_OK;
}

int ih_device_gradient(ih_case *c, int dimension_id, double *phi, double *grad) {
    if (!c) return fail(IH_EINVAL, "null case");
    Arg p{}, g{};
    int rc;
    if ((rc = arg_in(c, phi, IH_TMP1, p)) || (rc = arg_in(c, grad, IH_TMP2, g))) return rc;
    if ((rc = ih_gradient(c, dimension_id, (ih_field)p.field, (ih_field)g.field))) return rc;
    // the ghost fill of phi is part of the call, as in the reference
    if ((rc = arg_out(c, p, true)) || (rc = arg_out(c, g, true))) return rc;
    return IH_OK;
}

int ih_device_divergence2(ih_case *c, double *phi3, double *div) {
    int rc;
    if (!c || !phi3) return fail(IH_EINVAL, "null argument");
    const int64_t n = ih_field_elems(c);
    Arg v[3] = {}, d{};
    for (int q = 0; q < 3; ++q)
        if ((rc = arg_in(c, phi3 + q * n, (ih_field)(IH_TMP1 + q), v[q]))) return rc;
    if ((rc = arg_in(c, div, IH_TMP4, d))) return rc;
    if ((rc = ih_divergence2(c, (ih_field)v[0].field, (ih_field)v[1].field, (ih_field)v[2].field, (ih_field)d.field)))
        return rc;
    for (int q = 0; q < 3; ++q)
        if ((rc = arg_out(c, v[q], true))) return rc;
    return arg_out(c, d, true);
}

int ih_device_nonlinear_upwind(ih_case *c, int component_id, double *pp, double *vel_component, double *nl_component) {
    if (!c) return fail(IH_EINVAL, "null case");
    Arg a{}, v{}, nl{};
    int rc;
    if ((rc = arg_in(c, pp, IH_TMP1, a)) || (rc = arg_in(c, vel_component, IH_TMP2, v)) ||
        (rc = arg_in(c, nl_component, IH_TMP3, nl)))
        return rc;
    if ((rc = ih_nonlinear_upwind(c, component_id, (ih_field)a.field, (ih_field)v.field, (ih_field)nl.field))) return rc;
    return arg_out(c, nl, true);
}

int ih_device_helmholtz(ih_case *c, int component_id, double multL, double *vel_component, double *helm_component) {
    if (!c) return fail(IH_EINVAL, "null case");
    Arg v{}, h{};
    int rc;
    if ((rc = arg_in(c, vel_component, IH_TMP1, v)) || (rc = arg_in(c, helm_component, IH_TMP2, h))) return rc;
    if ((rc = ih_helmholtz(c, component_id, multL, (ih_field)v.field, (ih_field)h.field))) return rc;
    return arg_out(c, h, true);
}

int ih_device_axpby(ih_case *c, double a, double b, double *x, double *y, double *z) {
    if (!c) return fail(IH_EINVAL, "null case");
    Arg ax{}, ay{}, az{};
    int rc;
    if ((rc = arg_in(c, x, IH_TMP1, ax)) || (rc = arg_in(c, y, IH_TMP2, ay)) ||
        (rc = arg_in(c, z, IH_TMP3, az, false)))
        return rc;
    if ((rc = ih_axpby(c, a, b, (ih_field)ax.field, (ih_field)ay.field, (ih_field)az.field))) return rc;
    return arg_out(c, az, true);
}

int ih_device_interpolate(ih_case *c, int dimension_id, double *phi, double *inter) {
    if (!c) return fail(IH_EINVAL, "null case");
    Arg p{}, i{};
    int rc;
    if ((rc = arg_in(c, phi, IH_TMP1, p)) || (rc = arg_in(c, inter, IH_TMP2, i))) return rc;
    if ((rc = ih_interpolate2_pre_vel(c, dimension_id, (ih_field)p.field, (ih_field)i.field))) return rc;
    return arg_out(c, i, true);
}

int ih_device_interpolate2(ih_case *c, int dimension_id, double *phi, double *inter) {
    if (!c) return fail(IH_EINVAL, "null case");
    Arg p{}, i{};
    int rc;
    if ((rc = arg_in(c, phi, IH_TMP1, p)) || (rc = arg_in(c, inter, IH_TMP2, i))) return rc;
    if ((rc = ih_interpolate2_vel_pre(c, dimension_id, (ih_field)p.field, (ih_field)i.field))) return rc;
    return arg_out(c, i, true);
}

// device_jacobi_solver (src/jacobi_solver_host_v4.f90:5-717) in the CPU
// relaxation's out-of-place form.  Each trip applies the substep-3 Dirichlet
// pressure (:302-305); every 20th trip (iter mod 20 == 1) compares err_tot =
// sum |rel - rel_old| over the box (fixed reduction tree) with the previous
// value and stops when |log(err_tot_old / err_tot)| is within the threshold
// (:357-414); otherwise rel_old = rel and one sweep follows.  rel_old and
// err_tot_old are undefined at the reference's entry, so its first test
// compares with undefined values; here the first test only records them.  A
// later err_tot of 0 (a fixed point, where the reference's log is undefined and
// it never stops) stops; so does the sweep limit, with IH_ENOTCONV.  grid_id
// must be 1.
int ih_device_jacobi_solver(ih_case *c, int grid_id, double omega, int SS1, int SS2, int SS3, int NN1, int NN2,
                            int NN3, double *rel, double *bb) {
    int rc = need_coeffs(c);
    if (rc) return rc;
    if (grid_id != 1) return fail(IH_EINVAL, "only the finest grid (grid_id 1) is supported");
    if (!rel || !bb) return fail(IH_EINVAL, "null argument");
    const int ss[3] = {SS1, SS2, SS3}, nn[3] = {NN1, NN2, NN3};
    Box b;
    if ((rc = to_box(c, ss, nn, b))) return rc;
    Arg ar{}, ab{};
    if ((rc = arg_in(c, rel, IH_TMP1, ar)) || (rc = arg_in(c, bb, IH_TMP2, ab))) return rc;
    double *X = F(c, (ih_field)ar.field);
    const int mode_j = 20;
    const double thresh = 20. * mode_j * 8. * 1. / 3.0 * 0.8 * 3.1415 * 3.1415 *
                          (1. / (double)(c->g.N[0] * c->g.N[0]) + 1. / (double)(c->g.N[1] * c->g.N[1]) +
                           1. / (double)(c->g.N[2] * c->g.N[2]));
    const int substep = c->ct.substep;
    const double dtime = c->ct.dtime;
    double err_tot_old = 0.0;
    launch_fill(c->stream, F(c, IH_TMP3), 0.0, c->g.fe);
    const int64_t nb = red_blocks(b);
    if ((rc = ensure_part(c, nb))) return rc;
    const int max_iter = 200000;
    bool stopped = false;
    for (int iter = 1; iter <= max_iter; ++iter) {
        if ((rc = jacobi_pressure_bc(c, X, substep, dtime))) return rc;
        if (iter % mode_j == 1) {
            launch_absdiff_partial(c->stream, c->g, b, X, F(c, IH_TMP3), c->part);
            launch_final_sum(c->stream, c->part, nb, c->dscal + S_SUM);
            LAUNCHCHK();
            if ((rc = read_scalars(c, S_SUM, 1))) return rc;
            const double err_tot = c->hscal[S_SUM];
            if (iter > 1 && (err_tot == 0.0 || std::fabs(std::log(err_tot_old / err_tot)) <= thresh)) {
                stopped = true;
                break;
            }
            HIPCHK(hipMemcpyAsync(F(c, IH_TMP3), X, sizeof(double) * (size_t)c->g.fe,
                                  hipMemcpyDeviceToDevice, c->stream));
            err_tot_old = err_tot;
        }
        if ((rc = jor_run(c, 1, omega, b, (ih_field)ar.field, (ih_field)ab.field, 0, 0.0))) return rc;
    }
    if (!stopped && (rc = jacobi_pressure_bc(c, X, substep, dtime))) return rc;
    launch_exchange_relax(c->stream, c->g, X);
    LAUNCHCHK();
    if ((rc = arg_out(c, ar, true))) return rc;
    if (!stopped) return fail(IH_ENOTCONV, "device_jacobi_solver: no stop after %d sweeps", max_iter);
    return IH_OK;
}

// Launches of one kernel of the step, back to back between two events on the
// case's stream, for benchmarks and counter profiles.  The BiCGstab kernels run
// on the case's Krylov fields with the iteration flagged active (and scalars
// of 0.5); their results are meaningless.
int ih_bench_kernel(ih_case *c, const char *name, int launches, double *ms_per_launch) {
    int rc = need_coeffs(c);
    if (rc) return rc;
    if (!name || launches < 1 || !ms_per_launch) return fail(IH_EINVAL, "bad argument");
    HIPCHK(hipSetDevice(c->device));
    const Box b = pbox(c->g);
    hipStream_t s = c->stream;
    const std::string n(name);
    static const std::pair<const char *, int> parts[] = {{"p_update", KPART_P_UPDATE}, {"apply_w", KPART_APPLY_W},
                                                         {"s_update", KPART_S_UPDATE}, {"apply_self", KPART_APPLY_SELF},
                                                         {"x_update", KPART_X_UPDATE}};
    int part = -1;
    for (const auto &pp : parts)
        if (n == pp.first) part = pp.second;
    std::function<void()> one;
    KrylovPtrs k{F(c, IH_RR), F(c, IH_RH), F(c, IH_PK), F(c, IH_AP), F(c, IH_AR), F(c, IH_TMP4), c->part, c->part2,
                 c->ksc, c->kfl};
    k.part3 = c->part3;
    k.part4 = c->part4;
    if (n == "jor") {
        const JorLaunch sh = jor_shape(c, b, false, [&](JorLaunch l) {
            launch_jor(s, c->g, c->c, b, 0.8, F(c, IH_PRE), F(c, IH_RES), F(c, IH_COMP), l);
        });
        one = [&]() { launch_jor(s, c->g, c->c, b, 0.8, F(c, IH_PRE), F(c, IH_RES), F(c, IH_COMP), sh); };
    } else if (n == "jor_f32") {
        // one FP32 sweep of the mixed-precision preconditioner, on the case's float fields
        const JorLaunch sh = jor_shape(c, b, true, [&](JorLaunch l) {
            launch_jor_f32(s, c->g, c->c, b, 0.8f, c->f32[0], c->f32[1], c->f32[2], nullptr, l);
        });
        one = [&]() { launch_jor_f32(s, c->g, c->c, b, 0.8f, c->f32[0], c->f32[1], c->f32[2], nullptr, sh); };
    } else if (n == "apply") {
        one = [&]() {
            launch_apply_A(s, c->g, c->c, b, F(c, IH_PK), F(c, IH_AP), nullptr, KDOT_NONE, nullptr, nullptr, nullptr);
        };
    } else if (n == "nl_helm_rhs0" || n == "nl_helm_rhs1" || n == "nl_helm_rhs2") {
        // one component of the substep's fused nonlinear + Helmholtz + rhs (with the nl term of the last substep)
        const int q = n.back() - '0';
        one = [&, q]() {
            launch_nl_helm_rhs(s, c->g, c->c, q, F(c, IH_WORK1), F(c, IH_WORK2), F(c, IH_WORK3), F(c, IH_VEL1 + q),
                               F(c, IH_NL1 + q), F(c, IH_RHS1 + q), 0.5, 0.25, 0.125, 1);
        };
    } else if (n == "divergence2") {
        one = [&]() { launch_divergence2(s, c->g, c->c, F(c, IH_VEL1), F(c, IH_VEL2), F(c, IH_VEL3), F(c, IH_TMP1)); };
    } else if (n == "interp_vel_pre3") {
        one = [&]() {
            launch_interp_vel_pre3(s, c->g, c->c, F(c, IH_VEL1), F(c, IH_VEL2), F(c, IH_VEL3), F(c, IH_WORK1),
                                   F(c, IH_WORK2), F(c, IH_WORK3));
        };
    } else if (n == "grad_correct0") {
        one = [&]() { launch_grad_correct(s, c->g, c->c, 0, F(c, IH_PRE), F(c, IH_VEL1)); };
    } else if (part >= 0) {
        if ((rc = ensure_part(c, red_blocks(b)))) return rc;
        for (int q = 0; q < KS_NSLOTS; ++q) c->hksc[q] = 0.5;
        for (int q = 0; q < KF_NSLOTS; ++q) c->hkfl[q] = 0;
        c->hkfl[KF_ACTIVE] = 1;
        c->hkfl[KF_MODE] = 1;
        HIPCHK(hipMemcpyAsync(c->ksc, c->hksc, sizeof(double) * KS_NSLOTS, hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(c->kfl, c->hkfl, sizeof(int) * KF_NSLOTS, hipMemcpyHostToDevice, s));
        // the update kernels reduce the partials of the operator form before them
        if (part == KPART_S_UPDATE) launch_krylov_part(s, c->g, c->c, b, k, KPART_APPLY_W);
        if (part == KPART_X_UPDATE) launch_krylov_part(s, c->g, c->c, b, k, KPART_APPLY_SELF);
        one = [&]() { launch_krylov_part(s, c->g, c->c, b, k, part); };
    } else {
        return fail(IH_EINVAL,
                    "unknown kernel '%s' (jor, jor_f32, apply, apply_w, apply_self, p_update, s_update, x_update, "
                    "nl_helm_rhs0/1/2, divergence2, interp_vel_pre3, grad_correct0)",
                    name);
    }
    one();  // warm
    HIPCHK(hipEventRecord(c->ev0, s));
    for (int q = 0; q < launches; ++q) one();
    HIPCHK(hipEventRecord(c->ev1, s));
    LAUNCHCHK();
    HIPCHK(hipEventSynchronize(c->ev1));
    float ms = 0.0f;
    HIPCHK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    *ms_per_launch = (double)ms / launches;
    return IH_OK;
}

// ---------------------------------------------------------------------------
int ih_timer_start(ih_case *c) {
    if (!c) return fail(IH_EINVAL, "null case");
    HIPCHK(hipEventRecord(c->ev0, c->stream));
    return IH_OK;
}

int ih_timer_stop(ih_case *c, double *ms) {
    if (!c || !ms) return fail(IH_EINVAL, "null argument");
    HIPCHK(hipEventRecord(c->ev1, c->stream));
    HIPCHK(hipEventSynchronize(c->ev1));
    float f = 0.f;
    HIPCHK(hipEventElapsedTime(&f, c->ev0, c->ev1));
    *ms = f;
    return IH_OK;
}

int ih_kernel_timing_enable(ih_case *c, int on) {
    if (!c) return fail(IH_EINVAL, "null case");
    resolve_timings(c);
    c->ktime = on != 0;
    c->ktimes.clear();
    return IH_OK;
}

int ih_kernel_timing_get(ih_case *c, const char *family, double *total_ms, int64_t *launches) {
    if (!c || !family || !total_ms || !launches) return fail(IH_EINVAL, "null argument");
    resolve_timings(c);
    auto it = c->ktimes.find(family);
    if (it == c->ktimes.end()) {
        *total_ms = 0.0;
        *launches = 0;
        return IH_OK;
    }
    *total_ms = it->second.total_ms;
    *launches = it->second.launches;
    return IH_OK;
}

}  // extern "C"

// ===========================================================================
// Batched relaxation of many cases (uneven grids) in single launches.
// ===========================================================================
struct ih_batch {
    std::vector<ih_case *> cases;
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    int tpb = 256, zchunk = 16, layout = 0;
    int ntiles[2] = {}, row_tiles[2] = {};  // [fp32]: the two types tile differently
    JorBatchItem *items[2][2] = {};        // [fp32][direction]
    int *tile_item[2] = {};                // [fp32]: item of every global tile
    int tile_cap[2] = {};
    bool any_periodic_x = false;           // some case is periodic in x
    // Cases are relaxed in groups whose fields (rel, out, bb) fit group_bytes,
    // all sweeps of a group before the next: a group small enough to stay in
    // the device's last-level cache is swept from there.  0: one group.
    int64_t group_bytes = 0;
    std::vector<int> group;                // first case of every group, then the case count
    std::vector<int> gtile[2];             // [fp32]: first global tile of every group, then the tile count
    std::vector<int> grows[2];             // [fp32]: row tiles of every group (its largest case's)
    int64_t g_group = -1;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int g_sweeps = -1, g_fp32 = -1, g_tpb = -1, g_zchunk = -1, g_layout = -1;
    double g_omega = 0.0;
    // the batched step's Poisson solve (ih_batch_timestep): the Krylov case
    // table and tile -> case map, the FP32 preconditioner sweeps' item tables
    // ([0] M pk -> Z1, [1] M rr -> Z2; the middle sweeps by parity), the flags
    // of all cases gathered in one place, and the captured iteration graph
    ih::KBatchCase *kcases = nullptr;
    int *ktile = nullptr, kntiles = 0;
    bool k_pz = false, k_nonpz = false, k_px = false;
    ih::JorFirstItem *pfirst[2] = {};
    int *pfirst_tile = nullptr, pfirst_ntiles = 0;
    ih::JorBatchItem *pmid[2] = {}, *plast[2] = {};
    int *pjor_tile = nullptr, pjor_ntiles = 0;
    int k_sweeps = -1;
    std::vector<const void *> k_sig;  // the pointers the tables were built from
    int *dflags = nullptr, *hflags = nullptr;
    double *dscal = nullptr, *hscal = nullptr;
    hipGraph_t kgraph = nullptr;
    hipGraphExec_t kexec = nullptr;
    int kg_iters = -1, kg_sweeps = -1;
    float kg_omega = 0.0f;
};

namespace {

void batch_drop_step(ih_batch *b) {
    if (b->kexec) (void)hipGraphExecDestroy(b->kexec);
    if (b->kgraph) (void)hipGraphDestroy(b->kgraph);
    b->kexec = nullptr;
    b->kgraph = nullptr;
    b->kg_iters = -1;
    for (void *q : {(void *)b->kcases, (void *)b->ktile, (void *)b->pfirst[0], (void *)b->pfirst[1],
                    (void *)b->pfirst_tile, (void *)b->pmid[0], (void *)b->pmid[1], (void *)b->plast[0],
                    (void *)b->plast[1], (void *)b->pjor_tile, (void *)b->dflags, (void *)b->dscal})
        if (q) (void)hipFree(q);
    if (b->hflags) (void)hipHostFree(b->hflags);
    if (b->hscal) (void)hipHostFree(b->hscal);
    b->kcases = nullptr;
    b->ktile = b->pfirst_tile = b->pjor_tile = nullptr;
    b->pfirst[0] = b->pfirst[1] = nullptr;
    b->pmid[0] = b->pmid[1] = b->plast[0] = b->plast[1] = nullptr;
    b->dflags = b->hflags = nullptr;
    b->dscal = b->hscal = nullptr;
    b->k_sig.clear();
    b->k_sweeps = -1;
}

void batch_drop_graph(ih_batch *b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
    b->g_sweeps = -1;
}

// item tables for the current tile shape, one numbering per type (2 doubles
// or 4 floats per lane)
int batch_build(ih_batch *b) {
    b->any_periodic_x = false;
    for (ih_case *c : b->cases) b->any_periodic_x = b->any_periodic_x || c->g.per[0] != 0;
    b->group.assign(1, 0);
    int64_t used = 0;
    for (size_t q = 0; q < b->cases.size(); ++q) {
        const int64_t ws = 3 * (int64_t)sizeof(double) * b->cases[q]->g.fe;
        if (b->group_bytes > 0 && used > 0 && used + ws > b->group_bytes) {
            b->group.push_back((int)q);
            used = 0;
        }
        used += ws;
    }
    b->group.push_back((int)b->cases.size());
    for (int p = 0; p < 2; ++p) {
        std::vector<JorBatchItem> h[2];
        int t0 = 0;
        b->row_tiles[p] = 0;
        for (ih_case *c : b->cases) {
            const Box pb = pbox(c->g);
            JorBatchItem it{};
            const int nt = jor_batch_tiles(pb, c->g, p ? 4 : 2, b->tpb, b->zchunk, it.sh, it.ntxy, it.nbz);
            b->row_tiles[p] = std::max(b->row_tiles[p], nt);
            it.g = c->g;
            it.r = Range3{pb.lo[0], pb.lo[1], pb.lo[2], pb.hi[0] - pb.lo[0] + 1, pb.hi[1] - pb.lo[1] + 1,
                          pb.hi[2] - pb.lo[2] + 1};
            for (int d = 0; d < 3; ++d) {
                it.dg[d] = c->c.dg[d];
                it.dgf[d] = c->c.dgf[d];
            }
            it.tile0 = t0;
            t0 += nt;
            for (int dir = 0; dir < 2; ++dir) {
                JorBatchItem a = it;
                if (p == 0) {
                    a.rel = dir == 0 ? F(c, IH_PRE) : F(c, IH_COMP);
                    a.out = dir == 0 ? F(c, IH_COMP) : F(c, IH_PRE);
                    a.bb = F(c, IH_RES);
                } else {
                    a.rel = dir == 0 ? (void *)c->f32[0] : (void *)c->f32[2];
                    a.out = dir == 0 ? (void *)c->f32[2] : (void *)c->f32[0];
                    a.bb = c->f32[1];
                }
                h[dir].push_back(a);
            }
        }
        b->ntiles[p] = t0;
        b->gtile[p].clear();
        b->grows[p].clear();
        for (size_t gi = 0; gi + 1 < b->group.size(); ++gi) {
            b->gtile[p].push_back(h[0][(size_t)b->group[gi]].tile0);
            int rows = 0;
            for (int q = b->group[gi]; q < b->group[gi + 1]; ++q) {
                const int e = q + 1 < (int)b->cases.size() ? h[0][(size_t)q + 1].tile0 : t0;
                rows = std::max(rows, e - h[0][(size_t)q].tile0);
            }
            b->grows[p].push_back(rows);
        }
        b->gtile[p].push_back(t0);
        std::vector<int> ti((size_t)t0);
        for (size_t q = 0; q < b->cases.size(); ++q) {
            const int a = h[0][q].tile0, e = q + 1 < b->cases.size() ? h[0][q + 1].tile0 : t0;
            for (int t = a; t < e; ++t) ti[(size_t)t] = (int)q;
        }
        if (t0 > b->tile_cap[p]) {
            if (b->tile_item[p]) (void)hipFree(b->tile_item[p]);
            b->tile_item[p] = nullptr;
            HIPCHK(hipMalloc(&b->tile_item[p], sizeof(int) * (size_t)t0));
            b->tile_cap[p] = t0;
        }
        if (t0 > 0)
            HIPCHK(hipMemcpy(b->tile_item[p], ti.data(), sizeof(int) * (size_t)t0, hipMemcpyHostToDevice));
        for (int dir = 0; dir < 2; ++dir) {
            if (!b->items[p][dir]) HIPCHK(hipMalloc(&b->items[p][dir], sizeof(JorBatchItem) * b->cases.size()));
            HIPCHK(hipMemcpy(b->items[p][dir], h[dir].data(), sizeof(JorBatchItem) * b->cases.size(),
                             hipMemcpyHostToDevice));
        }
    }
    batch_drop_graph(b);
    return IH_OK;
}

// n sweeps of every case, group after group; dir >= 0 keeps one direction
// (PRE -> COMP, for timing without changing PRE), else the sweeps alternate
int batch_run(ih_batch *b, int n, double omega, bool fp32, int layout, int dir) {
    const int p = fp32 ? 1 : 0;
    for (size_t gi = 0; gi + 1 < b->group.size(); ++gi) {
        const int c0 = b->group[gi], c1 = b->group[gi + 1];
        const int t0 = b->gtile[p][gi], nt = b->gtile[p][gi + 1] - t0;
        if (nt == 0) continue;
        for (int q = 0; q < n; ++q) {
            const JorBatchItem *it = b->items[p][dir >= 0 ? dir : (q & 1)];
            if (launch_jor_batch(b->stream, layout == 1 ? it + c0 : it, b->tile_item[p], t0, nt, c1 - c0,
                                 b->grows[p][gi], layout, b->tpb, fp32, omega, b->any_periodic_x))
                return fail(IH_EINVAL, "unsupported tile shape");
        }
    }
    LAUNCHCHK();
    return IH_OK;
}

int batch_sweeps(ih_batch *b, int n, double omega, bool fp32) { return batch_run(b, n, omega, fp32, b->layout, -1); }

}  // namespace

extern "C" {

int ih_batch_create(ih_case *const *cases, int n, ih_batch **out) {
    if (!cases || n <= 0 || !out) return fail(IH_EINVAL, "bad argument");
    *out = nullptr;
    for (int q = 0; q < n; ++q) {
        if (!cases[q]) return fail(IH_EINVAL, "case %d is null", q);
        if (!cases[q]->has_coeffs) return fail(IH_ESTATE, "case %d has no coefficients", q);
        if (cases[q]->device != cases[0]->device) return fail(IH_EINVAL, "cases of one batch share a device");
        for (int r = 0; r < q; ++r)
            if (cases[r] == cases[q]) return fail(IH_EINVAL, "case %d appears twice", q);
    }
    ih_batch *b = new ih_batch();
    b->cases.assign(cases, cases + n);
    b->device = cases[0]->device;
    int rc;
    if (hipSetDevice(b->device) != hipSuccess || hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&b->ev0) != hipSuccess || hipEventCreate(&b->ev1) != hipSuccess) {
        ih_batch_destroy(b);
        return fail(IH_EHIP, "stream/event creation failed");
    }
    if ((rc = batch_build(b))) {
        ih_batch_destroy(b);
        return rc;
    }
    *out = b;
    return IH_OK;
}

void ih_batch_destroy(ih_batch *b) {
    if (!b) return;
    (void)hipSetDevice(b->device);
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    batch_drop_graph(b);
    batch_drop_step(b);
    for (auto &pp : b->items)
        for (auto *q : pp)
            if (q) (void)hipFree(q);
    for (int *t : b->tile_item)
        if (t) (void)hipFree(t);
    if (b->ev0) (void)hipEventDestroy(b->ev0);
    if (b->ev1) (void)hipEventDestroy(b->ev1);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
}

int ih_batch_set_shape(ih_batch *b, int tpb, int zchunk, int layout) {
    if (!b || (tpb != 128 && tpb != 256 && tpb != 512) || zchunk < 1 || zchunk > 64 || (layout != 0 && layout != 1))
        return fail(IH_EINVAL, "bad launch shape");
    HIPCHK(hipSetDevice(b->device));
    HIPCHK(hipStreamSynchronize(b->stream));
    b->tpb = tpb;
    b->zchunk = zchunk;
    b->layout = layout;
    return batch_build(b);
}

int ih_batch_autotune(ih_batch *b, int sweeps, int *tpb, int *zchunk, int *layout, double *ms_best) {
    if (!b || sweeps < 1) return fail(IH_EINVAL, "bad argument");
    HIPCHK(hipSetDevice(b->device));
    for (ih_case *c : b->cases) HIPCHK(hipStreamSynchronize(c->stream));
    static const int tpbs[3] = {128, 256, 512}, zcs[5] = {8, 12, 16, 24, 32};
    static const int64_t gbs[4] = {0, (int64_t)192 << 20, (int64_t)96 << 20, (int64_t)48 << 20};
    double best = 1e300;
    int bt = b->tpb, bz = b->zchunk, bl = b->layout, rc;
    int64_t bg = b->group_bytes;
    for (int64_t gb : gbs)
        for (int lay = 0; lay < 2; ++lay)
            for (int a : tpbs)
                for (int z : zcs) {
                    b->tpb = a;
                    b->zchunk = z;
                    b->group_bytes = gb;
                    if ((rc = batch_build(b))) return rc;
                    // direction 0 only (PRE -> COMP): the relaxed field is left untouched
                    if ((rc = batch_run(b, 2, 0.8, false, lay, 0))) return rc;
                    HIPCHK(hipEventRecord(b->ev0, b->stream));
                    if ((rc = batch_run(b, sweeps, 0.8, false, lay, 0))) return rc;
                    HIPCHK(hipEventRecord(b->ev1, b->stream));
                    HIPCHK(hipEventSynchronize(b->ev1));
                    float ms = 0.f;
                    HIPCHK(hipEventElapsedTime(&ms, b->ev0, b->ev1));
                    if (ms < best) {
                        best = ms;
                        bt = a;
                        bz = z;
                        bl = lay;
                        bg = gb;
                    }
                }
    b->tpb = bt;
    b->zchunk = bz;
    b->layout = bl;
    b->group_bytes = bg;
    if ((rc = batch_build(b))) return rc;
    if (tpb) *tpb = bt;
    if (zchunk) *zchunk = bz;
    if (layout) *layout = bl;
    if (ms_best) *ms_best = best / sweeps;
    return IH_OK;
}

int ih_batch_set_groups(ih_batch *b, int64_t bytes) {
    if (!b || bytes < 0) return fail(IH_EINVAL, "bad argument");
    HIPCHK(hipSetDevice(b->device));
    HIPCHK(hipStreamSynchronize(b->stream));
    b->group_bytes = bytes;
    return batch_build(b);
}

int ih_batch_get_groups(ih_batch *b, int64_t *bytes, int *count) {
    if (!b) return fail(IH_EINVAL, "bad argument");
    if (bytes) *bytes = b->group_bytes;
    if (count) *count = (int)b->group.size() - 1;
    return IH_OK;
}

}  // extern "C"

namespace {

// ih_batch_jor's work enqueued on the batch's stream (no synchronisation)
int batch_jor_enqueue(ih_batch *b, int n_sweeps, double omega, int fp32, int use_graph) {
    int rc;
    hipStream_t s = b->stream;
    if (fp32) {
        // defect correction per case: r = RES - A_cdg PRE in FP64, A e = r in FP32 from e = 0
        for (ih_case *c : b->cases) {
            const Box pb = pbox(c->g);
            if ((rc = ensure_part(c, red_blocks(pb)))) return rc;
            launch_exchange_relax(s, c->g, F(c, IH_PRE));
            launch_residual_cdg(s, c->g, c->c, pb, F(c, IH_PRE), F(c, IH_RES), F(c, IH_TMP1), c->part);
            launch_to_f32(s, c->g, pb, F(c, IH_TMP1), c->f32[1]);
            launch_fill_f32(s, c->f32[0], 0.0f, c->g.fe);
            launch_fill_f32(s, c->f32[2], 0.0f, c->g.fe);
        }
        LAUNCHCHK();
    }
    if (n_sweeps > 0) {
        if (use_graph) {
            if (!b->exec || b->g_sweeps != n_sweeps || b->g_fp32 != fp32 || b->g_omega != omega ||
                b->g_tpb != b->tpb || b->g_zchunk != b->zchunk || b->g_layout != b->layout ||
                b->g_group != b->group_bytes) {
                batch_drop_graph(b);
                HIPCHK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
                rc = batch_sweeps(b, n_sweeps, omega, fp32 != 0);
                hipGraph_t gr = nullptr;
                const hipError_t e = hipStreamEndCapture(s, &gr);
                if (rc) {
                    if (gr) (void)hipGraphDestroy(gr);
                    return rc;
                }
                if (e != hipSuccess) return fail(IH_EHIP, "graph capture: %s", hipGetErrorString(e));
                b->graph = gr;
                HIPCHK(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
                b->g_sweeps = n_sweeps;
                b->g_fp32 = fp32;
                b->g_omega = omega;
                b->g_tpb = b->tpb;
                b->g_zchunk = b->zchunk;
                b->g_layout = b->layout;
                b->g_group = b->group_bytes;
            }
            HIPCHK(hipGraphLaunch(b->exec, s));
        } else if ((rc = batch_sweeps(b, n_sweeps, omega, fp32 != 0))) {
            return rc;
        }
    }
    for (ih_case *c : b->cases) {
        const Box pb = pbox(c->g);
        if (fp32) {
            launch_correct_f32(s, c->g, pb, (n_sweeps & 1) ? c->f32[2] : c->f32[0], F(c, IH_PRE));
        } else if (n_sweeps & 1) {
            launch_copy_box(s, c->g, pb, F(c, IH_COMP), F(c, IH_PRE));
        }
        launch_exchange_relax(s, c->g, F(c, IH_PRE));
    }
    LAUNCHCHK();
    return IH_OK;
}

}  // namespace

extern "C" {

int ih_batch_jor(ih_batch *b, int n_sweeps, double omega, int fp32, int use_graph) {
    if (!b || n_sweeps < 0) return fail(IH_EINVAL, "bad argument");
    HIPCHK(hipSetDevice(b->device));
    for (ih_case *c : b->cases) HIPCHK(hipStreamSynchronize(c->stream));
    int rc = batch_jor_enqueue(b, n_sweeps, omega, fp32, use_graph);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(b->stream));
    return IH_OK;
}

}  // extern "C"

// ===========================================================================
// Batched time step: ih_timestep's sequence for every case of a batch, with
// the Poisson solves of all cases in shared launches.
// ===========================================================================
namespace {

// FP32 preconditioner sweeps in the batch's JOR tiling (16-byte lanes, 4 floats)
constexpr int PRE_TPB = 256, PRE_ZCHUNK = 10;

// the pointers the device tables hold: a change (a re-allocated partial buffer,
// a swapped velocity field does not matter) means the tables are rebuilt
std::vector<const void *> step_signature(const ih_batch *b, int sweeps) {
    std::vector<const void *> sig;
    sig.push_back((const void *)(intptr_t)sweeps);
    for (const ih_case *c : b->cases) {
        for (int f : {IH_RR, IH_RH, IH_PK, IH_AP, IH_AR, IH_PRE, IH_Z1, IH_Z2, IH_RES}) sig.push_back(c->fields[f]);
        for (const void *q : {(const void *)c->part, (const void *)c->part2, (const void *)c->part3,
                              (const void *)c->part4, (const void *)c->ksc, (const void *)c->kfl,
                              (const void *)c->f32[0], (const void *)c->f32[1], (const void *)c->f32[2],
                              (const void *)c->dcoef, (const void *)c->dcoef32})
            sig.push_back(q);
    }
    return sig;
}

template <typename T>
int upload_table(T **dev, const std::vector<T> &h) {
    if (*dev) (void)hipFree(*dev);
    *dev = nullptr;
    if (h.empty()) return IH_OK;
    HIPCHK(hipMalloc(dev, sizeof(T) * h.size()));
    HIPCHK(hipMemcpy(*dev, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice));
    return IH_OK;
}

// the Krylov and preconditioner tables of the batch (cases' partial buffers sized first)
int batch_step_tables(ih_batch *b, int sweeps) {
    int rc;
    for (ih_case *c : b->cases)
        if ((rc = ensure_part(c, red_blocks(pbox(c->g))))) return rc;
    std::vector<const void *> sig = step_signature(b, sweeps);
    if (b->kcases && sig == b->k_sig) return IH_OK;
    HIPCHK(hipStreamSynchronize(b->stream));
    batch_drop_step(b);
    const int n = (int)b->cases.size();
    std::vector<KBatchCase> kc((size_t)n);
    std::vector<int> ktile;
    std::vector<JorFirstItem> first[2];
    std::vector<int> ftile;
    std::vector<JorBatchItem> mid[2], last[2];
    std::vector<int> jtile;
    b->k_pz = b->k_nonpz = b->k_px = false;
    for (int q = 0; q < n; ++q) {
        ih_case *c = b->cases[(size_t)q];
        const Geo &g = c->g;
        const Box pb = pbox(g);
        KBatchCase &K = kc[(size_t)q];
        K.g = g;
        K.c = c->c;
        K.r = Range3{pb.lo[0], pb.lo[1], pb.lo[2], pb.hi[0] - pb.lo[0] + 1, pb.hi[1] - pb.lo[1] + 1,
                     pb.hi[2] - pb.lo[2] + 1};
        K.nbx = (K.r.n0 + RED_TX - 1) / RED_TX;
        K.nby = (K.r.n1 + RED_TY - 1) / RED_TY;
        K.nbz = (K.r.n2 + RED_KC - 1) / RED_KC;
        K.nb = (int64_t)K.nbx * K.nby * K.nbz;
        K.tile0 = (int)ktile.size();
        for (int64_t t = 0; t < K.nb; ++t) ktile.push_back(q);
        K.rr = F(c, IH_RR);
        K.rh = F(c, IH_RH);
        K.pk = F(c, IH_PK);
        K.Ap = F(c, IH_AP);
        K.Ar = F(c, IH_AR);
        K.phi = F(c, IH_PRE);
        K.z1 = F(c, IH_Z1);
        K.z2 = F(c, IH_Z2);
        K.bb = F(c, IH_RES);
        K.part1 = c->part;
        K.part2 = c->part2;
        K.part3 = c->part3;
        K.part4 = c->part4;
        K.sc = c->ksc;
        K.fl = c->kfl;
        K.b32 = sweeps > 0 ? c->f32[1] : nullptr;
        K.e32 = sweeps > 0 ? c->f32[0] : nullptr;
        (g.per[2] ? b->k_pz : b->k_nonpz) = true;
        b->k_px = b->k_px || g.per[0] != 0;
        if (sweeps > 0) {
            const int *act = c->kfl + KF_ACTIVE;
            // both applications of M (to pk and to rr) share the case's tiles
            JorFirstItem f{};
            f.g = g;
            for (int d = 0; d < 3; ++d) f.dgf[d] = c->c.dgf[d];
            const int ntf = jor_first_tiles(pb, g, f);
            f.tile0 = (int)ftile.size();
            f.b32 = c->f32[1];
            f.active = act;
            for (int t = 0; t < ntf; ++t) ftile.push_back(q);
            for (int w = 0; w < 2; ++w) {
                JorFirstItem fw = f;
                fw.v = w ? F(c, IH_RR) : F(c, IH_PK);
                fw.e = sweeps == 1 ? (void *)(w ? F(c, IH_Z2) : F(c, IH_Z1)) : (void *)c->f32[0];
                first[w].push_back(fw);
            }
            JorBatchItem it{};
            const int nt = jor_batch_tiles(pb, g, 4, PRE_TPB, PRE_ZCHUNK, it.sh, it.ntxy, it.nbz);
            it.g = g;
            it.r = K.r;
            for (int d = 0; d < 3; ++d) {
                it.dg[d] = c->c.dg[d];
                it.dgf[d] = c->c.dgf[d];
            }
            it.tile0 = (int)jtile.size();
            it.active = act;
            for (int t = 0; t < nt; ++t) jtile.push_back(q);
            float *e0 = c->f32[0], *e1 = c->f32[2];
            for (int par = 0; par < 2; ++par) {
                JorBatchItem m = it;
                m.rel = par ? e1 : e0;
                m.out = par ? e0 : e1;
                m.bb = c->f32[1];
                mid[par].push_back(m);
            }
            // the last sweep reads e0 after an even number of middle sweeps
            float *xl = ((sweeps - 2) % 2 == 0) ? e0 : e1;
            for (int w = 0; w < 2; ++w) {
                JorBatchItem m = it;
                m.rel = xl;
                m.out = w ? (void *)F(c, IH_Z2) : (void *)F(c, IH_Z1);
                m.bb = c->f32[1];
                last[w].push_back(m);
            }
        }
    }
    // every item's tiles map back to it: a block derives its tile from
    // t - tile0, which must lie in 0 .. its tile count - 1
    auto owns = [&](const std::vector<int> &map, int q, int t0, int64_t cnt) {
        if (t0 < 0 || cnt < 0 || (int64_t)t0 + cnt > (int64_t)map.size()) return false;
        for (int64_t t = t0; t < t0 + cnt; ++t)
            if (map[(size_t)t] != q) return false;
        return true;
    };
    for (int q = 0; q < n; ++q) {
        const KBatchCase &K = kc[(size_t)q];
        bool ok = owns(ktile, q, K.tile0, K.nb);
        if (sweeps > 0)
            for (int w = 0; w < 2; ++w) {
                const JorFirstItem &f = first[w][(size_t)q];
                const JorBatchItem &m = mid[w][(size_t)q], &l = last[w][(size_t)q];
                ok = ok && owns(ftile, q, f.tile0, (int64_t)f.ntile * f.nbz) &&
                     owns(jtile, q, m.tile0, (int64_t)m.ntxy * m.nbz) && owns(jtile, q, l.tile0, (int64_t)l.ntxy * l.nbz);
            }
        if (!ok) return fail(IH_EINVAL, "batched step: tile table of case %d is inconsistent", q);
    }
    b->kntiles = (int)ktile.size();
    b->pfirst_ntiles = (int)ftile.size();
    b->pjor_ntiles = (int)jtile.size();
    if ((rc = upload_table(&b->kcases, kc)) || (rc = upload_table(&b->ktile, ktile))) return rc;
    if (sweeps > 0) {
        for (int w = 0; w < 2; ++w) {
            if ((rc = upload_table(&b->pfirst[w], first[w])) || (rc = upload_table(&b->pmid[w], mid[w])) ||
                (rc = upload_table(&b->plast[w], last[w])))
                return rc;
        }
        if ((rc = upload_table(&b->pfirst_tile, ftile)) || (rc = upload_table(&b->pjor_tile, jtile))) return rc;
    }
    HIPCHK(hipMalloc(&b->dflags, sizeof(int) * KF_NSLOTS * (size_t)n));
    HIPCHK(hipMalloc(&b->dscal, sizeof(double) * KS_NSLOTS * (size_t)n));
    HIPCHK(hipHostMalloc(&b->hflags, sizeof(int) * KF_NSLOTS * (size_t)n, hipHostMallocDefault));
    HIPCHK(hipHostMalloc(&b->hscal, sizeof(double) * KS_NSLOTS * (size_t)n, hipHostMallocDefault));
    b->k_sweeps = sweeps;
    b->k_sig = sig;
    return IH_OK;
}

// z = M v for every case (v = pk -> Z1 with w = 0, v = rr -> Z2 with w = 1): the
// FP32 sweeps of launch_precond, batched; first_done: the update that produced v
// has already formed the first sweep
void batch_precond(ih_batch *b, int w, int sweeps, float omega, bool first_done) {
    hipStream_t s = b->stream;
    if (!first_done)
        launch_jor_f32_first_batch(s, b->pfirst[w], b->pfirst_tile, b->pfirst_ntiles, omega, sweeps == 1);
    if (sweeps < 2) return;
    for (int q = 1; q < sweeps - 1; ++q)
        launch_jor_batch_f32(s, b->pmid[(q - 1) & 1], b->pjor_tile, b->pjor_ntiles, omega, false, b->k_px);
    launch_jor_batch_f32(s, b->plast[w], b->pjor_tile, b->pjor_ntiles, omega, true, b->k_px);
}

// one BiCGstab iteration of every case (launch_krylov_iteration's sequence)
void batch_iteration(ih_batch *b, int sweeps, float omega) {
    hipStream_t s = b->stream;
    const bool pre = sweeps > 0;
    // two or more sweeps: the p and s updates also form M's first sweep
    const bool fuse = sweeps >= 2;
    launch_p_update_batch(s, b->kcases, b->ktile, b->kntiles, fuse, omega);
    if (pre) batch_precond(b, 0, sweeps, omega, fuse);
    launch_apply_batch(s, b->kcases, b->ktile, b->kntiles, KB_APPLY_W, pre, b->k_pz, b->k_nonpz);
    launch_s_update_batch(s, b->kcases, b->ktile, b->kntiles, fuse, omega);
    if (pre) {
        batch_precond(b, 1, sweeps, omega, fuse);
        launch_apply_batch(s, b->kcases, b->ktile, b->kntiles, KB_APPLY_PAIR, pre, b->k_pz, b->k_nonpz);
    } else {
        launch_apply_batch(s, b->kcases, b->ktile, b->kntiles, KB_APPLY_SELF, pre, b->k_pz, b->k_nonpz);
    }
    launch_x_update_batch(s, b->kcases, b->ktile, b->kntiles, pre);
    launch_ctl_end_batch(s, b->kcases, (int)b->cases.size());
}

// BiCGstab_orig (preconditioner 0, or right-preconditioned by `sweeps` FP32 JOR
// sweeps) of RES for PRE in every case, from PRE as the guess; the per-case
// statistics into bst[]
int batch_bicgstab(ih_batch *b, const ih_step_params *p, int sweeps, std::vector<ih_bicgstab_stats> &bst) {
    hipStream_t s = b->stream;
    const int n = (int)b->cases.size();
    const float omega = (float)p->jor_omega;
    int rc;
    if ((rc = batch_step_tables(b, sweeps))) return rc;
    for (ih_case *c : b->cases) {
        if (sweeps > 0) {
            launch_fill_f32(s, c->f32[0], 0.0f, c->g.fe);
            launch_fill_f32(s, c->f32[2], 0.0f, c->g.fe);
        }
        // apply_div_grad's ghost refresh of phi
        for (int m = 1; m <= 3; ++m)
            if ((rc = exch(c, m, 0, F(c, IH_PRE)))) return rc;
    }
    // rr = A phi, then rr = bb - rr, rh = rr, eps = tol * max|bb| and the first head
    launch_apply_batch(s, b->kcases, b->ktile, b->kntiles, KB_APPLY_PHI, false, b->k_pz, b->k_nonpz);
    launch_krylov_init_batch(s, b->kcases, n, b->ktile, b->kntiles, p->bicg_rel_tol, p->bicg_max_it);
    LAUNCHCHK();
    // iterations enqueued per convergence check, replayed from a captured graph
    const int iters = 8;
    if (!b->kexec || b->kg_iters != iters || b->kg_sweeps != sweeps || b->kg_omega != omega) {
        if (b->kexec) (void)hipGraphExecDestroy(b->kexec);
        if (b->kgraph) (void)hipGraphDestroy(b->kgraph);
        b->kexec = nullptr;
        b->kgraph = nullptr;
        HIPCHK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        for (int q = 0; q < iters; ++q) batch_iteration(b, sweeps, omega);
        hipGraph_t gr = nullptr;
        const hipError_t e = hipStreamEndCapture(s, &gr);
        if (e != hipSuccess) return fail(IH_EHIP, "graph capture: %s", hipGetErrorString(e));
        b->kgraph = gr;
        HIPCHK(hipGraphInstantiate(&b->kexec, b->kgraph, nullptr, nullptr, 0));
        b->kg_iters = iters;
        b->kg_sweeps = sweeps;
        b->kg_omega = omega;
    }
    for (;;) {
        HIPCHK(hipGraphLaunch(b->kexec, s));
        launch_gather_batch(s, b->kcases, n, b->dflags, b->dscal);
        LAUNCHCHK();
        HIPCHK(hipMemcpyAsync(b->hflags, b->dflags, sizeof(int) * KF_NSLOTS * (size_t)n, hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        bool all = true;
        for (int q = 0; q < n; ++q) {
            const int *f = b->hflags + (size_t)q * KF_NSLOTS;
            if (f[KF_DONE]) continue;
            all = false;
            if (f[KF_RESET]) {
                // rhr == 0: new reference vector rh = rr for that case; its next
                // iteration is in the next replay
                ih_case *c = b->cases[(size_t)q];
                KrylovPtrs k{F(c, IH_RR), F(c, IH_RH), F(c, IH_PK), F(c, IH_AP), F(c, IH_AR), F(c, IH_PRE),
                             c->part, c->part2, c->ksc, c->kfl};
                launch_krylov_resume(s, c->g, pbox(c->g), k);
                LAUNCHCHK();
            }
        }
        if (all) break;
    }
    launch_gather_batch(s, b->kcases, n, b->dflags, b->dscal);
    HIPCHK(hipMemcpyAsync(b->hflags, b->dflags, sizeof(int) * KF_NSLOTS * (size_t)n, hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(b->hscal, b->dscal, sizeof(double) * KS_NSLOTS * (size_t)n, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    bst.assign((size_t)n, ih_bicgstab_stats{});
    for (int q = 0; q < n; ++q) {
        const int *f = b->hflags + (size_t)q * KF_NSLOTS;
        const double *sc = b->hscal + (size_t)q * KS_NSLOTS;
        bst[(size_t)q].iterations = f[KF_COUNTER];
        bst[(size_t)q].norm_inf_initial = sc[KS_NORM0];
        bst[(size_t)q].norm_inf_final = sc[KS_NORM];
        bst[(size_t)q].exit_reason = f[KF_EXIT];
    }
    return IH_OK;
}

// poisson_solve for every case of the batch (the FP32 or FP64 JOR sweeps and the
// plain or FP32-preconditioned BiCGstab, with the plain FP64 redo of a
// preconditioned solve that stops abnormally)
int batch_poisson(ih_batch *b, const ih_step_params *p, std::vector<ih_bicgstab_stats> &bst) {
    hipStream_t s = b->stream;
    int rc;
    if (p->jor_sweeps > 0 && (rc = batch_jor_enqueue(b, p->jor_sweeps, p->jor_omega, p->jor_fp32 ? 1 : 0, 1)))
        return rc;
    const int sweeps = p->precond_sweeps;
    if (sweeps > 0)
        for (ih_case *c : b->cases)
            HIPCHK(hipMemcpyAsync(F(c, IH_COMP), F(c, IH_PRE), sizeof(double) * (size_t)c->g.fe,
                                  hipMemcpyDeviceToDevice, s));
    if ((rc = batch_bicgstab(b, p, sweeps, bst))) return rc;
    if (sweeps > 0) {
        for (size_t q = 0; q < b->cases.size(); ++q) {
            if (bst[q].exit_reason == 0) continue;
            ih_case *c = b->cases[q];
            const Box pb = pbox(c->g);
            const int ss[3] = {pb.lo[0], pb.lo[1], pb.lo[2]}, nn[3] = {pb.hi[0], pb.hi[1], pb.hi[2]};
            const double eps = b->hscal[q * KS_NSLOTS + KS_EPS];
            HIPCHK(hipMemcpyAsync(F(c, IH_PRE), F(c, IH_COMP), sizeof(double) * (size_t)c->g.fe,
                                  hipMemcpyDeviceToDevice, s));
            const int mixed_its = bst[q].iterations;
            if ((rc = ih_bicgstab(c, eps, p->bicg_max_it, ss, nn, IH_RES, IH_PRE, &bst[q]))) return rc;
            bst[q].iterations += mixed_its;
        }
    }
    return IH_OK;
}

// every case's work on the batch's stream for the duration of a batched step
struct StreamSwap {
    ih_batch *b;
    std::vector<hipStream_t> saved;
    explicit StreamSwap(ih_batch *bb) : b(bb) {
        for (ih_case *c : b->cases) {
            saved.push_back(c->stream);
            c->stream = b->stream;
        }
    }
    ~StreamSwap() {
        for (size_t q = 0; q < b->cases.size(); ++q) b->cases[q]->stream = saved[q];
    }
};

}  // namespace

extern "C" {

int ih_batch_timestep(ih_batch *b, const ih_step_params *p, double time, ih_step_stats *st) {
    if (!b || !p) return fail(IH_EINVAL, "bad argument");
    int rc;
    for (ih_case *c : b->cases)
        if ((rc = step_check(c, p))) return rc;
    if (p->precond_poisson != 0) return fail(IH_EINVAL, "the batched step has no multigrid preconditioner");
    if (p->pressure_bc) return fail(IH_EINVAL, "the batched step has no Dirichlet pressure in the Jacobi sweeps");
    if (p->precond_sweeps < 0) return fail(IH_EINVAL, "bad preconditioner sweep count");
    if (p->precond_sweeps > 0 && p->precond_fp64)
        return fail(IH_EINVAL, "the batched step's preconditioner sweeps are FP32");
    HIPCHK(hipSetDevice(b->device));
    for (ih_case *c : b->cases) HIPCHK(hipStreamSynchronize(c->stream));
    StreamSwap guard(b);
    const int n = (int)b->cases.size();
    std::vector<ih_bicgstab_stats> bst;
    for (int sub = 0; sub < 3; ++sub) {
        for (ih_case *c : b->cases)
            if ((rc = step_pre(c, p, time, sub))) return rc;
        if ((rc = batch_poisson(b, p, bst))) return rc;
        if (st)
            for (int q = 0; q < n; ++q) {
                st[q].bicg_iterations[sub] = bst[(size_t)q].iterations;
                st[q].bicg_final[sub] = bst[(size_t)q].norm_inf_final;
            }
        for (ih_case *c : b->cases)
            if ((rc = step_post(c, p, sub))) return rc;
    }
    for (int q = 0; q < n; ++q)
        if ((rc = step_end(b->cases[(size_t)q], p, time, st ? &st[q] : nullptr))) return rc;
    HIPCHK(hipStreamSynchronize(b->stream));
    return IH_OK;
}

}  // extern "C"
