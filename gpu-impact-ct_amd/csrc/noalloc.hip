// ih_dot_noalloc: the fixed-tree dot product of ih_dot for callers that cannot
// allocate device memory (code under HIP graph capture, tight serving loops,
// harnesses that forbid hipMalloc).  ih_dot keeps its block partials in the
// case's partial buffer, which it grows with hipMalloc when a box needs more
// partials than it holds; this entry point puts the partials and the final sum
// in `scratch`, a file-scope device array of fixed size, and allocates nothing.
//
// `scratch` is one array per device for the whole process: every call on a
// device shares it and it keeps the last call's contents.  Calls that overlap on
// one device (from several host threads or streams) race on it.  Every partial a
// call reads is written by the same call's partial kernel first, so nothing has
// to be cleared between calls.
#include <hip/hip_runtime.h>

#include <mutex>

#include "ih_host.h"

namespace {

using namespace ih;

// SCRATCH_DOUBLES doubles (2 MiB a device): the partials of boxes of up to
// SCRATCH_MAX_PARTIALS reduction tiles (16 x 16 columns x 32-plane chunks, e.g.
// 2048 x 1024 x 992 points), then the final sum at SCRATCH_RESULT
constexpr int64_t SCRATCH_DOUBLES = 262144;
constexpr int64_t SCRATCH_RESULT = SCRATCH_DOUBLES - 32;
constexpr int64_t SCRATCH_MAX_PARTIALS = SCRATCH_RESULT;
static __device__ __align__(256) double scratch[SCRATCH_DOUBLES];

// The device address of `scratch` on each device, looked up once per device
// with that device current.  In host code the name `scratch` is only a host-side
// stand-in: the kernels are always given the address looked up here.
constexpr int MAX_DEVICES = 64;
double *g_scratch[MAX_DEVICES] = {};
std::mutex g_mu;

int scratch_address(int device, double **out) {
    if (device < 0 || device >= MAX_DEVICES) return set_fail(IH_EINVAL, "device index out of range");
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g_scratch[device]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) return set_fail(IH_EHIP, "hipGetSymbolAddress(scratch) failed");
        g_scratch[device] = static_cast<double *>(p);
    }
    *out = g_scratch[device];
    return IH_OK;
}

}  // namespace

extern "C" int ih_dot_noalloc(ih_case *c, const int ss[3], const int nn[3], ih_field a, ih_field b, double *out) {
    if (!c || !ss || !nn || !out || (int)a < 0 || (int)a >= IH_NFIELDS || (int)b < 0 || (int)b >= IH_NFIELDS)
        return set_fail(IH_EINVAL, "bad argument");
    Box bx;
    int rc = box_from(c, ss, nn, bx);
    if (rc) return rc;
    const int64_t nb = red_blocks(bx);
    if (nb > SCRATCH_MAX_PARTIALS)
        return set_fail(IH_EINVAL,
                        "box needs more reduction partials than the no-allocation scratch holds (262112)");
    if (hipSetDevice(c->device) != hipSuccess) return set_fail(IH_EHIP, "hipSetDevice failed");
    double *sp = nullptr;
    if ((rc = scratch_address(c->device, &sp))) return rc;
    launch_dot_partial(c->stream, c->g, bx, c->fields[a], c->fields[b], sp);
    launch_final_sum(c->stream, sp, nb, sp + SCRATCH_RESULT);
    if (hipGetLastError() != hipSuccess) return set_fail(IH_EHIP, "kernel launch failed");
    // through the case's pinned scalar mirror (allocated with the case)
    if (hipMemcpyAsync(c->hscal + HSCAL_NOALLOC, sp + SCRATCH_RESULT, sizeof(double), hipMemcpyDeviceToHost,
                       c->stream) != hipSuccess ||
        hipStreamSynchronize(c->stream) != hipSuccess)
        return set_fail(IH_EHIP, "reading the result failed");
    *out = c->hscal[HSCAL_NOALLOC];
    return IH_OK;
}
