// Serving loop: a trace of simulation requests (stenosed tubes of different
// sizes and lengths) admitted FIFO by an HBM budget and a cap on concurrent
// simulations, each driven by its own host thread on its own case (and so its
// own stream), so that the work of admitted requests overlaps on the device.
// Built on the public C ABI only.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <thread>
#include <vector>

#include "../../include/impact_hip.h"

namespace {

using Clock = std::chrono::steady_clock;

double seconds_since(Clock::time_point t0) {
    return std::chrono::duration<double>(Clock::now() - t0).count();
}

// Radius of the stenosed tube at axial position x: a cosine constriction of
// length 2*ell centred at x0 whose throat area is (1 - stenosis) of the tube's.
double tube_radius(double x, double R, double stenosis, double x0, double ell) {
    const double dr = 1.0 - std::sqrt(1.0 - stenosis);
    if (std::fabs(x - x0) > ell) return R;
    return R * (1.0 - dr / 2.0 * (1.0 + std::cos(M_PI * (x - x0) / ell)));
}

struct Slot {
    std::thread th;
    std::atomic<int> done{0};
    int index = -1;
    int64_t bytes = 0;
};

void run_request(const ih_serve_config &cfg, const ih_request &rq, ih_request_result &res, Clock::time_point t0) {
    res.start_s = seconds_since(t0);
    res.status = IH_OK;
    res.bicg_iterations = 0;
    ih_grid_desc d{};
    d.version = IH_ABI_VERSION;
    for (int q = 0; q < 3; ++q) d.N[q] = rq.N[q];
    for (int q = 0; q < 6; ++q) d.bc[q] = 1;
    d.L[0] = 2.0;
    d.L[1] = d.L[2] = 1.0;
    d.stretch[0] = 0.0;
    d.stretch[1] = d.stretch[2] = 1.5;
    ih_case *c = nullptr;
    int rc = ih_case_create(&d, cfg.device, &c);
    if (!rc) rc = ih_case_generate_coeffs(c);
    if (!rc) {
        std::vector<double> xu(rq.N[0] + 7), xp(rq.N[0] + 7), ru(rq.N[0] + 7), rp(rq.N[0] + 7);
        rc = ih_case_get_coords(c, 0, 1, xu.data());
        if (!rc) rc = ih_case_get_coords(c, 0, 0, xp.data());
        for (int i = 0; i < rq.N[0] + 7; ++i) {
            ru[i] = tube_radius(xu[i], rq.radius, rq.stenosis, 0.5 * d.L[0], 0.4);
            rp[i] = tube_radius(xp[i], rq.radius, rq.stenosis, 0.5 * d.L[0], 0.4);
        }
        if (!rc) rc = ih_set_tube_mask(c, 0.5, 0.5, ru.data(), rp.data());
    }
    // plug flow inside the tube, at rest outside; no history, zero pressure
    if (!rc) rc = ih_field_fill(c, IH_VEL1, 1.0);
    for (int f : {IH_VEL2, IH_VEL3, IH_NL1, IH_NL2, IH_NL3, IH_PRE})
        if (!rc) rc = ih_field_fill(c, (ih_field)f, 0.0);
    if (!rc) rc = ih_apply_mask(c);
    ih_step_stats st{};
    for (int s = 0; !rc && s < rq.steps; ++s) {
        rc = ih_timestep(c, &cfg.step, s * cfg.step.dtime, &st);
        if (!rc) res.bicg_iterations += st.bicg_iterations[0] + st.bicg_iterations[1] + st.bicg_iterations[2];
    }
    if (!rc) {
        int ss[3], nn[3], r48[48];
        rc = ih_case_get_ranges(c, r48);
        for (int q = 0; q < 3; ++q) {
            ss[q] = r48[q];
            nn[q] = r48[3 + q];
        }
        double a = 0.0, b = 0.0;
        if (!rc) rc = ih_dot(c, ss, nn, IH_VEL1, IH_VEL1, &a);
        if (!rc) rc = ih_dot(c, ss, nn, IH_PRE, IH_PRE, &b);
        res.checksum_vel = a;
        res.checksum_pre = b;
    }
    if (c) ih_case_destroy(c);
    res.status = rc;
    res.finish_s = seconds_since(t0);
}

}  // namespace

extern "C" int ih_serve(const ih_serve_config *cfg, const ih_request *req, int n, ih_request_result *out,
                        ih_serve_summary *sum) {
    if (!cfg || (!req && n > 0) || (!out && n > 0) || n < 0 || cfg->max_active < 1 || cfg->hbm_budget <= 0)
        return IH_EINVAL;
    std::vector<int> order(n);
    for (int q = 0; q < n; ++q) order[q] = q;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return req[a].arrival_s < req[b].arrival_s; });
    std::vector<int64_t> need(n);
    for (int q = 0; q < n; ++q) {
        ih_grid_desc d{};
        d.version = IH_ABI_VERSION;
        for (int k = 0; k < 3; ++k) d.N[k] = req[q].N[k];
        for (int k = 0; k < 6; ++k) d.bc[k] = 1;
        need[q] = ih_grid_bytes(&d);
        out[q] = ih_request_result{};
        out[q].status = IH_OK;
    }
    const double scale = cfg->time_scale > 0.0 ? cfg->time_scale : 1.0;
    std::vector<Slot> slots(cfg->max_active);
    int next = 0, running = 0, finished = 0;
    int64_t used = 0, peak = 0;
    int peak_active = 0;
    const Clock::time_point t0 = Clock::now();
    while (finished < n) {
        // retire
        for (Slot &s : slots)
            if (s.index >= 0 && s.done.load(std::memory_order_acquire)) {
                s.th.join();
                used -= s.bytes;
                s.index = -1;
                --running;
                ++finished;
            }
        // admit, first come first served
        const double now = seconds_since(t0);
        while (next < n) {
            const int q = order[next];
            if (req[q].arrival_s * scale > now) break;
            if (need[q] <= 0 || need[q] > cfg->hbm_budget) {  // can never fit
                out[q].status = IH_ENOMEM;
                out[q].admit_s = out[q].start_s = out[q].finish_s = now;
                out[q].latency_s = now - req[q].arrival_s * scale;
                ++next;
                ++finished;
                continue;
            }
            if (running >= cfg->max_active || used + need[q] > cfg->hbm_budget) break;
            Slot *s = nullptr;
            for (Slot &x : slots)
                if (x.index < 0) {
                    s = &x;
                    break;
                }
            s->index = q;
            s->bytes = need[q];
            s->done.store(0);
            used += need[q];
            peak = std::max(peak, used);
            ++running;
            peak_active = std::max(peak_active, running);
            out[q].admit_s = now;
            ih_request_result *r = &out[q];
            const ih_request *rq = &req[q];
            std::atomic<int> *flag = &s->done;
            s->th = std::thread([cfg, rq, r, t0, flag]() {
                run_request(*cfg, *rq, *r, t0);
                flag->store(1, std::memory_order_release);
            });
            ++next;
        }
        std::this_thread::sleep_for(std::chrono::microseconds(200));
    }
    const double makespan = seconds_since(t0);
    int rc = IH_OK;
    for (int q = 0; q < n; ++q) {
        out[q].latency_s = out[q].finish_s - req[q].arrival_s * scale;
        if (out[q].status != IH_OK) rc = out[q].status;
    }
    if (sum) {
        sum->makespan_s = makespan;
        sum->peak_bytes = peak;
        sum->peak_active = peak_active;
    }
    return rc;
}
