// Serving loop: a trace of simulation requests (stenosed tubes of different
// sizes and lengths) run on a pool of resident case slots.
//
// Set-up (before the first request is timed): the trace's grids are grouped
// into size classes, and each class gets slots -- cases created, coefficients
// generated, multigrid levels built and launch shapes tuned -- one class at a
// time in turn while the pool fits the HBM budget, up to max_active slots or
// the class's request count.  Serving then allocates and frees nothing on the
// device: a request is admitted as soon as a slot of its class is free and
// fewer than max_active run (classes do not share slots, so a request never
// waits behind one of another class); its thread resets the slot's fields,
// builds its mask and advances it on the slot's stream, so the kernels of the
// admitted requests overlap on the device.  Built on the public C ABI only.
#include <cstdio>
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <deque>
#include <map>
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

ih_grid_desc tube_grid(const int N[3]) {
    ih_grid_desc d{};
    d.version = IH_ABI_VERSION;
    for (int q = 0; q < 3; ++q) d.N[q] = N[q];
    for (int q = 0; q < 6; ++q) d.bc[q] = 1;
    d.L[0] = 2.0;
    d.L[1] = d.L[2] = 1.0;
    d.stretch[0] = 0.0;
    d.stretch[1] = d.stretch[2] = 1.5;
    return d;
}

struct Slot {
    ih_case *c = nullptr;
    int cls = -1;
    int64_t bytes = 0;
    std::vector<double> xu, xp;  // x coordinates of the u and pressure grids
    std::thread th;
    std::atomic<int> done{0};
    int index = -1;  // request running on the slot, -1 when free
};

// a slot's case, ready to serve: coefficients, the step's multigrid levels and
// launch shapes (a JOR sweep on the zero fields tunes them without changing them)
int prepare_slot(const ih_serve_config &cfg, const int N[3], Slot &s) {
    ih_grid_desc d = tube_grid(N);
    int rc = ih_case_create(&d, cfg.device, &s.c);
    if (!rc) rc = ih_case_generate_coeffs(s.c);
    s.xu.assign(N[0] + 7, 0.0);
    s.xp.assign(N[0] + 7, 0.0);
    if (!rc) rc = ih_case_get_coords(s.c, 0, 1, s.xu.data());
    if (!rc) rc = ih_case_get_coords(s.c, 0, 0, s.xp.data());
    int r48[48] = {};
    if (!rc) rc = ih_case_get_ranges(s.c, r48);
    if (!rc && cfg.step.jor_sweeps > 0) rc = ih_jor(s.c, 1, cfg.step.jor_omega, r48, r48 + 3, IH_PRE, IH_RES);
    if (!rc && cfg.step.precond_poisson == 1 && cfg.step.precond_sweeps == 0) {
        int G = 0, NN[45];
        rc = ih_mg_levels(s.c, &G, NN);
        // the levels are built on first use: one cycle of zeros builds them
        if (!rc) rc = ih_multigrid_v(s.c, IH_TMP1, IH_TMP2);
    }
    if (!rc) rc = ih_synchronize(s.c);
    if (!rc) s.bytes = ih_case_bytes(s.c);
    return rc;
}

void run_request(const ih_serve_config &cfg, const ih_request &rq, Slot &slot, ih_request_result &res,
                 Clock::time_point t0) {
    res.start_s = seconds_since(t0);
    res.status = IH_OK;
    res.bicg_iterations = 0;
    ih_case *c = slot.c;
    const int n = (int)slot.xu.size();
    std::vector<double> ru(n), rp(n);
    for (int i = 0; i < n; ++i) {
        ru[i] = tube_radius(slot.xu[i], rq.radius, rq.stenosis, 0.5 * 2.0, 0.4);
        rp[i] = tube_radius(slot.xp[i], rq.radius, rq.stenosis, 0.5 * 2.0, 0.4);
    }
    int rc = ih_set_tube_mask(c, 0.5, 0.5, ru.data(), rp.data());
    // plug flow inside the tube, at rest outside; no history, zero pressure:
    // every field a step reads before writing starts as it would in a new case
    if (!rc) rc = ih_field_fill(c, IH_VEL1, 1.0);
    for (int f = IH_VEL2; !rc && f < IH_NFIELDS; ++f) rc = ih_field_fill(c, (ih_field)f, 0.0);
    if (!rc) rc = ih_apply_mask(c);
    ih_step_stats st{};
    for (int s = 0; !rc && s < rq.steps; ++s) {
        rc = ih_timestep(c, &cfg.step, s * cfg.step.dtime, &st);
        if (!rc) res.bicg_iterations += st.bicg_iterations[0] + st.bicg_iterations[1] + st.bicg_iterations[2];
    }
    if (!rc) {
        int r48[48];
        rc = ih_case_get_ranges(c, r48);
        double a = 0.0, b = 0.0;
        if (!rc) rc = ih_dot(c, r48, r48 + 3, IH_VEL1, IH_VEL1, &a);
        if (!rc) rc = ih_dot(c, r48, r48 + 3, IH_PRE, IH_PRE, &b);
        res.checksum_vel = a;
        res.checksum_pre = b;
    }
    res.status = rc;
    res.finish_s = seconds_since(t0);
}

}  // namespace

extern "C" int ih_serve(const ih_serve_config *cfg, const ih_request *req, int n, ih_request_result *out,
                        ih_serve_summary *sum) {
    if (!cfg || (!req && n > 0) || (!out && n > 0) || n < 0 || cfg->max_active < 1 || cfg->hbm_budget <= 0)
        return IH_EINVAL;
    if (sum) *sum = ih_serve_summary{};
    std::vector<int> order(n);
    for (int q = 0; q < n; ++q) order[q] = q;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return req[a].arrival_s < req[b].arrival_s; });
    // size classes in order of first arrival
    std::map<std::vector<int>, int> cls_of;
    std::vector<std::vector<int>> cls_N;
    std::vector<int> cls(n), demand;
    for (int q : order) {
        const std::vector<int> key{req[q].N[0], req[q].N[1], req[q].N[2]};
        auto it = cls_of.find(key);
        if (it == cls_of.end()) {
            it = cls_of.emplace(key, (int)cls_N.size()).first;
            cls_N.push_back(key);
            demand.push_back(0);
        }
        cls[q] = it->second;
        ++demand[it->second];
        out[q] = ih_request_result{};
        out[q].status = IH_OK;
    }
    const int C = (int)cls_N.size();
    std::vector<int64_t> est(C);
    for (int k = 0; k < C; ++k) {
        const ih_grid_desc d = tube_grid(cls_N[k].data());
        est[k] = ih_grid_bytes_step(&d, &cfg->step);
    }
    // the pool: one more slot per class in turn while it fits the budget
    const Clock::time_point ts = Clock::now();
    std::deque<Slot> slots;  // stable addresses: each running request's thread holds its slot
    std::vector<int> have(C, 0);
    int64_t pool = 0;
    int status = IH_OK;
    for (bool grew = true; grew && status == IH_OK;) {
        grew = false;
        for (int k = 0; k < C && status == IH_OK; ++k) {
            if (have[k] >= std::min(demand[k], cfg->max_active) || est[k] <= 0 || pool + est[k] > cfg->hbm_budget)
                continue;
            slots.emplace_back();
            Slot &s = slots.back();
            s.cls = k;
            status = prepare_slot(*cfg, cls_N[k].data(), s);
            if (status) break;
            pool += std::max(s.bytes, est[k]);
            ++have[k];
            grew = true;
        }
    }
    const double setup_s = seconds_since(ts);
    int next = 0, running = 0, finished = 0, peak_active = 0;
    int64_t used = 0, peak = 0;
    double admit_wait_max = 0.0, admit_wait_sum = 0.0;
    const double scale = cfg->time_scale > 0.0 ? cfg->time_scale : 1.0;
    std::vector<int> queue;  // arrived, not yet admitted, in arrival order
    const Clock::time_point t0 = Clock::now();
    while (status == IH_OK && finished < n) {
        // retire
        for (Slot &s : slots)
            if (s.index >= 0 && s.done.load(std::memory_order_acquire)) {
                s.th.join();
                used -= s.bytes;
                s.index = -1;
                --running;
                ++finished;
                if (cfg->progress_every > 0 && finished % cfg->progress_every == 0) {
                    std::fprintf(stderr, "ih_serve: %d/%d requests finished, %.1f s into the trace\n", finished, n,
                                 seconds_since(t0));
                    std::fflush(stderr);
                }
            }
        const double now = seconds_since(t0);
        while (next < n && req[order[next]].arrival_s * scale <= now) {
            const int q = order[next++];
            if (have[cls[q]] == 0) {  // no slot of its size fits the budget
                out[q].status = IH_ENOMEM;
                out[q].admit_s = out[q].start_s = out[q].finish_s = now;
                ++finished;
                continue;
            }
            queue.push_back(q);
        }
        // admit in arrival order; a request whose class has no free slot does not hold back other classes
        for (size_t w = 0; w < queue.size() && running < cfg->max_active;) {
            const int q = queue[w];
            Slot *s = nullptr;
            for (Slot &x : slots)
                if (x.cls == cls[q] && x.index < 0) {
                    s = &x;
                    break;
                }
            if (!s) {
                ++w;
                continue;
            }
            queue.erase(queue.begin() + (long)w);
            s->index = q;
            s->done.store(0);
            used += s->bytes;
            peak = std::max(peak, used);
            ++running;
            peak_active = std::max(peak_active, running);
            out[q].admit_s = now;
            const double wait = now - req[q].arrival_s * scale;
            admit_wait_max = std::max(admit_wait_max, wait);
            admit_wait_sum += wait;
            ih_request_result *r = &out[q];
            const ih_request *rq = &req[q];
            std::atomic<int> *flag = &s->done;
            s->th = std::thread([cfg, rq, s, r, t0, flag]() {
                run_request(*cfg, *rq, *s, *r, t0);
                flag->store(1, std::memory_order_release);
            });
        }
        std::this_thread::sleep_for(std::chrono::microseconds(200));
    }
    const double makespan = seconds_since(t0);
    for (Slot &s : slots)
        if (s.th.joinable()) s.th.join();
    for (Slot &s : slots)
        if (s.c) ih_case_destroy(s.c);
    int rc = status;
    int admitted = 0;
    for (int q = 0; q < n; ++q) {
        out[q].latency_s = out[q].finish_s - req[q].arrival_s * scale;
        if (rc == IH_OK && out[q].status != IH_OK) rc = out[q].status;
        if (out[q].status == IH_OK) ++admitted;
    }
    if (sum) {
        sum->makespan_s = makespan;
        sum->peak_bytes = peak;
        sum->peak_active = peak_active;
        sum->setup_s = setup_s;
        sum->pool_bytes = pool;
        sum->pool_slots = (int)slots.size();
        sum->admit_wait_max_s = admit_wait_max;
        sum->admit_wait_mean_s = admitted ? admit_wait_sum / admitted : 0.0;
    }
    return rc;
}
