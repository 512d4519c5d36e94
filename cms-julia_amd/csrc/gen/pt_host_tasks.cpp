// pt_host_tasks - the reference runner's threading model over the host-buffer
// entry points, without an interpreter in between: n host threads, each with
// its own TrackSOA, each making tuples of one event at a time
// (CAHitNtuplet.produce, ca_hit_ntuplet.jl:10-16) for a fixed wall time, either
// through a context of its own (pt_make_tuples, the per-task scratch of
// ca_hit_ntuplet_generator.jl:9) or, with max_batch > 0, through one batcher
// that all threads share (pt_batcher_make_tuples).
//
//   pt_host_tasks <threads> <seconds> [events=100] [first_seed=1] [max_batch=0]
//
// The events are PU50-like (1500 prompt helices from 50 vertices, pT power law
// of index 2.5 above 0.3 GeV, 5% noise hits), from the seeded generator.
// Prints one JSON line.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <thread>
#include <vector>

#include "pixeltrack.h"
#include "pixeltrack_gen.h"

namespace {

constexpr double kBField = 0.0114256972711507;  // bf of CAHitNtuplet (1/field in GeV)
constexpr uint32_t kModules = 1856;

struct TrackBuffers {
    std::vector<int32_t> quality;
    std::vector<float> chi2, eta, pt, state, cov;
    std::vector<uint32_t> hit_off, det_off;
    std::vector<uint16_t> hit_bins, det_bins;
    pt_track_soa view{};
    TrackBuffers()
        : quality(PT_MAX_TUPLES), chi2(PT_MAX_TUPLES), eta(PT_MAX_TUPLES), pt(PT_MAX_TUPLES),
          state(5 * PT_MAX_TUPLES), cov(15 * PT_MAX_TUPLES), hit_off(PT_MAX_TUPLES + 1), det_off(PT_MAX_TUPLES + 1),
          hit_bins(PT_MAX_TUPLE_HITS), det_bins(PT_MAX_TUPLE_HITS) {
        view.quality = quality.data(), view.chi2 = chi2.data(), view.eta = eta.data(), view.pt = pt.data();
        view.state = state.data(), view.covariance = cov.data();
        view.hit_off = hit_off.data(), view.hit_bins = hit_bins.data();
        view.det_off = det_off.data(), view.det_bins = det_bins.data();
    }
};

pt_hits_view view_of(const ptg_event *e) {
    pt_hits_view v{};
    v.xg = e->xg, v.yg = e->yg, v.zg = e->zg, v.rg = e->rg, v.xerr = e->xerr, v.yerr = e->yerr;
    v.iphi = e->iphi, v.ysize = e->ysize, v.det_ind = e->det_ind;
    v.hist_off = e->hist_off, v.hist_bins = e->hist_bins;
    v.n_hits = e->n_hits;
    std::memcpy(v.layer_start, e->layer_start, sizeof(v.layer_start));
    return v;
}

int fail(const char *what) {
    std::fprintf(stderr, "pt_host_tasks: %s: %s\n", what, pt_last_error());
    return 1;
}

}  // namespace

int main(int argc, char **argv) {
    if (argc < 3) {
        std::fprintf(stderr, "usage: %s <threads> <seconds> [events=100] [first_seed=1] [max_batch=0]\n", argv[0]);
        return 2;
    }
    const int n_threads = std::atoi(argv[1]);
    const double seconds = std::atof(argv[2]);
    const int n_events = argc > 3 ? std::atoi(argv[3]) : 100;
    const int seed0 = argc > 4 ? std::atoi(argv[4]) : 1;
    const int max_batch = argc > 5 ? std::atoi(argv[5]) : 0;
    if (n_threads < 1 || n_threads > 256 || seconds <= 0.0 || n_events < 1 || max_batch < 0 || max_batch > 64) {
        std::fprintf(stderr, "pt_host_tasks: bad arguments\n");
        return 2;
    }

    std::vector<ptg_event *> events(n_events, nullptr);
    std::vector<pt_hits_view> views(n_events);
    for (int i = 0; i < n_events; ++i) {
        ptg_config c;
        ptg_config_default(&c);
        c.n_helices = 1500, c.n_vertices = 50, c.noise_frac = 0.05;
        c.pt_mode = PTG_PT_POWERLAW, c.pt_min = 0.3, c.pt_max = 0.0, c.pt_index = 2.5;
        c.seed = (uint64_t)(seed0 + i);
        if (ptg_generate_event(&c, &events[i]) != 0) {
            std::fprintf(stderr, "pt_host_tasks: event generation failed\n");
            return 1;
        }
        views[i] = view_of(events[i]);
    }
    std::vector<float> rot(9 * kModules);
    if (ptg_frames(rot.data(), kModules) != 0) {
        std::fprintf(stderr, "pt_host_tasks: module frames failed\n");
        return 1;
    }

    pt_params params;
    pt_params_default(&params);
    std::vector<pt_ctx *> ctxs(max_batch ? 0 : n_threads, nullptr);
    pt_batcher *batcher = nullptr;
    std::vector<TrackBuffers> outs(n_threads);
    if (max_batch && pt_batcher_create(&batcher, &params, 0, (uint32_t)max_batch, rot.data(), kModules) != PT_OK)
        return fail("pt_batcher_create");
    auto make = [&](int t, const pt_hits_view *v, pt_counters *cnt) {
        return batcher ? pt_batcher_make_tuples(batcher, v, kBField, &outs[t].view, 0, cnt)
                       : pt_make_tuples(ctxs[t], v, kBField, &outs[t].view, 0, cnt);
    };
    for (int t = 0; t < n_threads; ++t) {
        if (!batcher) {
            if (pt_ctx_create(&ctxs[t], &params, 0, 1) != PT_OK) return fail("pt_ctx_create");
            if (pt_ctx_set_frames(ctxs[t], rot.data(), kModules) != PT_OK) return fail("pt_ctx_set_frames");
        }
        pt_counters cnt;
        // first call: staging allocation and code loading, outside the timing
        if (make(t, &views[0], &cnt) != PT_OK) return fail("make tuples");
    }

    std::atomic<bool> stop{false};
    std::atomic<int> errors{0};
    std::vector<long> done(n_threads, 0);
    std::vector<std::thread> pool;
    const auto t0 = std::chrono::steady_clock::now();
    for (int t = 0; t < n_threads; ++t)
        pool.emplace_back([&, t] {
            pt_counters cnt;
            for (long k = t; !stop.load(std::memory_order_relaxed); k += n_threads) {
                if (make(t, &views[k % n_events], &cnt) != PT_OK) {
                    std::fprintf(stderr, "pt_host_tasks: make tuples: %s\n", pt_last_error());
                    errors.fetch_add(1);
                    return;
                }
                ++done[t];
            }
        });
    std::this_thread::sleep_for(std::chrono::duration<double>(seconds));
    stop.store(true);
    for (auto &th : pool) th.join();
    const double wall = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();

    long total = 0;
    for (long d : done) total += d;
    for (auto *c : ctxs) pt_ctx_destroy(c);
    pt_batcher_destroy(batcher);
    for (auto *e : events) ptg_free_event(e);
    if (errors.load()) return 1;
    std::printf("{\"threads\": %d, \"max_batch\": %d, \"events\": %ld, \"seconds\": %.3f, \"events_per_s\": %.1f, "
                "\"events_per_s_per_thread\": %.1f}\n",
                n_threads, max_batch, total, wall, total / wall, total / wall / n_threads);
    return 0;
}
