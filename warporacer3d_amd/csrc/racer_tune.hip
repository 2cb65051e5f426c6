// Launch-shape autotuners: the lidar's wavefronts per car, the fused or
// two-launch sim step, and the forward kernel of the Runner step.
//
// The reference fixes its launch shapes (racer/device.mojo:56-98). Here every
// candidate computes bitwise the same results, so the choice is made by time
// alone: each candidate's steps are captured as a graph and replayed from the
// same saved sim state (the state, obs, actions, reward and done buffers are
// copied aside once and back before every replay), so all candidates run the
// same cars through the same steps with the caches as a real step leaves them.
// The best of three replays counts; the saved state is put back at the end.
#include <algorithm>
#include <functional>

#include "racer_internal.h"

namespace {

constexpr int TUNE_REPLAYS = 3;

// A device copy of some sims' step buffers.
struct Saved {
    racer_device* dev = nullptr;
    float* copy = nullptr;
    std::vector<racer_buf*> bufs;
    Saved() = default;
    Saved(const Saved&) = delete;
    Saved& operator=(const Saved&) = delete;
    ~Saved() {
        if (copy) (void)hipFree(copy);
    }

    int save(racer_device* d, const std::vector<racer_sim*>& sims) {
        dev = d;
        size_t total = 0;
        for (racer_sim* s : sims)
            for (racer_buf* b : {s->state, s->obs, s->actions, s->reward, s->done}) {
                bufs.push_back(b);
                total += b->len;
            }
        RACER_TRY(hipMalloc(&copy, total * sizeof(float)));
        size_t at = 0;
        for (racer_buf* b : bufs) {
            RACER_TRY(hipMemcpyAsync(copy + at, b->ptr, b->len * sizeof(float), hipMemcpyDeviceToDevice,
                                     dev->stream));
            at += b->len;
        }
        return 0;
    }

    int restore() {
        if (!copy) return 0;
        size_t at = 0;
        for (racer_buf* b : bufs) {
            RACER_TRY(hipMemcpyAsync(b->ptr, copy + at, b->len * sizeof(float), hipMemcpyDeviceToDevice,
                                     dev->stream));
            at += b->len;
        }
        return 0;
    }
};

// The best of TUNE_REPLAYS timed replays of g (milliseconds), after one
// untimed replay; the saved state is restored before each.
int time_replays(racer_graph* g, racer_device* dev, Saved& saved, racer_event_pair& ev, float* best) {
    if (saved.restore() || racer_graph_launch(g, 0)) return 1;
    *best = 1e30f;
    for (int r = 0; r < TUNE_REPLAYS; ++r) {
        if (saved.restore()) return 1;
        RACER_TRY(hipEventRecord(ev.a, dev->stream));
        if (racer_graph_launch(g, 0)) return 1;
        RACER_TRY(hipEventRecord(ev.b, dev->stream));
        RACER_TRY(hipEventSynchronize(ev.b));
        float ms = 0.0f;
        RACER_TRY(hipEventElapsedTime(&ms, ev.a, ev.b));
        *best = std::min(*best, ms);
    }
    return 0;
}

// Times the graph `make` captures, destroying it afterwards.
int time_candidate(racer_device* dev, Saved& saved, racer_event_pair& ev,
                   const std::function<int(racer_graph**)>& make, float* ms) {
    racer_graph* g = nullptr;
    if (make(&g)) return 1;
    const int rc = time_replays(g, dev, saved, ev, ms);
    racer_graph_destroy(g);
    return rc;
}

int tune_preconditions(racer_device* dev) {
    RACER_REQUIRE(!dev->capturing, "cannot autotune while capturing a graph");
    RACER_REQUIRE(!dev->profiling, "turn the profiler off before autotuning");
    return 0;
}

int make_events(racer_event_pair& ev) {
    RACER_TRY(hipEventCreate(&ev.a));
    RACER_TRY(hipEventCreate(&ev.b));
    return 0;
}

// The forward kernels of racer_policy_act worth timing for this policy.
std::vector<int> act_candidates(const racer_policy* p) {
    if (p->precision == 1 && p->hidden <= 256) return {RACER_ACT_FUSED, RACER_ACT_WIDE, RACER_ACT_TILED};
    if (p->precision == 0) return {RACER_ACT_FUSED, RACER_ACT_TILED};
    return {RACER_ACT_AUTO};
}

}  // namespace

extern "C" int racer_sim_autotune(racer_sim* sim, int* chosen) {
    racer_clear_error();
    RACER_REQUIRE(sim, "null argument");
    racer_device* dev = sim->dev;
    if (tune_preconditions(dev)) return 1;
    racer_event_pair ev;
    if (make_events(ev)) return 1;
    const bool counting = sim->count_iterations;
    const int waves0 = sim->waves_per_car;
    const bool fused0 = sim->fused_step, coop0 = sim->lidar_coop;
    sim->count_iterations = false;
    int best_waves = waves0;
    bool best_fused = fused0, best_coop = coop0;
    float best = 1e30f;
    Saved saved;
    auto run = [&]() -> int {
        if (saved.save(dev, {sim})) return 1;
        for (int w : {1, 2, 4})
            for (int shape : {0, 1, 2}) {  // cooperative stores, per-wavefront stores, fused step
                const bool f = shape == 2, c = shape == 0;
                sim->waves_per_car = w;
                sim->fused_step = f;
                sim->lidar_coop = c;
                if (f && !racer_sim_fused_step(sim)) continue;
                float ms = 0.0f;
                if (time_candidate(dev, saved, ev, [&](racer_graph** g) { return racer_graph_sim_steps(sim, 8, g); },
                                   &ms))
                    return 1;
                if (ms < best) {
                    best = ms;
                    best_waves = w;
                    best_fused = f;
                    best_coop = c;
                }
            }
        return 0;
    };
    const int failed = run();
    const int restored = saved.restore();
    sim->count_iterations = counting;
    sim->waves_per_car = failed ? waves0 : best_waves;
    sim->fused_step = failed ? fused0 : best_fused;
    sim->lidar_coop = failed ? coop0 : best_coop;
    if (failed || restored) return 1;
    RACER_TRY(hipStreamSynchronize(dev->stream));
    if (chosen) *chosen = best_waves;
    return 0;
}

extern "C" int racer_runner_autotune(racer_sim* sim, racer_policy* pol, racer_acts* acts, racer_buf* scratch,
                                     int steps, int* waves_per_car, int* fused_step, int* act_kernel) {
    racer_clear_error();
    RACER_REQUIRE(sim && pol && acts && scratch, "null argument");
    RACER_REQUIRE(steps >= 1, "autotuning needs at least one step");
    RACER_REQUIRE(pol->dev == sim->dev, "the policy lives on another device");
    racer_device* dev = sim->dev;
    if (tune_preconditions(dev)) return 1;
    racer_event_pair ev;
    if (make_events(ev)) return 1;
    const bool counting = sim->count_iterations;
    const int waves0 = sim->waves_per_car, kernel0 = pol->act_kernel;
    const bool fused0 = sim->fused_step, coop0 = sim->lidar_coop;
    sim->count_iterations = false;
    int best_waves = waves0, best_kernel = kernel0;
    bool best_fused = fused0, best_coop = coop0;
    float best = 1e30f;
    Saved saved;
    auto run = [&]() -> int {
        if (saved.save(dev, {sim})) return 1;
        for (int k : act_candidates(pol))
            for (int w : {1, 2, 4})
                for (int shape : {0, 1, 2}) {  // cooperative stores, per-wavefront stores, fused step
                    const bool f = shape == 2, c = shape == 0;
                    pol->act_kernel = k;
                    sim->waves_per_car = w;
                    sim->fused_step = f;
                    sim->lidar_coop = c;
                    if (f && !racer_sim_fused_step(sim)) continue;
                    float ms = 0.0f;
                    if (time_candidate(
                            dev, saved, ev,
                            [&](racer_graph** g) { return racer_graph_runner(sim, pol, acts, scratch, steps, 0, g); },
                            &ms))
                        return 1;
                    if (ms < best) {
                        best = ms;
                        best_waves = w;
                        best_fused = f;
                        best_coop = c;
                        best_kernel = k;
                    }
                }
        return 0;
    };
    const int failed = run();
    const int restored = saved.restore();
    sim->count_iterations = counting;
    sim->waves_per_car = failed ? waves0 : best_waves;
    sim->fused_step = failed ? fused0 : best_fused;
    sim->lidar_coop = failed ? coop0 : best_coop;
    pol->act_kernel = failed ? kernel0 : best_kernel;
    if (failed || restored) return 1;
    RACER_TRY(hipStreamSynchronize(dev->stream));
    if (waves_per_car) *waves_per_car = best_waves;
    if (fused_step) *fused_step = best_fused ? 1 : 0;
    if (act_kernel) *act_kernel = best_kernel;
    return 0;
}

extern "C" int racer_batch_autotune(racer_batch* b, int* chosen) {
    racer_clear_error();
    RACER_REQUIRE(b, "null argument");
    racer_device* dev = b->dev;
    if (tune_preconditions(dev)) return 1;
    racer_event_pair ev;
    if (make_events(ev)) return 1;
    const bool counting = b->count_iterations;
    const int waves0 = b->waves_per_car, fused0 = b->fused_step;
    b->count_iterations = false;
    int best_waves = waves0, best_fused = fused0;
    float best = 1e30f;
    Saved saved;
    auto run = [&]() -> int {
        if (saved.save(dev, b->sims)) return 1;
        for (int w : {1, 2, 4})
            for (int f : {0, 1, 2}) {
                int pose_at = 0;
                if (f && !racer_batch_fused_shmem(b, w, &pose_at)) continue;
                b->waves_per_car = w;
                b->fused_step = f;
                float ms = 0.0f;
                if (time_candidate(dev, saved, ev, [&](racer_graph** g) { return racer_graph_batch_steps(b, 8, g); },
                                   &ms))
                    return 1;
                if (ms < best) {
                    best = ms;
                    best_waves = w;
                    best_fused = f;
                }
            }
        return 0;
    };
    const int failed = run();
    const int restored = saved.restore();
    b->count_iterations = counting;
    b->waves_per_car = failed ? waves0 : best_waves;
    b->fused_step = failed ? fused0 : best_fused;
    if (failed || restored) return 1;
    RACER_TRY(hipStreamSynchronize(dev->stream));
    if (chosen) *chosen = best_waves;
    return 0;
}

extern "C" int racer_batch_runner_autotune(racer_batch* b, racer_policy* pol, int steps, int* waves_per_car,
                                           int* fused_step) {
    racer_clear_error();
    RACER_REQUIRE(b && pol, "null argument");
    RACER_REQUIRE(steps >= 1, "autotuning needs at least one step");
    RACER_REQUIRE(pol->dev == b->dev, "the policy lives on another device");
    racer_device* dev = b->dev;
    if (tune_preconditions(dev)) return 1;
    racer_event_pair ev;
    if (make_events(ev)) return 1;
    const bool counting = b->count_iterations;
    const int waves0 = b->waves_per_car, fused0 = b->fused_step, kernel0 = pol->act_kernel;
    b->count_iterations = false;
    int best_waves = waves0, best_fused = fused0, best_kernel = kernel0;
    float best = 1e30f;
    // the batched bf16 forward has one kernel; otherwise each instance runs racer_policy_act
    const bool batched_forward = pol->precision == 1 && pol->hidden <= 256;
    const std::vector<int> kernels = batched_forward ? std::vector<int>{kernel0} : act_candidates(pol);
    Saved saved;
    auto run = [&]() -> int {
        if (saved.save(dev, b->sims)) return 1;
        for (int k : kernels)
            for (int w : {1, 2, 4})
                for (int f : {0, 1, 2}) {
                    int pose_at = 0;
                    if (f && !racer_batch_fused_shmem(b, w, &pose_at)) continue;
                    pol->act_kernel = k;
                    b->waves_per_car = w;
                    b->fused_step = f;
                    float ms = 0.0f;
                    if (time_candidate(dev, saved, ev,
                                       [&](racer_graph** g) { return racer_graph_batch(b, pol, steps, 0, g); }, &ms))
                        return 1;
                    if (ms < best) {
                        best = ms;
                        best_waves = w;
                        best_fused = f;
                        best_kernel = k;
                    }
                }
        return 0;
    };
    const int failed = run();
    const int restored = saved.restore();
    b->count_iterations = counting;
    b->waves_per_car = failed ? waves0 : best_waves;
    b->fused_step = failed ? fused0 : best_fused;
    pol->act_kernel = failed ? kernel0 : best_kernel;
    if (failed || restored) return 1;
    RACER_TRY(hipStreamSynchronize(dev->stream));
    if (waves_per_car) *waves_per_car = best_waves;
    if (fused_step) *fused_step = best_fused;
    return 0;
}

extern "C" int racer_policy_set_act_kernel(racer_policy* pol, int kernel) {
    racer_clear_error();
    RACER_REQUIRE(pol, "null argument");
    RACER_REQUIRE(kernel >= RACER_ACT_AUTO && kernel <= RACER_ACT_TILED,
                  "act kernel must be RACER_ACT_AUTO, _FUSED, _WIDE or _TILED");
    pol->act_kernel = kernel;
    return 0;
}

extern "C" int racer_policy_act_kernel(const racer_policy* pol) { return pol->act_kernel; }
