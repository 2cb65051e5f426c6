// Several sims stepped together: one launch per stage for all of them.
//
// The reference steps one Sim per Device.run (racer/sim.mojo:314-321); a
// batch holds up to RACER_BATCH_MAX sims of any track and car count with the
// same beam count, and runs the policy forward with sampling, the physics and
// the lidar of all of them as one launch each, every kernel finding its
// instance from prefix sums of its work items. Each instance keeps its own
// buffers, seed and track, so its results are bitwise those of stepping it
// alone; the batch only changes how the work is launched.
#include "racer_internal.h"

extern "C" int racer_batch_create(racer_device* dev, racer_sim** sims, int count, const racer_config* cfg,
                                  racer_batch** out) {
    racer_clear_error();
    *out = nullptr;
    RACER_REQUIRE(dev && sims, "null argument");
    RACER_REQUIRE(count >= 1 && count <= RACER_BATCH_MAX, "a batch holds 1 to 16 sims");
    racer_config def;
    racer_config_default(&def);
    if (!cfg) cfg = &def;
    for (int k = 0; k < count; ++k) {
        RACER_REQUIRE(sims[k] && sims[k]->dev == dev, "every sim of a batch must live on the batch's device");
        RACER_REQUIRE(sims[k]->beams == sims[0]->beams, "every sim of a batch must have the same beam count");
        RACER_REQUIRE(sims[k]->beams == cfg->beams, "the configuration's beam count differs from the sims'");
        for (int j = 0; j < k; ++j) RACER_REQUIRE(sims[j] != sims[k], "a sim appears twice in the batch");
    }
    auto* b = new racer_batch();
    b->dev = dev;
    racer_device_retain(dev);
    b->beams = sims[0]->beams;
    b->hidden = cfg->hidden;
    b->car_start.push_back(0);
    for (int k = 0; k < count; ++k) {
        b->sims.push_back(sims[k]);
        b->cars += sims[k]->n;
        b->car_start.push_back(b->cars);
        racer_acts* a = nullptr;
        if (racer_acts_create(dev, sims[k]->n, cfg, &a)) {
            racer_batch_destroy(b);
            return 1;
        }
        b->acts.push_back(a);
    }
    if (racer_alloc(dev, (size_t)(2 * b->cars), 0.0f, &b->scratch)) {
        racer_batch_destroy(b);
        return 1;
    }
    if (hipMalloc(&b->counter, 2 * sizeof(unsigned long long)) != hipSuccess) {
        racer_set_error("batch counter allocation failed");
        racer_batch_destroy(b);
        return 1;
    }
    *out = b;
    return 0;
}

extern "C" void racer_batch_destroy(racer_batch* b) {
    if (!b) return;
    if (b->dev) (void)hipStreamSynchronize(b->dev->stream);
    for (auto* a : b->acts) racer_acts_destroy(a);
    racer_free(b->scratch);
    if (b->counter) (void)hipFree(b->counter);
    racer_device* dev = b->dev;
    delete b;
    racer_device_release(dev);
}

extern "C" int racer_batch_count(const racer_batch* b) { return (int)b->sims.size(); }
extern "C" int64_t racer_batch_cars(const racer_batch* b) { return b->cars; }
extern "C" racer_buf* racer_batch_scratch(racer_batch* b) { return b->scratch; }

extern "C" int racer_batch_step(racer_batch* b) {
    racer_clear_error();
    return racer_launch_batch_step(b);
}

// Runner.step (main.mojo:183-191) for every instance: forward and sample (one
// fused launch when the policy runs the fused bf16 forward, else one forward
// per instance), then the batched physics and lidar.
int racer_batch_runner_enqueue(racer_batch* b, racer_policy* p, int deterministic, int64_t t) {
    RACER_REQUIRE(p->dev == b->dev, "the policy lives on another device");
    RACER_REQUIRE(p->beams == b->beams && p->hidden == b->hidden, "batch and policy shapes differ");
    const int count = (int)b->sims.size();
    if (p->precision == 1 && p->hidden <= 256 && p->frag0) {
        if (racer_policy_refresh(p)) return 1;
        racer_policy_table tab{};
        tab.count = count;
        tab.start[0] = 0;
        for (int k = 0; k < count; ++k) {
            const racer_sim* sim = b->sims[k];
            const int64_t n = sim->n;
            racer_policy_slot& q = tab.slot[k];
            q.x = sim->obs->ptr;
            q.ldx = n;
            q.cols = n;
            q.h1 = b->acts[k]->h1->ptr;
            q.h2 = b->acts[k]->h2->ptr;
            q.o = b->acts[k]->o->ptr;
            q.act = sim->actions->ptr;
            q.act_cols = n;
            q.logp = b->scratch->ptr + 2 * b->car_start[k];
            q.value = q.logp + n;
            q.seed = sim->seed;
            tab.start[k + 1] = tab.start[k] + (n + 15) / 16;
        }
        if (racer_policy_fused_batch(p, tab, deterministic, t)) return 1;
    } else {
        for (int k = 0; k < count; ++k) {
            racer_sim* sim = b->sims[k];
            float* logp = b->scratch->ptr + 2 * b->car_start[k];
            if (racer_policy_act(p, sim->obs->ptr, sim->n, b->acts[k], sim->n, sim->actions->ptr, logp,
                                 logp + sim->n, deterministic, t, sim->seed))
                return 1;
        }
    }
    return racer_launch_batch_step(b);
}

extern "C" int racer_batch_runner_step(racer_batch* b, racer_policy* p, int deterministic, int64_t t) {
    racer_clear_error();
    RACER_REQUIRE(b && p, "null argument");
    return racer_batch_runner_enqueue(b, p, deterministic, t);
}

extern "C" int racer_batch_count_iterations(racer_batch* b, int on) {
    b->count_iterations = on != 0;
    return 0;
}
extern "C" int64_t racer_batch_last_iterations(const racer_batch* b) { return b->last_iterations; }

extern "C" int racer_batch_set_waves_per_car(racer_batch* b, int waves) {
    racer_clear_error();
    RACER_REQUIRE(b && (waves == 1 || waves == 2 || waves == 4), "waves per car must be 1, 2 or 4");
    b->waves_per_car = waves;
    return 0;
}
extern "C" int racer_batch_waves_per_car(const racer_batch* b) { return b->waves_per_car; }

extern "C" int racer_batch_set_fused_step(racer_batch* b, int mode) {
    racer_clear_error();
    RACER_REQUIRE(b && mode >= 0 && mode <= 2, "the fused step mode must be 0, 1 or 2");
    b->fused_step = mode;
    return 0;
}

extern "C" int racer_batch_fused_step(const racer_batch* b) {
    int pose_at = 0;
    return b->fused_step && racer_batch_fused_shmem(b, b->waves_per_car, &pose_at) != 0 ? b->fused_step : 0;
}

