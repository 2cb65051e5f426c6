// Runner steps and Trainer rollouts captured once as HIP graphs and replayed.
//
// The reference enqueues every launch from the host: three or more launches
// per environment step (main.mojo:183-191) and ROLLOUT of those per PPO
// iteration plus the bootstrap forward, GAE and a sum (racer/ppo.mojo:187-232).
// Capturing the sequence once removes the per-launch host work and lets the
// runtime submit the whole chain at once. The captured kernels are the same
// kernels with the same arguments, so results are bitwise those of the
// enqueued calls; the only argument that changes between replays, the Philox
// offset of the Gaussian sample, is read by the sampling kernels from a
// device scalar that racer_graph_launch writes (with a one-thread kernel, not
// a memset node) before the graph.
#include "racer_internal.h"

struct racer_graph {
    racer_device* dev = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    racer_policy* policy = nullptr;  // its bf16 copies are refreshed before each replay
    racer_trainer* trainer = nullptr;
    int64_t rollout = 0;
};

namespace {

__global__ void k_set_base(int64_t* base, int64_t value) { *base = value; }

int set_base(racer_device* dev, int64_t value) {
    hipLaunchKernelGGL(k_set_base, dim3(1), dim3(1), 0, dev->stream, dev->graph_base, value);
    return racer_launch_check("graph base");
}

// Captures what `body` enqueues on the device's stream into g.
template <class F>
int capture(racer_device* dev, racer_policy* pol, racer_graph* g, F body) {
    RACER_REQUIRE(!dev->profiling, "turn the profiler off before capturing a graph");
    RACER_REQUIRE(!dev->capturing, "a capture is already in progress on this device");
    if (!dev->graph_base) RACER_TRY(hipMalloc(&dev->graph_base, sizeof(int64_t)));
    // weights first, outside the graph: a replay refreshes them itself when stale
    if (pol && racer_policy_refresh(pol)) return 1;
    RACER_TRY(hipStreamSynchronize(dev->stream));
    dev->capturing = true;
    dev->step_base = dev->graph_base;
    hipError_t begun = hipStreamBeginCapture(dev->stream, hipStreamCaptureModeRelaxed);
    if (begun != hipSuccess) {
        dev->capturing = false;
        dev->step_base = nullptr;
        RACER_TRY(begun);
    }
    int failed = body();
    hipGraph_t graph = nullptr;
    hipError_t ended = hipStreamEndCapture(dev->stream, &graph);
    dev->capturing = false;
    dev->step_base = nullptr;
    if (failed) {
        if (graph) (void)hipGraphDestroy(graph);
        return 1;
    }
    RACER_TRY(ended);
    g->graph = graph;
    RACER_TRY(hipGraphInstantiate(&g->exec, graph, nullptr, nullptr, 0));
    return 0;
}

}  // namespace

extern "C" int racer_graph_runner(racer_sim* sim, racer_policy* pol, racer_acts* acts, racer_buf* scratch, int steps,
                                  int deterministic, racer_graph** out) {
    racer_clear_error();
    *out = nullptr;
    RACER_REQUIRE(sim && pol && acts && scratch, "null argument");
    RACER_REQUIRE(steps >= 1, "a graph needs at least one step");
    RACER_REQUIRE(!sim->count_iterations, "turn the lidar iteration counter off before capturing");
    auto* g = new racer_graph();
    g->dev = sim->dev;
    racer_device_retain(g->dev);
    g->policy = pol;
    int rc = capture(sim->dev, pol, g, [&]() -> int {
        for (int t = 0; t < steps; ++t)
            if (racer_runner_step(sim, pol, acts, scratch, deterministic, t)) return 1;
        return 0;
    });
    if (rc) {
        racer_graph_destroy(g);
        return 1;
    }
    *out = g;
    return 0;
}

extern "C" int racer_graph_batch(racer_batch* b, racer_policy* pol, int steps, int deterministic,
                                 racer_graph** out) {
    racer_clear_error();
    *out = nullptr;
    RACER_REQUIRE(b && pol, "null argument");
    RACER_REQUIRE(steps >= 1, "a graph needs at least one step");
    RACER_REQUIRE(!b->count_iterations, "turn the lidar iteration counter off before capturing");
    auto* g = new racer_graph();
    g->dev = b->dev;
    racer_device_retain(g->dev);
    g->policy = pol;
    int rc = capture(b->dev, pol, g, [&]() -> int {
        for (int t = 0; t < steps; ++t)
            if (racer_batch_runner_enqueue(b, pol, deterministic, t)) return 1;
        return 0;
    });
    if (rc) {
        racer_graph_destroy(g);
        return 1;
    }
    *out = g;
    return 0;
}

// Sim steps (physics and lidar, no policy) as one graph, for the autotuners.
int racer_graph_sim_steps(racer_sim* sim, int steps, racer_graph** out) {
    *out = nullptr;
    RACER_REQUIRE(steps >= 1, "a graph needs at least one step");
    RACER_REQUIRE(!sim->count_iterations, "turn the lidar iteration counter off before capturing");
    auto* g = new racer_graph();
    g->dev = sim->dev;
    racer_device_retain(g->dev);
    int rc = capture(sim->dev, nullptr, g, [&]() -> int {
        for (int t = 0; t < steps; ++t)
            if (racer_launch_step(sim, sim->actions->ptr, sim->reward->ptr, sim->done->ptr, sim->obs->ptr, sim->n))
                return 1;
        return 0;
    });
    if (rc) {
        racer_graph_destroy(g);
        return 1;
    }
    *out = g;
    return 0;
}

// Batched sim steps (racer_batch_step) as one graph, for the autotuners.
int racer_graph_batch_steps(racer_batch* b, int steps, racer_graph** out) {
    *out = nullptr;
    RACER_REQUIRE(steps >= 1, "a graph needs at least one step");
    RACER_REQUIRE(!b->count_iterations, "turn the lidar iteration counter off before capturing");
    auto* g = new racer_graph();
    g->dev = b->dev;
    racer_device_retain(g->dev);
    int rc = capture(b->dev, nullptr, g, [&]() -> int {
        for (int t = 0; t < steps; ++t)
            if (racer_launch_batch_step(b)) return 1;
        return 0;
    });
    if (rc) {
        racer_graph_destroy(g);
        return 1;
    }
    *out = g;
    return 0;
}

extern "C" int racer_graph_rollout(racer_trainer* tr, racer_sim* sim, racer_policy* pol, racer_acts* acts,
                                   racer_graph** out) {
    racer_clear_error();
    *out = nullptr;
    RACER_REQUIRE(tr && sim && pol && acts, "null argument");
    RACER_REQUIRE(!sim->count_iterations, "turn the lidar iteration counter off before capturing");
    auto* g = new racer_graph();
    g->dev = tr->dev;
    racer_device_retain(g->dev);
    g->policy = pol;
    g->trainer = tr;
    g->rollout = tr->rollout;
    int rc = capture(tr->dev, pol, g, [&]() -> int { return racer_rollout_enqueue(tr, sim, pol, acts, 0); });
    if (rc) {
        racer_graph_destroy(g);
        return 1;
    }
    *out = g;
    return 0;
}

extern "C" int racer_graph_launch(racer_graph* g, int64_t t0) {
    racer_clear_error();
    RACER_REQUIRE(g && g->exec, "not a graph");
    if (g->policy && racer_policy_refresh(g->policy)) return 1;
    if (set_base(g->dev, t0)) return 1;
    RACER_TRY(hipGraphLaunch(g->exec, g->dev->stream));
    return 0;
}

extern "C" int racer_trainer_rollout_graph(racer_trainer* tr, racer_graph* g, int iteration, float* mean_reward) {
    racer_clear_error();
    RACER_REQUIRE(g && g->trainer == tr, "not a rollout graph of this trainer");
    if (racer_graph_launch(g, (int64_t)iteration * g->rollout)) return 1;
    return racer_trainer_mean_reward(tr, mean_reward);
}

extern "C" void racer_graph_destroy(racer_graph* g) {
    if (!g) return;
    if (g->exec) (void)hipGraphExecDestroy(g->exec);
    if (g->graph) (void)hipGraphDestroy(g->graph);
    racer_device* dev = g->dev;
    delete g;
    racer_device_release(dev);
}
