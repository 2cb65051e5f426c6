"""Captured HIP graphs of Runner steps and of Trainer.rollout against the same
calls enqueued one by one. Every replay is checked, with the program's own
eager reads and synchronizations between replays."""
import numpy as np
import pytest

import warporacer3d_amd as R
from helpers import track

pytestmark = pytest.mark.gpu

SIM_BUFFERS = ("state", "obs", "actions", "reward", "done")


def _runner(dev, precision, name="ramp", n=256, beams=108, hidden=64):
    cfg = R.Config(beams=beams, hidden=hidden, precision=precision, heading_noise=0.1)
    return R.Runner(track(name), n, cfg, seed=42, device=dev)


@pytest.mark.parametrize("precision", [0, 1, 2])
def test_runner_graph_replays_match_eager_steps(dev, precision):
    print(f"runner graph precision={precision}", flush=True)
    eager, graphed = _runner(dev, precision), _runner(dev, precision)
    S = 5
    graph = graphed.capture(S)
    t = 0
    for replay in range(3):
        for k in range(S):
            eager.step(False, t + k)
        graph.launch(t)
        dev.synchronize()
        for name in SIM_BUFFERS:
            a = getattr(eager.sim, name).read()
            b = getattr(graphed.sim, name).read()
            assert np.array_equal(a, b), (replay, name)
        assert np.array_equal(eager.scratch.read(), graphed.scratch.read()), replay
        t += S
    # eager steps on the graphed runner continue from the replayed state
    eager.step(False, t)
    graphed.step(False, t)
    for name in SIM_BUFFERS:
        assert np.array_equal(getattr(eager.sim, name).read(), getattr(graphed.sim, name).read()), name


def test_runner_graph_deterministic_and_lidar_schedules(dev):
    print("runner graph deterministic, 1081 beams", flush=True)
    eager, graphed = _runner(dev, 1, beams=1081, hidden=256, n=128), _runner(dev, 1, beams=1081, hidden=256, n=128)
    graphed.sim.set_waves_per_car(4)
    eager.sim.set_waves_per_car(1)
    graph = graphed.capture(3, deterministic=True)
    for replay in range(2):
        for k in range(3):
            eager.step(True, 10 * replay + k)
        graph.launch(10 * replay)
        dev.synchronize()
        for name in SIM_BUFFERS:
            assert np.array_equal(getattr(eager.sim, name).read(), getattr(graphed.sim, name).read()), (replay, name)


def test_capture_refuses_profiler_and_counter(dev):
    run = _runner(dev, 0, n=64)
    dev.profile(True)
    try:
        with pytest.raises(R.RacerError, match="profiler"):
            run.capture(1)
    finally:
        dev.profile(False)
        dev.profile_reset()
    run.sim.count_iterations(True)
    try:
        with pytest.raises(R.RacerError, match="counter"):
            run.capture(1)
    finally:
        run.sim.count_iterations(False)
    run.step(False, 0)  # the device still works after refused captures
    dev.synchronize()


TRAINER_BUFFERS = ("obs", "actions", "logp", "value", "reward", "done", "advantage")


@pytest.mark.parametrize("precision", [0, 1])
def test_rollout_graph_matches_eager_rollout_across_updates(dev, precision):
    print(f"rollout graph precision={precision}", flush=True)
    cfg = R.Config(beams=108, hidden=64, rollout=8, minibatches=4, precision=precision, heading_noise=0.1)
    arr = track("bridge")

    def setup():
        sim = R.Sim(dev, arr, 128, 42, cfg)
        sim.reset()
        pol = R.Policy(dev, 42, cfg)
        acts = R.Activations(dev, 128, cfg)
        tr = R.Trainer(dev, 128, cfg)
        return sim, pol, acts, tr

    ea, ga = setup(), setup()
    graph = ga[3].capture_rollout(ga[0], ga[1], ga[2])
    for iteration in range(3):
        r_e = ea[3].rollout(ea[0], ea[1], ea[2], iteration)
        r_g = ga[3].rollout_graph(graph, iteration)
        assert r_e == r_g, iteration
        for name in TRAINER_BUFFERS:
            assert np.array_equal(getattr(ea[3], name).read(), getattr(ga[3], name).read()), (iteration, name)
        for name in SIM_BUFFERS:
            assert np.array_equal(getattr(ea[0], name).read(), getattr(ga[0], name).read()), (iteration, name)
        # the update changes the weights: the next replay must see them
        kl_e, kl_g = ea[3].update(ea[1]), ga[3].update(ga[1])
        assert kl_e == kl_g
        assert np.array_equal(ea[1].theta.read(), ga[1].theta.read())
