"""Batched launches (several sims, one launch per stage) against each sim
stepped alone with the same policy: bitwise the same state, observations,
actions, rewards, done codes, log probabilities and values."""
import numpy as np
import pytest

import warporacer3d_amd as R
from helpers import track

pytestmark = pytest.mark.gpu

# uneven instances: 1- and 2-layer tracks, car counts that are not multiples of 64
SPECS = [("flat", 100, 3), ("ramp10", 37, 4), ("bridge8", 256, 5), ("bank10", 64, 6), ("random0", 129, 7)]
SIM_BUFFERS = ("state", "obs", "actions", "reward", "done")


def _instances(dev, cfg):
    sims = []
    for name, n, seed in SPECS:
        s = R.Sim(dev, track(name), n, seed, cfg)
        s.reset()
        sims.append(s)
    return sims


def _assert_same(alone, together, what):
    for k, (a, b) in enumerate(zip(alone, together)):
        for name in SIM_BUFFERS:
            assert np.array_equal(getattr(a, name).read(), getattr(b, name).read()), (what, k, name)


@pytest.mark.parametrize("precision,beams,hidden", [(1, 108, 64), (0, 108, 64), (1, 1081, 256), (2, 64, 32)])
def test_batch_runner_steps_match_separate(dev, precision, beams, hidden):
    print(f"batch runner precision={precision} beams={beams} hidden={hidden}", flush=True)
    cfg = R.Config(beams=beams, hidden=hidden, precision=precision, heading_noise=0.1)
    pol = R.Policy(dev, 42, cfg)
    alone, together = _instances(dev, cfg), _instances(dev, cfg)
    acts = [R.Activations(dev, s.n, cfg) for s in alone]
    scratch = [dev.alloc(2 * s.n) for s in alone]
    batch = R.Batch(dev, together, cfg)
    assert batch.cars == sum(n for _, n, _ in SPECS)
    for t in range(4):
        for s, a, sc in zip(alone, acts, scratch):
            R.runner_step(s, pol, a, sc, False, t)
        batch.runner_step(pol, False, t)
        _assert_same(alone, together, f"step {t}")
        for k, s in enumerate(alone):
            lp, v = batch.logp_value(k)
            sc = scratch[k].read()
            assert np.array_equal(sc[: s.n], lp) and np.array_equal(sc[s.n:], v), (t, k)


def test_batch_step_matches_sim_step(dev):
    print("batch sim step", flush=True)
    cfg = R.Config(beams=200, hidden=16)
    alone, together = _instances(dev, cfg), _instances(dev, cfg)
    batch = R.Batch(dev, together, cfg)
    rng = np.random.default_rng(1)
    for t in range(3):
        for a, b in zip(alone, together):
            act = rng.uniform(-1, 1, (2, a.n)).astype(np.float32)
            a.actions.write(act)
            b.actions.write(act)
            a.step()
        batch.step()
        _assert_same(alone, together, f"step {t}")


def test_batch_graph_replays_and_lidar_waves(dev):
    print("batch graph", flush=True)
    cfg = R.Config(beams=300, hidden=64, precision=1, heading_noise=0.1)
    pol = R.Policy(dev, 42, cfg)
    eager_sims, graphed_sims = _instances(dev, cfg), _instances(dev, cfg)
    eager, graphed = R.Batch(dev, eager_sims, cfg), R.Batch(dev, graphed_sims, cfg)
    eager.set_waves_per_car(1)
    graphed.set_waves_per_car(4)
    graph = graphed.capture(pol, 3)
    for replay in range(3):
        for k in range(3):
            eager.runner_step(pol, False, 10 * replay + k)
        graph.launch(10 * replay)
        dev.synchronize()
        _assert_same(eager_sims, graphed_sims, f"replay {replay}")
        assert np.array_equal(eager.scratch.read(), graphed.scratch.read())


def test_batch_iteration_count_and_autotune(dev):
    print("batch iterations", flush=True)
    cfg = R.Config(beams=128, hidden=16)
    alone, together = _instances(dev, cfg), _instances(dev, cfg)
    batch = R.Batch(dev, together, cfg)
    chosen = batch.autotune()
    assert chosen in (1, 2, 4) and batch.waves_per_car == chosen
    total = 0
    for s in alone:
        s.count_iterations(True)
        s.step()
        s.count_iterations(False)
        total += s.last_iterations
    batch.count_iterations(True)
    batch.step()
    batch.count_iterations(False)
    assert batch.last_iterations == total > 0
    _assert_same(alone, together, "counted step")


@pytest.mark.parametrize("waves", [1, 2])
@pytest.mark.parametrize("mode", [1, 2])
def test_batch_fused_step_matches_separate(dev, mode, waves):
    """The fused batched step (one workgroup per 8 cars, or a persistent grid)
    on 1- and 2-layer tracks with car counts that are not multiples of 8."""
    print(f"batch fused step mode={mode} waves={waves}", flush=True)
    cfg = R.Config(beams=300, hidden=64, precision=0, heading_noise=0.1)
    pol = R.Policy(dev, 42, cfg)
    alone, together = _instances(dev, cfg), _instances(dev, cfg)
    acts = [R.Activations(dev, s.n, cfg) for s in alone]
    scratch = [dev.alloc(2 * s.n) for s in alone]
    batch = R.Batch(dev, together, cfg)
    batch.set_waves_per_car(waves)
    batch.set_fused_step(mode)
    assert batch.fused_step == mode
    for t in range(4):
        for s, a, sc in zip(alone, acts, scratch):
            R.runner_step(s, pol, a, sc, False, t)
        batch.runner_step(pol, False, t)
        _assert_same(alone, together, f"mode {mode} step {t}")
    total = 0
    for s in alone:
        s.count_iterations(True)
        s.step()
        s.count_iterations(False)
        total += s.last_iterations
    batch.count_iterations(True)
    batch.step()
    batch.count_iterations(False)
    assert batch.last_iterations == total > 0
    _assert_same(alone, together, f"mode {mode} counted step")


def test_batch_fused_step_applies_to_one_or_two_waves(dev):
    cfg = R.Config(beams=64, hidden=16)
    batch = R.Batch(dev, _instances(dev, cfg), cfg)
    batch.set_fused_step(2)
    batch.set_waves_per_car(4)
    assert batch.fused_step == 0
    batch.set_waves_per_car(2)
    assert batch.fused_step == 2
    with pytest.raises(R.RacerError, match="mode"):
        batch.set_fused_step(3)


def test_batch_errors(dev):
    cfg = R.Config(beams=32)
    a = R.Sim(dev, track("flat"), 8, 1, cfg)
    b = R.Sim(dev, track("flat"), 8, 2, R.Config(beams=16))
    with pytest.raises(R.RacerError, match="beam count"):
        R.Batch(dev, [a, b], cfg)
    with pytest.raises(R.RacerError, match="twice"):
        R.Batch(dev, [a, a], cfg)
    with pytest.raises(R.RacerError, match="1 to 16"):
        R.Batch(dev, [R.Sim(dev, track("flat"), 4, k, cfg) for k in range(17)], cfg)
    batch = R.Batch(dev, [a], cfg)
    with pytest.raises(R.RacerError, match="shapes differ"):
        batch.runner_step(R.Policy(dev, 1, R.Config(beams=16)), False, 0)
