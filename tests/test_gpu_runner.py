"""Runner.step (main.mojo:183-191: policy forward, sample, physics, lidar) on
the GPU against the oracle: with the float32 policy the whole trajectory is
bit-identical, so crash/respawn/progress events agree over 200 steps."""
import numpy as np
import pytest

import oracle as O
import warporacer3d_amd as R
from helpers import track

pytestmark = pytest.mark.gpu


def test_runner_200_steps_bit_identical(dev):
    print("runner 200 steps ramp10 h256 b1081", flush=True)
    O.set_threads(8)
    n, beams, hidden = 64, 1081, 256
    cfg = R.Config(beams=beams, hidden=hidden, heading_noise=0.1)
    tr = track("ramp10")
    run = R.Runner(tr, n, cfg, seed=42, device=dev)
    osim = O.Sim(tr, n, 42, beams=beams, heading_noise=0.1)
    osim.reset()
    opol = O.Policy(42, beams, hidden)
    logp, value = np.zeros(n, np.float32), np.zeros(n, np.float32)
    events = 0
    for t in range(200):
        run.step(False, t)
        _, _, out = opol.forward(osim.obs)
        osim.sample(out, opol.log_std(), osim.actions, logp, value, False, t)
        osim.step()
        d = run.sim.done.read()
        assert np.array_equal(d, osim.done), f"events differ at step {t}"
        events += int((d != 0).sum())
    assert np.array_equal(run.sim.state.read().reshape(11, n), osim.state)
    assert np.array_equal(run.sim.obs.read().reshape(beams + 5, n), osim.obs)
    assert events > 0


def test_benchmark_protocol_runs(dev):
    run = R.Runner(track("flat"), 256, R.Config(), device=dev)
    ms = run.benchmark(20)
    assert 0 < ms < 1000
    assert np.isfinite(run.sim.obs.read()).all()
