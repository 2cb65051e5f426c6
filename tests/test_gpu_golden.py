"""The HIP path against the committed golden vectors (tests/golden): the same
seeds and actions give the same bits on the GPU."""
from pathlib import Path

import numpy as np
import pytest

import warporacer3d_amd as R
from fixtures import corridor
from helpers import track

pytestmark = pytest.mark.gpu

GOLDEN = Path(__file__).resolve().parent / "golden"
TRACKS = {"sim_corridor.npz": corridor, "sim_ramp10.npz": lambda: track("ramp10"),
          "sim_bridge8.npz": lambda: track("bridge8")}


@pytest.mark.parametrize("name", sorted(TRACKS))
@pytest.mark.parametrize("waves", [1, 2])
def test_sim_golden(dev, name, waves):
    print(f"golden {name} waves/car {waves}", flush=True)
    g = np.load(GOLDEN / name, allow_pickle=False)
    n, beams = int(g["n"]), int(g["beams"])
    sim = R.Sim(dev, TRACKS[name](), n, int(g["seed"]), R.Config(beams=beams, heading_noise=float(g["noise"])))
    sim.set_waves_per_car(waves)
    sim.reset()
    assert np.array_equal(sim.state.read().reshape(11, n), g["reset_state"])
    assert np.array_equal(sim.obs.read().reshape(beams + 5, n), g["reset_obs"])
    for k in range(g["actions"].shape[0]):
        sim.actions.write(g["actions"][k])
        sim.count_iterations(True)
        sim.step()
        sim.count_iterations(False)
        assert sim.last_iterations == g["iterations"][k]
        for label, shape in (("state", (11, n)), ("obs", (beams + 5, n)), ("reward", (n,)), ("done", (n,))):
            got = getattr(sim, label).read().reshape(shape)
            assert np.array_equal(got, g[label][k]), (name, k, label)


def test_policy_golden(dev):
    g = np.load(GOLDEN / "policy_108x64.npz", allow_pickle=False)
    cfg = R.Config(beams=int(g["beams"]), hidden=int(g["hidden"]))
    pol = R.Policy(dev, int(g["seed"]), cfg)
    assert np.array_equal(pol.theta.read(), g["theta"])
    cols = g["x"].shape[1]
    xb = dev.upload(g["x"])
    acts = R.Activations(dev, cols, cfg)
    pol.forward(R.mat(xb, cfg.IN, cols), acts)
    assert np.array_equal(acts.h1.read().reshape(cfg.HA, cols), g["h1"])
    assert np.array_equal(acts.h2.read().reshape(cfg.HA, cols), g["h2"])
    assert np.array_equal(acts.o.read().reshape(3, cols), g["o"])


def test_sample_golden(dev):
    g = np.load(GOLDEN / "sample.npz", allow_pickle=False)
    sim = R.Sim(dev, corridor(), 8, int(g["seed"]), R.Config(beams=16))
    out, ls = dev.upload(g["out"]), dev.upload(g["log_std"])
    act, logp, value = dev.alloc(16), dev.alloc(8), dev.alloc(8)
    sim.sample(R.mat(out, 3, 8), R.mat(ls, 1, 2), R.mat(act, 2, 8), R.mat(logp, 1, 8), R.mat(value, 1, 8), False,
               int(g["step"]))
    assert np.array_equal(act.read().reshape(2, 8), g["actions"])
    assert np.array_equal(logp.read(), g["logp"]) and np.array_equal(value.read(), g["value"])


def test_trainer_golden(dev):
    g = np.load(GOLDEN / "trainer_corridor.npz", allow_pickle=False)
    n = int(g["n"])
    cfg = R.Config(beams=int(g["beams"]), hidden=int(g["hidden"]), rollout=int(g["rollout"]), epochs=int(g["epochs"]),
                   minibatches=int(g["minibatches"]), heading_noise=0.1)
    pol = R.Policy(dev, 42, cfg)
    sim = R.Sim(dev, corridor(), n, 42, cfg)
    sim.reset()
    acts = R.Activations(dev, n, cfg)
    tr = R.Trainer(dev, n, cfg)
    assert np.float32(tr.rollout(sim, pol, acts, 0)) == g["mean_reward"]
    for name in ("obs", "actions", "logp", "value", "reward", "done", "advantage"):
        assert np.array_equal(getattr(tr, name).read(), g[name].ravel()), name
    kl = tr.update(pol)
    # the update's float64 reductions may differ from the oracle's in the last bit
    assert abs(kl - float(g["kl"])) <= 1e-3 * max(abs(float(g["kl"])), 1e-6) + 1e-7
    np.testing.assert_allclose(pol.theta.read(), g["theta"], rtol=1e-4, atol=1e-6)
    assert tr.step == int(g["steps"])
