"""The benchmark workloads on the GPU against the oracle, at their own shapes:

* the flat R = 10 m ring, 1024 cars x 108 beams, float32 2 x 64 policy:
  200 Runner steps (main.mojo:183-191), state, observations and done codes
  bit-identical at every step;
* the ramp ring, 4096 cars x 1081 beams, float32 2 x 256 policy: full-size
  Runner steps bit-identical;
* PPO on the two-lobe bridge (racer/ppo.mojo:187-337) with 128-step rollouts,
  4 epochs x 8 minibatches, 2 x 256: float32 rollout bit-identical and the
  update's weights and KL against the oracle's; bf16 within a stated bound; a
  16384-car iteration's KL in [0, 0.1) (tests/test_mojo.mojo:213-214);
* eight uneven instances (256 .. 8192 cars, 1081 beams) in one batched launch
  per stage: every instance bit-identical to the oracle's own Sim and Policy.
"""
import numpy as np
import pytest

import oracle as O
import warporacer3d_amd as R
from helpers import track

pytestmark = pytest.mark.gpu


def _oracle_runner(arr, n, beams, hidden, seed, noise):
    osim = O.Sim(arr, n, seed, beams=beams, heading_noise=noise)
    osim.reset()
    return osim, O.Policy(seed, beams, hidden)


def _oracle_step(osim, opol, t, logp, value):
    _, _, out = opol.forward(osim.obs)
    osim.sample(out, opol.log_std(), osim.actions, logp, value, False, t)
    osim.step()


def test_ring10_1024_cars_200_steps_bit_identical(dev):
    print("ring10 1024 x 108, 2 x 64 float32, 200 steps", flush=True)
    O.set_threads(16)
    n, beams, hidden = 1024, 108, 64
    cfg = R.Config(beams=beams, hidden=hidden, precision=0, heading_noise=0.1)
    arr = track("ring10")
    run = R.Runner(arr, n, cfg, seed=42, device=dev)
    osim, opol = _oracle_runner(arr, n, beams, hidden, 42, 0.1)
    assert np.array_equal(run.sim.state.read().reshape(11, n), osim.state)
    logp, value = np.zeros(n, np.float32), np.zeros(n, np.float32)
    events = 0
    for t in range(200):
        run.step(False, t)
        _oracle_step(osim, opol, t, logp, value)
        d = run.sim.done.read()
        assert np.array_equal(d, osim.done), f"done codes differ at step {t}"
        assert np.array_equal(run.sim.state.read().reshape(11, n), osim.state), f"state differs at step {t}"
        assert np.array_equal(run.sim.obs.read().reshape(beams + 5, n), osim.obs), f"obs differ at step {t}"
        sc = run.scratch.read()
        assert np.array_equal(sc[:n], logp) and np.array_equal(sc[n:], value), f"logp/value differ at step {t}"
        events += int((d != 0).sum())
    assert events > 0


def test_ramp10_4096_cars_1081_beams_bit_identical(dev):
    print("ramp10 4096 x 1081, 2 x 256 float32, 3 steps", flush=True)
    O.set_threads(16)
    n, beams, hidden = 4096, 1081, 256
    cfg = R.Config(beams=beams, hidden=hidden, precision=0, heading_noise=0.1)
    arr = track("ramp10")
    run = R.Runner(arr, n, cfg, seed=42, device=dev)
    osim, opol = _oracle_runner(arr, n, beams, hidden, 42, 0.1)
    assert np.array_equal(run.sim.obs.read().reshape(beams + 5, n), osim.obs)
    logp, value = np.zeros(n, np.float32), np.zeros(n, np.float32)
    for t in range(3):
        run.step(False, t)
        _oracle_step(osim, opol, t, logp, value)
        assert np.array_equal(run.sim.done.read(), osim.done), t
        assert np.array_equal(run.sim.state.read().reshape(11, n), osim.state), t
        assert np.array_equal(run.sim.obs.read().reshape(beams + 5, n), osim.obs), t
        sc = run.scratch.read()
        assert np.array_equal(sc[:n], logp) and np.array_equal(sc[n:], value), t


# ---------------- PPO on the two-lobe bridge
PPO = dict(rollout=128, epochs=4, minibatches=8)


def _trainer_pair(dev, n, beams, precision, seed=42):
    cfg = R.Config(beams=beams, hidden=256, precision=precision, heading_noise=0.1, **PPO)
    arr = track("bridge8")
    gsim = R.Sim(dev, arr, n, seed, cfg)
    gsim.reset()
    gpol = R.Policy(dev, seed, cfg)
    gacts = R.Activations(dev, n, cfg)
    gtr = R.Trainer(dev, n, cfg)
    osim = O.Sim(arr, n, seed, beams=beams, heading_noise=0.1)
    osim.reset()
    opol = O.Policy(seed, beams, 256)
    otr = O.Trainer(n, opol, PPO["rollout"], PPO["epochs"], PPO["minibatches"])
    return gsim, gpol, gacts, gtr, osim, opol, otr


@pytest.fixture(scope="module")
def bridge_fp32(dev):
    print("bridge8 trainer 64 cars x 108 beams, R 128, 4 x 8, float32", flush=True)
    O.set_threads(16)
    gsim, gpol, gacts, gtr, osim, opol, otr = _trainer_pair(dev, 64, 108, 0)
    theta0 = opol.theta.copy()
    grew = gtr.rollout(gsim, gpol, gacts, 0)
    orew = otr.rollout(osim, 0)
    rollout = {name: (getattr(gtr, name).read(), getattr(otr, name).ravel().copy())
               for name in ("obs", "actions", "logp", "value", "reward", "done", "advantage")}
    gkl = gtr.update(gpol)
    okl = otr.update()
    return dict(rollout=rollout, grew=grew, orew=orew, gkl=gkl, okl=okl, gtheta=gpol.theta.read(),
                otheta=opol.theta.copy(), theta0=theta0, gstep=gtr.step, ostep=otr.step, otr=otr)


def test_bridge8_rollout_bit_identical(bridge_fp32):
    for name, (got, ref) in bridge_fp32["rollout"].items():
        assert np.array_equal(got, ref), name
    assert abs(bridge_fp32["grew"] - bridge_fp32["orew"]) <= 1e-6 * max(1.0, abs(bridge_fp32["orew"]))
    assert int((bridge_fp32["rollout"]["done"][0] != 0).sum()) > 0, "episodes end inside the rollout"


def test_bridge8_update_matches_oracle(bridge_fp32):
    b = bridge_fp32
    assert b["gstep"] == b["ostep"] == 32
    assert np.isfinite(b["gkl"]) and 0 <= b["gkl"] < 0.1
    np.testing.assert_allclose(b["gtheta"], b["otheta"], rtol=1e-4, atol=1e-6)
    assert abs(b["gkl"] - b["okl"]) <= 1e-3 * abs(b["okl"]) + 1e-7
    assert np.abs(b["otheta"] - b["theta0"]).max() > 1e-3, "the update moved the weights"


def test_bridge8_bf16_update_within_bound(dev, bridge_fp32):
    """bf16 products (8-bit significands, float32 accumulation) on the same
    rollout start: after 32 Adam steps the weights move in the direction the
    float32 oracle's update takes them (cosine of the two updates >= 0.99, the
    update norms within 5 %), and the last minibatch's KL is within 10 % of the
    oracle's. The rollout itself is bf16 too, so the two runs also see
    slightly different actions."""
    print("bridge8 trainer 64 cars x 108 beams, R 128, 4 x 8, bf16", flush=True)
    gsim, gpol, gacts, gtr, _, _, _ = _trainer_pair(dev, 64, 108, 1)
    gtr.rollout(gsim, gpol, gacts, 0)
    kl = gtr.update(gpol)
    b = bridge_fp32
    d_bf = gpol.theta.read().astype(np.float64) - b["theta0"]
    d_fp = b["otheta"].astype(np.float64) - b["theta0"]
    cosine = float(d_bf @ d_fp / (np.linalg.norm(d_bf) * np.linalg.norm(d_fp)))
    ratio = float(np.linalg.norm(d_bf) / np.linalg.norm(d_fp))
    print(f"bf16 vs float32 update: cosine {cosine:.5f}, norm ratio {ratio:.4f}, KL {kl:.6g} vs {b['okl']:.6g}",
          flush=True)
    assert cosine >= 0.99 and abs(ratio - 1) <= 0.05
    assert np.isfinite(kl) and abs(kl - b["okl"]) <= 0.1 * b["okl"]


@pytest.mark.parametrize("beams", [108, 1081])
def test_bridge8_16384_cars_iteration(dev, beams):
    """The full-size iteration (16384 cars, bf16): every output finite and the
    KL after the update in [0, 0.1) (tests/test_mojo.mojo:213-214)."""
    print(f"bridge8 trainer 16384 cars x {beams} beams, bf16 iteration", flush=True)
    cfg = R.Config(beams=beams, hidden=256, precision=1, heading_noise=0.1, **PPO)
    sim = R.Sim(dev, track("bridge8"), 16384, 42, cfg)
    sim.reset()
    pol = R.Policy(dev, 42, cfg)
    acts = R.Activations(dev, 16384, cfg)
    tr = R.Trainer(dev, 16384, cfg)
    reward = tr.rollout(sim, pol, acts, 0)
    kl = tr.update(pol)
    assert np.isfinite(reward) and np.isfinite(kl) and 0 <= kl < 0.1
    assert np.isfinite(pol.theta.read()).all()
    assert np.isfinite(tr.advantage.read()).all()


# ---------------- eight uneven instances in one batched launch per stage
BATCH_TRACKS = ("ring10", "ramp10", "bank10", "bridge8", "random0", "random1", "random2", "random3")
BATCH_CARS = (256, 512, 1024, 2048, 4096, 4096, 8192, 8192)


def test_eight_instance_batch_matches_oracle(dev):
    print("eight-instance batch, 28416 cars x 1081 beams, float32 2 x 256, 3 steps", flush=True)
    O.set_threads(16)
    beams, hidden = 1081, 256
    cfg = R.Config(beams=beams, hidden=hidden, precision=0, heading_noise=0.1)
    pol = R.Policy(dev, 42, cfg)
    opol = O.Policy(42, beams, hidden)
    sims, osims = [], []
    for k, (name, n) in enumerate(zip(BATCH_TRACKS, BATCH_CARS)):
        arr = track(name)
        s = R.Sim(dev, arr, n, 100 + k, cfg)
        s.reset()
        sims.append(s)
        o = O.Sim(arr, n, 100 + k, beams=beams, heading_noise=0.1)
        o.reset()
        osims.append(o)
    batch = R.Batch(dev, sims, cfg)
    assert batch.cars == sum(BATCH_CARS)
    for t in range(3):
        batch.runner_step(pol, False, t)
        for k, (s, o) in enumerate(zip(sims, osims)):
            logp, value = np.zeros(o.n, np.float32), np.zeros(o.n, np.float32)
            _oracle_step(o, opol, t, logp, value)
            assert np.array_equal(s.done.read(), o.done), (t, k, "done")
            assert np.array_equal(s.state.read().reshape(11, o.n), o.state), (t, k, "state")
            assert np.array_equal(s.obs.read().reshape(beams + 5, o.n), o.obs), (t, k, "obs")
            lp, v = batch.logp_value(k)
            assert np.array_equal(lp, logp) and np.array_equal(v, value), (t, k, "logp/value")


@pytest.mark.parametrize("mode", [1, 2])
def test_eight_instance_fused_batch_matches_two_launches(dev, mode):
    """The eight-instance shapes (3552 workgroups of 8 cars, several rounds of
    the persistent grid) in each fused mode against the two-launch batch,
    which the test above pins to the oracle."""
    print(f"eight-instance batch, fused mode {mode}, 28416 cars x 256 beams", flush=True)
    cfg = R.Config(beams=256, hidden=64, precision=0, heading_noise=0.1)
    pol = R.Policy(dev, 42, cfg)

    def make():
        sims = []
        for k, (name, n) in enumerate(zip(BATCH_TRACKS, BATCH_CARS)):
            s = R.Sim(dev, track(name), n, 100 + k, cfg)
            s.reset()
            sims.append(s)
        return sims, R.Batch(dev, sims, cfg)

    sims0, b0 = make()
    sims1, b1 = make()
    b0.set_fused_step(0)
    b1.set_fused_step(mode)
    assert b1.fused_step == mode
    for t in range(3):
        b0.runner_step(pol, False, t)
        b1.runner_step(pol, False, t)
        for k, (a, b) in enumerate(zip(sims0, sims1)):
            for name in ("state", "obs", "actions", "reward", "done"):
                assert np.array_equal(getattr(a, name).read(), getattr(b, name).read()), (t, k, name)
