"""HIP trainer against the oracle and the reference's ppo_tests
(tests/test_mojo.mojo:162-215)."""
import numpy as np
import pytest

import oracle as O
import warporacer3d_amd as R
from helpers import host_gae, host_loss, track

pytestmark = pytest.mark.gpu

N, RO, MB, BEAMS, H = 4, 32, 4, 108, 64
BATCH = N * RO // MB


@pytest.fixture(scope="module")
def pair(dev):
    print("ppo rollout pair", flush=True)
    cfg = R.Config(beams=BEAMS, hidden=H, rollout=RO, minibatches=MB)
    tr_arr = track("flat")
    gsim, osim = R.Sim(dev, tr_arr, N, 42, cfg), O.Sim(tr_arr, N, 42, beams=BEAMS)
    gsim.reset()
    osim.reset()
    s = gsim.state.read().reshape(11, N)
    for i in range(N):
        s[6, i] = 3000 - 5 - 7 * i
    gsim.state.write(s)
    osim.state[:] = s
    gpol, opol = R.Policy(dev, 42, cfg), O.Policy(42, BEAMS, H)
    gacts = R.Activations(dev, N, cfg)
    gtr, otr = R.Trainer(dev, N, cfg), O.Trainer(N, opol, RO, 4, MB)
    grew = gtr.rollout(gsim, gpol, gacts, 0)
    orew = otr.rollout(osim, 0)
    return dict(cfg=cfg, gsim=gsim, osim=osim, gpol=gpol, opol=opol, gacts=gacts, gtr=gtr, otr=otr, grew=grew,
                orew=orew)


def test_rollout_bit_identical(pair):
    g, o = pair["gtr"], pair["otr"]
    for name in ("obs", "actions", "logp", "value", "reward", "done", "advantage"):
        got = getattr(g, name).read()
        ref = getattr(o, name).ravel()
        assert np.array_equal(got, ref), name
    assert abs(pair["grew"] - pair["orew"]) <= 1e-6 * max(1.0, abs(pair["orew"]))


def test_gae_matches_host_recurrence(pair):
    g = pair["gtr"]
    last = pair["gacts"].o.read().reshape(3, N)[2]
    adv = host_gae(g.reward.read(), g.done.read(), g.value.read(), last, N, RO)
    np.testing.assert_allclose(g.advantage.read(), adv, atol=1e-4, rtol=0)
    assert int((g.done.read() != 0).sum()) == N


def test_gradient_matches_finite_differences(pair):
    g, pol = pair["gtr"], pair["gpol"]
    g.minibatch(pol, 0, 0.0, 1.0)
    analytic = g.grad.read()
    w = pol.theta.read().astype(np.float64)
    args = (BEAMS, H, N, MB, BATCH, g.obs.read(), g.actions.read(), g.logp.read(), g.value.read(),
            g.advantage.read())
    largest = 0.0
    i = 0
    while i < pol.P:
        w[i] += 1e-4
        plus = host_loss(w, *args)
        w[i] -= 2e-4
        numeric = (plus - host_loss(w, *args)) / 2e-4
        w[i] += 1e-4
        assert abs(numeric - analytic[i]) <= 3e-2 * abs(numeric) + 1e-3, (i, analytic[i], numeric)
        largest = max(largest, abs(numeric))
        i += 293 if i < pol.LOGSTD - 293 else 1
    assert largest > 1e-3


def test_minibatch_gradient_matches_oracle(pair):
    g, pol, o = pair["gtr"], pair["gpol"], pair["otr"]
    g.minibatch(pol, 1, 0.1, 0.9)
    ref = o.minibatch(1, 0.1, 0.9)
    got = g.grad.read()
    assert np.array_equal(got[: pair["opol"].LOGSTD], ref["grad"][: pair["opol"].LOGSTD])
    np.testing.assert_allclose(got[pair["opol"].LOGSTD:], ref["grad"][pair["opol"].LOGSTD:], rtol=1e-5, atol=1e-8)
    stats = g.stats.read()
    np.testing.assert_allclose(stats[3], ref["norm_sq"], rtol=1e-5)


def test_update_kl_and_theta_against_oracle(pair):
    g, pol, o = pair["gtr"], pair["gpol"], pair["otr"]
    kl = g.update(pol)
    okl = o.update()
    assert np.isfinite(kl) and 0 <= kl < 0.1
    assert g.step == o.step == 16
    np.testing.assert_allclose(pol.theta.read(), pair["opol"].theta, rtol=1e-4, atol=1e-6)
    assert abs(kl - okl) <= 1e-3 * max(abs(okl), 1e-6) + 1e-7


def test_checkpoint_round_trip(dev, tmp_path, pair):
    pol, cfg = pair["gpol"], pair["cfg"]
    path = tmp_path / "agent.npy"
    R.save(pol, path, 123)
    arr = np.load(path, allow_pickle=False)
    assert arr.dtype == np.float32 and arr.shape == (pol.P + 1,) and arr[0] == 123
    assert np.array_equal(arr[1:], pol.theta.read())
    fresh = R.Policy(dev, 7, cfg)
    assert R.load(fresh, path) == 123
    assert np.array_equal(fresh.theta.read(), arr[1:])
    np.save(tmp_path / "bad.npy", np.zeros(10, np.float32))
    with pytest.raises(R.RacerError, match="checkpoint does not match the network size"):
        R.load(fresh, tmp_path / "bad.npy")


def test_bf16_trainer_refuses_minibatches_beyond_32bit_offsets(dev):
    """The bf16 forward reads minibatch inputs through 32-bit buffer offsets: a
    minibatch of 1086 x 1048576 floats (4.2 GiB) is refused at creation, before
    anything is allocated, instead of wrapping its offsets."""
    cfg = R.Config(beams=1081, hidden=256, rollout=256, minibatches=4, precision=1)
    with pytest.raises(R.RacerError, match="4 GiB"):
        R.Trainer(dev, 16384, cfg)
