"""The oracle's PPO against the reference's ppo_tests (tests/test_mojo.mojo:162-215):
GAE recomputed on the host, the analytic minibatch gradient against central
finite differences of a float64 host loss, and the KL after one update."""
import numpy as np
import pytest

import oracle as O
from helpers import host_gae, host_loss, track

N, R, MB, BEAMS, H = 4, 32, 4, 108, 64
BATCH = N * R // MB


@pytest.fixture(scope="module")
def rolled():
    sim = O.Sim(track("flat"), N, 42, beams=BEAMS)
    sim.reset()
    for i in range(N):
        sim.state[6, i] = 3000 - 5 - 7 * i  # episodes end mid-rollout (tests/test_mojo.mojo:168-171)
    pol = O.Policy(42, BEAMS, H)
    tr = O.Trainer(N, pol, R, 4, MB)
    reward = tr.rollout(sim, 0)
    return sim, pol, tr, reward


def test_gae_matches_host_recurrence(rolled):
    _, _, tr, reward = rolled
    assert np.isfinite(reward)
    adv = host_gae(tr.reward, tr.done, tr.value, tr.last_o[2], N, R)
    np.testing.assert_allclose(tr.advantage, adv, atol=1e-4, rtol=0)
    ended = int((tr.done != 0).sum())
    assert ended == N, "every car's episode ended once during the rollout"


def test_gradient_matches_finite_differences(rolled):
    _, pol, tr, _ = rolled
    res = tr.minibatch(0, 0.0, 1.0)
    analytic = res["grad"]
    w = pol.theta.astype(np.float64)
    args = (BEAMS, H, N, MB, BATCH, tr.obs, tr.actions, tr.logp, tr.value, tr.advantage)
    worst = largest = 0.0
    i = 0
    while i < pol.P:
        w[i] += 1e-4
        plus = host_loss(w, *args)
        w[i] -= 2e-4
        numeric = (plus - host_loss(w, *args)) / 2e-4
        w[i] += 1e-4
        err = abs(numeric - analytic[i])
        assert err <= 5e-3 * abs(numeric) + 1e-5, (i, analytic[i], numeric)
        worst, largest = max(worst, err), max(largest, abs(numeric))
        i += 293 if i < pol.LOGSTD - 293 else 1
    assert largest > 1e-3, "the rollout produces a nontrivial gradient"


def test_kl_after_update(rolled):
    _, pol, tr, _ = rolled
    theta = pol.theta.copy()
    kl = tr.update()
    assert np.isfinite(kl) and 0 <= kl < 0.1
    assert tr.step == 16 and not np.array_equal(theta, pol.theta)
    assert pol.theta[pol.LOGSTD:].min() >= -2 and pol.theta[pol.LOGSTD:].max() <= 0


def test_policy_init_layout_and_quirk():
    """ppo.mojo:74-89: W0 bias column zero, log std -0.5, and the absolute-index
    bias tests of W1/W2 (zeroed column i % HA == H of the absolute index)."""
    pol = O.Policy(42, BEAMS, H)
    IN, HA = BEAMS + 5, H + 1
    W0 = pol.theta[: H * IN].reshape(H, IN)
    assert np.all(W0[:, IN - 1] == 0) and np.all(W0[:, : IN - 1] != 0)
    assert np.all(pol.theta[pol.LOGSTD:] == -0.5)
    w1_idx = np.arange(pol.W1, pol.W2)
    zeroed = w1_idx[pol.theta[pol.W1:pol.W2] == 0]
    assert np.all(zeroed % HA == H)
    assert abs(W0[:, :-1].std() - np.sqrt(2 / (BEAMS + 4))) < 0.01


def test_adam_textbook_step():
    """ppo.mojo:325-336 against a float64 Adam step with the global-norm clip."""
    pol = O.Policy(1, 8, 4)
    rng = np.random.default_rng(0)
    g = rng.normal(size=pol.P).astype(np.float32)
    first = np.zeros(pol.P, np.float32)
    second = np.zeros(pol.P, np.float32)
    th0 = pol.theta.astype(np.float64).copy()
    norm_sq = float((g.astype(np.float64) ** 2).sum())
    O.adam(pol, g, first, second, norm_sq, 1)
    clip = min(1.0, 0.5 / (np.sqrt(norm_sq) + 1e-8))
    gg = g * clip
    m, v = 0.1 * gg, 0.001 * gg * gg
    ref = th0 - 3e-4 * (m / 0.1) / (np.sqrt(v / 0.001) + 1e-5)
    ref[pol.LOGSTD:] = np.clip(ref[pol.LOGSTD:], -2, 0)
    np.testing.assert_allclose(pol.theta, ref, rtol=1e-5, atol=1e-7)
