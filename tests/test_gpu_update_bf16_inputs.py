"""The bf16 update's permuted minibatch inputs kept in bf16 (the conversion
the bf16 kernels apply on load, done once by the permute) against float32
copies: the same weights and KL bit for bit over two PPO iterations
(racer/ppo.mojo:234-315)."""
import numpy as np
import pytest

import warporacer3d_amd as R
from helpers import track

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("beams,hidden,n,mb", [(108, 64, 4096, 2), (1081, 256, 2048, 1)])
def test_bf16_inputs_give_the_same_update(dev, beams, hidden, n, mb):
    print(f"bf16 update inputs beams={beams} H={hidden} n={n} minibatches={mb}", flush=True)
    cfg = R.Config(beams=beams, hidden=hidden, rollout=32, epochs=2, minibatches=mb, precision=1,
                   heading_noise=0.1)
    runs = []
    for x16 in (True, False):
        sim = R.Sim(dev, track("bridge8"), n, 42, cfg)
        sim.reset()
        pol = R.Policy(dev, 42, cfg)
        acts = R.Activations(dev, n, cfg)
        tr = R.Trainer(dev, n, cfg)
        tr.set_bf16_inputs(x16)
        assert tr.bf16_inputs == x16
        kls = [0.0, 0.0]
        for it in range(2):
            tr.rollout(sim, pol, acts, it)
            kls[it] = tr.update(pol)
        runs.append((pol.theta.read().copy(), kls))
        del tr, acts, pol, sim
    (theta16, kl16), (theta32, kl32) = runs
    assert np.isfinite(theta16).all()
    assert np.array_equal(theta16, theta32)
    assert kl16 == kl32


def test_bf16_inputs_only_where_they_apply(dev):
    small = R.Config(beams=108, hidden=64, rollout=8, minibatches=2, precision=1)
    tr = R.Trainer(dev, 256, small)  # 1024-column minibatches take the 16-column forward
    assert not tr.bf16_inputs
    tr.set_bf16_inputs(True)
    assert not tr.bf16_inputs
    fp32 = R.Config(beams=108, hidden=64, rollout=32, minibatches=2, precision=0)
    assert not R.Trainer(dev, 4096, fp32).bf16_inputs
    big = R.Config(beams=108, hidden=64, rollout=32, minibatches=2, precision=1)
    tr = R.Trainer(dev, 4096, big)
    assert tr.bf16_inputs
    tr.set_bf16_inputs(False)
    assert not tr.bf16_inputs
