"""The allocation-free backward (racer_policy_backward_noalloc) against the
allocating one (racer_policy_backward, racer/ppo.mojo:113-138) through the C
ABI: the same kernels in the same order, so d2, d1 and the weight gradients
are bitwise the same, on two calls in a row with the second smaller (the
fixed scratch keeps the first call's partials), in float32 and bf16; and a
shape whose partials exceed the fixed scratch is refused."""
import numpy as np
import pytest

import warporacer3d_amd as R

pytestmark = pytest.mark.gpu

BEAMS, HIDDEN = 1081, 256


def _case(dev, pol, cfg, cols, seed):
    rng = np.random.default_rng(seed)
    IN, HA = cfg.IN, cfg.HA
    x = dev.upload(rng.uniform(0, 1, (IN, cols)).astype(np.float32))
    acts = R.Activations(dev, cols, cfg)
    pol.forward(R.mat(x, IN, cols), acts)
    d3 = dev.upload(rng.normal(0, 1e-3, (3, cols)).astype(np.float32))
    return x, acts, d3


def _backward(dev, pol, cfg, cols, x, acts, d3, noalloc):
    d2, d1, grad = dev.alloc(cfg.HA * cols, 7.0), dev.alloc(cfg.HA * cols, 7.0), dev.alloc(pol.P, 7.0)
    fn = pol.backward_noalloc if noalloc else pol.backward
    fn(R.mat(x, cfg.IN, cols), acts, R.mat(d3, 3, cols), d2, d1, grad)
    dev.synchronize()
    return d2.read(), d1.read(), grad.read()


@pytest.mark.parametrize("precision", [0, 1])
def test_noalloc_backward_matches_backward(dev, precision):
    print(f"no-alloc backward, precision {precision}", flush=True)
    cfg = R.Config(beams=BEAMS, hidden=HIDDEN, precision=precision)
    pol = R.Policy(dev, 42, cfg)
    for call, cols in enumerate((16384 if precision == 0 else 40000, 3000)):
        x, acts, d3 = _case(dev, pol, cfg, cols, 10 * precision + call)
        ref = _backward(dev, pol, cfg, cols, x, acts, d3, noalloc=False)
        got = _backward(dev, pol, cfg, cols, x, acts, d3, noalloc=True)
        for name, a, b in zip(("d2", "d1", "grad"), ref, got):
            assert np.array_equal(a, b), (precision, call, cols, name)
        assert np.abs(ref[2][: HIDDEN * cfg.IN]).max() > 0


def test_noalloc_backward_refuses_beyond_fixed_scratch(dev):
    print("no-alloc backward, too many columns", flush=True)
    cfg = R.Config(beams=BEAMS, hidden=HIDDEN, precision=0)
    pol = R.Policy(dev, 42, cfg)
    cols = 65536  # 16 chunks x 256 x 1086 partial floats > 4194304
    x, acts, d3 = _case(dev, pol, cfg, cols, 3)
    d2, d1, grad = dev.alloc(cfg.HA * cols), dev.alloc(cfg.HA * cols), dev.alloc(pol.P)
    with pytest.raises(R.RacerError, match="allocation-free backward needs"):
        pol.backward_noalloc(R.mat(x, cfg.IN, cols), acts, R.mat(d3, 3, cols), d2, d1, grad)
