"""HIP policy products against the oracle and float64: the float32 forward is
bit-identical; bf16 MFMA products stay within the bf16 rounding bound
|C - C64| <= 2^-8 sum_k |a_k b_k| (+ float32 slack)."""
import numpy as np
import pytest

import oracle as O
import warporacer3d_amd as R

pytestmark = pytest.mark.gpu


def _policy_pair(dev, beams, hidden, precision=0, seed=42):
    cfg = R.Config(beams=beams, hidden=hidden, precision=precision)
    return R.Policy(dev, seed, cfg), O.Policy(seed, beams, hidden), cfg


@pytest.mark.parametrize("beams,hidden,cols", [(108, 64, 200), (1081, 256, 300), (13, 7, 5)])
def test_init_and_forward_float32_bit_identical(dev, beams, hidden, cols):
    print(f"forward f32 {beams}/{hidden}/{cols}", flush=True)
    g, o, cfg = _policy_pair(dev, beams, hidden)
    assert np.array_equal(g.theta.read(), o.theta), "init differs"
    x = np.random.default_rng(0).uniform(-1, 1, (cfg.IN, cols)).astype(np.float32)
    x[-1] = 1
    xb = dev.upload(x)
    acts = R.Activations(dev, cols, cfg)
    g.forward(R.mat(xb, cfg.IN, cols), acts)
    h1, h2, out = o.forward(x)
    assert np.array_equal(acts.h1.read().reshape(cfg.HA, cols), h1)
    assert np.array_equal(acts.h2.read().reshape(cfg.HA, cols), h2)
    assert np.array_equal(acts.o.read().reshape(3, cols), out)


def test_forward_bf16_within_rounding_bound(dev):
    print("forward bf16 1081/256", flush=True)
    beams, hidden, cols = 1081, 256, 384
    g, o, cfg = _policy_pair(dev, beams, hidden, precision=1)
    x = np.random.default_rng(1).uniform(0, 1, (cfg.IN, cols)).astype(np.float32)
    x[-1] = 1
    xb = dev.upload(x)
    acts = R.Activations(dev, cols, cfg)
    g.forward(R.mat(xb, cfg.IN, cols), acts)
    th = o.theta.astype(np.float64)
    W0 = th[: hidden * cfg.IN].reshape(hidden, cfg.IN)
    pre = W0 @ x.astype(np.float64)
    bound = 2.0**-8 * (np.abs(W0) @ np.abs(x.astype(np.float64))) + 1e-5
    h1 = acts.h1.read().reshape(cfg.HA, cols)
    assert np.all(np.abs(h1[:hidden] - np.tanh(pre)) <= bound)
    assert np.all(h1[hidden] == 1)


def _gemm_case(dev, M, N, K, precision, epi, rng):
    """One call of the policy forward restricted to W0 (M = hidden, K = IN)."""
    beams = K - 5
    cfg = R.Config(beams=beams, hidden=M, precision=precision)
    pol = R.Policy(dev, 3, cfg)
    theta = pol.theta.read()
    W = rng.normal(0, 0.2, (M, K)).astype(np.float32)
    theta[: M * K] = W.ravel()
    pol.theta.write(theta)
    pol.transpose()
    x = rng.uniform(-1, 1, (K, N)).astype(np.float32)
    xb = dev.upload(x)
    acts = R.Activations(dev, N, cfg)
    pol.forward(R.mat(xb, K, N), acts)
    return W.astype(np.float64), x.astype(np.float64), acts.h1.read().reshape(M + 1, N)[:M]


@pytest.mark.parametrize("M,N,K", [(64, 64, 32), (17, 100, 37), (256, 513, 1086), (130, 70, 260), (3, 1, 9)])
@pytest.mark.parametrize("precision", [0, 1])
def test_ragged_products(dev, M, N, K, precision):
    print(f"gemm {M}x{N}x{K} p{precision}", flush=True)
    rng = np.random.default_rng(M * 7 + N + K)
    W, x, h = _gemm_case(dev, M, N, K, precision, "tanh", rng)
    ref = np.tanh(W @ x)
    if precision == 0:
        bound = 1e-5 * (np.abs(W) @ np.abs(x)) + 1e-6
    else:
        bound = 2.0**-8 * (np.abs(W) @ np.abs(x)) + 1e-5
    assert np.all(np.abs(h - ref) <= bound)


@pytest.mark.parametrize("precision", [0, 1])
def test_backward_against_oracle(dev, precision):
    print(f"backward p{precision}", flush=True)
    beams, hidden, cols = 108, 64, 640
    g, o, cfg = _policy_pair(dev, beams, hidden, precision)
    rng = np.random.default_rng(2)
    x = rng.uniform(-1, 1, (cfg.IN, cols)).astype(np.float32)
    d3 = rng.normal(0, 0.01, (3, cols)).astype(np.float32)
    xb, d3b = dev.upload(x), dev.upload(d3)
    acts = R.Activations(dev, cols, cfg)
    g.forward(R.mat(xb, cfg.IN, cols), acts)
    d2, d1, grad = dev.alloc(cfg.HA * cols), dev.alloc(cfg.HA * cols), dev.alloc(cfg.params())
    g.backward(R.mat(xb, cfg.IN, cols), acts, R.mat(d3b, 3, cols), d2, d1, grad)
    h1 = acts.h1.read().reshape(cfg.HA, cols)
    h2 = acts.h2.read().reshape(cfg.HA, cols)
    rd2, rd1, rgrad = o.backward(x, h1, h2, d3)
    gd2 = d2.read().reshape(cfg.HA, cols)
    gd1 = d1.read().reshape(cfg.HA, cols)
    gg = grad.read()
    if precision == 0:
        assert np.array_equal(gd2, rd2) and np.array_equal(gd1, rd1)
        assert np.array_equal(gg[: o.LOGSTD], rgrad[: o.LOGSTD])
    else:
        np.testing.assert_array_equal(gd2, rd2)
        scale = np.abs(rgrad[: o.LOGSTD]).max()
        assert np.max(np.abs(gg[: o.LOGSTD] - rgrad[: o.LOGSTD])) <= 2e-2 * scale
