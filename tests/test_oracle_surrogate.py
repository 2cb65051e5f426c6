"""Pins of the oracle's surrogate pieces: the float32 MLP and its gradient
against a float64 numpy MLP and central finite differences; the float64 RK4
step (the surrogate's float64 reference) against a closed form and against the float32
step of the reference restatement."""
import numpy as np
import pytest

import oracle as O
from warporacer3d_amd import track as T

W = O.SURROGATE_WIDTHS


def _mlp64(theta, x):
    """Float64 forward: tanh hidden layers, linear output, bias as the last weight column."""
    a = x.astype(np.float64)
    o = 0
    for l in range(len(W) - 1):
        rows, k = W[l + 1], W[l] + 1
        w = theta[o: o + rows * k].astype(np.float64).reshape(rows, k)
        o += rows * k
        z = w @ a
        if l < len(W) - 2:
            a = np.vstack([np.tanh(z), np.ones((1, a.shape[1]))])
        else:
            a = z
    return a


def _loss64(theta, x, y, mask):
    out = _mlp64(theta, x)
    return float(np.sum(mask * np.sum((out - y) ** 2, axis=0)) / (W[-1] * x.shape[1]))


def _batch(seed, cols=40):
    rng = np.random.default_rng(seed)
    theta = rng.normal(0, 0.25, O.surrogate_params()).astype(np.float32)
    x = rng.normal(0, 1, (12, cols)).astype(np.float32)
    x[11] = 1
    y = rng.normal(0, 1, (5, cols)).astype(np.float32)
    mask = (rng.uniform(size=cols) > 0.2).astype(np.float32)
    return theta, x, y, mask


def test_mlp_forward_matches_float64():
    theta, x, _, _ = _batch(0)
    out = O.mlp_forward(theta, x)[-1].astype(np.float64)
    ref = _mlp64(theta, x)
    assert np.max(np.abs(out - ref)) <= 1e-5 * max(1.0, np.max(np.abs(ref)))


def test_mlp_gradient_matches_finite_differences():
    theta, x, y, mask = _batch(1)
    loss, grad = O.mlp_gradient(theta, x, y, mask)
    assert abs(loss - _loss64(theta, x, y, mask)) <= 1e-5 * loss
    t64 = theta.astype(np.float64)
    rng = np.random.default_rng(3)
    # every layer's weights and biases, sampled
    idx = np.concatenate([rng.choice(O.surrogate_params(), 60, replace=False), [11, 12 * 128 - 1]])
    for i in idx:
        h = 1e-4
        t64[i] += h
        plus = _loss64(t64, x, y, mask)
        t64[i] -= 2 * h
        minus = _loss64(t64, x, y, mask)
        t64[i] += h
        numeric = (plus - minus) / (2 * h)
        assert abs(numeric - grad[i]) <= 2e-3 * abs(numeric) + 2e-5, (i, grad[i], numeric)


def test_masked_samples_do_not_contribute():
    theta, x, y, mask = _batch(2)
    _, g = O.mlp_gradient(theta, x, y, mask)
    y2 = y.copy()
    y2[:, mask == 0] += 100.0  # change only the masked samples' targets
    _, g2 = O.mlp_gradient(theta, x, y2, mask)
    assert np.array_equal(g, g2)


def test_f64_step_closed_form_straight_line():
    """Flat road, steer 0, no throttle: the car keeps its speed and heading and
    moves v SUBSTEPS DT along it (exact in float64 up to rounding)."""
    arr = T.cached("ring10")
    n = 3
    st = np.zeros((6, n))
    st[0] = [10.0, -10.0, 0.0]
    st[1] = [0.0, 0.2, 10.0]
    st[3] = [np.pi / 2, -np.pi / 2, np.pi]
    st[4] = [1.5, 3.0, 0.7]
    start = st.copy()
    O.step_f64(arr, st, np.ones(n, np.float32), np.zeros((2, n), np.float32))
    dist = st[4] * 4 / 240.0
    assert np.allclose(st[0], start[0] + dist * np.cos(start[3]), atol=1e-12)
    assert np.allclose(st[1], start[1] + dist * np.sin(start[3]), atol=1e-12)
    assert np.allclose(st[4], start[4]) and np.allclose(st[5], 0.0)


def test_f64_step_agrees_with_float32_step():
    """On the banked ring, one float64 step lands within float32 rounding of the
    float32 restatement for every car that does not crash."""
    O.set_threads(8)
    arr = T.cached("bank10")
    sim = O.Sim(arr, 256, 11, beams=16, heading_noise=0.1)
    sim.reset()
    rng = np.random.default_rng(4)
    for _ in range(5):
        sim.actions[...] = rng.uniform(-1, 1, (2, 256)).astype(np.float32)
        sim.step()
    sim.actions[...] = rng.uniform(-1, 1, (2, 256)).astype(np.float32)
    s0 = sim.state.copy()
    st = np.ascontiguousarray(s0[[0, 1, 2, 3, 4, 5]].astype(np.float64))
    O.step_f64(arr, st, s0[9], sim.actions)
    sim.step()
    ok = sim.done == 0
    assert ok.sum() > 200
    s1 = sim.state
    assert np.max(np.abs(st[0][ok] - s1[0][ok])) < 2e-5
    assert np.max(np.abs(st[1][ok] - s1[1][ok])) < 2e-5
    dh = np.angle(np.exp(1j * (st[3][ok] - s1[3][ok])))
    assert np.max(np.abs(dh)) < 2e-5
    assert np.max(np.abs(st[4][ok] - s1[4][ok])) < 2e-5
    assert np.max(np.abs(st[5][ok] - s1[5][ok])) < 1e-6
