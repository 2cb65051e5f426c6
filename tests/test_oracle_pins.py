"""Independent pins of the oracle functions that the reference's own tests do
not reach: spawn (racer/sim.mojo:170-181), the sample's log-density
(racer/sim.mojo:337-348), the surrogate's inputs and targets, and the
advantage moments (racer/ppo.mojo:242-250). Each check comes from outside the
oracle's code: the route table itself, the Gaussian density in float64, the
kinematics of a straight line and of a constant-curvature arc, a planar slope,
and numpy's float64 sums."""
import numpy as np
import pytest

import oracle as O
from helpers import track

HALF_LOG_2PI = 0.5 * np.log(2 * np.pi)
DT, SUBSTEPS, WHEELBASE = 1.0 / 240.0, 4, 0.3302


def _route(arr):
    nx, ny, layers, W = int(arr[0]), int(arr[1]), int(arr[7]), int(arr[5])
    start = 12 + 3 * layers * nx * ny
    return arr[start: start + 4 * W].reshape(W, 4)


@pytest.mark.parametrize("name,noise", [("ring10", 0.0), ("bridge8", 0.0), ("ramp10", 0.1)])
def test_spawn_places_cars_on_route_rows(name, noise):
    """A spawned car stands on a route row w: x, y, z and (noise aside) the heading
    are row w's fields, PROGRESS = w, FRICTION in [0.85, 1.15), EPISODE + 1, and
    every other row is zero whatever it held before."""
    arr = track(name)
    route = _route(arr)
    W = route.shape[0]
    n = 600
    rng = np.random.default_rng(5)
    s = rng.uniform(-50, 50, (11, n)).astype(np.float32)
    s[8] = rng.integers(0, 7, n).astype(np.float32)
    episode = s[8].copy()
    for i in range(n):
        O.spawn(arr, s, i, 42, noise)
    rows = np.empty(n, np.int64)
    for i in range(n):
        # the route row the car stands on, found from the table alone
        hit = np.flatnonzero((route[:, 0] == s[0, i]) & (route[:, 1] == s[1, i]) & (route[:, 2] == s[2, i]))
        assert hit.size >= 1, f"car {i} is not on a route row"
        rows[i] = hit[0]
    assert np.array_equal(s[7], rows.astype(np.float32)), "PROGRESS is the row index"
    dh = s[3] - route[rows, 3]
    if noise == 0:
        assert np.array_equal(dh, np.zeros(n, np.float32)), "the heading is the row's"
    else:
        assert np.all(np.abs(dh) <= noise * (1 + 1e-6)) and np.std(dh) > noise / 3, "heading noise U(-a, a)"
    assert np.all(s[9] >= 0.85) and np.all(s[9] < 1.15 + 1e-6) and np.std(s[9]) > 0.05, "friction U(0.85, 1.15)"
    assert np.array_equal(s[8], episode + 1), "EPISODE counts spawns"
    for r in (4, 5, 6, 10):
        assert np.array_equal(s[r], np.zeros(n, np.float32)), f"row {r} is zeroed"
    # the start row is uniform over the route: every decile of it is used about equally
    counts = np.bincount(rows * 10 // W, minlength=10)
    assert counts.min() > n / 10 * 0.5 and counts.max() < n / 10 * 1.6, counts


def test_spawn_row_and_friction_follow_the_uniform_stream():
    """w = min(floor(u0 W), W - 1) and friction = 0.85 + 0.3 u1 with (u0, u1) the
    first two uniforms of Philox(seed, subsequence = car, offset = new episode)."""
    arr = track("ring10")
    W = int(arr[5])
    n = 64
    s = np.zeros((11, n), np.float32)
    s[8] = np.arange(n) % 3
    for i in range(n):
        O.spawn(arr, s, i, 7)
        u = O.uniform4(7, i, int(s[8, i]))
        assert s[7, i] == min(int(np.floor(np.float32(u[0]) * np.float32(W))), W - 1)
        assert s[9, i] == np.float32(np.float32(0.85) + np.float32(0.3) * u[1])


def test_sample_deterministic_is_the_mean():
    n = 257
    rng = np.random.default_rng(1)
    out = rng.normal(0, 1, (3, n)).astype(np.float32)
    log_std = np.array([-0.5, -1.25], np.float32)
    act, logp, value = np.zeros((2, n), np.float32), np.zeros(n, np.float32), np.zeros(n, np.float32)
    O.Sim(track("ring10"), n, 3).sample(out, log_std, act, logp, value, True, 9)
    assert np.array_equal(act, out[:2]) and np.array_equal(value, out[2])
    expect = -(log_std.astype(np.float64) + HALF_LOG_2PI).sum()
    assert np.all(np.abs(logp - expect) <= 2e-6 * abs(expect))


@pytest.mark.parametrize("log_std", [(-0.5, -0.5), (-2.0, 0.0), (-1.0, -0.3)])
def test_sample_log_density_of_the_emitted_action(log_std):
    """Stochastic samples: logp is the float64 Gaussian log-density of the action
    the sample emitted, N(mean, e^(2 log_std)), per car; and the implied
    standard normal draws have mean 0 and variance 1."""
    n = 4096
    rng = np.random.default_rng(2)
    out = rng.normal(0, 0.5, (3, n)).astype(np.float32)
    ls = np.array(log_std, np.float32)
    act, logp, value = np.zeros((2, n), np.float32), np.zeros(n, np.float32), np.zeros(n, np.float32)
    O.Sim(track("ring10"), n, 11).sample(out, ls, act, logp, value, False, 123)
    sigma = np.exp(ls.astype(np.float64))[:, None]
    z = (act.astype(np.float64) - out[:2].astype(np.float64)) / sigma
    density = -np.sum(0.5 * z * z + ls.astype(np.float64)[:, None] + HALF_LOG_2PI, axis=0)
    # float32 rounding of the action (|a| <= ~6) perturbs z by ulp(a) / sigma
    tol = 1e-5 * (1 + np.abs(density)) + np.sum(np.abs(z) * 5e-7 * (1 + np.abs(act)) / sigma, axis=0)
    assert np.all(np.abs(logp - density) <= tol)
    assert abs(z.mean()) < 0.05 and abs(z.var() - 1) < 0.06
    assert np.array_equal(value, out[2])


def _car(x, y, heading, speed, steer, n=1):
    s = np.zeros((11, n), np.float32)
    s[0], s[1], s[3], s[4], s[5], s[9] = x, y, heading, speed, steer, 1.0
    return s


def test_surrogate_features_on_flat_ground():
    """Inputs: pose, speed, the yaw rate v tan(steer) / wheelbase, steer, zero
    slopes on the flat ring, the clamped throttle and steer commands, then 1."""
    arr = track("ring10")
    s = _car([10.0, -10.0, 0.3], [0.0, 0.1, 10.0], [1.5707964, -1.5707964, 3.0], [2.0, -1.0, 4.5],
             [0.1, -0.3, 0.0], n=3)
    act = np.array([[0.5, 2.0, -3.0], [1.5, -0.25, 0.0]], np.float32)
    f = O.surrogate_features(arr, s, act).astype(np.float64)
    sp, st = s[4].astype(np.float64), s[5].astype(np.float64)
    expect = np.stack([s[0], s[1], s[2], s[3], sp, sp * np.tan(st) / WHEELBASE, st, np.zeros(3), np.zeros(3),
                       np.clip(act[1], -1, 1), np.clip(act[0], -1, 1), np.ones(3)]).astype(np.float64)
    np.testing.assert_allclose(f, expect, rtol=2e-6, atol=1e-7)


def test_surrogate_slope_inputs_on_the_banked_ring():
    """bank10 rises 0.1 m per metre outward: a car facing along the ring sees no
    slope ahead and the bank to its side (outward is to its right when it drives
    anticlockwise, so the rise to its left is -0.1); facing outward, 0.1 ahead."""
    arr = track("bank10")
    s = _car([10.0, 10.0, 0.0], [0.0, 0.0, 10.0], [1.5707964, 0.0, 3.1415925], [1.0, 1.0, 1.0], [0.0, 0.0, 0.0],
             n=3)
    s[2] = [O.Map(arr).surface(0, float(s[0, i]), float(s[1, i]))[0] for i in range(3)]
    f = O.surrogate_features(arr, s, np.zeros((2, 3), np.float32))
    np.testing.assert_allclose(f[7], [0.0, 0.1, 0.0], atol=3e-3)
    np.testing.assert_allclose(f[8], [-0.1, 0.0, -0.1], atol=3e-3)


def _targets_after_one_step(arr, s, actions):
    sim = O.Sim(arr, s.shape[1], 1, beams=8)
    sim.state[:] = s
    sim.actions[:] = actions
    prev = sim.state.copy()
    sim.step()
    assert np.all(sim.done == 0)
    return O.surrogate_targets(prev, sim.state, sim.done)


def test_surrogate_targets_straight_line():
    """Flat road, steer 0, no throttle: the car moves SUBSTEPS DT v straight ahead
    and nothing else changes (the RK4 step is exact on a straight line)."""
    arr = track("ring10")
    s = _car([10.0, -10.0, 0.0], [0.0, 0.0, 10.0], [1.5707964, -1.5707964, 3.1415925], [2.0, 3.0, 0.75],
             [0.0, 0.0, 0.0], n=3)
    tgt, mask = _targets_after_one_step(arr, s, np.zeros((2, 3), np.float32))
    expect = np.zeros((5, 3))
    expect[0] = s[4].astype(np.float64) * SUBSTEPS * DT
    np.testing.assert_allclose(tgt, expect, atol=2e-6)
    assert np.array_equal(mask, np.ones(3, np.float32))


def test_surrogate_targets_constant_curvature_arc():
    """Flat road, fixed steer, no throttle: a circular arc of yaw rate
    w = v tan(steer) / wheelbase; over T = SUBSTEPS DT the car-frame targets are
    (v/w sin wT, v/w (1 - cos wT), wT, 0, 0), with a left turn to the left."""
    arr = track("ring10")
    steer = np.array([0.1, -0.2, 0.3], np.float32)
    s = _car([10.0, -10.0, 0.0], [0.0, 0.0, 10.0], [1.5707964, -1.5707964, 3.1415925], [2.0, 1.5, 1.0], steer, n=3)
    tgt, _ = _targets_after_one_step(arr, s, np.zeros((2, 3), np.float32))
    v, st = s[4].astype(np.float64), steer.astype(np.float64)
    w = v * np.tan(st) / WHEELBASE
    T = SUBSTEPS * DT
    expect = np.stack([v / w * np.sin(w * T), v / w * (1 - np.cos(w * T)), w * T, 0 * v, 0 * v])
    np.testing.assert_allclose(tgt, expect, atol=3e-6)
    assert np.all(np.sign(tgt[1]) == np.sign(st)) and np.all(np.sign(tgt[2]) == np.sign(st))


def test_surrogate_targets_mask_respawns():
    prev = _car([10.0, 10.0], [0.0, 0.0], [0.0, 0.0], [1.0, 1.0], [0.0, 0.0], n=2)
    tgt, mask = O.surrogate_targets(prev, prev.copy(), np.array([0.0, 1.0], np.float32))
    assert np.array_equal(mask, np.array([1.0, 0.0], np.float32))
    assert np.array_equal(tgt, np.zeros((5, 2), np.float32))


@pytest.mark.parametrize("size", [1, 1000, 262144])
def test_moments_are_float64_sums(size):
    rng = np.random.default_rng(size)
    adv = (rng.normal(3, 40, size) * rng.choice([1e-3, 1, 1e3], size)).astype(np.float32)
    s, sq = O.moments(adv)
    a = adv.astype(np.float64)
    assert abs(s - a.sum()) <= 1e-12 * np.abs(a).sum()
    assert abs(sq - (a * a).sum()) <= 1e-12 * (a * a).sum()
