"""The learned dynamics surrogate on the GPU: its inputs and targets and its
float32 gradient bit for bit against the oracle, the bf16 gradient within its
rounding bound, the float64 reference step against the oracle's, the fallback to the
exact step, and what training gains: error against the float64 RK4 far below
the untrained network's while the network serves most car-steps."""
import numpy as np
import pytest

import oracle as O
import warporacer3d_amd as R
from helpers import track

pytestmark = pytest.mark.gpu


def _f64_buffer(dev, st: np.ndarray):
    return dev.upload(np.ascontiguousarray(st, np.float64).view(np.float32).ravel())


def _read_f64(buf, n):
    return buf.read().view(np.float64).reshape(6, n)


def _setup(dev, n=512, beams=64, hidden=64, seed=3, name="bank10"):
    cfg = R.Config(beams=beams, hidden=hidden, precision=1, heading_noise=0.1)
    sim = R.Sim(dev, track(name), n, seed, cfg)
    sim.reset()
    pol = R.Policy(dev, 42, cfg)
    acts = R.Activations(dev, n, cfg)
    scratch = dev.alloc(2 * n)
    return cfg, sim, pol, acts, scratch


def test_recorded_inputs_and_targets_match_oracle(dev):
    print("surrogate dataset vs oracle", flush=True)
    _, sim, pol, acts, scratch = _setup(dev)
    n = sim.n
    data = R.Dataset(dev, 4 * n)
    data.record(sim, pol, acts, scratch, 2, 0)  # warm the cars up a little
    before = sim.state.read().reshape(11, n).copy()
    data.record(sim, pol, acts, scratch, 1, 2)
    after = sim.state.read().reshape(11, n)
    act = sim.actions.read().reshape(2, n)
    done = sim.done.read()
    cap = data.capacity
    feat = data.features.read().reshape(12, cap)[:, 2 * n: 3 * n]
    tgt = data.targets.read().reshape(5, cap)[:, 2 * n: 3 * n]
    mask = data.mask.read()[2 * n: 3 * n]
    assert np.array_equal(feat, O.surrogate_features(track("bank10"), before, act))
    otgt, omask = O.surrogate_targets(before, after, done)
    assert np.array_equal(tgt, otgt) and np.array_equal(mask, omask)
    assert data.count == 3 * n


@pytest.mark.parametrize("cols", [300, 5000])
def test_float32_gradient_bit_identical_and_bf16_bounded(dev, cols):
    print(f"surrogate gradient cols={cols}", flush=True)
    rng = np.random.default_rng(cols)
    theta = rng.normal(0, 0.2, O.surrogate_params()).astype(np.float32)
    x = rng.normal(0, 1, (12, cols)).astype(np.float32)
    x[11] = 1
    y = rng.normal(0, 1, (5, cols)).astype(np.float32)
    mask = (rng.uniform(size=cols) > 0.1).astype(np.float32)
    loss, ref = O.mlp_gradient(theta, x, y, mask)
    xb, yb, mb = dev.upload(x), dev.upload(y), dev.upload(mask)
    for precision in (0, 1):
        sur = R.Surrogate(dev, precision=precision, cols=cols)
        sur.theta.write(theta)
        sur.mark_updated()
        sur.batch_gradient(R.mat(xb, 12, cols), R.mat(yb, 5, cols), R.mat(mb, 1, cols))
        got = sur.grad.read()
        if precision == 0:
            assert np.array_equal(got, ref)
        else:
            assert np.max(np.abs(got - ref)) <= 0.03 * np.max(np.abs(ref))
            assert np.dot(got, ref) / (np.linalg.norm(got) * np.linalg.norm(ref)) > 0.999


def test_f64_reference_matches_oracle(dev):
    print("f64 reference step", flush=True)
    _, sim, pol, acts, scratch = _setup(dev, n=300, name="bridge8")
    for t in range(4):
        R.runner_step(sim, pol, acts, scratch, False, t)
    n = sim.n
    s = sim.state.read().reshape(11, n)
    st = np.ascontiguousarray(s[[0, 1, 2, 3, 4, 5]].astype(np.float64))
    act = np.random.default_rng(5).uniform(-1, 1, (2, n)).astype(np.float32)
    buf, ab = _f64_buffer(dev, st), dev.upload(act)
    sim.step_f64(buf, R.mat(ab, 2, n))
    got = _read_f64(buf, n)
    O.step_f64(track("bridge8"), st, s[9], act)
    # same algorithm in float64 on both sides; libm and device math may differ in the last bits
    assert np.max(np.abs(got - st)) < 1e-10


def test_all_fallback_is_the_exact_step(dev):
    print("surrogate fallback", flush=True)
    _, a, pol, acts, scratch = _setup(dev)
    _, b, _, acts_b, scratch_b = _setup(dev)
    sur = R.Surrogate(dev, precision=1, cols=a.n)
    fallback = dev.alloc(a.n)
    for t in range(3):
        R.runner_step(a, pol, acts, scratch, False, t)
        sur.runner_step(b, pol, acts_b, scratch_b, False, t, min_clearance=float("inf"), fallback=fallback)
        for name in ("state", "obs", "actions", "reward", "done"):
            assert np.array_equal(getattr(a, name).read(), getattr(b, name).read()), (t, name)
        assert np.all(fallback.read() == 1.0)


@pytest.mark.parametrize("name,n", [("bank10", 1000), ("bridge", 777)])
def test_fused_step_matches_separate_kernels(dev, name, n):
    """The one-launch bf16 step against the separate inputs, bf16 products,
    float32 output product and step kernels: the same bits, over ragged car
    counts, both map kinds, cars served by the network and cars that fall back."""
    print(f"surrogate fused vs separate {name} n={n}", flush=True)
    _, sim, pol, acts, scratch = _setup(dev, n=n, name=name)
    data = R.Dataset(dev, 24 * n)
    data.record(sim, pol, acts, scratch, 24, 0)
    sur = R.Surrogate(dev, seed=2, precision=1, cols=n)
    sur.fit_normalization(data)
    sur.train(data, epochs=1, batch=n, lr=2e-3)
    _, a, _, _, _ = _setup(dev, n=n, name=name, seed=8)
    _, b, _, _, _ = _setup(dev, n=n, name=name, seed=8)
    fa, fb = dev.alloc(n), dev.alloc(n)
    rng = np.random.default_rng(5)
    served = []
    for t, clearance in enumerate((0.1, 1.0, 2.0, 4.0)):  # from mostly served to mostly fallen back
        act = rng.uniform(-1.2, 1.2, (2, n)).astype(np.float32)
        a.actions.write(act)
        b.actions.write(act)
        sur.set_fused(True)
        sur.step(a, min_clearance=clearance, fallback=fa)
        sur.set_fused(False)
        sur.step(b, min_clearance=clearance, fallback=fb)
        for buf in ("state", "obs", "reward", "done"):
            assert np.array_equal(getattr(a, buf).read(), getattr(b, buf).read()), (t, buf)
        assert np.array_equal(fa.read(), fb.read()), t
        served.append(float((fa.read() == 0).mean()))
    sur.set_fused(True)
    print(f"served by the network per step: {served}", flush=True)
    assert max(served) > 0.5 and min(served) < 0.9, served


def _one_step_error(dev, sur, sim, n, rng):
    """Position error (m) of one surrogate step against the float64 reference over
    cars that neither fall back nor crash."""
    s = sim.state.read().reshape(11, n)
    st = np.ascontiguousarray(s[[0, 1, 2, 3, 4, 5]].astype(np.float64))
    act = rng.uniform(-1, 1, (2, n)).astype(np.float32)
    sim.actions.write(act)
    buf, ab = _f64_buffer(dev, st), dev.upload(act)
    sim.step_f64(buf, R.mat(ab, 2, n))
    ref = _read_f64(buf, n)
    fallback = dev.alloc(n)
    sur.step(sim, min_clearance=0.1, fallback=fallback)
    s1 = sim.state.read().reshape(11, n)
    ok = (fallback.read() == 0) & (sim.done.read() == 0)
    err = np.hypot(s1[0] - ref[0], s1[1] - ref[1])[ok]
    return float(np.sqrt(np.mean(err ** 2))), float(ok.mean())


def test_training_gains_over_untrained_network(dev):
    print("surrogate training", flush=True)
    n = 2048
    _, sim, pol, acts, scratch = _setup(dev, n=n, beams=64, hidden=64)
    data = R.Dataset(dev, 80 * n)
    data.record(sim, pol, acts, scratch, 80, 0)
    trained = R.Surrogate(dev, seed=1, precision=1, cols=n)
    untrained = R.Surrogate(dev, seed=1, precision=1, cols=n)
    trained.fit_normalization(data)
    untrained.fit_normalization(data)
    before = trained.loss(data)
    trained.train(data, epochs=4, batch=2048, lr=2e-3)
    after = trained.loss(data)
    print(f"loss before {before:.4g} after {after:.4g}", flush=True)
    assert after < 0.05 * before, (before, after)
    rng = np.random.default_rng(9)
    _, sim_a, _, _, _ = _setup(dev, n=n, beams=64, hidden=64, seed=21)
    _, sim_b, _, _, _ = _setup(dev, n=n, beams=64, hidden=64, seed=21)
    for t in range(6):  # the same start states for both networks
        for s in (sim_a, sim_b):
            s.actions.write(np.full((2, n), 0.3, np.float32))
            s.step()
    e_trained, served = _one_step_error(dev, trained, sim_a, n, np.random.default_rng(10))
    e_untrained, _ = _one_step_error(dev, untrained, sim_b, n, np.random.default_rng(10))
    print(f"one-step position error: trained {e_trained:.3g} m, untrained {e_untrained:.3g} m, "
          f"served by the network {served:.3f}", flush=True)
    assert e_trained < 0.1 * e_untrained, (e_trained, e_untrained)
    assert served > 0.9, served


ROWS = ("state", "reward", "done")


def _obs_rows(sim, beams):
    """obs rows 0-3 and the bias row: what a step writes besides the lidar's ranges."""
    o = sim.obs.read().reshape(beams + 5, sim.n)
    return np.concatenate([o[:4], o[-1:]])


@pytest.mark.parametrize("name,n", [("bank10", 1000), ("bridge", 333)])
def test_imagined_steps_match_single_steps(dev, name, n):
    """Eight open-loop imagined steps in one launch (weights loaded once)
    against eight single fused surrogate steps with the same actions: the
    same state, reward, done, obs rows and fallback cars, bit for bit."""
    print(f"surrogate imagine {name} n={n}", flush=True)
    beams = 64
    _, sim, pol, acts, scratch = _setup(dev, n=n, beams=beams, name=name)
    data = R.Dataset(dev, 24 * n)
    data.record(sim, pol, acts, scratch, 24, 0)
    sur = R.Surrogate(dev, seed=2, precision=1, cols=n)
    sur.fit_normalization(data)
    sur.train(data, epochs=1, batch=n, lr=2e-3)
    rng = np.random.default_rng(6)
    actions = rng.uniform(-1.2, 1.2, (8, 2, n)).astype(np.float32)
    for clearance in (0.1, 2.0):
        _, a, _, _, _ = _setup(dev, n=n, beams=beams, name=name, seed=8)
        _, b, _, _, _ = _setup(dev, n=n, beams=beams, name=name, seed=8)
        fa, fb = dev.alloc(n), dev.alloc(n)
        sur.imagine(a, R.mat(dev.upload(actions), 16, n), 8, min_clearance=clearance, fallback=fa)
        fell = np.zeros(n, bool)
        for k in range(8):
            b.actions.write(actions[k])
            sur.step(b, min_clearance=clearance, fallback=fb)
            fell |= fb.read() > 0
        for buf in ROWS:
            assert np.array_equal(getattr(a, buf).read(), getattr(b, buf).read()), (clearance, buf)
        assert np.array_equal(_obs_rows(a, beams), _obs_rows(b, beams)), clearance
        assert np.array_equal(fa.read() > 0, fell), clearance
        print(f"min clearance {clearance}: fell back {fell.mean():.3f}", flush=True)


@pytest.mark.parametrize("name,n", [("ramp10", 700), ("bridge8", 1024)])
def test_sim_imagine_matches_steps(dev, name, n):
    """Eight exact physics steps in one launch against eight Sim.step calls."""
    print(f"sim imagine {name} n={n}", flush=True)
    beams = 64
    cfg = R.Config(beams=beams, heading_noise=0.1)
    a, b = R.Sim(dev, track(name), n, 4, cfg), R.Sim(dev, track(name), n, 4, cfg)
    a.reset()
    b.reset()
    actions = np.random.default_rng(7).uniform(-1.2, 1.2, (8, 2, n)).astype(np.float32)
    a.imagine(R.mat(dev.upload(actions), 16, n), 8)
    for k in range(8):
        b.actions.write(actions[k])
        b.step()
    for buf in ROWS:
        assert np.array_equal(getattr(a, buf).read(), getattr(b, buf).read()), buf
    assert np.array_equal(_obs_rows(a, beams), _obs_rows(b, beams))


def test_imagine_errors(dev):
    cfg = R.Config(beams=32)
    sim = R.Sim(dev, track("flat"), 64, 1, cfg)
    acts = dev.alloc(2 * 64)
    with pytest.raises(R.RacerError):
        sim.imagine(R.mat(acts, 2, 64), 2)  # two steps need four action rows
    sur = R.Surrogate(dev, seed=1, precision=0, cols=64)
    with pytest.raises(R.RacerError, match="fused bf16"):
        sur.imagine(sim, R.mat(acts, 2, 64), 1)
