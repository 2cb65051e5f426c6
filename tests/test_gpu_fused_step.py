"""The fused step (physics and lidar in one launch, each car's track window
staged in LDS) against the oracle and against the two-launch step
(racer/sim.mojo:314-321): bit-identical state, obs, rewards and done codes."""
import numpy as np
import pytest

import oracle as O
import warporacer3d_amd as R
from helpers import track

pytestmark = pytest.mark.gpu


def _assert_same(g: R.Sim, o: O.Sim, what: str):
    for label, gb, ob in (("state", g.state, o.state), ("obs", g.obs, o.obs), ("reward", g.reward, o.reward),
                          ("done", g.done, o.done)):
        got = gb.read().reshape(ob.shape)
        if not np.array_equal(got, ob, equal_nan=True):
            bad = np.argwhere(got != ob)
            raise AssertionError(f"{what}: {label} differs at {len(bad)} places, first {bad[:3].tolist()} "
                                 f"gpu {got[tuple(bad[0])]} oracle {ob[tuple(bad[0])]}")


@pytest.mark.parametrize("name,beams,waves,noise", [("flat", 108, 1, 0.0), ("ramp10", 1081, 2, 0.1),
                                                    ("bridge8", 1081, 2, 0.0), ("bridge", 64, 1, 0.0),
                                                    ("bank10", 256, 2, 0.0), ("ring10", 108, 2, 0.1)])
def test_fused_step_matches_oracle(dev, name, beams, waves, noise):
    n = 100  # the last workgroup holds four cars
    print(f"fused step {name} beams={beams} waves={waves}", flush=True)
    O.set_threads(8)
    tr = track(name)
    cfg = R.Config(beams=beams, heading_noise=noise)
    g, o = R.Sim(dev, tr, n, 42, cfg), O.Sim(tr, n, 42, beams=beams, heading_noise=noise)
    g.set_waves_per_car(waves)
    g.set_fused_step(True)
    assert g.fused_step
    g.reset()
    o.reset()
    rng = np.random.default_rng(5)
    for t in range(6):
        a = rng.uniform(-1.2, 1.2, (2, n)).astype(np.float32)
        g.actions.write(a)
        o.actions[:] = a
        g.step()
        o.step()
        _assert_same(g, o, f"step {t}")


def test_fused_step_outside_the_window(dev):
    """Cars whose step leaves the LDS window (speeds far above the clamp, so
    the first substep jumps several cells), cars at the map's edges and off
    it, and cars at rest: the probes that fall outside the window read the
    track array, with the same results."""
    print("fused step outside the window", flush=True)
    O.set_threads(8)
    name, beams, n = "bridge8", 128, 64
    tr = track(name)
    cfg = R.Config(beams=beams)
    g, o = R.Sim(dev, tr, n, 7, cfg), O.Sim(tr, n, 7, beams=beams)
    g.set_fused_step(True)
    g.reset()
    o.reset()
    s = o.state.copy()
    hdr = tr[:12]
    nx, ny, x0, y0, cell = int(hdr[0]), int(hdr[1]), hdr[2], hdr[3], hdr[4]
    xs = np.float32(x0) + np.float32(cell) * np.array([0, 0.4, nx - 1.2, nx - 0.6, nx + 3, -5], np.float32)
    ys = np.float32(y0) + np.float32(cell) * np.array([0, ny - 1.1, 0.2, ny - 0.5, 2, ny / 2], np.float32)
    s[0, :6], s[1, :6] = xs, ys
    s[4, 6:20] = np.linspace(-60, 60, 14, dtype=np.float32)  # speed
    s[3, 6:20] = np.linspace(-3, 3, 14, dtype=np.float32)
    s[4, 20:24] = 0.0
    g.state.write(s)
    o.state[:] = s
    rng = np.random.default_rng(11)
    for t in range(3):
        a = rng.uniform(-1, 1, (2, n)).astype(np.float32)
        g.actions.write(a)
        o.actions[:] = a
        g.step()
        o.step()
        _assert_same(g, o, f"step {t}")


@pytest.mark.parametrize("name", ["ramp10", "bridge8"])
def test_fused_equals_pair_with_counts(dev, name):
    """Fused and two-launch steps give the same bits and the same counted
    lidar march iterations, at the benchmark's beam count."""
    print(f"fused vs pair {name}", flush=True)
    tr = track(name)
    cfg = R.Config(beams=1081, heading_noise=0.1)
    n = 1000
    a, b = R.Sim(dev, tr, n, 3, cfg), R.Sim(dev, tr, n, 3, cfg)
    b.set_fused_step(True)
    assert not a.fused_step and b.fused_step
    a.reset()
    b.reset()
    a.count_iterations(True)
    b.count_iterations(True)
    rng = np.random.default_rng(2)
    for t in range(20):
        act = rng.uniform(-1, 1, (2, n)).astype(np.float32)
        a.actions.write(act)
        b.actions.write(act)
        a.step()
        b.step()
        assert a.last_iterations == b.last_iterations > 0, t
    for x, y in ((a.state, b.state), (a.obs, b.obs), (a.reward, b.reward), (a.done, b.done)):
        assert np.array_equal(x.read(), y.read(), equal_nan=True)


def test_fused_step_applies_where_stated(dev):
    tr1, tr2 = track("ramp10"), track("bridge8")
    cfg = R.Config(beams=1081)
    s = R.Sim(dev, tr1, 64, 1, cfg)
    assert not s.fused_step
    s.set_fused_step(True)
    for w, want in ((1, True), (2, True), (4, False)):
        s.set_waves_per_car(w)
        assert s.fused_step == want, w
    s.set_waves_per_car(2)
    s.set_schedule(1)
    assert not s.fused_step
    s2 = R.Sim(dev, tr2, 64, 1, cfg)
    s2.set_fused_step(True)
    assert s2.fused_step
