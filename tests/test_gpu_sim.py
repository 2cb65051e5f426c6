"""HIP simulation against the oracle: bit-identical state, rewards, done codes
and lidar ranges on the same seeds (racer/sim.mojo:170-348)."""
import numpy as np
import pytest

import oracle as O
import warporacer3d_amd as R
from helpers import corridor_state, track

pytestmark = pytest.mark.gpu


def _pair(dev, name, n, seed, beams, noise=0.0):
    tr = track(name)
    cfg = R.Config(beams=beams, heading_noise=noise)
    return R.Sim(dev, tr, n, seed, cfg), O.Sim(tr, n, seed, beams=beams, heading_noise=noise)


def _assert_same(g: R.Sim, o: O.Sim, what: str):
    for label, gb, ob in (("state", g.state, o.state), ("obs", g.obs, o.obs), ("reward", g.reward, o.reward),
                          ("done", g.done, o.done)):
        got = gb.read().reshape(ob.shape)
        if not np.array_equal(got, ob, equal_nan=True):
            bad = np.argwhere(got != ob)
            raise AssertionError(f"{what}: {label} differs at {len(bad)} places, first {bad[:3].tolist()} "
                                 f"gpu {got[tuple(bad[0])]} oracle {ob[tuple(bad[0])]}")


@pytest.mark.parametrize("name,beams,noise", [("flat", 108, 0.0), ("ramp10", 1081, 0.1), ("bridge8", 108, 0.0),
                                              ("bridge", 64, 0.0), ("bank10", 256, 0.0)])
def test_reset_and_steps_bit_identical(dev, name, beams, noise):
    print(f"sim parity {name} beams={beams}", flush=True)
    O.set_threads(8)
    g, o = _pair(dev, name, 96, 42, beams, noise)
    g.reset()
    o.reset()
    _assert_same(g, o, "reset")
    rng = np.random.default_rng(1)
    for t in range(5):
        a = rng.uniform(-1, 1, (2, 96)).astype(np.float32)
        g.actions.write(a)
        o.actions[:] = a
        g.step()
        o.step()
        _assert_same(g, o, f"step {t}")


def test_corridor_suite_on_gpu(dev):
    """tests/test_mojo.mojo:58-92 through the HIP path."""
    print("corridor suite", flush=True)
    beams = 108
    sim = R.Sim(dev, track("corridor"), 4, 1, R.Config(beams=beams))
    s, actions = corridor_state()
    sim.state.write(s)
    sim.actions.write(actions)
    sim.step()
    o = sim.obs.read().reshape(beams + 5, 4)
    assert abs(o[4 + beams // 2, 0] * 10 - 7.1) <= 0.1
    assert abs(o[4 + beams // 2, 1] * 10 - 0.73) <= 0.1
    crashed = False
    for _ in range(60):
        sim.step()
        d = sim.done.read()
        assert d[0] == 0 and d[2] == 0 and d[3] == 0
        if d[1] == 1 and not crashed:
            st = sim.state.read().reshape(11, 4)
            assert st[6, 1] == 0 and st[8, 1] == 1
            crashed = True
    assert crashed
    st = sim.state.read().reshape(11, 4)
    o = sim.obs.read().reshape(beams + 5, 4)
    assert st[0, 0] == 1 and st[4, 0] == 0 and st[3, 0] == 0
    assert st[4, 2] == 5.0 and st[0, 2] > 3.5 and abs(st[1, 2]) < 1e-4
    assert abs(st[10, 2] - 10 * (st[0, 2] - 1)) <= 0.01
    assert st[1, 3] > 0.1 and st[0, 3] < st[0, 2]
    assert o[1, 2] == 1 and o[beams + 4, 2] == 1 and np.isfinite(o).all()


def test_200_steps_events_identical(dev):
    """Crash, respawn and progress events (and the whole state) identical over
    200 steps on the ramp ribbon at 1081 beams, seeded random actions."""
    print("200-step events ramp10 1081 beams", flush=True)
    O.set_threads(8)
    n, beams = 128, 1081
    g, o = _pair(dev, "ramp10", n, 7, beams, 0.1)
    g.reset()
    o.reset()
    rng = np.random.default_rng(11)
    events = 0
    for t in range(200):
        a = rng.uniform(-1, 1, (2, n)).astype(np.float32)
        a[1] = np.abs(a[1])  # mostly forward, so cars reach walls and crash
        g.actions.write(a)
        o.actions[:] = a
        g.step()
        o.step()
        gd = g.done.read()
        assert np.array_equal(gd, o.done), f"done codes differ at step {t}"
        events += int((gd != 0).sum())
        if t % 50 == 49:
            _assert_same(g, o, f"step {t}")
    _assert_same(g, o, "step 200")
    assert events > 0, "the run should contain crash/respawn events"


def test_sample_bit_identical(dev):
    print("sample parity", flush=True)
    n = 300
    g, o = _pair(dev, "flat", n, 42, 16)
    rng = np.random.default_rng(5)
    out = rng.normal(size=(3, n)).astype(np.float32)
    log_std = np.array([-0.5, -0.3], np.float32)
    ob = dev.upload(out)
    lb = dev.upload(log_std)
    act, lp, val = dev.alloc(2 * n), dev.alloc(n), dev.alloc(n)
    for det in (False, True):
        g.sample(R.mat(ob, 3, n), R.mat(lb, 1, 2), R.mat(act, 2, n), R.mat(lp, 1, n), R.mat(val, 1, n), det, 17)
        oa, ol, ov = np.zeros((2, n), np.float32), np.zeros(n, np.float32), np.zeros(n, np.float32)
        o.sample(out, log_std, oa, ol, ov, det, 17)
        assert np.array_equal(act.read().reshape(2, n), oa)
        assert np.array_equal(lp.read(), ol) and np.array_equal(val.read(), ov)


@pytest.mark.parametrize("name,beams", [("ramp10", 1081), ("bridge8", 108)])
def test_march_iteration_count(dev, name, beams):
    """The instrumented lidar counts exactly the oracle's march iterations."""
    print(f"iteration count {name}", flush=True)
    O.set_threads(8)
    g, o = _pair(dev, name, 64, 3, beams)
    g.reset()
    o.reset()
    g.count_iterations(True)
    g.step()
    it = o.step()
    g.count_iterations(False)
    assert g.last_iterations == it > 0


def test_views_and_errors(dev):
    sim = R.Sim(dev, track("flat"), 8, 1, R.Config(beams=16))
    with pytest.raises(R.RacerError, match="rows"):
        sim.step(obs=R.mat(sim.obs, 20, 8))
    with pytest.raises(R.RacerError, match="not a track array"):
        R.Sim(dev, np.zeros(5, np.float32), 8, 1)
    with pytest.raises(R.RacerError, match="exceeds"):
        sim.step(reward=R.mat(sim.reward, 1, 8, offset=4))


def test_reference_schedule_identical(dev):
    """The thread-per-(car, beam) schedule gives the same bits as the wave-per-car one."""
    print("lidar schedules", flush=True)
    tr = track("bridge8")
    cfg = R.Config(beams=300)
    a, b = R.Sim(dev, tr, 64, 9, cfg), R.Sim(dev, tr, 64, 9, cfg)
    b.set_schedule(1)
    a.reset()
    b.reset()
    for _ in range(3):
        a.step()
        b.step()
    assert np.array_equal(a.obs.read(), b.obs.read())
    assert np.array_equal(a.state.read(), b.state.read())


@pytest.mark.parametrize("waves", [1, 2, 4])
def test_waves_per_car_identical(dev, waves):
    """Splitting a car's beams over 1, 2 or 4 wavefronts gives the oracle's bits."""
    print(f"waves per car {waves}", flush=True)
    O.set_threads(8)
    g, o = _pair(dev, "bridge8", 40, 13, 1081)
    g.set_waves_per_car(waves)
    g.reset()
    o.reset()
    for _ in range(2):
        g.step()
        o.step()
    _assert_same(g, o, f"waves {waves}")


def test_autotune_keeps_results(dev):
    g, o = _pair(dev, "ramp10", 64, 21, 1081)
    g.reset()
    o.reset()
    chosen = g.autotune()
    assert g.schedule == 0 and chosen in (1, 2, 4) and g.waves_per_car == chosen
    _assert_same(g, o, "after autotune")
    g.step()
    o.step()
    _assert_same(g, o, "step after autotune")


@pytest.mark.parametrize("name", ["ramp10", "bridge8"])
def test_lidar_quotient_paths_bit_identical(dev, name):
    """The march step takes fl / descent from the beam's reciprocal only for
    cars whose height and slopes keep that correctly rounded, and divides for
    the others. Cars pushed past those bounds (|z| > 2^40, |along| + |across|
    > 2^20) and cars just inside them share one launch; every range against
    the oracle's scan of the same beam."""
    print(f"lidar quotient paths {name}", flush=True)
    n, beams = 64, 256
    g, _ = _pair(dev, name, n, 5, beams)
    g.reset()
    s = g.state.read().reshape(11, n).copy()
    obs = g.obs.read().reshape(beams + 5, n).copy()
    s[2, 0:8] += 2.0**41  # z beyond 2^40: divides
    obs[2, 8:16] = 2.0**21  # along beyond 2^20: divides
    obs[3, 16:20] = -(2.0**21)  # across beyond 2^20: divides
    s[2, 20:28] += 2.0**39  # z inside: reciprocal, with large dividends
    obs[2, 28:32] = 2.0**19  # steep beams inside: reciprocal, with small divisors
    g.state.write(s)
    g.obs.write(obs)
    g.scan()  # the lidar alone on the current state
    got = g.obs.read().reshape(beams + 5, n)[4:4 + beams]
    tr, s, obs = O.f32(track(name)), O.f32(s), O.f32(obs)
    want = np.empty((beams, n), np.float32)
    for i in range(n):
        for b in range(beams):
            r = O.lib().o_scan_one(O.ptr(tr), tr.size, O.ptr(s), n, O.ptr(obs), n, i, b, beams, None)
            want[b, i] = np.float32(r) / np.float32(10.0)  # RACER_RANGE
    bad = np.argwhere(got != want)
    assert len(bad) == 0, f"{len(bad)} ranges differ, first (beam, car) {bad[:3].tolist()}"


def test_padded_probe_floor_instruction(dev):
    """The padded lidar probe takes the floor and the integer conversion of
    its nearest cell as one instruction: for every float x with |x| < 2^23,
    both signs (2.1 G values), the same integer as (int)floorf(x)."""
    print("floor-and-convert instruction over |x| < 2^23", flush=True)
    assert dev.check_floor_index(1) == 0


def test_objects_outlive_the_device_handle():
    """Destroying the device handle first (a garbage collector may finalise a
    cycle in any order) leaves the objects made on it working; the device goes
    with the last of them."""
    print("device handle destroyed before its objects", flush=True)
    d = R.Device(True)
    cfg = R.Config(beams=32, hidden=16, heading_noise=0.1)
    sim = R.Sim(d, track("ramp"), 64, 3, cfg)
    sim.reset()
    sur = R.Surrogate(d, precision=1, cols=64)
    buf = d.alloc(64)
    R.lib().racer_device_destroy(d.handle)
    d.handle = None
    sim.step()
    sur.step(sim, 0.1, buf)
    state = sim.state.read()
    assert np.all(np.isfinite(state))
    del sur, buf
    sim.step()
    del sim


@pytest.mark.parametrize("name,beams,waves,coop", [
    ("ramp10", 2000, 1, True),    # 32 pieces in one share: more than the order covers (natural order)
    ("ramp10", 1600, 4, True),    # shares of 400 beams, each starting inside a 64-beam chunk
    ("bridge8", 65, 2, False),    # a one-beam second piece; per-wavefront stores
    ("ring10", 2, 2, True),       # the fewest beams: one beam a wavefront, 63 spare records each
    ("bank10", 1081, 4, False),   # shares of 271 beams, 5 pieces each, per-wavefront stores
])
def test_lidar_piece_order_edge_shapes(dev, name, beams, waves, coop):
    """The lidar's piece order (longest first by the previous step's keys,
    spare records past a piece's end) at share and piece shapes the benchmark
    never uses: every step bit-identical to the oracle over steps whose keys
    come from the steps before."""
    print(f"lidar piece order {name} beams={beams} waves={waves} coop={coop}", flush=True)
    g, o = _pair(dev, name, 40, 5, beams, 0.1)
    g.set_waves_per_car(waves)
    g.set_lidar_coop(coop)
    g.reset()
    o.reset()
    _assert_same(g, o, "reset")
    rng = np.random.default_rng(beams)
    for t in range(4):
        a = rng.uniform(-1, 1, (2, 40)).astype(np.float32)
        g.actions.write(a)
        o.actions[...] = a
        g.step()
        o.step()
        _assert_same(g, o, f"step {t}")
