"""Launch-shape choices and the autotuners that pick among them: every choice
(lidar wavefronts per car, cooperative or per-wavefront range stores, fused or
two-launch sim step, forward kernel of the Runner step) gives bitwise the same
Runner steps, and the autotuners leave the sim's buffers as they found them."""
import numpy as np
import pytest

import warporacer3d_amd as R
from helpers import track

pytestmark = pytest.mark.gpu

BUFFERS = ("state", "obs", "actions", "reward", "done")


def _snap(run: R.Runner) -> dict:
    out = {k: getattr(run.sim, k).read().copy() for k in BUFFERS}
    out["scratch"] = run.scratch.read().copy()
    return out


def _same(a: dict, b: dict, what: str, keys=BUFFERS + ("scratch",)):
    for k in keys:
        assert np.array_equal(a[k], b[k], equal_nan=True), f"{what}: {k} differs"


@pytest.mark.parametrize("precision,n,beams,hidden", [(0, 1000, 1081, 256), (1, 1000, 1081, 256),
                                                      (1, 333, 108, 64), (0, 512, 108, 64)])
def test_every_launch_choice_gives_the_same_steps(dev, precision, n, beams, hidden):
    print(f"launch choices precision={precision} n={n} beams={beams} H={hidden}", flush=True)
    cfg = R.Config(beams=beams, hidden=hidden, precision=precision, heading_noise=0.1)
    ref = None
    kernels = (R.ACT_AUTO, R.ACT_FUSED, R.ACT_WIDE, R.ACT_TILED) if precision == 1 else \
        (R.ACT_AUTO, R.ACT_FUSED, R.ACT_TILED)
    for kernel in kernels:
        for waves in (1, 2, 4):
            for shape in ("coop", "per-wavefront", "fused"):
                run = R.Runner(track("ramp10"), n, cfg, seed=9, device=dev)
                for t in range(3):
                    run.step(False, t)
                run.policy.set_act_kernel(kernel)
                run.sim.set_waves_per_car(waves)
                run.sim.set_fused_step(shape == "fused")
                run.sim.set_lidar_coop(shape == "coop")
                for t in range(3, 7):
                    run.step(False, t)
                got = _snap(run)
                if ref is None:
                    ref = got
                else:
                    _same(ref, got, f"kernel {kernel} waves {waves} {shape}")
                del run


@pytest.mark.parametrize("precision", [0, 1])
def test_runner_autotune_restores_the_sim(dev, precision):
    print(f"runner autotune precision={precision}", flush=True)
    cfg = R.Config(beams=1081, hidden=256, precision=precision, heading_noise=0.1)
    a = R.Runner(track("bridge8"), 1024, cfg, seed=4, device=dev)
    b = R.Runner(track("bridge8"), 1024, cfg, seed=4, device=dev)
    for t in range(3):
        a.step(False, t)
        b.step(False, t)
    before = _snap(a)
    choice = a.autotune(steps=4)
    assert choice["waves_per_car"] in (1, 2, 4) and a.sim.waves_per_car == choice["waves_per_car"]
    assert a.sim.fused_step == choice["fused_step"] and a.policy.act_kernel == choice["act_kernel"]
    assert a.sim.lidar_coop == choice["lidar_coop"]
    assert choice["act_kernel"] in ((R.ACT_FUSED, R.ACT_WIDE, R.ACT_TILED) if precision else
                                    (R.ACT_FUSED, R.ACT_TILED))
    _same(before, _snap(a), "after autotune", BUFFERS)
    for t in range(3, 8):
        a.step(False, t)
        b.step(False, t)
    _same(_snap(b), _snap(a), "steps after autotune")


def test_sim_autotune_restores_the_sim(dev):
    print("sim autotune", flush=True)
    cfg = R.Config(beams=512)
    a = R.Sim(dev, track("ramp10"), 2048, 5, cfg)
    b = R.Sim(dev, track("ramp10"), 2048, 5, cfg)
    a.reset()
    b.reset()
    rng = np.random.default_rng(3)
    act = rng.uniform(-1, 1, (2, 2048)).astype(np.float32)
    a.actions.write(act)
    b.actions.write(act)
    before = {k: getattr(a, k).read().copy() for k in BUFFERS}
    chosen = a.autotune()
    assert chosen == a.waves_per_car in (1, 2, 4)
    for k in BUFFERS:
        assert np.array_equal(before[k], getattr(a, k).read(), equal_nan=True), k
    for _ in range(4):
        a.step()
        b.step()
    for k in BUFFERS:
        assert np.array_equal(getattr(a, k).read(), getattr(b, k).read(), equal_nan=True), k


def test_batch_autotune_restores_every_instance(dev):
    print("batch autotune", flush=True)
    cfg = R.Config(beams=1081, hidden=16)
    names, cars = ("ring10", "ramp10", "bank10", "bridge8"), (256, 512, 700, 1000)
    sims = [R.Sim(dev, track(k), n, 10 + i, cfg) for i, (k, n) in enumerate(zip(names, cars))]
    alone = [R.Sim(dev, track(k), n, 10 + i, cfg) for i, (k, n) in enumerate(zip(names, cars))]
    for s in sims + alone:
        s.reset()
    batch = R.Batch(dev, sims, cfg)
    before = [{k: getattr(s, k).read().copy() for k in BUFFERS} for s in sims]
    chosen = batch.autotune()
    assert chosen == batch.waves_per_car in (1, 2, 4)
    assert batch.fused_step in ((0, 1, 2) if chosen < 4 else (0,))
    for s, b in zip(sims, before):
        for k in BUFFERS:
            assert np.array_equal(b[k], getattr(s, k).read(), equal_nan=True), k
    batch.step()
    for s in alone:
        s.step()
    for s, o in zip(sims, alone):
        for k in BUFFERS:
            assert np.array_equal(getattr(s, k).read(), getattr(o, k).read(), equal_nan=True), k


@pytest.mark.parametrize("precision", [0, 1])
def test_batch_runner_autotune_restores_every_instance(dev, precision):
    print(f"batch runner autotune precision={precision}", flush=True)
    cfg = R.Config(beams=1081, hidden=64, precision=precision, heading_noise=0.1)
    pol = R.Policy(dev, 42, cfg)
    names, cars = ("ring10", "ramp10", "bridge8"), (256, 333, 700)
    sims = [R.Sim(dev, track(k), n, 10 + i, cfg) for i, (k, n) in enumerate(zip(names, cars))]
    alone = [R.Sim(dev, track(k), n, 10 + i, cfg) for i, (k, n) in enumerate(zip(names, cars))]
    for s in sims + alone:
        s.reset()
    batch = R.Batch(dev, sims, cfg)
    before = [{k: getattr(s, k).read().copy() for k in BUFFERS} for s in sims]
    chosen = batch.autotune_runner(pol, 4)
    assert chosen["waves_per_car"] == batch.waves_per_car in (1, 2, 4)
    assert chosen["fused_step"] == batch.fused_step in (0, 1, 2)
    for s, b in zip(sims, before):
        for k in BUFFERS:
            assert np.array_equal(b[k], getattr(s, k).read(), equal_nan=True), k
    acts = [R.Activations(dev, s.n, cfg) for s in alone]
    scratch = [dev.alloc(2 * s.n) for s in alone]
    for t in range(2):
        batch.runner_step(pol, False, t)
        for s, a, sc in zip(alone, acts, scratch):
            R.runner_step(s, pol, a, sc, False, t)
    for s, o in zip(sims, alone):
        for k in BUFFERS:
            assert np.array_equal(getattr(s, k).read(), getattr(o, k).read(), equal_nan=True), k


def test_act_kernel_rejects_unknown(dev):
    pol = R.Policy(dev, 1, R.Config(beams=32, hidden=16))
    with pytest.raises(R.RacerError):
        pol.set_act_kernel(7)
    assert pol.act_kernel == R.ACT_AUTO
