"""Benchmark of the racer hot path on one MI355X.

Workload (one step = Runner.step of main.mojo:183-191 for every car: policy
forward, Gaussian sample, RK4 physics with crash/respawn/progress, lidar):
ramp ribbon of radius 10 m with road height 0.3 sin(3 theta), 0.025 m grids,
4096 cars x 1081 beams over 270 degrees, random-init 2 x 256 tanh policy,
spawn heading noise U(-0.1, 0.1). Inputs are resident in HBM before timing.

The headline `value` is the float32 step (float32 policy products, bitwise
the CPU restatement); the bf16-policy step is reported beside it (`bf16`).
Prints one JSON line with env-steps/s, the roofline of the dominant kernel
(measured with HIP events on the library's stream right after the timed
steps), the policy forward's MFMA roofline and the CPU baseline (the oracle's
float32 restatement of the same workload, timed on the host's threads and on
one).
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "env-steps/s at 4096 cars x 1081 beams; PPO iter seconds; % HBM/MFMA peak"
HBM_PEAK_GBS = 8000.0
MFMA_BF16_PEAK_TFLOPS = 2500.0
FP32_PEAK_TFLOPS = 157.3


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def lidar_bytes(iterations: float, cars: int, beams: int, layers: int) -> float:
    """Algorithmic bytes of one lidar launch: one packed (clearance, height)
    record of 8 B per layer per march iteration, the 4 B range store per beam,
    and 6 floats of car state read per car."""
    return iterations * 8.0 * layers + 4.0 * cars * beams + 24.0 * cars


def policy_flops(cars: int, beams: int, hidden: int) -> float:
    IN, HA = beams + 5, hidden + 1
    return 2.0 * cars * (hidden * IN + hidden * HA + 3 * HA)


def cpu_baseline(track, beams: int, hidden: int, cars: int, threads: int, steps: int, steps_1core: int,
                 track_name: str = "ramp10") -> dict:
    """The oracle's Runner step (float32 policy forward, sample, physics,
    lidar) on the host at the benchmark's own size, timed like the reference's
    `racer benchmark ... cpu` (main.mojo:225-235: a warm-up step, then timed
    steps): `steps` steps on `threads` threads, then `steps_1core` on one."""
    import numpy as np

    import oracle as O

    sim = O.Sim(track, cars, 42, beams=beams, heading_noise=0.1)
    O.set_threads(threads)
    sim.reset()
    pol = O.Policy(42, beams, hidden)
    logp, value = np.zeros(cars, np.float32), np.zeros(cars, np.float32)

    def step(t):
        _, _, out = pol.forward(sim.obs)
        sim.sample(out, pol.log_std(), sim.actions, logp, value, False, t)
        sim.step()

    def timed(n, t0_step):
        t0 = time.perf_counter()
        for k in range(n):
            step(t0_step + k)
        return time.perf_counter() - t0

    step(0)
    dt = timed(steps, 1)
    O.set_threads(1)
    dt1 = timed(steps_1core, 1 + steps)
    O.set_threads(threads)
    return {"value": cars * steps / dt, "unit": "env-steps/s", "cores": threads, "nproc": os.cpu_count(),
            "kind": "port", "value_1core": cars * steps_1core / dt1,
            "ms_per_step": dt * 1000.0 / steps, "ms_per_step_1core": dt1 * 1000.0 / steps_1core,
            "sample": f"the benchmark workload itself ({cars} cars x {beams} beams, 2x{hidden} float32 policy, "
                      f"{track_name}): {steps} Runner steps on {threads} host threads ({dt:.1f} s) and "
                      f"{steps_1core} on 1 ({dt1:.1f} s), after one warm-up step; host has {os.cpu_count()} "
                      f"logical CPUs"}


def ppo_flops(cars: int, beams: int, hidden: int, rollout: int, epochs: int) -> float:
    """Flops of one PPO iteration: the rollout's (R + 1) forwards per car and the
    update's epochs x R x n samples of forward plus backward."""
    IN, HA, H = beams + 5, hidden + 1, hidden
    fwd = 2.0 * (H * IN + H * HA + 3 * HA)
    bwd = 2.0 * HA * (3 + H) + fwd
    return (rollout + 1) * cars * fwd + epochs * rollout * cars * (fwd + bwd)


def forward_precision(R, dev, pol, tr, cfg, cars: int, t: int = 64) -> dict:
    """Error of the policy outputs (steer mean, throttle mean, value) on the
    observations of rollout step t (all cars): the bf16 forward and a float32
    forward of the same weights, each against a float64 restatement of the
    three layers (numpy, in this function)."""
    import numpy as np

    IN, H, HA = cfg.IN, cfg.hidden, cfg.hidden + 1
    x = R.mat(tr.obs, IN, cars, t * IN * cars)
    acts_b = R.Activations(dev, cars, cfg)
    pol.forward(x, acts_b)
    cfg32 = R.Config(beams=cfg.beams, hidden=cfg.hidden, precision=0, heading_noise=0.1)
    pol32 = R.Policy(dev, 42, cfg32)
    pol32.theta.write(pol.theta.read())
    acts_f = R.Activations(dev, cars, cfg32)
    pol32.forward(x, acts_f)
    dev.synchronize()
    o_b = acts_b.o.read()[:3 * cars].reshape(3, cars).astype(np.float64)
    o_f = acts_f.o.read()[:3 * cars].reshape(3, cars).astype(np.float64)
    th = pol.theta.read().astype(np.float64)
    W0 = th[:H * IN].reshape(H, IN)
    W1 = th[H * IN:H * IN + H * HA].reshape(H, HA)
    W2 = th[H * IN + H * HA:H * IN + H * HA + 3 * HA].reshape(3, HA)
    X = x.read().astype(np.float64).reshape(IN, cars)
    ones = np.ones((1, cars))
    h1 = np.tanh(W0 @ X)
    h2 = np.tanh(W1 @ np.vstack([h1, ones]))
    o = W2 @ np.vstack([h2, ones])
    out = {}
    for r, name in enumerate(("steer_mean", "throttle_mean", "value")):
        scale = float(np.sqrt(np.mean(o[r] ** 2)))
        out[name] = {"rms_float64": scale,
                     "bf16": {"rms_error": float(np.sqrt(np.mean((o_b[r] - o[r]) ** 2))),
                              "max_error": float(np.max(np.abs(o_b[r] - o[r])))},
                     "float32": {"rms_error": float(np.sqrt(np.mean((o_f[r] - o[r]) ** 2))),
                                 "max_error": float(np.max(np.abs(o_f[r] - o[r])))}}
    return {"columns": cars, "rollout_step": t, "outputs": out}


def ppo_leg(R, T, dev, beams: int, iterations: int, cars: int = 16384, precision: int = 1) -> dict:
    """PPO iterations on the bridge (two 8 m lobes, deck 0.5 m above the
    crossing, two layers): 16384 cars, 128-step rollouts replayed from a HIP
    graph, 4 epochs x 8 minibatches, 2 x 256 policy, Adam 3e-4."""
    arr = T.cached("bridge8")
    cfg = R.Config(beams=beams, hidden=256, rollout=128, epochs=4, minibatches=8, precision=precision,
                   heading_noise=0.1)
    sim = R.Sim(dev, arr, cars, 42, cfg)
    sim.reset()
    pol = R.Policy(dev, 42, cfg)
    acts = R.Activations(dev, cars, cfg)
    tr = R.Trainer(dev, cars, cfg)
    R.autotune_runner(sim, pol, acts, dev.alloc(2 * cars))  # launch shapes of the rollout's steps
    graph = tr.capture_rollout(sim, pol, acts)
    tr.rollout_graph(graph, 0)  # untimed first iteration
    tr.update(pol)
    dev.synchronize()
    roll, upd, rewards, kls = [], [], [], []
    for it in range(1, iterations + 1):
        t0 = time.perf_counter()
        rewards.append(tr.rollout_graph(graph, it))
        t1 = time.perf_counter()
        kls.append(tr.update(pol))
        t2 = time.perf_counter()
        roll.append(t1 - t0)
        upd.append(t2 - t1)
    # one more iteration, eager and with an event pair per launch: kernel shares
    dev.profile_only(None)
    dev.profile_reset()
    dev.profile(True)
    tr.rollout(sim, pol, acts, iterations + 1)
    tr.update(pol)
    dev.profile(False)
    prof = dev.profile_results()
    top = sorted(prof.items(), key=lambda kv: -kv[1][1])[:10]
    precision_check = forward_precision(R, dev, pol, tr, cfg, cars) if precision == 1 else None
    it_s = (sum(roll) + sum(upd)) / iterations
    flops = ppo_flops(cars, beams, 256, 128, 4)
    del graph
    return {"iteration_s": it_s, "rollout_s": sum(roll) / iterations, "update_s": sum(upd) / iterations,
            "iterations": iterations, "env_steps_per_s_in_rollout": 128 * cars / (sum(roll) / iterations),
            "flops_per_iteration": flops, "tflops": flops / it_s / 1e12,
            "mean_reward": rewards, "kl": kls,
            "config": {"track": "bridge8 (two 8 m lobes, 2 layers)", "cars": cars, "beams": beams, "hidden": 256,
                       "rollout": 128, "epochs": 4, "minibatches": 8,
                       "precision": {0: "fp32", 1: "bf16+fp32", 2: "plain fp32"}[precision],
                       "lidar_waves_per_car": sim.waves_per_car, "lidar_coop_stores": sim.lidar_coop},
            "kernels_ms_per_iteration": {k: tot for k, (n, tot) in top},
            "forward_error_vs_float64": precision_check}


def train_parity_leg(R, T, dev, iterations: int, seeds: int, beams: int = 108, cars: int = 16384,
                     last: int = 20) -> dict:
    """The PPO training curve in bf16 (bf16 MFMA products, float32 accumulation)
    against float32 products, on the bridge with the cfg-3 settings (16384
    cars, 128-step rollouts, 4 epochs x 8 minibatches, 2 x 256, Adam 3e-4):
    the reward per step and the KL that the reference's training loop reports
    each iteration (main.mojo:60-68), from the same seeds. The reference
    computes no loss value; the statistic is the mean reward per step over
    the last `last` iterations, whose bf16 - fp32 difference is set beside the
    spread between seeds within one precision."""
    arr = T.cached("bridge8")
    runs = {}
    for precision in (1, 0):
        name = "bf16" if precision == 1 else "fp32"
        for seed in range(42, 42 + seeds):
            cfg = R.Config(beams=beams, hidden=256, rollout=128, epochs=4, minibatches=8, precision=precision,
                           heading_noise=0.1)
            sim = R.Sim(dev, arr, cars, seed, cfg)
            sim.reset()
            pol = R.Policy(dev, seed, cfg)
            acts = R.Activations(dev, cars, cfg)
            tr = R.Trainer(dev, cars, cfg)
            graph = tr.capture_rollout(sim, pol, acts)
            rewards, kls = [], []
            t0 = time.perf_counter()
            for it in range(iterations):
                rewards.append(tr.rollout_graph(graph, it))
                kls.append(tr.update(pol))
                if (it + 1) % 50 == 0:
                    log(f"train parity {name} seed {seed}: iteration {it + 1}, reward/step {rewards[-1]:.5f}, "
                        f"KL {kls[-1]:.5f}")
            dev.synchronize()
            seconds = time.perf_counter() - t0
            del graph
            tail = rewards[-last:]
            runs[f"{name}_seed{seed}"] = {
                "precision": name, "seed": seed, "seconds": seconds,
                "mean_reward_last": sum(tail) / len(tail), "mean_kl_last": sum(kls[-last:]) / len(tail),
                "reward_per_step": rewards[::10] + [rewards[-1]], "finite": all(k == k for k in kls)}
    bf = [r["mean_reward_last"] for r in runs.values() if r["precision"] == "bf16"]
    fp = [r["mean_reward_last"] for r in runs.values() if r["precision"] == "fp32"]
    per_seed = [abs(b - f) / max(abs(f), 1e-12) for b, f in zip(bf, fp)]
    mean_bf, mean_fp = sum(bf) / len(bf), sum(fp) / len(fp)
    spread = (max(fp) - min(fp)) / max(abs(mean_fp), 1e-12) if len(fp) > 1 else None
    return {"iterations": iterations, "seeds": seeds, "statistic": f"mean reward per step over the last {last} "
            "iterations", "beams": beams, "cars": cars, "track": "bridge8 (two 8 m lobes, 2 layers)",
            "bf16_mean": mean_bf, "fp32_mean": mean_fp,
            "relative_difference": abs(mean_bf - mean_fp) / max(abs(mean_fp), 1e-12),
            "relative_difference_per_seed": per_seed, "fp32_seed_spread": spread, "runs": runs}


BATCH_TRACKS = ("ring10", "ramp10", "bank10", "bridge8", "random0", "random1", "random2", "random3")
BATCH_CARS = (256, 512, 1024, 2048, 4096, 4096, 8192, 8192)


STEP_MODES = ("physics and lidar launches", "fused step", "fused step, persistent grid")


def batch_leg(R, T, dev, steps: int, separate: bool = True) -> dict:
    """Eight uneven instances (flat, ramp, bank, bridge and four random routes;
    256 .. 8192 cars, 1081 beams, one 2 x 256 bf16 policy): Runner steps of all
    of them from one captured graph of batched launches, against the same
    instances replayed from one graph each, one after another."""
    cfg = R.Config(beams=1081, hidden=256, precision=1, heading_noise=0.1)
    pol = R.Policy(dev, 42, cfg)

    def instances():
        sims = []
        for k, (name, n) in enumerate(zip(BATCH_TRACKS, BATCH_CARS)):
            s = R.Sim(dev, T.cached(name), n, 100 + k, cfg)
            s.reset()
            sims.append(s)
        return sims

    total = sum(BATCH_CARS)
    chunk = math.gcd(steps, 50)
    # batched: one launch per stage for all eight instances
    sims = instances()
    batch = R.Batch(dev, sims, cfg)
    for t in range(5):
        batch.runner_step(pol, False, t)
    # every launch shape of the batched step: 10 Runner steps replayed from a
    # graph, three times from the same saved sim states (the cost of a step
    # drifts as the cars spread over the tracks, so shapes are compared from
    # one state), the best replay counting
    names = ("state", "obs", "actions", "reward", "done")
    dev.synchronize()
    saved = [{k: getattr(s, k).read().copy() for k in names} for s in sims]

    def restore():
        for s, d in zip(sims, saved):
            for k in names:
                getattr(s, k).write(d[k])
        dev.synchronize()

    modes = {}
    for mode in (0, 1, 2):
        for w in (1, 2, 4):
            batch.set_waves_per_car(w)
            batch.set_fused_step(mode)
            if batch.fused_step != mode:
                continue
            g = batch.capture(pol, 10)
            best = math.inf
            for r in range(4):
                restore()
                t0 = time.perf_counter()
                g.launch(5)
                dev.synchronize()
                if r:
                    best = min(best, time.perf_counter() - t0)
            modes[f"{STEP_MODES[mode]}, {w} wavefronts per car"] = best * 1000.0 / 10
            del g
    log("batch step launch shapes (ms per Runner step, same state):", {k: round(v, 4) for k, v in modes.items()})
    restore()
    tuned = batch.autotune_runner(pol)
    wpc, mode, t = tuned["waves_per_car"], tuned["fused_step"], 5
    log("batch lidar wavefronts per car and step mode (library autotuner):", wpc, STEP_MODES[mode])
    graph = batch.capture(pol, chunk)
    graph.launch(t)
    dev.synchronize()
    t0 = time.perf_counter()
    for k in range(steps // chunk):
        graph.launch(t + chunk * (k + 1))
    dev.synchronize()
    batched_ms = (time.perf_counter() - t0) * 1000.0 / steps
    del graph, batch, sims
    result = {"instances": [{"track": t, "cars": n} for t, n in zip(BATCH_TRACKS, BATCH_CARS)], "beams": 1081,
              "cars": total, "steps": steps, "lidar_waves_per_car": wpc, "step_mode": STEP_MODES[mode],
              "launch_shapes_ms_per_step": modes,
              "batched_graph": {"ms_per_step": batched_ms, "env_steps_per_s": total * 1000.0 / batched_ms}}
    if not separate:
        return result
    # separate: each instance's own graph of Runner steps, replayed in turn
    sims = instances()
    own, t_own = [], 5
    for s in sims:
        acts = R.Activations(dev, s.n, cfg)
        scratch = dev.alloc(2 * s.n)
        for t in range(5):
            R.runner_step(s, pol, acts, scratch, False, t)
        R.autotune_runner(s, pol, acts, scratch)
        own.append(R.capture_runner(s, pol, acts, scratch, chunk))
    for g in own:
        g.launch(t_own)
    dev.synchronize()
    t0 = time.perf_counter()
    for k in range(steps // chunk):
        for g in own:
            g.launch(t_own + chunk * (k + 1))
    dev.synchronize()
    separate_ms = (time.perf_counter() - t0) * 1000.0 / steps
    del own, sims
    result["separate_graphs"] = {"ms_per_step": separate_ms, "env_steps_per_s": total * 1000.0 / separate_ms}
    result["speedup"] = separate_ms / batched_ms
    return result


def surrogate_leg(R, T, dev, transitions: int, epochs: int, cars: int = 16384) -> dict:
    """The learned dynamics surrogate on the bank track (0.1 m rise per metre
    outward): transitions from Runner steps of 16384 cars x 1081 beams with a
    random-init 2 x 256 bf16 policy, bf16 training of the 3 x 128 residual MLP,
    then 8-step imagined rollouts of 16384 cars (exact lidar, RK4 fallback below
    0.1 m clearance or on a layer change) measured against float64 RK4 on the GPU."""
    import numpy as np

    arr = T.cached("bank10")
    cfg = R.Config(beams=1081, hidden=256, precision=1, heading_noise=0.1)
    pol = R.Policy(dev, 42, cfg)
    sim = R.Sim(dev, arr, cars, 5, cfg)
    sim.reset()
    acts = R.Activations(dev, cars, cfg)
    scratch = dev.alloc(2 * cars)
    steps = max(1, transitions // cars)
    data = R.Dataset(dev, steps * cars)
    dev.synchronize()
    t0 = time.perf_counter()
    data.record(sim, pol, acts, scratch, steps, 0)
    dev.synchronize()
    record_s = time.perf_counter() - t0
    sur = R.Surrogate(dev, seed=1, precision=1, cols=cars)
    sur.fit_normalization(data)
    loss0 = sur.loss(data)
    t0 = time.perf_counter()
    sur.train(data, epochs=epochs, batch=cars, lr=2e-3)
    dev.synchronize()
    train_s = time.perf_counter() - t0
    loss1 = sur.loss(data)
    minibatches = epochs * (data.count // cars)

    # 8-step windows from the recorded sim's current state, the same actions for every method
    rng = np.random.default_rng(17)
    actions = rng.uniform(-1, 1, (8, 2, cars)).astype(np.float32)
    twins = []
    for _ in range(2):
        s = R.Sim(dev, arr, cars, 5, cfg)
        dev.copy(s.state, 0, sim.state, 0, 11 * cars)
        dev.copy(s.obs, 0, sim.obs, 0, (cfg.beams + 5) * cars)
        twins.append(s)
    imagined, exact = twins
    start = sim.state.read().reshape(11, cars)
    st = np.ascontiguousarray(start[[0, 1, 2, 3, 4, 5]].astype(np.float64))
    ref64 = dev.upload(st.view(np.float32).ravel())
    fallback = dev.alloc(cars)
    fell = np.zeros(cars, bool)
    respawned = np.zeros(cars, bool)
    errors = {}
    for k in range(8):
        ab = dev.upload(actions[k])
        imagined.actions.write(actions[k])
        exact.actions.write(actions[k])
        sur.step(imagined, 0.1, fallback)
        exact.step()
        sim.step_f64(ref64, R.mat(ab, 2, cars))
        fell |= fallback.read() > 0
        respawned |= (imagined.done.read() != 0) | (exact.done.read() != 0)
        if k in (0, 7):
            j = ref64.read().view(np.float64).reshape(6, cars)
            a = imagined.state.read().reshape(11, cars)
            b = exact.state.read().reshape(11, cars)
            ok = ~respawned
            def rms(v):
                return float(np.sqrt(np.mean(v[ok] ** 2)))
            errors[f"{k + 1}_step"] = {
                "cars": int(ok.sum()),
                "surrogate": {"position_m": rms(np.hypot(a[0] - j[0], a[1] - j[1])),
                              "heading_rad": rms(np.angle(np.exp(1j * (a[3] - j[3])))),
                              "speed_mps": rms(a[4] - j[4]), "steer_rad": rms(a[5] - j[5])},
                "float32_rk4": {"position_m": rms(np.hypot(b[0] - j[0], b[1] - j[1])),
                                "heading_rad": rms(np.angle(np.exp(1j * (b[3] - j[3])))),
                                "speed_mps": rms(b[4] - j[4]), "steer_rad": rms(b[5] - j[5])}}
    # physics time per step at 16384 cars: the RK4 kernel against the fused surrogate step
    # (inputs, MLP and pose update in one launch) and against its separate kernels
    def physics_us(fused: bool) -> dict:
        sur.set_fused(fused)
        dev.profile_only(None)
        dev.profile_reset()
        dev.profile(True)
        for k in range(20):
            exact.actions.write(actions[k % 8])
            exact.step()
            imagined.actions.write(actions[k % 8])
            sur.step(imagined, 0.1, None)
        dev.synchronize()
        dev.profile(False)
        return dev.profile_results()
    def per(prof, names):
        return sum(prof[k][1] for k in names if k in prof) * 1000.0 / 20
    separate = physics_us(False)
    fused = physics_us(True)
    rk4_us = per(fused, ["physics"])
    sur_us = per(fused, ["surrogate_fused"])
    sep_us = per(separate, ["surrogate_features", "gemm_bf16_tanh", "gemm_f32", "physics_surrogate"])

    def restart(s):  # both twins back to the recorded sim's state before each timed window
        dev.copy(s.state, 0, sim.state, 0, 11 * cars)
        dev.copy(s.obs, 0, sim.obs, 0, (cfg.beams + 5) * cars)

    def best_us(setup, fn, reps=5):
        best = None
        for _ in range(reps):
            setup()
            dev.synchronize()
            t0 = time.perf_counter()
            fn()
            dev.synchronize()
            dt = (time.perf_counter() - t0) * 1e6
            best = dt if best is None else min(best, dt)
        return best

    # closed loop: 8 Runner steps (policy forward with the sample, physics, exact lidar)
    sur_acts = R.Activations(dev, cars, cfg)
    window = {
        "surrogate_ms": best_us(lambda: restart(imagined), lambda: [
            sur.runner_step(imagined, pol, sur_acts, scratch, False, k) for k in range(8)]) / 1000.0,
        "rk4_ms": best_us(lambda: restart(exact), lambda: [
            R.runner_step(exact, pol, sur_acts, scratch, False, k) for k in range(8)]) / 1000.0,
    }
    window["speedup"] = window["rk4_ms"] / window["surrogate_ms"]
    # open loop: 8 steps of the physics alone (no lidar, no policy), the actions given
    acts8 = dev.upload(actions)
    view8 = R.mat(acts8, 16, cars)
    opened = {
        "surrogate_one_launch_us": best_us(lambda: restart(imagined), lambda: sur.imagine(imagined, view8, 8)),
        "rk4_one_launch_us": best_us(lambda: restart(exact), lambda: exact.imagine(view8, 8)),
        "rk4_eight_launches_us": best_us(lambda: restart(exact), lambda: [
            exact.imagine(R.mat(acts8, 2, cars, 2 * k * cars), 1) for k in range(8)]),
    }
    opened["speedup_vs_rk4_one_launch"] = opened["rk4_one_launch_us"] / opened["surrogate_one_launch_us"]
    log("surrogate 8-step windows:", window, opened)
    return {"track": "bank10", "cars": cars, "beams": 1081, "transitions": int(data.count),
            "record_s": record_s, "train_s": train_s, "epochs": epochs, "minibatch": cars, "adam_steps": minibatches,
            "network": "11 -> 3 x 128 tanh -> 5, residual (pose change), bf16 MFMA hidden layers",
            "normalised_loss": {"before": loss0, "after": loss1},
            "fallback_fraction_8_steps": float(fell.mean()), "error_vs_float64_rk4": errors,
            "physics_us_per_step": {"rk4_float32": rk4_us, "surrogate_fused": sur_us,
                                    "surrogate_separate_kernels": sep_us,
                                    "speedup": rk4_us / sur_us if sur_us else None},
            "runner_window_8_steps": window, "open_loop_8_steps": opened}


def plain_baseline(R, dev, track, args, steps: int) -> dict:
    """The same Runner step in the reference's own GPU decomposition, on the
    same GPU: one thread per car for physics, one thread per (car, beam) for the
    lidar, three float32 products and a separate sample kernel."""
    cfg = R.Config(beams=args.beams, hidden=args.hidden, precision=2, heading_noise=0.1)
    run = R.Runner(track, args.cars, cfg, seed=42, device=dev)
    run.sim.set_schedule(1)
    for t in range(5):
        run.step(False, t)
    dev.synchronize()
    t0 = time.perf_counter()
    for t in range(steps):
        run.step(False, 5 + t)
    dev.synchronize()
    dt = time.perf_counter() - t0
    # kernel durations in a separate profiled pass
    dev.profile_only(None)
    dev.profile_reset()
    dev.profile(True)
    for t in range(steps):
        run.step(False, 5 + steps + t)
    dev.synchronize()
    dev.profile(False)
    prof = dev.profile_results()
    ms = dt * 1000.0 / steps
    return {"value": args.cars * 1000.0 / ms, "unit": "env-steps/s", "ms_per_step": ms, "steps": steps,
            "kernels": {k: {"launches": n, "avg_us": tot / max(n, 1) * 1000.0} for k, (n, tot) in prof.items()},
            "schedule": "physics 1 thread/car; lidar 1 thread/(car, beam); 3 float32 products; sample kernel"}


ACT_NAMES = {0: "auto", 1: "fused", 2: "wide", 3: "tiled"}
SIM_KERNELS = ("physics", "lidar", "step_fused", "lidar_flat")


def physics_bytes(cars: int, layers: int) -> float:
    """Algorithmic bytes of the physics of one step: the state read and written
    (11 floats each way), the two actions, reward and done, five obs rows, and
    per car five probes of five packed records (8 B per layer), the nearest
    waypoint cell and the waypoint row it names."""
    return cars * (88.0 + 8.0 + 8.0 + 20.0 + 5 * 5 * 8.0 * layers + 4.0 * layers + 16.0)


def runner_leg(R, dev, arr, args, precision: str, layers: int) -> dict:
    """Runner steps (policy forward with sample, physics, lidar) at the
    benchmark's size: the launch shapes come from the library's autotuner
    (graph replays of every choice from one saved state), the timed region is
    `steps` steps replayed from a captured HIP graph, bracketed by device
    synchronizations; kernel durations come from HIP events in eager passes of
    the same steps right after."""
    import numpy as np

    cfg = R.Config(beams=args.beams, hidden=args.hidden, precision=1 if precision == "bf16" else 0,
                   heading_noise=0.1)
    run = R.Runner(arr, args.cars, cfg, seed=42, device=dev)
    for t in range(args.warmup):
        run.step(False, t)
    dev.synchronize()
    tune = run.autotune(10)
    if args.lidar_waves:
        # a pinned launch shape (profiler passes: the autotuner's timings under a
        # counter profiler need not pick the timed run's shape)
        run.sim.set_waves_per_car(args.lidar_waves)
        run.sim.set_lidar_coop(args.lidar_coop != 0)
        run.sim.set_fused_step(False)
        tune = dict(tune, waves_per_car=args.lidar_waves, lidar_coop=args.lidar_coop != 0, fused_step=False)
    t = args.warmup
    log(f"[{precision}] launch shapes (library autotuner): lidar wavefronts per car {tune['waves_per_car']}, "
        f"{'cooperative' if tune['lidar_coop'] else 'per-wavefront'} range stores, "
        f"{'fused' if tune['fused_step'] else 'two-launch'} sim step, {ACT_NAMES[tune['act_kernel']]} forward")
    sim_kernel = "step_fused" if run.sim.fused_step else "lidar"
    # march iterations of the lidar, counted by the instrumented variant on the warm state (untimed)
    run.sim.count_iterations(True)
    counts, trips = [], []
    for k in range(3):
        run.step(False, t + k)
        counts.append(run.sim.last_iterations)
        trips.append(run.sim.last_trips)
    run.sim.count_iterations(False)
    dev.synchronize()
    t += 3
    chunk = math.gcd(args.steps, 100)
    graph = run.capture(chunk)
    graph.launch(t)  # untimed first replay
    t += chunk
    dev.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps // chunk):
        graph.launch(t + k * chunk)
    dev.synchronize()
    dt = time.perf_counter() - t0
    t += args.steps
    ms = dt * 1000.0 / args.steps
    value = args.cars * 1000.0 / ms
    log(f"[{precision}] graph replay: {ms:.4f} ms/step, {value:.4g} env-steps/s")
    assert np.isfinite(run.sim.obs.read()).all(), "non-finite observations"
    del graph
    # The same number of eager steps with HIP events around the dominant sim
    # launch only (one event pair per step): its launch duration.
    dev.profile_reset()
    dev.profile_only(sim_kernel)
    dev.profile(True)
    t0 = time.perf_counter()
    for k in range(args.steps):
        run.step(False, t + k)
    dev.synchronize()
    eager_ms = (time.perf_counter() - t0) * 1000.0 / args.steps
    dev.profile(False)
    t += args.steps
    sim_n, sim_ms = dev.profile_results().get(sim_kernel, (0, 0.0))
    sim_us = sim_ms / max(sim_n, 1) * 1000.0
    # untimed: iteration count on the final state, to bracket the measured steps
    run.sim.count_iterations(True)
    run.step(False, t)
    t += 1
    counts.append(run.sim.last_iterations)
    run.sim.count_iterations(False)
    # every kernel, in a shorter profiled pass (an event pair per launch); the
    # forward's time per step is that of every launch that is not the sim's
    dev.profile_only(None)
    dev.profile_reset()
    dev.profile(True)
    prof_steps = min(args.steps, 100)
    for k in range(prof_steps):
        run.step(False, t + k)
    dev.synchronize()
    dev.profile(False)
    prof = dev.profile_results()
    for k, (n, tot) in sorted(prof.items(), key=lambda kv: -kv[1][1]):
        log(f"  {k:20s} launches {n:6d} avg {tot / max(n, 1) * 1000:9.2f} us  total {tot:9.2f} ms")
    policy_names = sorted(k for k in prof if k not in SIM_KERNELS)
    pol_us = sum(tot for k, (n, tot) in prof.items() if k not in SIM_KERNELS) * 1000.0 / prof_steps
    log(f"[{precision}] eager steps: {eager_ms:.4f} ms/step; {sim_kernel} {sim_us:.2f} us, forward "
        f"({' + '.join(policy_names)}) {pol_us:.2f} us")
    iters = float(np.mean(counts))
    lbytes = lidar_bytes(iters, args.cars, args.beams, layers)
    if sim_kernel == "step_fused":
        lbytes += physics_bytes(args.cars, layers)
    flops = policy_flops(args.cars, args.beams, args.hidden)
    traffic = None
    tfile = ROOT / "profiles" / "lidar_traffic.json"
    if tfile.exists() and sim_kernel == "lidar":
        try:
            traffic = json.loads(tfile.read_text()).get("bytes_per_launch")
        except Exception:
            traffic = None
    roofline = {
        "bound": "hbm", "kernel": sim_kernel, "achieved": lbytes / (sim_us * 1e-6) / 1e9, "peak": HBM_PEAK_GBS,
        "unit": "GB/s", "frac": lbytes / (sim_us * 1e-6) / 1e9 / HBM_PEAK_GBS, "traffic": traffic,
        "algorithmic_bytes_per_launch": lbytes, "march_iterations_per_launch": iters,
        "wavefront_trips_per_launch": float(np.mean(trips)) if trips else None,
        "lane_utilisation": iters / (64.0 * float(np.mean(trips))) if trips and np.mean(trips) > 0 else None,
        "avg_launch_us": sim_us,
    }
    peak = MFMA_BF16_PEAK_TFLOPS if precision == "bf16" else FP32_PEAK_TFLOPS
    achieved = flops / (pol_us * 1e-6) / 1e12 if pol_us > 0 else None
    policy = {"bound": "mfma", "kernel": "+".join(policy_names), "achieved": achieved, "peak": peak,
              "unit": "TFLOP/s", "frac": achieved / peak if achieved else None, "flops_per_launch": flops,
              "avg_launch_us": pol_us,
              "peak_is": "dense bf16 MFMA" if precision == "bf16" else
                         "dense float32 MFMA (the K=1 float32 MFMA the kernel runs on)"}
    kernels = {k: {"launches": n, "avg_us": tot / max(n, 1) * 1000.0} for k, (n, tot) in prof.items()}
    return {"value": value, "ms_per_step": ms, "eager_ms_per_step": eager_ms, "roofline": roofline,
            "policy_roofline": policy, "kernels": kernels, "lidar_waves_per_car": tune["waves_per_car"],
            "lidar_coop_stores": tune["lidar_coop"],
            "fused_step": tune["fused_step"], "act_kernel": ACT_NAMES[tune["act_kernel"]],
            "lidar_wavefront_trips": trips}


def flat_leg(R, T, dev, args) -> dict:
    """The reference's plain `racer benchmark` workload (main.mojo:225-235 on
    sim.mojo's defaults): the flat ribbon (circle of 10 m, 0.025 m grids, one
    layer), 1024 cars, the default 108 beams, a 2 x 64 float32 tanh policy,
    1000 steps. Timed three ways on the same box: the library's Runner step
    (autotuned launch shapes, graph replays), the plain GPU decomposition, and
    the oracle on the host cores."""
    from types import SimpleNamespace

    import numpy as np

    arr = T.cached("ring10")
    cars, beams, hidden, steps = 1024, 108, 64, args.flat_steps
    cfg = R.Config(beams=beams, hidden=hidden, precision=0, heading_noise=0.1)
    run = R.Runner(arr, cars, cfg, seed=42, device=dev)
    for t in range(20):
        run.step(False, t)
    dev.synchronize()
    tune = run.autotune(10)
    t = 20
    chunk = math.gcd(steps, 100)
    graph = run.capture(chunk)
    graph.launch(t)
    t += chunk
    dev.synchronize()
    t0 = time.perf_counter()
    for k in range(steps // chunk):
        graph.launch(t + k * chunk)
    dev.synchronize()
    dt = time.perf_counter() - t0
    t += steps
    del graph
    assert np.isfinite(run.sim.obs.read()).all(), "non-finite observations"
    ms = dt * 1000.0 / steps
    dev.profile_only(None)
    dev.profile_reset()
    dev.profile(True)
    for k in range(100):
        run.step(False, t + k)
    dev.synchronize()
    dev.profile(False)
    prof = dev.profile_results()
    gpu = {"value": cars * 1000.0 / ms, "unit": "env-steps/s", "ms_per_step": ms, "steps": steps,
           "lidar_waves_per_car": tune["waves_per_car"], "lidar_coop_stores": tune["lidar_coop"],
           "fused_step": tune["fused_step"], "act_kernel": ACT_NAMES[tune["act_kernel"]],
           "kernels": {k: {"launches": n, "avg_us": tot / max(n, 1) * 1000.0} for k, (n, tot) in prof.items()}}
    log("flat benchmark, GPU Runner step:", gpu)
    plain = plain_baseline(R, dev, arr, SimpleNamespace(cars=cars, beams=beams, hidden=hidden), min(steps, 200))
    log("flat benchmark, plain GPU decomposition:", plain)
    cpu = None
    if not args.no_cpu:
        threads = int(os.environ.get("OMP_NUM_THREADS") or 0) or (os.cpu_count() or 1)
        cpu = cpu_baseline(arr, beams, hidden, cars, threads, args.flat_cpu_steps, args.flat_cpu_steps_1core,
                           track_name="ring10")
        log("flat benchmark, oracle on the host:", cpu)
    return {"workload": f"ring10 flat ribbon R=10 m, 0.025 m grids; {cars} cars x {beams} beams; 2x{hidden} tanh "
                        "policy, float32 products; heading noise U(-0.1, 0.1)",
            "gpu": gpu, "plain_gpu": plain, "cpu": cpu,
            "speedup_vs_plain_gpu": gpu["value"] / plain["value"],
            "speedup_vs_cpu": gpu["value"] / cpu["value"] if cpu else None}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--cars", type=int, default=4096)
    ap.add_argument("--beams", type=int, default=1081)
    ap.add_argument("--hidden", type=int, default=256)
    ap.add_argument("--track", default="ramp10")
    ap.add_argument("--no-bf16", action="store_true", help="skip the bf16-policy Runner step")
    ap.add_argument("--cpu-steps", type=int, default=30, help="oracle Runner steps on all host threads")
    ap.add_argument("--cpu-steps-1core", type=int, default=3, help="oracle Runner steps on one thread")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true", help="skip the plain GPU decomposition baseline")
    ap.add_argument("--lidar-waves", type=int, default=0,
                    help="pin the Runner step's lidar to this many wavefronts per car after the autotuner (0: its choice)")
    ap.add_argument("--lidar-coop", type=int, default=1, help="with --lidar-waves: cooperative range stores (1) or not (0)")
    ap.add_argument("--flat-steps", type=int, default=1000,
                    help="timed steps of the flat 1024-car x 108-beam benchmark workload (0 skips it)")
    ap.add_argument("--flat-cpu-steps", type=int, default=1000, help="its oracle steps on all host threads")
    ap.add_argument("--flat-cpu-steps-1core", type=int, default=100, help="its oracle steps on one thread")
    ap.add_argument("--surrogate-transitions", type=int, default=10_000_000,
                    help="transitions recorded for the learned dynamics surrogate (0 skips it)")
    ap.add_argument("--surrogate-epochs", type=int, default=2)
    ap.add_argument("--batch-steps", type=int, default=100,
                    help="timed steps of the eight-instance batched launch (0 skips it)")
    ap.add_argument("--no-batch-separate", action="store_true",
                    help="skip the batch leg's comparison with each instance's own graph")
    ap.add_argument("--ppo-iterations", type=int, default=3,
                    help="timed PPO iterations per beam count on the bridge (0 skips the PPO leg)")
    ap.add_argument("--ppo-legs", default="1081,108,108fp32",
                    help="PPO legs to run: beam counts, bf16 policy, or fp32 with the suffix fp32")
    ap.add_argument("--train-parity", type=int, default=0,
                    help="PPO iterations of the bf16 / fp32 training-curve comparison (0: skip; minutes per 300)")
    ap.add_argument("--train-parity-seeds", type=int, default=2)
    ap.add_argument("--calibrate-copy", type=int, default=0,
                    help="MiB copied once device to device before the run (known-size traffic for counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this benchmark measures one GPU: the build runs on one GPU only, so "
                                   "--gpus above 1 is not supported; run with --gpus 1"}))
        return 2

    import warporacer3d_amd as R
    from warporacer3d_amd import track as T

    arr = T.cached(args.track)
    layers = int(arr[7])
    dev = R.Device(True)
    log("device:", dev.name)
    if args.calibrate_copy:
        count = args.calibrate_copy * (1 << 20) // 4
        src, dst = dev.alloc(count, 1.0), dev.alloc(count)
        dev.copy(dst, 0, src, 0, count)
        dev.synchronize()
        log(f"calibration copy: {count * 4} bytes read and written")
        del src, dst
    fp32 = runner_leg(R, dev, arr, args, "fp32", layers)
    bf16 = None if args.no_bf16 else runner_leg(R, dev, arr, args, "bf16", layers)
    value, ms = fp32["value"], fp32["ms_per_step"]
    plain = None
    if not args.no_plain:
        plain = plain_baseline(R, dev, arr, args, min(args.steps, 50))
        log("plain GPU baseline:", plain)
    surrogate = None
    if args.surrogate_transitions > 0:
        surrogate = surrogate_leg(R, T, dev, args.surrogate_transitions, args.surrogate_epochs)
        log("surrogate:", surrogate)
    flat = None
    if args.flat_steps > 0:
        flat = flat_leg(R, T, dev, args)
    batched = None
    if args.batch_steps > 0:
        batched = batch_leg(R, T, dev, args.batch_steps, not args.no_batch_separate)
        log("eight-instance batch:", batched)
    ppo = None
    if args.ppo_iterations > 0:
        ppo = {}
        legs = [(int(leg[:-4]), 0) if leg.endswith("fp32") else (int(leg), 1)
                for leg in args.ppo_legs.split(",") if leg]
        for beams, precision in legs:
            key = f"beams{beams}" + ("" if precision == 1 else "_fp32")
            ppo[key] = ppo_leg(R, T, dev, beams, args.ppo_iterations, precision=precision)
            log(f"PPO iteration (bridge8, 16384 cars, {beams} beams, precision {precision}):",
                {k: v for k, v in ppo[key].items() if k != "kernels_ms_per_iteration"})
    parity = None
    if args.train_parity > 0:
        parity = train_parity_leg(R, T, dev, args.train_parity, args.train_parity_seeds)
        log("training parity:", {k: v for k, v in parity.items() if k != "runs"})
    cpu = None
    if not args.no_cpu:
        threads = int(os.environ.get("OMP_NUM_THREADS") or 0) or (os.cpu_count() or 1)
        log(f"cpu baseline: {args.cars} cars x {args.beams} beams on {threads} threads and on 1 ...")
        cpu = cpu_baseline(arr, args.beams, args.hidden, args.cars, threads, args.cpu_steps, args.cpu_steps_1core)
        log("cpu baseline:", cpu)
    out = {
        "metric": METRIC, "value": value, "unit": "env-steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "fp32",
        "data": "synthetic (procedural ramp ribbon, random-init policy)",
        "config": {"workload": f"Runner step: {args.track} ribbon R=10 m, height 0.3 sin 3t, 0.025 m grid; "
                               f"{args.cars} cars x {args.beams} beams; 2x{args.hidden} tanh policy, float32 "
                               "products (bitwise the CPU restatement), float32 dynamics and lidar",
                   "cars": args.cars, "beams": args.beams, "hidden": args.hidden, "track": args.track,
                   "policy_precision": "fp32", "parallelism": "single GPU",
                   "lidar_waves_per_car": fp32["lidar_waves_per_car"], "lidar_coop_stores": fp32["lidar_coop_stores"],
                   "fused_step": fp32["fused_step"], "act_kernel": fp32["act_kernel"]},
        "eager_ms_per_step": fp32["eager_ms_per_step"],
        "timing": "graph replays of captured Runner steps; kernel durations from HIP events in eager passes of "
                  "the same steps right after",
        "roofline": fp32["roofline"], "policy_roofline": fp32["policy_roofline"], "kernels": fp32["kernels"],
        "value_bf16": bf16["value"] if bf16 else None,
        "bf16": ({k: bf16[k] for k in ("value", "ms_per_step", "eager_ms_per_step", "roofline", "policy_roofline",
                                       "kernels", "lidar_waves_per_car", "lidar_coop_stores", "fused_step",
                                       "act_kernel")}
                 if bf16 else None),
        "cpu_baseline": cpu,
        "plain_gpu_baseline": plain, "speedup_vs_plain_gpu": value / plain["value"] if plain else None,
        "speedup_bf16_vs_plain_gpu": bf16["value"] / plain["value"] if (plain and bf16) else None,
        "ppo_iteration": ppo, "batch": batched, "surrogate": surrogate, "flat_benchmark": flat,
    }
    if parity is not None:
        out["train_parity"] = parity
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
