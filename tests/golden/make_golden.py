"""Writes the golden vectors of this directory from the CPU oracle.

The reference ships no data files: its tests build procedural fixtures (the
corridor of tests/test_mojo.mojo:15-40) and check closed forms. These vectors
are seeded runs of the oracle, small enough to commit: the oracle tests
(tests/test_oracle_golden.py) keep the oracle reproducing them bit for bit, and
the GPU tests (tests/test_gpu_golden.py) hold the HIP path to the same bits.

    python tests/golden/make_golden.py        # rewrites tests/golden/*.npz
"""
from __future__ import annotations

import hashlib
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent.parent
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "tests"))

import oracle as O  # noqa: E402
from fixtures import corridor  # noqa: E402
from warporacer3d_amd import track as T  # noqa: E402


def track_digest(arr: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(arr, np.float32).tobytes()).hexdigest()


def sim_run(arr, n, seed, beams, noise, steps, action_seed):
    """Reset, then `steps` steps with seeded actions in [-1.2, 1.2] (clamping included)."""
    sim = O.Sim(arr, n, seed, beams=beams, heading_noise=noise)
    sim.reset()
    rng = np.random.default_rng(action_seed)
    actions = rng.uniform(-1.2, 1.2, (steps, 2, n)).astype(np.float32)
    out = {"reset_state": sim.state.copy(), "reset_obs": sim.obs.copy(), "actions": actions}
    states, obs, rewards, dones, iters = [], [], [], [], []
    for k in range(steps):
        sim.actions[...] = actions[k]
        iters.append(sim.step())
        states.append(sim.state.copy())
        obs.append(sim.obs.copy())
        rewards.append(sim.reward.copy())
        dones.append(sim.done.copy())
    out.update(state=np.array(states), obs=np.array(obs), reward=np.array(rewards), done=np.array(dones),
               iterations=np.array(iters, np.int64))
    return out


def main() -> None:
    O.set_threads(8)
    # corridor: 12 cars, 32 beams, 8 steps
    arr = corridor()
    data = sim_run(arr, 12, 5, 32, 0.1, 8, 11)
    np.savez_compressed(HERE / "sim_corridor.npz", track_sha256=track_digest(arr), n=12, seed=5, beams=32,
                        noise=np.float32(0.1), **data)
    # ramp10 (the benchmark track): 4 cars, 1081 beams, 2 steps
    arr = T.cached("ramp10")
    data = sim_run(arr, 4, 42, 1081, 0.1, 2, 12)
    np.savez_compressed(HERE / "sim_ramp10.npz", track_sha256=track_digest(arr), n=4, seed=42, beams=1081,
                        noise=np.float32(0.1), **data)
    # bridge8 (two layers): 6 cars, 108 beams, 3 steps
    arr = T.cached("bridge8")
    data = sim_run(arr, 6, 7, 108, 0.0, 3, 13)
    np.savez_compressed(HERE / "sim_bridge8.npz", track_sha256=track_digest(arr), n=6, seed=7, beams=108,
                        noise=np.float32(0.0), **data)

    # policy: init (with the absolute-index bias quirk) and float32 forward on 16 columns
    pol = O.Policy(42, 108, 64)
    x = np.random.default_rng(3).uniform(-1, 1, (pol.IN, 16)).astype(np.float32)
    x[-1] = 1
    h1, h2, o = pol.forward(x)
    np.savez_compressed(HERE / "policy_108x64.npz", seed=42, beams=108, hidden=64, theta=pol.theta, x=x, h1=h1,
                        h2=h2, o=o)

    # sample: Gaussian actions, log probabilities, values for 8 cars at step 17
    out = np.random.default_rng(4).normal(0, 0.5, (3, 8)).astype(np.float32)
    sim = O.Sim(corridor(), 8, 9, beams=16)
    act = np.zeros((2, 8), np.float32)
    logp = np.zeros(8, np.float32)
    value = np.zeros(8, np.float32)
    log_std = np.array([-0.5, -0.7], np.float32)
    sim.sample(out, log_std, act, logp, value, False, 17)
    np.savez_compressed(HERE / "sample.npz", seed=9, step=17, out=out, log_std=log_std, actions=act, logp=logp,
                        value=value)

    # trainer: one rollout and one update on the corridor (4 cars, R = 8, 2 epochs x 2 minibatches)
    pol = O.Policy(42, 16, 8)
    sim = O.Sim(corridor(), 4, 42, beams=16, heading_noise=0.1)
    sim.reset()
    tr = O.Trainer(4, pol, rollout=8, epochs=2, minibatches=2)
    reward = tr.rollout(sim, 0)
    kl = tr.update()
    np.savez_compressed(HERE / "trainer_corridor.npz", beams=16, hidden=8, n=4, rollout=8, epochs=2, minibatches=2,
                        mean_reward=np.float32(reward), kl=np.float32(kl), obs=tr.obs, actions=tr.actions,
                        logp=tr.logp, value=tr.value, reward=tr.reward, done=tr.done, advantage=tr.advantage,
                        theta=pol.theta, steps=tr.step)
    for f in sorted(HERE.glob("*.npz")):
        print(f.name, f.stat().st_size, "bytes")


if __name__ == "__main__":
    main()
