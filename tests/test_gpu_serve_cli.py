"""The viewer's step loop (main.mojo:94-124) and the command-line driver
(main.mojo:127-149) on the GPU.

* serve: each POST /step answered by the library equals, bit for bit, the
  oracle stepping the same actions (keyboard steps, a reset, and a policy
  step with a checkpoint driving).
* racer benchmark / train / eval with tracks given by name and as a map YAML;
  `train` with a resume checkpoint continues the iteration count and
  restarts Adam and the sim (main.mojo:197-211), checked against the same
  resume done through the Python API: the saved weights are bit-identical.
"""
import os
import subprocess
import sys
import threading
import urllib.request
from pathlib import Path

import numpy as np
import pytest

import oracle as O
import warporacer3d_amd as R
from helpers import track
from warporacer3d_amd import serve as S

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parent.parent
CLI = ROOT / "warporacer3d_amd" / "racer"


def _post(port, body):
    req = urllib.request.Request(f"http://127.0.0.1:{port}/step", data=body.encode(), method="POST")
    with urllib.request.urlopen(req, timeout=60) as r:
        return np.frombuffer(r.read(), np.float32)


def _rows(snap, n, IN):
    rows = snap.reshape(-1, n)
    return rows[:11], rows[11: 11 + IN], rows[11 + IN: 13 + IN], rows[13 + IN], rows[14 + IN]


def test_serve_steps_match_oracle(dev, tmp_path):
    print("serve loop vs oracle", flush=True)
    n, beams, hidden = 48, 32, 16
    cfg = R.Config(beams=beams, hidden=hidden)
    arr = track("ring10")
    runner = R.Runner(arr, n, cfg, seed=42, device=dev)
    ckpt = tmp_path / "agent.npy"
    R.save(runner.policy, ckpt, 7)
    session = S.Session(runner, "ring10", trained=True)
    bridge = S.Bridge(0, S.INDEX, session.info(), arr.tobytes())
    stop = threading.Event()
    loop = threading.Thread(target=S.serve_forever, args=(session, bridge, stop), daemon=True)
    loop.start()
    try:
        osim = O.Sim(arr, n, 42, beams=beams)
        osim.reset()
        opol = O.Policy(42, beams, hidden)
        IN = beams + 5
        for body, tick in (("50 100 0 0", 1), ("-30 -40 0 0", 2), ("0 0 1 0", 3), ("0 0 0 1", 4), ("0 0 0 1", 5)):
            steer, throttle, reset, policy = (int(w) for w in body.split())
            if reset:
                osim.reset()
            if policy:
                _, _, out = opol.forward(osim.obs)
                logp, value = np.zeros(n, np.float32), np.zeros(n, np.float32)
                osim.sample(out, opol.log_std(), osim.actions, logp, value, True, tick)
            else:
                osim.actions[0] = np.float32(steer) / np.float32(100)
                osim.actions[1] = np.float32(throttle) / np.float32(100)
            osim.step()
            state, obs, actions, reward, done = _rows(_post(bridge.port, body), n, IN)
            assert np.array_equal(state, osim.state), body
            assert np.array_equal(obs, osim.obs), body
            assert np.array_equal(actions, osim.actions), body
            assert np.array_equal(reward, osim.reward) and np.array_equal(done, osim.done), body
    finally:
        stop.set()
        loop.join(timeout=10)
        bridge.close()


def _cli(*args, env_extra=None):
    env = dict(os.environ, RACER_PYTHON=sys.executable)
    env.update(env_extra or {})
    print("racer", *args, flush=True)
    res = subprocess.run([str(CLI), *map(str, args)], capture_output=True, text=True, timeout=300, env=env)
    return res.returncode, res.stdout, res.stderr


def test_cli_benchmark_by_name_and_yaml(tmp_path):
    rc, out, err = _cli("benchmark", "ring10", "gpu", 256, 20)
    assert rc == 0, err
    assert "256 cars;" in out and "transitions/s" in out
    t = np.linspace(0, 2 * np.pi, 90, endpoint=False)
    np.savetxt(tmp_path / "route.txt", np.column_stack([5 * np.cos(t), 3 * np.sin(t), 0 * t]))
    (tmp_path / "map.yaml").write_text("route: route.txt\nresolution: 0.05\nhalf_width: 1.0\n")
    rc, out, err = _cli("benchmark", tmp_path / "map.yaml", "auto", 64, 10)
    assert rc == 0, err
    assert "64 cars;" in out
    rc, _, err = _cli("benchmark", "ring10", "cpu", 64, 10)
    assert rc != 0 and "GPU only" in err


def test_cli_train_resume_and_eval(dev, tmp_path):
    first, second = tmp_path / "a.npy", tmp_path / "b.npy"
    rc, out, err = _cli("train", "ring10", "gpu", 64, 2, first)
    assert rc == 0, err
    assert "iteration 1 " in out and "iteration 2 " in out
    saved = np.load(first, allow_pickle=False)
    assert saved[0] == 2
    rc, out, err = _cli("train", "ring10", "gpu", 64, 2, second, first)
    assert rc == 0, err
    assert "iteration 3 " in out and "iteration 4 " in out and "iteration 2 " not in out
    resumed = np.load(second, allow_pickle=False)
    assert resumed[0] == 4
    # the same resume through the Python API: a fresh sim (seed 42), the
    # checkpoint's weights, a fresh trainer (Adam restarted), iterations 2 and 3
    cfg = R.Config()
    arr = track("ring10")
    sim = R.Sim(dev, arr, 64, 42, cfg)
    sim.reset()
    pol = R.Policy(dev, 42, cfg)
    assert R.load(pol, first) == 2
    acts = R.Activations(dev, 64, cfg)
    tr = R.Trainer(dev, 64, cfg)
    for it in (2, 3):
        tr.rollout(sim, pol, acts, it)
        tr.update(pol)
    assert np.array_equal(pol.theta.read(), resumed[1:])
    assert not np.array_equal(saved[1:], resumed[1:])
    rc, out, err = _cli("eval", "ring10", "gpu", 64, 10, second)
    assert rc == 0, err
    assert "reward/step" in out and "crashes/car" in out
