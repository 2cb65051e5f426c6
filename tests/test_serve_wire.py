"""The viewer's wire format (racer/serve.py:33-45, main.mojo:94-124) over a
stand-in runner whose buffers live in host memory: /info, /map, POST /step
bodies and the snapshot's row order, without a GPU."""
import json
import threading
import urllib.error
import urllib.request

import numpy as np
import pytest

from warporacer3d_amd import serve as S


class _Buf:
    def __init__(self, a):
        self.a = a

    def write(self, v):
        self.a[...] = np.asarray(v, np.float32).reshape(self.a.shape)


class _Sim:
    """Host buffers shaped as the sim's; step() records what it was given."""

    def __init__(self, n, beams):
        self.n, self.beams, self.IN = n, beams, beams + 5
        self.state = np.zeros((11, n), np.float32)
        self.obs = np.zeros((self.IN, n), np.float32)
        self.actions = _Buf(np.zeros((2, n), np.float32))
        self.reward = np.zeros(n, np.float32)
        self.done = np.zeros(n, np.float32)
        self.resets = self.steps = 0

    def reset(self):
        self.resets += 1
        self.state[8] += 1

    def step(self):
        self.steps += 1
        self.state[0] += self.actions.a[1]
        self.reward[:] = self.steps
        self.obs[:] = self.steps

    def snapshot(self):
        return np.concatenate([self.state.ravel(), self.obs.ravel(), self.actions.a.ravel(), self.reward,
                               self.done])


class _Runner:
    def __init__(self, n, beams):
        self.sim = _Sim(n, beams)
        self.policy_steps = []

    def step(self, deterministic, t):
        self.policy_steps.append((deterministic, t))
        self.sim.steps += 1


@pytest.fixture()
def server():
    runner = _Runner(5, 7)
    session = S.Session(runner, "flat", trained=True)
    track = np.arange(30, dtype=np.float32)
    bridge = S.Bridge(0, S.INDEX, session.info(), track.tobytes())
    stop = threading.Event()
    loop = threading.Thread(target=S.serve_forever, args=(session, bridge, stop), daemon=True)
    loop.start()
    yield runner, bridge, track
    stop.set()
    loop.join(timeout=5)
    bridge.close()


def _get(bridge, path):
    with urllib.request.urlopen(f"http://127.0.0.1:{bridge.port}{path}", timeout=10) as r:
        return r.read(), r.headers["Content-Type"]


def _post(bridge, body):
    req = urllib.request.Request(f"http://127.0.0.1:{bridge.port}/step", data=body.encode(), method="POST",
                                 headers={"Content-Type": "text/plain"})
    with urllib.request.urlopen(req, timeout=10) as r:
        return np.frombuffer(r.read(), np.float32)


def test_info_and_map(server):
    runner, bridge, track = server
    info = json.loads(_get(bridge, "/info")[0])
    assert info == {"name": "flat", "cars": 5, "policy": True, "state": 11, "obs": 12, "beams": 7, "range": 10.0,
                    "mount": [float(np.float32(0.27)), 0.0, float(np.float32(0.23))]}
    body, kind = _get(bridge, "/map")
    assert kind == "application/octet-stream" and body == track.tobytes()
    assert b"/step" in _get(bridge, "/")[0]
    with pytest.raises(urllib.error.HTTPError):
        _get(bridge, "/nothing")


def test_keyboard_steps_reset_and_snapshot_layout(server):
    runner, bridge, _ = server
    n, IN = 5, 12
    snap = _post(bridge, "-25 100 0 0")
    assert snap.size == (11 + IN + 4) * n
    rows = snap.reshape(-1, n)
    actions = rows[11 + IN: 11 + IN + 2]
    assert np.all(actions[0] == np.float32(-25) / np.float32(100)) and np.all(actions[1] == 1.0)
    assert np.all(rows[11 + IN + 2] == 1) and np.all(rows[11] == 1)  # reward, then the first obs row
    assert runner.sim.resets == 0 and runner.sim.steps == 1
    _post(bridge, "0 0 1 0")
    assert runner.sim.resets == 1 and runner.sim.steps == 2


def test_policy_steps_count_ticks_and_bad_requests(server):
    runner, bridge, _ = server
    _post(bridge, "0 0 0 1")
    _post(bridge, "10 10 0 1")
    assert runner.policy_steps == [(True, 1), (True, 2)]
    with pytest.raises(urllib.error.HTTPError) as e:
        _post(bridge, "1 2")
    assert e.value.code == 400
    _post(bridge, "0 0 0 1")  # the loop serves on after a bad request, which took no step
    assert runner.policy_steps[-1] == (True, 3)
