"""The viewer's server: an HTTP front end in a thread and the step loop over
libracer_hip (racer/serve.py:15-48, main.mojo:94-124).

Wire format, as the reference's viewer reads it:
  GET /       the page given with --page (the reference's viewer.html), or a
              short text description of these endpoints
  GET /info   JSON: name, cars, policy (a checkpoint is driving), state (11),
              obs (beams + 5), beams, range (10), mount ([0.27, 0, 0.23])
  GET /map    the float32 track array's raw bytes
  POST /step  body "steer% throttle% reset policy" (integers): reset > 0
              respawns every car first; policy > 0 with a checkpoint drives
              every car with the policy's mean action (Runner.step,
              deterministic), otherwise every car takes steer / 100 and
              throttle / 100; the reply is one float32 snapshot of the state
              [11, n], obs [beams + 5, n], actions [2, n], reward [n] and done
              [n] buffers, back to back (main.mojo:117-122).

Requests are handled one at a time in arrival order.

    python -m warporacer3d_amd.serve TRACK [auto|gpu|cpu] [cars] [port] [checkpoint.npy|-] [--page PAGE]
"""
from __future__ import annotations

import http.server
import json
import queue
import sys
import threading
from pathlib import Path

import numpy as np

STATE = 11
MOUNT = (0.27, 0.0, 0.23)  # lidar mount (forward, left, up), float32 in the simulation

INDEX = b"""<!doctype html><title>racer</title><pre>
GET  /info   scene description (JSON)
GET  /map    float32 track array
POST /step   "steer% throttle% reset policy" -> float32 snapshot
             (state, obs, actions, reward, done rows)
Start the server with --page viewer.html to serve a viewer page here.
</pre>"""


class Bridge:
    """HTTP on 127.0.0.1:port in a daemon thread; POST /step bodies go to
    `requests`, and each is answered with the next bytes put on `replies`."""

    def __init__(self, port: int, page: bytes, info: dict, track: bytes):
        self.requests: queue.Queue = queue.Queue()
        self.replies: queue.Queue = queue.Queue()
        body = {"/": (page, "text/html; charset=utf-8"), "/info": (json.dumps(info).encode(), "application/json"),
                "/map": (track, "application/octet-stream")}
        bridge = self

        class Handler(http.server.BaseHTTPRequestHandler):
            def log_message(self, *args):
                pass

            def _send(self, data: bytes, kind: str) -> None:
                self.send_response(200)
                self.send_header("Content-Type", kind)
                self.send_header("Content-Length", str(len(data)))
                self.end_headers()
                self.wfile.write(data)

            def do_GET(self):
                if self.path in body:
                    self._send(*body[self.path])
                else:
                    self.send_error(404)

            def do_POST(self):
                if self.path != "/step":
                    self.send_error(404)
                    return
                bridge.requests.put(self.rfile.read(int(self.headers.get("Content-Length", 0))).decode())
                reply = bridge.replies.get()
                if isinstance(reply, Exception):
                    self.send_error(400, str(reply))
                else:
                    self._send(reply, "application/octet-stream")

        self.server = http.server.HTTPServer(("127.0.0.1", port), Handler)
        self.port = self.server.server_address[1]
        self.thread = threading.Thread(target=self.server.serve_forever, daemon=True)
        self.thread.start()

    def close(self) -> None:
        self.server.shutdown()
        self.server.server_close()


class Session:
    """The step loop's state: a Runner, whether a checkpoint drives, and the
    step counter that seeds the policy's Philox stream."""

    def __init__(self, runner, name: str, trained: bool):
        self.runner, self.name, self.trained = runner, name, trained
        self.tick = 0

    def info(self) -> dict:
        sim = self.runner.sim
        return {"name": self.name, "cars": sim.n, "policy": self.trained, "state": STATE, "obs": sim.IN,
                "beams": sim.beams, "range": 10.0, "mount": [float(np.float32(v)) for v in MOUNT]}

    def handle(self, request: str) -> bytes:
        """One POST /step (main.mojo:107-123)."""
        words = request.split()
        if len(words) != 4:
            raise ValueError(f"a step request is 'steer% throttle% reset policy', got {request!r}")
        steer, throttle, reset, policy = (int(w) for w in words)
        sim = self.runner.sim
        if reset > 0:
            sim.reset()
        self.tick += 1
        if policy > 0 and self.trained:
            self.runner.step(True, self.tick)
        else:
            n = sim.n
            hundred = np.float32(100)
            actions = np.stack([np.full(n, np.float32(steer) / hundred, np.float32),
                                np.full(n, np.float32(throttle) / hundred, np.float32)])
            sim.actions.write(actions)
            sim.step()
        return sim.snapshot().tobytes()


def serve_forever(session: Session, bridge: Bridge, stop: threading.Event | None = None) -> None:
    """Answer step requests until `stop` is set (or for ever)."""
    while stop is None or not stop.is_set():
        try:
            request = bridge.requests.get(timeout=0.1)
        except queue.Empty:
            continue
        try:
            bridge.replies.put(session.handle(request))
        except Exception as e:  # a bad request: answer it and keep serving
            bridge.replies.put(e)


def main(argv: list[str]) -> int:
    import argparse

    ap = argparse.ArgumentParser(prog="racer serve", description="Serve the viewer's step loop (main.mojo:94-124).")
    ap.add_argument("track", help="flat, ramp, bank, bridge, ring10 .. random<k>, or a map YAML")
    ap.add_argument("device", nargs="?", default="auto", choices=["auto", "gpu", "cpu"])
    ap.add_argument("cars", nargs="?", type=int, default=1024)
    ap.add_argument("port", nargs="?", type=int, default=8765)
    ap.add_argument("checkpoint", nargs="?", default="agent.npy", help="checkpoint driving the cars, or - for none")
    ap.add_argument("--page", help="HTML page served at / (the reference's viewer.html)")
    a = ap.parse_args(argv)
    if a.device == "cpu":
        print("error: libracer_hip runs on the GPU only; the CPU device is not provided", file=sys.stderr)
        return 1
    import warporacer3d_amd as R
    from warporacer3d_amd import track as T

    arr = T.load(a.track)
    runner = R.Runner(arr, a.cars, R.Config.from_env(), seed=42)
    print("Device:", runner.device.name, flush=True)
    trained = a.checkpoint != "-"
    if trained:
        R.load(runner.policy, a.checkpoint)
    session = Session(runner, a.track, trained)
    page = Path(a.page).read_bytes() if a.page else INDEX
    bridge = Bridge(a.port, page, session.info(), arr.tobytes())
    print(f"http://127.0.0.1:{bridge.port} (Ctrl+C to stop)", flush=True)
    try:
        serve_forever(session, bridge)
    except KeyboardInterrupt:
        pass
    finally:
        bridge.close()
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
