"""TEST INFRASTRUCTURE ONLY: ctypes binding of the CPU oracle (racer_oracle.c).

Imported by tests/, __graft_entry__.smoke() (to compare against) and bench.py's
cpu_baseline leg. Arrays are numpy float32, feature-major [rows, cols]; every
function here restates one reference function and works on host memory only.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
_lib = None

F = C.c_float
L = C.c_long
I = C.c_int
P = C.c_void_p


def lib():
    global _lib
    if _lib is None:
        path = HERE / "liboracle.so"
        if not path.exists():
            from . import build

            build.build()
        _lib = C.CDLL(str(path))
        sig = {
            "o_sim_step": (L, [P, L, P, L, P, P, P, L, I, P, C.c_size_t, I, F]),
            "o_sim_reset": (None, [P, P, P, P, P, L, I, P, C.c_size_t, I, F]),
            "o_sim_sample": (None, [L, P, L, P, P, L, P, P, I, C.c_int64, I]),
            "o_param_count": (L, [I, I]),
            "o_policy_init": (None, [P, I, I, I]),
            "o_forward": (None, [P, I, I, P, L, L, P, P, P]),
            "o_backward": (None, [P, I, I, P, L, P, P, P, P, P, P]),
            "o_gae": (None, [L, I, P, P, P, P, P]),
            "o_moments": (None, [L, P, P, P]),
            "o_minibatch": (None, [P, I, I, L, I, I, P, P, P, P, P, I, F, F, P, P, P, P, P, P, P, P, P]),
            "o_adam": (None, [P, L, L, P, P, P, F, I]),
            "o_bias_correction": (F, [F, I]),
            "o_set_threads": (None, [I]),
            "o_get_threads": (I, []),
            "o_sin": (F, [F]),
            "o_cos": (F, [F]),
            "o_exp": (F, [F]),
            "o_log": (F, [F]),
            "o_tanh": (F, [F]),
            "o_tanh_mlp": (F, [F]),
            "o_philox": (None, [P, P, P]),
            "o_normal4": (None, [C.c_uint64, C.c_uint64, C.c_uint64, P]),
            "o_uniform4": (None, [C.c_uint64, C.c_uint64, C.c_uint64, P]),
            "o_derivative": (None, [P, F, F, F, F, F, P]),
            "om_init": (I, [P, P, C.c_size_t]),
            "o_probe": (I, [P, C.c_size_t, F, F, F, P]),
            "o_surface": (I, [P, C.c_size_t, I, F, F, P]),
            "o_progress": (F, [P, C.c_size_t, I, F, F]),
            "o_scan_one": (F, [P, C.c_size_t, P, L, P, L, L, I, I, P]),
            "o_spawn_one": (I, [P, C.c_size_t, P, L, L, I, F]),
            "o_check_nearest_index": (L, [F, F, F]),
            "o_check_nearest_index_lim": (L, [F, F, F]),
            "o_check_div_by_reciprocal": (L, [L, C.c_uint64]),
            "o_check_nearest_by_reciprocal": (L, [C.c_float, L]),
            "o_check_nearest_by_reciprocal1": (L, [C.c_float, L]),
            "o_nearest_form_bad": (L, [C.c_float, C.c_int, C.c_int, C.c_int]),
            "o_check_tanh_select": (L, [L]),
            "o_surrogate_features": (None, [P, C.c_size_t, P, L, P, L, P, L]),
            "o_surrogate_targets": (None, [P, P, P, L, P, L, P]),
            "o_mlp_forward": (None, [P, P, I, P, L, P]),
            "o_mlp_gradient": (C.c_double, [P, P, I, P, P, P, L, P]),
            "o_step_f64": (None, [P, C.c_size_t, L, P, P, P, L]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(_lib, name)
            fn.restype = res
            fn.argtypes = args
    return _lib


def ptr(a: np.ndarray):
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.c_void_p)


def f32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float32)


def set_threads(n: int) -> None:
    lib().o_set_threads(int(n))


def param_count(beams: int, hidden: int) -> int:
    return int(lib().o_param_count(beams, hidden))


class Sim:
    """Reference Sim on the host (racer/sim.mojo:269-348)."""

    def __init__(self, track: np.ndarray, n: int, seed: int, beams: int = 108, heading_noise: float = 0.0):
        self.track = f32(track)
        self.n, self.seed, self.beams, self.noise = n, seed, beams, float(heading_noise)
        self.IN = beams + 5
        self.state = np.zeros((11, n), np.float32)
        self.obs = np.zeros((self.IN, n), np.float32)
        self.actions = np.zeros((2, n), np.float32)
        self.reward = np.zeros(n, np.float32)
        self.done = np.zeros(n, np.float32)
        if len(self.track) < 12:
            raise ValueError("not a track array")

    def reset(self) -> None:
        lib().o_sim_reset(ptr(self.state), ptr(self.actions), ptr(self.reward), ptr(self.done), ptr(self.obs), self.n,
                          self.beams, ptr(self.track), self.track.size, self.seed, self.noise)

    def step(self, actions=None, reward=None, done=None, obs=None) -> int:
        """Physics then lidar; returns the total number of lidar march iterations."""
        actions = self.actions if actions is None else actions
        reward = self.reward if reward is None else reward
        done = self.done if done is None else done
        obs = self.obs if obs is None else obs
        return int(lib().o_sim_step(ptr(self.state), self.n, ptr(actions), actions.shape[-1], ptr(reward), ptr(done),
                                    ptr(obs), obs.shape[-1], self.beams, ptr(self.track), self.track.size, self.seed,
                                    self.noise))

    def sample(self, out, log_std, actions, logp, value, deterministic: bool, step: int) -> None:
        lib().o_sim_sample(self.n, ptr(out), out.shape[-1], ptr(f32(log_std)), ptr(actions), actions.shape[-1],
                           ptr(logp), ptr(value), int(deterministic), int(step), self.seed)


class Policy:
    """Reference Policy on the host (racer/ppo.mojo:63-138)."""

    def __init__(self, seed: int, beams: int = 108, hidden: int = 64):
        self.beams, self.hidden = beams, hidden
        self.IN, self.HA = beams + 5, hidden + 1
        self.P = param_count(beams, hidden)
        self.W1 = hidden * self.IN
        self.W2 = self.W1 + hidden * self.HA
        self.LOGSTD = self.W2 + 3 * self.HA
        self.theta = np.zeros(self.P, np.float32)
        lib().o_policy_init(ptr(self.theta), beams, hidden, seed)

    def log_std(self) -> np.ndarray:
        return self.theta[self.LOGSTD:]

    def forward(self, x: np.ndarray):
        cols = x.shape[1]
        h1 = np.ones((self.HA, cols), np.float32)
        h2 = np.ones((self.HA, cols), np.float32)
        o = np.zeros((3, cols), np.float32)
        lib().o_forward(ptr(self.theta), self.beams, self.hidden, ptr(f32(x)), cols, cols, ptr(h1), ptr(h2), ptr(o))
        return h1, h2, o

    def backward(self, x, h1, h2, d3):
        cols = x.shape[1]
        d2 = np.zeros((self.HA, cols), np.float32)
        d1 = np.zeros((self.HA, cols), np.float32)
        grad = np.zeros(self.P, np.float32)
        lib().o_backward(ptr(self.theta), self.beams, self.hidden, ptr(f32(x)), cols, ptr(f32(h1)), ptr(f32(h2)),
                         ptr(f32(d3)), ptr(d2), ptr(d1), ptr(grad))
        return d2, d1, grad


def gae(n, rollout, reward, done, value, last_value) -> np.ndarray:
    adv = np.zeros(rollout * n, np.float32)
    lib().o_gae(n, rollout, ptr(f32(reward)), ptr(f32(done)), ptr(f32(value)), ptr(f32(last_value)), ptr(adv))
    return adv


def moments(advantage: np.ndarray):
    a, b = C.c_double(), C.c_double()
    adv = f32(advantage)
    lib().o_moments(adv.size, ptr(adv), C.byref(a), C.byref(b))
    return a.value, b.value


def minibatch(policy: Policy, n, rollout, minibatches, obs, actions, logp, value, advantage, m, adv_mean, adv_scale):
    batch = n * rollout // minibatches
    IN, HA = policy.IN, policy.HA
    x = np.zeros((IN, batch), np.float32)
    h1 = np.ones((HA, batch), np.float32)
    h2 = np.ones((HA, batch), np.float32)
    o = np.zeros((3, batch), np.float32)
    d3 = np.zeros((6, batch), np.float32)
    d2 = np.zeros((HA, batch), np.float32)
    d1 = np.zeros((HA, batch), np.float32)
    grad = np.zeros(policy.P, np.float32)
    norm = C.c_float()
    lib().o_minibatch(ptr(policy.theta), policy.beams, policy.hidden, n, rollout, minibatches, ptr(f32(obs)),
                      ptr(f32(actions)), ptr(f32(logp)), ptr(f32(value)), ptr(f32(advantage)), m, adv_mean, adv_scale,
                      ptr(x), ptr(h1), ptr(h2), ptr(o), ptr(d3), ptr(d2), ptr(d1), ptr(grad), C.byref(norm))
    return dict(x=x, h1=h1, h2=h2, o=o, d3=d3, d2=d2, d1=d1, grad=grad, norm_sq=norm.value)


def adam(policy: Policy, grad, first, second, norm_sq, step) -> None:
    lib().o_adam(ptr(policy.theta), policy.P, policy.LOGSTD, ptr(f32(grad)), ptr(first), ptr(second), float(norm_sq),
                 int(step))


def derivative(p, steer, accel, grip, along, across) -> np.ndarray:
    out = np.zeros(4, np.float32)
    lib().o_derivative(ptr(f32(p)), steer, accel, grip, along, across, ptr(out))
    return out


def philox(ctr, key) -> list[int]:
    c = np.ascontiguousarray(ctr, np.uint32)
    k = np.ascontiguousarray(key, np.uint32)
    out = np.zeros(4, np.uint32)
    lib().o_philox(ptr(c), ptr(k), ptr(out))
    return [int(v) for v in out]


def normal4(seed, subsequence, offset) -> np.ndarray:
    out = np.zeros(4, np.float32)
    lib().o_normal4(seed, subsequence, offset, ptr(out))
    return out


def uniform4(seed, subsequence, offset) -> np.ndarray:
    out = np.zeros(4, np.float32)
    lib().o_uniform4(seed, subsequence, offset, ptr(out))
    return out


class Map:
    """Reference Map lookups on the host (racer/sim.mojo:81-152)."""

    def __init__(self, track: np.ndarray):
        self.d = f32(track)
        if self.d.size < 12:
            raise ValueError("not a track array")
        h = self.d
        self.nx, self.ny, self.cell = int(h[0]), int(h[1]), float(h[4])
        self.waypoints, self.layers, self.slope = int(h[5]), int(h[7]), float(h[8])
        self.route = 12 + 3 * self.layers * self.nx * self.ny

    def waypoint(self, w: int, field: int) -> float:
        return float(self.d[self.route + 4 * w + field])

    def probe(self, x, y, z):
        out = np.zeros(3, np.float32)
        lib().o_probe(ptr(self.d), self.d.size, x, y, z, ptr(out))
        return int(out[0]), float(out[1]), float(out[2])

    def surface(self, layer, x, y):
        out = np.zeros(3, np.float32)
        lib().o_surface(ptr(self.d), self.d.size, layer, x, y, ptr(out))
        return float(out[0]), float(out[1]), float(out[2])

    def progress(self, layer, x, y) -> float:
        return float(lib().o_progress(ptr(self.d), self.d.size, layer, x, y))


def spawn(track, state: np.ndarray, i: int, seed: int, heading_noise: float = 0.0) -> None:
    """spawn (racer/sim.mojo:170-181) of car i of state [11, n], in place."""
    t = f32(track)
    assert state.dtype == np.float32 and state.flags["C_CONTIGUOUS"]
    if lib().o_spawn_one(ptr(t), t.size, ptr(state), state.shape[1], int(i), int(seed), float(heading_noise)):
        raise ValueError("not a track array")


def elementary(name: str):
    return getattr(lib(), "o_" + name)


class Trainer:
    """Reference Trainer on the host (racer/ppo.mojo:141-337), for small parity runs."""

    def __init__(self, n: int, policy: Policy, rollout: int = 32, epochs: int = 4, minibatches: int = 4):
        self.n, self.R, self.E, self.MB = n, rollout, epochs, minibatches
        self.policy = policy
        IN = policy.IN
        self.obs = np.zeros((rollout + 1, IN, n), np.float32)
        self.actions = np.zeros((rollout, 2, n), np.float32)
        self.logp = np.zeros((rollout, n), np.float32)
        self.value = np.zeros((rollout, n), np.float32)
        self.reward = np.zeros((rollout, n), np.float32)
        self.done = np.zeros((rollout, n), np.float32)
        self.advantage = np.zeros(rollout * n, np.float32)
        self.first = np.zeros(policy.P, np.float32)
        self.second = np.zeros(policy.P, np.float32)
        self.step = 0
        self.last_o = None

    def rollout(self, sim: Sim, iteration: int) -> float:
        """ppo.mojo:187-232"""
        n = self.n
        self.obs[0] = sim.obs
        for t in range(self.R):
            _, _, o = self.policy.forward(self.obs[t])
            sim.sample(o, self.policy.log_std(), self.actions[t], self.logp[t], self.value[t], False,
                       iteration * self.R + t)
            sim.step(self.actions[t], self.reward[t], self.done[t], self.obs[t + 1])
        sim.obs[...] = self.obs[self.R]
        _, _, o = self.policy.forward(self.obs[self.R])
        self.last_o = o
        self.advantage[:] = gae(n, self.R, self.reward, self.done, self.value, o[2])
        return float(self.reward.sum(dtype=np.float64) / (self.R * n))

    def update(self) -> float:
        """ppo.mojo:234-257, with the advantage moments in float64."""
        samples = self.R * self.n
        s, sq = moments(self.advantage)
        mean = s / samples
        scale = 1.0 / np.sqrt(max(sq / samples - mean * mean, 1e-8))
        res = None
        for epoch in range(self.E):
            for m in range(self.MB):
                res = self.minibatch((m + epoch) % self.MB, float(mean), float(scale))
                self.step += 1
                adam(self.policy, res["grad"], self.first, self.second, res["norm_sq"], self.step)
        return float(res["d3"][5].sum(dtype=np.float64) / res["d3"].shape[1])

    def minibatch(self, m, adv_mean, adv_scale):
        return minibatch(self.policy, self.n, self.R, self.MB, self.obs, self.actions, self.logp, self.value,
                         self.advantage, m, adv_mean, adv_scale)


# ---------------- learned dynamics surrogate (no reference counterpart)
SURROGATE_WIDTHS = (11, 128, 128, 128, 5)


def surrogate_features(track, state, actions) -> np.ndarray:
    """The surrogate's 11 inputs and constant row for every car, [12, n]."""
    t = f32(track)
    n = state.shape[-1]
    feat = np.zeros((12, n), np.float32)
    lib().o_surrogate_features(ptr(t), t.size, ptr(f32(state)), n, ptr(f32(actions)), actions.shape[-1], ptr(feat), n)
    return feat


def surrogate_targets(prev, state, done):
    n = state.shape[-1]
    tgt = np.zeros((5, n), np.float32)
    mask = np.zeros(n, np.float32)
    lib().o_surrogate_targets(ptr(f32(prev)), ptr(f32(state)), ptr(f32(done)), n, ptr(tgt), n, ptr(mask))
    return tgt, mask


def surrogate_params(widths=SURROGATE_WIDTHS) -> int:
    return sum(widths[l + 1] * (widths[l] + 1) for l in range(len(widths) - 1))


def mlp_forward(theta, x, widths=SURROGATE_WIDTHS) -> list:
    """Per-layer outputs (hidden ones with their constant row) of the float32 MLP."""
    w = np.array(widths, np.int32)
    cols = x.shape[1]
    layers = len(widths) - 1
    sizes = [(widths[l + 1] + (1 if l < layers - 1 else 0)) for l in range(layers)]
    acts = np.zeros(sum(sizes) * cols, np.float32)
    lib().o_mlp_forward(ptr(f32(theta)), ptr(w), layers, ptr(f32(x)), cols, ptr(acts))
    out, o = [], 0
    for r in sizes:
        out.append(acts[o: o + r * cols].reshape(r, cols))
        o += r * cols
    return out


def mlp_gradient(theta, x, y, mask, widths=SURROGATE_WIDTHS):
    """(loss, gradient) of the masked mean squared error, float32 orders of the HIP surrogate."""
    w = np.array(widths, np.int32)
    grad = np.zeros(surrogate_params(widths), np.float32)
    loss = lib().o_mlp_gradient(ptr(f32(theta)), ptr(w), len(widths) - 1, ptr(f32(x)), ptr(f32(y)), ptr(f32(mask)),
                                x.shape[1], ptr(grad))
    return float(loss), grad


def step_f64(track, st: np.ndarray, friction, actions) -> None:
    """One float64 RK4 environment step of st = [x, y, z, heading, speed, steer] rows (in place)."""
    t = f32(track)
    assert st.dtype == np.float64 and st.flags["C_CONTIGUOUS"]
    n = st.shape[1]
    lib().o_step_f64(ptr(t), t.size, n, ptr(st), ptr(f32(friction)), ptr(f32(actions)), actions.shape[-1])
