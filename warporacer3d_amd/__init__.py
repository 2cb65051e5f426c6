"""warporacer3d_amd: the racer hot path on MI355X (gfx950) behind the
reference's Device / Sim / Policy / Activations / Trainer API
(racer/device.mojo, racer/sim.mojo, racer/ppo.mojo, main.mojo).

This module is a thin ctypes binding of libracer_hip.so (include/racer.h).
There is no CPU fallback: importing works without a GPU, but creating a
Device raises unless a HIP device is present, and a missing library raises.

Buffers are float32 device memory, feature-major [rows, cols] as in the
reference; `mat(buffer, rows, cols, offset)` is the reference's view.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "libracer_hip.so"

# Constants of racer/sim.mojo:14-56 and racer/ppo.mojo:17-37.
STATE = 11
X, Y, Z, HEADING, SPEED, STEER, STEPS, PROGRESS, EPISODE, FRICTION, RETURN = range(11)
OUT = 3
CRASH, TIMEOUT = 1, 2
RANGE = 10.0
FOV = 4.71238898
HORIZON = 3000
# Forward kernels of Runner steps and rollouts (Policy.set_act_kernel; racer.h RACER_ACT_*).
ACT_AUTO, ACT_FUSED, ACT_WIDE, ACT_TILED = range(4)
SPEED_MAX = 5.0
STEER_MAX = 0.4189
GAMMA, LAMBDA = 0.99, 0.95
STATS = 5
REWARD_SLOT, ADV_SLOT, NORM_SLOT, KL_SLOT = 0, 1, 3, 4


class RacerError(RuntimeError):
    """A failed libracer_hip call (the reference raises Error(msg))."""


class Config(C.Structure):
    """Shape parameters that are compile-time constants in the reference."""

    _fields_ = [("beams", C.c_int), ("hidden", C.c_int), ("rollout", C.c_int), ("epochs", C.c_int),
                ("minibatches", C.c_int), ("precision", C.c_int), ("heading_noise", C.c_float)]

    def __init__(self, beams=108, hidden=64, rollout=32, epochs=4, minibatches=4, precision=0, heading_noise=0.0):
        super().__init__(beams, hidden, rollout, epochs, minibatches, precision, heading_noise)

    @property
    def IN(self) -> int:
        return self.beams + 5

    @property
    def HA(self) -> int:
        return self.hidden + 1

    def params(self) -> int:
        return self.hidden * self.IN + self.hidden * self.HA + OUT * self.HA + 2

    @classmethod
    def from_env(cls) -> "Config":
        """The shapes the reference fixes at compile time (-D BEAMS, racer/ppo.mojo:17-26), from the
        environment as the CLI reads them: RACER_BEAMS, RACER_HIDDEN, RACER_ROLLOUT, RACER_EPOCHS,
        RACER_MINIBATCHES and RACER_PRECISION (fp32 or bf16)."""
        e = os.environ
        return cls(beams=int(e.get("RACER_BEAMS", 108)), hidden=int(e.get("RACER_HIDDEN", 64)),
                   rollout=int(e.get("RACER_ROLLOUT", 32)), epochs=int(e.get("RACER_EPOCHS", 4)),
                   minibatches=int(e.get("RACER_MINIBATCHES", 4)),
                   precision=1 if e.get("RACER_PRECISION", "fp32") == "bf16" else 0)


class _View(C.Structure):
    _fields_ = [("buf", C.c_void_p), ("offset", C.c_size_t), ("rows", C.c_int64), ("cols", C.c_int64)]


_lib = None


def lib() -> C.CDLL:
    """The loaded libracer_hip.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise RacerError(f"{LIB_PATH} is missing: run warporacer3d_amd/build.py")
    L = C.CDLL(str(LIB_PATH))
    vp, i, i64, f, sz = C.c_void_p, C.c_int, C.c_int64, C.c_float, C.c_size_t
    cp = C.POINTER(C.c_void_p)
    VP = C.POINTER(_View)
    sigs = {
        "racer_version": (C.c_char_p, []),
        "racer_last_error": (C.c_char_p, []),
        "racer_config_default": (None, [C.POINTER(Config)]),
        "racer_device_create": (i, [i, cp]),
        "racer_device_destroy": (None, [vp]),
        "racer_device_name": (C.c_char_p, [vp]),
        "racer_sync": (i, [vp]),
        "racer_alloc": (i, [vp, sz, f, cp]),
        "racer_free": (None, [vp]),
        "racer_buf_len": (sz, [vp]),
        "racer_buf_ptr": (vp, [vp]),
        "racer_read": (i, [vp, vp]),
        "racer_write": (i, [vp, vp]),
        "racer_read_range": (i, [vp, sz, sz, vp]),
        "racer_write_range": (i, [vp, sz, sz, vp]),
        "racer_upload": (i, [vp, vp, sz, cp]),
        "racer_copy": (i, [vp, sz, vp, sz, sz]),
        "racer_sim_create": (i, [vp, vp, sz, i, i, C.POINTER(Config), cp]),
        "racer_sim_destroy": (None, [vp]),
        "racer_sim_reset": (i, [vp]),
        "racer_sim_step": (i, [vp, VP, VP, VP, VP]),
        "racer_sim_sample": (i, [vp, VP, VP, VP, VP, VP, i, i64]),
        "racer_sim_state": (vp, [vp]),
        "racer_sim_obs": (vp, [vp]),
        "racer_sim_actions": (vp, [vp]),
        "racer_sim_reward": (vp, [vp]),
        "racer_sim_done": (vp, [vp]),
        "racer_sim_cars": (i, [vp]),
        "racer_sim_beams": (i, [vp]),
        "racer_sim_snapshot": (i, [vp, vp, sz]),
        "racer_acts_create": (i, [vp, i, C.POINTER(Config), cp]),
        "racer_acts_destroy": (None, [vp]),
        "racer_acts_h1": (vp, [vp]),
        "racer_acts_h2": (vp, [vp]),
        "racer_acts_o": (vp, [vp]),
        "racer_policy_create": (i, [vp, i, C.POINTER(Config), cp]),
        "racer_policy_destroy": (None, [vp]),
        "racer_policy_theta": (vp, [vp]),
        "racer_policy_params": (sz, [vp]),
        "racer_policy_log_std": (i, [vp, VP]),
        "racer_policy_transpose": (i, [vp]),
        "racer_policy_forward": (i, [vp, VP, vp]),
        "racer_policy_backward": (i, [vp, VP, vp, VP, vp, vp, vp]),
        "racer_policy_backward_noalloc": (i, [vp, VP, vp, VP, vp, vp, vp]),
        "racer_trainer_create": (i, [vp, i, C.POINTER(Config), cp]),
        "racer_trainer_destroy": (None, [vp]),
        "racer_trainer_rollout": (i, [vp, vp, vp, vp, i, C.POINTER(C.c_float)]),
        "racer_trainer_update": (i, [vp, vp, C.POINTER(C.c_float)]),
        "racer_trainer_minibatch": (i, [vp, vp, i, f, f]),
        "racer_trainer_adam": (i, [vp, vp]),
        "racer_trainer_step_count": (i, [vp]),
        "racer_save": (i, [vp, C.c_char_p, i]),
        "racer_load": (i, [vp, C.c_char_p, C.POINTER(C.c_int)]),
        "racer_runner_step": (i, [vp, vp, vp, vp, i, i64]),
        "racer_benchmark": (i, [vp, vp, vp, vp, i, C.POINTER(C.c_double)]),
        "racer_profile_enable": (i, [vp, i]),
        "racer_profile_count": (i, [vp]),
        "racer_profile_get": (i, [vp, i, C.POINTER(C.c_char_p), C.POINTER(C.c_int64), C.POINTER(C.c_double)]),
        "racer_profile_reset": (i, [vp]),
        "racer_sim_count_iterations": (i, [vp, i]),
        "racer_sim_last_iterations": (i64, [vp]),
        "racer_sim_set_schedule": (i, [vp, i]),
        "racer_sim_set_waves_per_car": (i, [vp, i]),
        "racer_sim_waves_per_car": (i, [vp]),
        "racer_sim_set_lidar_coop": (i, [vp, i]),
        "racer_sim_lidar_coop": (i, [vp]),
        "racer_sim_set_fused_step": (i, [vp, i]),
        "racer_sim_scan": (i, [vp]),
        "racer_trainer_set_bf16_inputs": (i, [vp, i]),
        "racer_trainer_bf16_inputs": (i, [vp]),
        "racer_policy_set_act_kernel": (i, [vp, i]),
        "racer_policy_act_kernel": (i, [vp]),
        "racer_runner_autotune": (i, [vp, vp, vp, vp, i, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                      C.POINTER(C.c_int)]),
        "racer_sim_fused_step": (i, [vp]),
        "racer_sim_autotune": (i, [vp, C.POINTER(C.c_int)]),
        "racer_profile_only": (i, [vp, C.c_char_p]),
        "racer_sim_schedule": (i, [vp]),
        "racer_sim_last_trips": (i64, [vp]),
        "racer_check_floor_index": (i, [vp, C.c_uint32, C.POINTER(C.c_uint64)]),
        "racer_graph_runner": (i, [vp, vp, vp, vp, i, i, cp]),
        "racer_graph_rollout": (i, [vp, vp, vp, vp, cp]),
        "racer_graph_launch": (i, [vp, i64]),
        "racer_trainer_rollout_graph": (i, [vp, vp, i, C.POINTER(C.c_float)]),
        "racer_graph_destroy": (None, [vp]),
        "racer_graph_batch": (i, [vp, vp, i, i, cp]),
        "racer_surrogate_create": (i, [vp, i, i, i64, cp]),
        "racer_surrogate_destroy": (None, [vp]),
        "racer_surrogate_params": (i64, [vp]),
        "racer_surrogate_theta": (vp, [vp]),
        "racer_surrogate_mark_updated": (i, [vp]),
        "racer_surrogate_grad": (vp, [vp]),
        "racer_surrogate_stats": (vp, [vp]),
        "racer_surrogate_batch_gradient": (i, [vp, VP, VP, VP]),
        "racer_surrogate_outputs": (vp, [vp]),
        "racer_dataset_create": (i, [vp, i64, cp]),
        "racer_dataset_destroy": (None, [vp]),
        "racer_dataset_count": (i64, [vp]),
        "racer_dataset_features": (vp, [vp]),
        "racer_dataset_targets": (vp, [vp]),
        "racer_dataset_mask": (vp, [vp]),
        "racer_dataset_record": (i, [vp, vp, vp, vp, vp, i, i64]),
        "racer_surrogate_fit_normalization": (i, [vp, vp]),
        "racer_surrogate_train": (i, [vp, vp, i, i64, f]),
        "racer_surrogate_loss": (i, [vp, vp, C.POINTER(C.c_double)]),
        "racer_surrogate_step": (i, [vp, vp, f, vp]),
        "racer_surrogate_imagine": (i, [vp, vp, VP, i, f, vp]),
        "racer_sim_imagine": (i, [vp, VP, i]),
        "racer_surrogate_set_fused": (i, [vp, i]),
        "racer_sim_step_f64": (i, [vp, vp, VP]),
        "racer_surrogate_runner_step": (i, [vp, vp, vp, vp, vp, i, i64, f, vp]),
        "racer_batch_create": (i, [vp, C.POINTER(C.c_void_p), i, C.POINTER(Config), cp]),
        "racer_batch_destroy": (None, [vp]),
        "racer_batch_count": (i, [vp]),
        "racer_batch_cars": (i64, [vp]),
        "racer_batch_step": (i, [vp]),
        "racer_batch_runner_step": (i, [vp, vp, i, i64]),
        "racer_batch_scratch": (vp, [vp]),
        "racer_batch_set_waves_per_car": (i, [vp, i]),
        "racer_batch_waves_per_car": (i, [vp]),
        "racer_batch_autotune": (i, [vp, C.POINTER(C.c_int)]),
        "racer_batch_set_fused_step": (i, [vp, i]),
        "racer_batch_runner_autotune": (i, [vp, vp, i, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
        "racer_batch_fused_step": (i, [vp]),
        "racer_batch_count_iterations": (i, [vp, i]),
        "racer_batch_last_iterations": (i64, [vp]),
    }
    for name in ("obs", "actions", "logp", "value", "reward", "done", "advantage", "grad", "d3", "stats", "acts"):
        sigs[f"racer_trainer_{name}"] = (vp, [vp])
    for name, (res, args) in sigs.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def _check(rc: int) -> None:
    if rc != 0:
        raise RacerError(lib().racer_last_error().decode())


def exported_symbols() -> list[str]:
    """Names of the C ABI entry points declared in include/racer.h."""
    import re

    text = (HERE.parent / "include" / "racer.h").read_text()
    return sorted(set(re.findall(r"\b(racer_[a-z0-9_]+)\(", text)))


def _ptr(a: np.ndarray) -> C.c_void_p:
    return a.ctypes.data_as(C.c_void_p)


class Buffer:
    """Flat float32 device memory (racer/device.mojo:20); owned unless wrapped."""

    def __init__(self, device: "Device", handle: int, owned: bool = True):
        self.device, self.handle, self.owned = device, handle, owned

    def __len__(self) -> int:
        return int(lib().racer_buf_len(self.handle))

    @property
    def ptr(self) -> int:
        return int(lib().racer_buf_ptr(self.handle))

    def read(self) -> np.ndarray:
        return self.device.read(self)

    def write(self, values) -> None:
        self.device.write(self, values)

    def free(self) -> None:
        if self.owned and self.handle:
            lib().racer_free(self.handle)
        self.handle = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class View:
    """mat[rows](buffer, cols, offset) (racer/device.mojo:29-32)."""

    def __init__(self, buffer: Buffer, rows: int, cols: int, offset: int = 0):
        self.buffer, self.rows, self.cols, self.offset = buffer, rows, cols, offset

    @property
    def c(self):
        """Pointer to the C racer_view (valid while this View lives)."""
        self._c = _View(self.buffer.handle, self.offset, self.rows, self.cols)
        return C.byref(self._c)

    def read(self) -> np.ndarray:
        out = np.empty((self.rows, self.cols), np.float32)
        _check(lib().racer_read_range(self.buffer.handle, self.offset, out.size, _ptr(out)))
        return out


def mat(buffer: Buffer, rows: int, cols: int, offset: int = 0) -> View:
    return View(buffer, rows, cols, offset)


class Device:
    """Device(gpu) (racer/device.mojo:35-134). Only the GPU is provided."""

    def __init__(self, gpu: bool = True):
        h = C.c_void_p()
        _check(lib().racer_device_create(int(gpu), C.byref(h)))
        self.handle = h.value
        self.gpu = True

    @property
    def name(self) -> str:
        return lib().racer_device_name(self.handle).decode()

    def alloc(self, count: int, fill: float = 0.0) -> Buffer:
        h = C.c_void_p()
        _check(lib().racer_alloc(self.handle, int(count), float(fill), C.byref(h)))
        return Buffer(self, h.value)

    def upload(self, values) -> Buffer:
        a = np.ascontiguousarray(values, np.float32).ravel()
        h = C.c_void_p()
        _check(lib().racer_upload(self.handle, _ptr(a), a.size, C.byref(h)))
        return Buffer(self, h.value)

    def read(self, buffer: Buffer) -> np.ndarray:
        out = np.empty(len(buffer), np.float32)
        _check(lib().racer_read(buffer.handle, _ptr(out)))
        return out

    def write(self, buffer: Buffer, values) -> None:
        a = np.ascontiguousarray(values, np.float32).ravel()
        if a.size != len(buffer):
            raise RacerError(f"write of {a.size} floats into a buffer of {len(buffer)}")
        _check(lib().racer_write(buffer.handle, _ptr(a)))

    def write_range(self, buffer: Buffer, offset: int, values) -> None:
        a = np.ascontiguousarray(values, np.float32).ravel()
        _check(lib().racer_write_range(buffer.handle, offset, a.size, _ptr(a)))

    def copy(self, dst: Buffer, dst_offset: int, src: Buffer, src_offset: int, count: int) -> None:
        _check(lib().racer_copy(dst.handle, dst_offset, src.handle, src_offset, count))

    def synchronize(self) -> None:
        _check(lib().racer_sync(self.handle))

    def check_floor_index(self, stride: int = 1) -> int:
        """Floats x with |x| < 2^23 (every stride-th pattern, both signs) where
        the padded lidar probe's floor-and-convert instruction and
        (int)floorf(x) differ (racer_check_floor_index)."""
        bad = C.c_uint64(0)
        _check(lib().racer_check_floor_index(self.handle, stride, C.byref(bad)))
        return int(bad.value)

    # kernel timing with HIP events on the library's stream
    def profile(self, on: bool = True) -> None:
        _check(lib().racer_profile_enable(self.handle, int(on)))

    def profile_reset(self) -> None:
        _check(lib().racer_profile_reset(self.handle))

    def profile_only(self, name: str | None) -> None:
        """Time only the kernel called `name` (None: every kernel)."""
        _check(lib().racer_profile_only(self.handle, (name or "").encode()))

    def profile_results(self) -> dict:
        out = {}
        for i in range(lib().racer_profile_count(self.handle)):
            name, n, ms = C.c_char_p(), C.c_int64(), C.c_double()
            _check(lib().racer_profile_get(self.handle, i, C.byref(name), C.byref(n), C.byref(ms)))
            out[name.value.decode()] = (int(n.value), float(ms.value))
        return out

    def __del__(self):
        try:
            if self.handle:
                lib().racer_device_destroy(self.handle)
        except Exception:
            pass
        self.handle = None


def _wrap(device: Device, handle) -> Buffer:
    return Buffer(device, handle, owned=False)


class Sim:
    """Sim(device, track, n, seed) (racer/sim.mojo:269-348)."""

    def __init__(self, device: Device, track, n: int, seed: int, config: Config | None = None):
        self.device = device
        self.config = config or Config()
        arr = np.ascontiguousarray(track, np.float32).ravel()
        h = C.c_void_p()
        _check(lib().racer_sim_create(device.handle, _ptr(arr), arr.size, int(n), int(seed),
                                      C.byref(self.config), C.byref(h)))
        self.handle = h.value
        self.n, self.seed, self.beams = int(n), int(seed), self.config.beams
        self.IN = self.beams + 5
        L = lib()
        self.state = _wrap(device, L.racer_sim_state(self.handle))
        self.obs = _wrap(device, L.racer_sim_obs(self.handle))
        self.actions = _wrap(device, L.racer_sim_actions(self.handle))
        self.reward = _wrap(device, L.racer_sim_reward(self.handle))
        self.done = _wrap(device, L.racer_sim_done(self.handle))

    def reset(self) -> None:
        _check(lib().racer_sim_reset(self.handle))

    def step(self, actions: View | None = None, reward: View | None = None, done: View | None = None,
             obs: View | None = None) -> None:
        n = self.n
        actions = actions or mat(self.actions, 2, n)
        reward = reward or mat(self.reward, 1, n)
        done = done or mat(self.done, 1, n)
        obs = obs or mat(self.obs, self.IN, n)
        _check(lib().racer_sim_step(self.handle, actions.c, reward.c, done.c, obs.c))

    def sample(self, output: View, log_std: View, actions: View, logp: View, value: View, deterministic: bool,
               step: int) -> None:
        _check(lib().racer_sim_sample(self.handle, output.c, log_std.c, actions.c, logp.c, value.c,
                                      int(deterministic), int(step)))

    def snapshot(self) -> np.ndarray:
        """state, obs, actions, reward and done back to back, as the viewer reads them (main.mojo:117-122)."""
        out = np.empty((STATE + self.IN + 4) * self.n, np.float32)
        _check(lib().racer_sim_snapshot(self.handle, _ptr(out), out.size))
        return out

    def count_iterations(self, on: bool = True) -> None:
        _check(lib().racer_sim_count_iterations(self.handle, int(on)))

    def set_schedule(self, schedule: int) -> None:
        """0: wavefront per car with beam refill (default); 1: thread per (car, beam) as the reference
        launches it."""
        _check(lib().racer_sim_set_schedule(self.handle, int(schedule)))

    @property
    def schedule(self) -> int:
        return int(lib().racer_sim_schedule(self.handle))

    def set_waves_per_car(self, waves: int) -> None:
        _check(lib().racer_sim_set_waves_per_car(self.handle, int(waves)))

    @property
    def waves_per_car(self) -> int:
        return int(lib().racer_sim_waves_per_car(self.handle))

    def set_lidar_coop(self, on: bool) -> None:
        """Cooperative (workgroup) or per-wavefront range stores for the lidar; results are identical."""
        _check(lib().racer_sim_set_lidar_coop(self.handle, int(on)))

    @property
    def lidar_coop(self) -> bool:
        return bool(lib().racer_sim_lidar_coop(self.handle))

    def imagine(self, actions: View, steps: int) -> None:
        """`steps` exact physics steps in one launch (actions [2 steps, n]), no lidar: open-loop planning."""
        _check(lib().racer_sim_imagine(self.handle, actions.c, int(steps)))

    def scan(self) -> None:
        """The lidar pass alone on the current state and obs slope rows (into obs)."""
        _check(lib().racer_sim_scan(self.handle))

    def set_fused_step(self, on: bool) -> None:
        """Run physics and lidar as one launch (k_step_fused) where it applies; results are identical."""
        _check(lib().racer_sim_set_fused_step(self.handle, int(on)))

    @property
    def fused_step(self) -> bool:
        """Whether the next step runs as one launch."""
        return bool(lib().racer_sim_fused_step(self.handle))

    def autotune(self) -> int:
        """Time replays of captured sim steps with 1, 2 and 4 wavefronts per car, cooperative or
        per-wavefront range stores or fused, from the same saved state; keep the fastest (restoring the
        state) and return its wavefronts."""
        w = C.c_int()
        _check(lib().racer_sim_autotune(self.handle, C.byref(w)))
        return int(w.value)

    @property
    def last_iterations(self) -> int:
        return int(lib().racer_sim_last_iterations(self.handle))

    def step_f64(self, st: Buffer, actions: View) -> None:
        """One float64 RK4 step of the kinematics: st holds [6, n] doubles (x, y, z, heading, speed,
        steer) as 12 n floats; friction from this sim's state."""
        _check(lib().racer_sim_step_f64(self.handle, st.handle, actions.c))

    @property
    def last_trips(self) -> int:
        """Wavefront loop trips of the last counted lidar launch (schedule 0)."""
        return int(lib().racer_sim_last_trips(self.handle))

    def __del__(self):
        try:
            if self.handle:
                lib().racer_sim_destroy(self.handle)
        except Exception:
            pass
        self.handle = None


class Activations:
    """Activations(device, cols) (racer/ppo.mojo:47-60)."""

    def __init__(self, device: Device, cols: int, config: Config | None = None, handle=None):
        self.device = device
        self.config = config or Config()
        self.owned = handle is None
        if handle is None:
            h = C.c_void_p()
            _check(lib().racer_acts_create(device.handle, int(cols), C.byref(self.config), C.byref(h)))
            handle = h.value
        self.handle = handle
        self.cols = int(cols)
        L = lib()
        self.h1 = _wrap(device, L.racer_acts_h1(handle))
        self.h2 = _wrap(device, L.racer_acts_h2(handle))
        self.o = _wrap(device, L.racer_acts_o(handle))

    def __del__(self):
        try:
            if self.owned and self.handle:
                lib().racer_acts_destroy(self.handle)
        except Exception:
            pass
        self.handle = None


class Policy:
    """Policy(device, seed) (racer/ppo.mojo:63-138)."""

    def __init__(self, device: Device, seed: int, config: Config | None = None):
        self.device = device
        self.config = config or Config()
        h = C.c_void_p()
        _check(lib().racer_policy_create(device.handle, int(seed), C.byref(self.config), C.byref(h)))
        self.handle = h.value
        self.theta = _wrap(device, lib().racer_policy_theta(self.handle))
        cfg = self.config
        self.IN, self.HA, self.H = cfg.IN, cfg.HA, cfg.hidden
        self.W0 = 0
        self.W1 = cfg.hidden * cfg.IN
        self.W2 = self.W1 + cfg.hidden * cfg.HA
        self.LOGSTD = self.W2 + OUT * cfg.HA
        self.P = int(lib().racer_policy_params(self.handle))

    def log_std(self) -> View:
        return mat(self.theta, 1, 2, self.LOGSTD)

    def transpose(self) -> None:
        _check(lib().racer_policy_transpose(self.handle))

    def forward(self, x: View, acts: Activations) -> None:
        _check(lib().racer_policy_forward(self.handle, x.c, acts.handle))

    def backward(self, x: View, acts: Activations, d3: View, d2: Buffer, d1: Buffer, grad: Buffer) -> None:
        _check(lib().racer_policy_backward(self.handle, x.c, acts.handle, d3.c, d2.handle, d1.handle, grad.handle))

    def set_act_kernel(self, kernel: int) -> None:
        """The forward kernel of Runner steps and rollouts: ACT_AUTO, ACT_FUSED, ACT_WIDE or ACT_TILED
        (identical results)."""
        _check(lib().racer_policy_set_act_kernel(self.handle, int(kernel)))

    @property
    def act_kernel(self) -> int:
        return int(lib().racer_policy_act_kernel(self.handle))

    def backward_noalloc(self, x: View, acts: Activations, d3: View, d2: Buffer, d1: Buffer, grad: Buffer) -> None:
        """backward with the split-K partials in the GPU's fixed 16 MiB scratch: no allocation."""
        _check(lib().racer_policy_backward_noalloc(self.handle, x.c, acts.handle, d3.c, d2.handle, d1.handle,
                                                   grad.handle))

    def __del__(self):
        try:
            if self.handle:
                lib().racer_policy_destroy(self.handle)
        except Exception:
            pass
        self.handle = None


class Trainer:
    """Trainer(device, n) (racer/ppo.mojo:141-337)."""

    def __init__(self, device: Device, n: int, config: Config | None = None):
        self.device = device
        self.config = config or Config()
        h = C.c_void_p()
        _check(lib().racer_trainer_create(device.handle, int(n), C.byref(self.config), C.byref(h)))
        self.handle = h.value
        self.n = int(n)
        self.batch = self.n * self.config.rollout // self.config.minibatches
        L = lib()
        for name in ("obs", "actions", "logp", "value", "reward", "done", "advantage", "grad", "d3", "stats"):
            setattr(self, name, _wrap(device, getattr(L, f"racer_trainer_{name}")(self.handle)))
        self.acts = Activations(device, self.batch, self.config, handle=L.racer_trainer_acts(self.handle))

    @property
    def step(self) -> int:
        return int(lib().racer_trainer_step_count(self.handle))

    def rollout(self, sim: Sim, policy: Policy, acts: Activations, iteration: int) -> float:
        r = C.c_float()
        _check(lib().racer_trainer_rollout(self.handle, sim.handle, policy.handle, acts.handle, int(iteration),
                                           C.byref(r)))
        return float(r.value)

    def update(self, policy: Policy) -> float:
        kl = C.c_float()
        _check(lib().racer_trainer_update(self.handle, policy.handle, C.byref(kl)))
        return float(kl.value)

    def minibatch(self, policy: Policy, m: int, adv_mean: float, adv_scale: float) -> None:
        _check(lib().racer_trainer_minibatch(self.handle, policy.handle, int(m), float(adv_mean), float(adv_scale)))

    def adam(self, policy: Policy) -> None:
        _check(lib().racer_trainer_adam(self.handle, policy.handle))

    def set_bf16_inputs(self, on: bool) -> None:
        """Keep the update's permuted minibatch inputs in bf16 where the bf16 kernels read them so (default)."""
        _check(lib().racer_trainer_set_bf16_inputs(self.handle, int(on)))

    @property
    def bf16_inputs(self) -> bool:
        return bool(lib().racer_trainer_bf16_inputs(self.handle))

    def capture_rollout(self, sim: Sim, policy: Policy, acts: Activations) -> "Graph":
        """The whole rollout as one HIP graph (replay with rollout_graph)."""
        h = C.c_void_p()
        _check(lib().racer_graph_rollout(self.handle, sim.handle, policy.handle, acts.handle, C.byref(h)))
        return Graph(h.value, (self, sim, policy, acts))

    def rollout_graph(self, graph: "Graph", iteration: int) -> float:
        """Trainer.rollout by replaying a captured graph; same results."""
        r = C.c_float()
        _check(lib().racer_trainer_rollout_graph(self.handle, graph.handle, int(iteration), C.byref(r)))
        return float(r.value)

    def __del__(self):
        try:
            if self.handle:
                lib().racer_trainer_destroy(self.handle)
        except Exception:
            pass
        self.handle = None


def runner_step(sim: "Sim", policy: "Policy", acts: "Activations", scratch: Buffer, deterministic: bool,
                t: int) -> None:
    """Runner.step (main.mojo:183-191) on explicit objects: forward, sample into
    sim.actions and scratch (log probabilities, then values), physics, lidar."""
    _check(lib().racer_runner_step(sim.handle, policy.handle, acts.handle, scratch.handle, int(deterministic),
                                   int(t)))


def autotune_runner(sim: "Sim", policy: "Policy", acts: "Activations", scratch: Buffer, steps: int = 8) -> dict:
    """Time `steps` Runner steps replayed from graphs for every launch-shape
    choice (lidar wavefronts per car, range stores, fused sim step, forward kernel) from the
    same saved sim state; keep the fastest in sim and policy and restore the
    sim. Returns {"waves_per_car", "fused_step", "lidar_coop", "act_kernel"}."""
    w, f, k = C.c_int(), C.c_int(), C.c_int()
    _check(lib().racer_runner_autotune(sim.handle, policy.handle, acts.handle, scratch.handle, int(steps),
                                       C.byref(w), C.byref(f), C.byref(k)))
    return {"waves_per_car": int(w.value), "fused_step": bool(f.value), "lidar_coop": sim.lidar_coop,
            "act_kernel": int(k.value)}


def capture_runner(sim: "Sim", policy: "Policy", acts: "Activations", scratch: Buffer, steps: int,
                   deterministic: bool = False) -> "Graph":
    """`steps` runner_step calls on explicit objects as one HIP graph."""
    h = C.c_void_p()
    _check(lib().racer_graph_runner(sim.handle, policy.handle, acts.handle, scratch.handle, int(steps),
                                    int(deterministic), C.byref(h)))
    return Graph(h.value, (sim, policy, acts, scratch))


class Batch:
    """Sims stepped together, one launch per stage (bitwise as stepping each alone)."""

    def __init__(self, device: "Device", sims: list, config: Config | None = None):
        self.device = device
        self.config = config or Config()
        self.sims = list(sims)
        arr = (C.c_void_p * len(self.sims))(*[s.handle for s in self.sims])
        h = C.c_void_p()
        _check(lib().racer_batch_create(device.handle, arr, len(self.sims), C.byref(self.config), C.byref(h)))
        self.handle = h.value
        self.cars = int(lib().racer_batch_cars(self.handle))
        self.scratch = _wrap(device, lib().racer_batch_scratch(self.handle))
        self.starts = [0]
        for s in self.sims:
            self.starts.append(self.starts[-1] + s.n)

    def step(self) -> None:
        _check(lib().racer_batch_step(self.handle))

    def runner_step(self, policy: "Policy", deterministic: bool, t: int) -> None:
        _check(lib().racer_batch_runner_step(self.handle, policy.handle, int(deterministic), int(t)))

    def logp_value(self, k: int):
        """Log probabilities and values of sim k from the last runner step."""
        a = self.scratch.read()
        n, o = self.sims[k].n, 2 * self.starts[k]
        return a[o: o + n], a[o + n: o + 2 * n]

    def set_waves_per_car(self, waves: int) -> None:
        _check(lib().racer_batch_set_waves_per_car(self.handle, int(waves)))

    @property
    def waves_per_car(self) -> int:
        return int(lib().racer_batch_waves_per_car(self.handle))

    def set_fused_step(self, mode: int) -> None:
        """0: physics and lidar as two launches; 1: one fused launch, a workgroup per 8 cars;
        2: the fused launch on a persistent grid (bitwise the same results in every mode)."""
        _check(lib().racer_batch_set_fused_step(self.handle, int(mode)))

    @property
    def fused_step(self) -> int:
        """The step mode that runs (0 when the fused step does not apply to these sims)."""
        return int(lib().racer_batch_fused_step(self.handle))

    def autotune(self) -> int:
        """Time replays of batched sim steps for every wavefront count and step mode from the
        same saved states; keep the fastest (restoring the states) and return its wavefronts."""
        w = C.c_int()
        _check(lib().racer_batch_autotune(self.handle, C.byref(w)))
        return int(w.value)

    def autotune_runner(self, policy: "Policy", steps: int = 8) -> dict:
        """Time `steps` batched Runner steps with `policy` replayed from graphs for every
        wavefront count, step mode and (for per-instance forwards) forward kernel, from the
        same saved states; keep the fastest and restore the states."""
        w, f = C.c_int(), C.c_int()
        _check(lib().racer_batch_runner_autotune(self.handle, policy.handle, int(steps), C.byref(w), C.byref(f)))
        return {"waves_per_car": int(w.value), "fused_step": int(f.value), "act_kernel": policy.act_kernel}

    def count_iterations(self, on: bool = True) -> None:
        _check(lib().racer_batch_count_iterations(self.handle, int(on)))

    @property
    def last_iterations(self) -> int:
        return int(lib().racer_batch_last_iterations(self.handle))

    def capture(self, policy: "Policy", steps: int, deterministic: bool = False) -> "Graph":
        h = C.c_void_p()
        _check(lib().racer_graph_batch(self.handle, policy.handle, int(steps), int(deterministic), C.byref(h)))
        return Graph(h.value, (self, policy))

    def __del__(self):
        try:
            if self.handle:
                lib().racer_batch_destroy(self.handle)
        except Exception:
            pass
        self.handle = None


class Dataset:
    """Recorded transitions of a sim: inputs [12, capacity], targets [5, capacity], mask."""

    def __init__(self, device: "Device", capacity: int):
        self.device = device
        h = C.c_void_p()
        _check(lib().racer_dataset_create(device.handle, int(capacity), C.byref(h)))
        self.handle = h.value
        self.capacity = int(capacity)
        L = lib()
        self.features = _wrap(device, L.racer_dataset_features(self.handle))
        self.targets = _wrap(device, L.racer_dataset_targets(self.handle))
        self.mask = _wrap(device, L.racer_dataset_mask(self.handle))

    @property
    def count(self) -> int:
        return int(lib().racer_dataset_count(self.handle))

    def record(self, sim: "Sim", policy: "Policy", acts: "Activations", scratch: Buffer, steps: int, t0: int = 0):
        _check(lib().racer_dataset_record(self.handle, sim.handle, policy.handle, acts.handle, scratch.handle,
                                          int(steps), int(t0)))

    def __del__(self):
        try:
            if self.handle:
                lib().racer_dataset_destroy(self.handle)
        except Exception:
            pass
        self.handle = None


class Surrogate:
    """The learned dynamics surrogate (11 -> 3 x 128 tanh -> 5 residual MLP)."""

    INPUTS, HIDDEN, OUTPUTS = 11, 128, 5

    def __init__(self, device: "Device", seed: int = 1, precision: int = 1, cols: int = 16384):
        self.device = device
        h = C.c_void_p()
        _check(lib().racer_surrogate_create(device.handle, int(seed), int(precision), int(cols), C.byref(h)))
        self.handle = h.value
        self.cols = int(cols)
        L = lib()
        self.P = int(L.racer_surrogate_params(self.handle))
        self.theta = _wrap(device, L.racer_surrogate_theta(self.handle))
        self.grad = _wrap(device, L.racer_surrogate_grad(self.handle))
        self.stats = _wrap(device, L.racer_surrogate_stats(self.handle))
        self.outputs = _wrap(device, L.racer_surrogate_outputs(self.handle))

    def mark_updated(self) -> None:
        _check(lib().racer_surrogate_mark_updated(self.handle))

    def batch_gradient(self, x: View, y: View, mask: View) -> None:
        _check(lib().racer_surrogate_batch_gradient(self.handle, x.c, y.c, mask.c))

    def fit_normalization(self, data: Dataset) -> None:
        _check(lib().racer_surrogate_fit_normalization(self.handle, data.handle))

    def train(self, data: Dataset, epochs: int, batch: int, lr: float = 1e-3) -> None:
        _check(lib().racer_surrogate_train(self.handle, data.handle, int(epochs), int(batch), float(lr)))

    def loss(self, data: Dataset) -> float:
        v = C.c_double()
        _check(lib().racer_surrogate_loss(self.handle, data.handle, C.byref(v)))
        return float(v.value)

    def set_fused(self, fused: bool) -> None:
        """bf16 imagined steps as one fused launch (default) or as separate kernels (same results)."""
        _check(lib().racer_surrogate_set_fused(self.handle, int(fused)))

    def step(self, sim: "Sim", min_clearance: float = 0.1, fallback: Buffer | None = None) -> None:
        _check(lib().racer_surrogate_step(self.handle, sim.handle, float(min_clearance),
                                          fallback.handle if fallback is not None else None))

    def imagine(self, sim: "Sim", actions: View, steps: int, min_clearance: float = 0.1,
                fallback: Buffer | None = None) -> None:
        """`steps` open-loop imagined steps in one launch (actions [2 steps, n]); no lidar."""
        _check(lib().racer_surrogate_imagine(self.handle, sim.handle, actions.c, int(steps), float(min_clearance),
                                             fallback.handle if fallback is not None else None))

    def runner_step(self, sim: "Sim", policy: "Policy", acts: "Activations", scratch: Buffer, deterministic: bool,
                    t: int, min_clearance: float = 0.1, fallback: Buffer | None = None) -> None:
        _check(lib().racer_surrogate_runner_step(self.handle, sim.handle, policy.handle, acts.handle, scratch.handle,
                                                 int(deterministic), int(t), float(min_clearance),
                                                 fallback.handle if fallback is not None else None))

    def __del__(self):
        try:
            if self.handle:
                lib().racer_surrogate_destroy(self.handle)
        except Exception:
            pass
        self.handle = None


class Graph:
    """A captured HIP graph of Runner steps or of a rollout. Keeps the objects
    it was captured from alive, since the graph holds their buffers."""

    def __init__(self, handle: int, owners: tuple):
        self.handle = handle
        self._owners = owners

    def launch(self, t0: int) -> None:
        _check(lib().racer_graph_launch(self.handle, int(t0)))

    def __del__(self):
        try:
            if self.handle:
                lib().racer_graph_destroy(self.handle)
        except Exception:
            pass
        self.handle = None


def save(policy: Policy, path: str | os.PathLike, iteration: int) -> None:
    """save (racer/ppo.mojo:341-345): float32 .npy of [iteration, theta...]."""
    _check(lib().racer_save(policy.handle, str(path).encode(), int(iteration)))


def load(policy: Policy, path: str | os.PathLike) -> int:
    """load (racer/ppo.mojo:348-355): returns the saved iteration."""
    it = C.c_int()
    _check(lib().racer_load(policy.handle, str(path).encode(), C.byref(it)))
    return int(it.value)


class Runner:
    """A device, n cars on a track and a policy driving them (main.mojo:165-235)."""

    def __init__(self, track, n: int, config: Config | None = None, seed: int = 42, device: Device | None = None):
        self.device = device or Device(True)
        self.config = config or Config()
        self.sim = Sim(self.device, track, n, seed, self.config)
        self.policy = Policy(self.device, seed, self.config)
        self.acts = Activations(self.device, n, self.config)
        self.scratch = self.device.alloc(2 * n)
        self.sim.reset()

    def step(self, deterministic: bool, t: int) -> None:
        _check(lib().racer_runner_step(self.sim.handle, self.policy.handle, self.acts.handle, self.scratch.handle,
                                       int(deterministic), int(t)))

    def autotune(self, steps: int = 8) -> dict:
        """autotune_runner on this runner's objects."""
        return autotune_runner(self.sim, self.policy, self.acts, self.scratch, steps)

    def capture(self, steps: int, deterministic: bool = False) -> Graph:
        """`steps` Runner steps as one HIP graph; graph.launch(t0) runs steps
        t0 .. t0 + steps - 1, bitwise as step(deterministic, t) would."""
        h = C.c_void_p()
        _check(lib().racer_graph_runner(self.sim.handle, self.policy.handle, self.acts.handle, self.scratch.handle,
                                        int(steps), int(deterministic), C.byref(h)))
        return Graph(h.value, (self.sim, self.policy, self.acts, self.scratch))

    def benchmark(self, steps: int) -> float:
        """10 warm-up steps then `steps` timed steps; returns ms/step (main.mojo:225-235)."""
        ms = C.c_double()
        _check(lib().racer_benchmark(self.sim.handle, self.policy.handle, self.acts.handle, self.scratch.handle,
                                     int(steps), C.byref(ms)))
        return float(ms.value)
