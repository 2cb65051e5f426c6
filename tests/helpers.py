"""Shared test helpers: track cache, the reference test's corridor set-up,
a float64 restatement of the reference test's host_loss, and closed forms."""
from __future__ import annotations

import numpy as np

from fixtures import corridor
from warporacer3d_amd import track as T

_TRACKS: dict[str, np.ndarray] = {}


def track(name: str) -> np.ndarray:
    if name == "corridor":
        return corridor()
    if name not in _TRACKS:
        _TRACKS[name] = T.cached(name)
    return _TRACKS[name]


def corridor_state(n: int = 4) -> tuple[np.ndarray, np.ndarray]:
    """tests/test_mojo.mojo:61-68: four cars at x = 1 -- at rest facing the ramp,
    reversed into the wall at full throttle, full throttle, steering left."""
    s = np.zeros((11, n), np.float32)
    s[0], s[9], s[7] = 1, 1, np.float32(0.5 / 0.15)
    s[3, 1] = np.float32(np.pi)
    actions = np.array([0, 0, 0, 1, 0, 1, 1, 1], np.float32).reshape(2, 4)
    return s, actions


def host_loss(w: np.ndarray, beams: int, hidden: int, n: int, minibatches: int, batch: int, obs, actions, logp, value,
              advantage) -> float:
    """tests/test_mojo.mojo:126-159: minibatch 0's clipped PPO objective in
    float64 with raw advantages (vectorised over the batch)."""
    IN, HA, H = beams + 5, hidden + 1, hidden
    W0 = w[: H * IN].reshape(H, IN)
    W1 = w[H * IN : H * IN + H * HA].reshape(H, HA)
    W2 = w[H * IN + H * HA : H * IN + H * HA + 3 * HA].reshape(3, HA)
    ls = w[H * IN + H * HA + 3 * HA :]
    s = np.arange(batch) * minibatches
    t, i = s // n, s % n
    obs = np.asarray(obs, np.float64).reshape(-1, IN, n)
    x = obs[t, :, i].T  # [IN, batch]
    h1 = np.vstack([np.tanh(W0 @ x), np.ones((1, batch))])
    h2 = np.vstack([np.tanh(W1 @ h1), np.ones((1, batch))])
    o = W2 @ h2
    act = np.asarray(actions, np.float64).reshape(-1, 2, n)[t, :, i].T  # [2, batch]
    half_log_2pi = float(np.float32(0.5 * np.log(2 * np.pi)))
    new_logp = -np.sum(0.5 * (act - o[:2]) ** 2 * np.exp(-2 * ls)[:, None] + ls[:, None] + half_log_2pi, axis=0)
    lp = np.asarray(logp, np.float64).ravel()[s]
    adv = np.asarray(advantage, np.float64).ravel()[s]
    val = np.asarray(value, np.float64).ravel()[s]
    ratio = np.exp(new_logp - lp)
    total = -np.sum(np.minimum(ratio * adv, np.clip(ratio, 0.8, 1.2) * adv)) / batch
    e = np.abs(o[2] - (val + adv))
    total += np.sum(np.where(e <= 1, 0.25 * e * e, 0.5 * e - 0.25)) / batch
    return float(total)


def ray_circle_range(ox, oy, dx, dy, radii) -> float:
    """Distance along the ray (ox, oy) + t (dx, dy) to the nearest circle about the origin."""
    best = np.inf
    for r in radii:
        b = ox * dx + oy * dy
        c = ox * ox + oy * oy - r * r
        disc = b * b - c
        if disc < 0:
            continue
        for t in (-b - np.sqrt(disc), -b + np.sqrt(disc)):
            if t > 1e-9:
                best = min(best, t)
    return best


def host_gae(reward, done, value, last_value, n, rollout, gamma=0.99, lam=0.95):
    """tests/test_mojo.mojo:182-191: GAE recomputed on the host in float32."""
    adv = np.zeros(rollout * n, np.float32)
    reward, done, value = (np.asarray(a, np.float32).ravel() for a in (reward, done, value))
    g, gl = np.float32(gamma), np.float32(gamma * lam)
    for i in range(n):
        running = np.float32(0)
        nxt = np.float32(last_value[i])
        for t in range(rollout - 1, -1, -1):
            k = t * n + i
            alive = np.float32(1 if done[k] == 0 else 0)
            running = reward[k] + g * alive * nxt - value[k] + gl * alive * running
            adv[k] = running
            nxt = value[k]
    return adv
