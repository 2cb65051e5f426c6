"""Procedural fixtures restated from the reference's test file.

corridor(): the analytic 12 x 6 m corridor of tests/test_mojo.mojo:15-40 -- along
+x from (0, -3), flat then a 1:2 ramp face from x = 8, waypoints every 0.15 m
along y = 0 from x = 0.5; cell 0.05 m, NX = 281, NY = 161, 74 waypoints, one
layer, header slope 0.5. Values are computed in float32 as the reference does.
"""
import numpy as np

CELL = np.float32(0.05)
NX, NY, WP = 281, 161, 74


def ramp(x):
    x = np.float32(x)
    return np.maximum(x - np.float32(8), np.float32(0)) * np.float32(0.5)


def corridor() -> np.ndarray:
    ix = np.arange(NX, dtype=np.float32)
    iy = np.arange(NY, dtype=np.float32)
    x = np.float32(-1) + CELL * ix[None, :]
    y = np.float32(-4) + CELL * iy[:, None]
    x = np.broadcast_to(x, (NY, NX)).astype(np.float32)
    y = np.broadcast_to(y, (NY, NX)).astype(np.float32)
    nearest = np.clip(np.floor((x - np.float32(0.5)) / np.float32(0.15) + np.float32(0.5)).astype(np.int64), 0, WP - 1)
    clearance = np.minimum(x, np.minimum(np.float32(12) - x, np.float32(3) - np.abs(y)))
    height = ramp(x)
    header = np.array([NX, NY, -1, -4, CELL, WP, 0.15, 1, 0.5, 0, 0, 0], np.float32)
    w = np.arange(WP, dtype=np.float32)
    wx = np.float32(0.5) + np.float32(0.15) * w
    route = np.column_stack([wx, np.zeros(WP, np.float32), ramp(wx), np.zeros(WP, np.float32)]).astype(np.float32)
    return np.concatenate([header, clearance.ravel(), height.ravel(), nearest.astype(np.float32).ravel(),
                           route.ravel()]).astype(np.float32)
