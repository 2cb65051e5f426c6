"""Track compiler: closed routes packed into the float32 track array the
simulation uploads once (racer/track.py:15-18, 108-125).

Array layout (one float32 array): header[12] = nx, ny, x0, y0, cell, waypoints,
spacing, layers, steepest slope, 0, 0, 0; then for every layer clearance[ny*nx],
height[ny*nx], nearest waypoint[ny*nx]; then route[waypoints*4] = (x, y, z,
heading). Cell (ix, iy) is centred at (x0 + ix*cell, y0 + iy*cell).

Ribbons (racer/track.py:61-105): a road `half_width` either side of the route,
rising `bank` metres per metre outward; clearance is the analytic distance to
the ribbon's edges; where the route passes over itself, the raised pass and the
whole hump it belongs to form layer 1. Off the road, heights continue from the
nearest road cell and are blurred (Gaussian, 1.5 cells).

Built-in names: the reference's `flat`, `ramp`, `bank`, `bridge`
(racer/track.py:38-44), plus the benchmark workloads `ring10`
(R = 10 m, 0.1 m waypoints), `ramp10` (height 0.3 sin 3t), `bank10`
(0.1 m rise per metre outward), `bridge8` (figure eight of two 8 m lobes, deck
0.5 m up) and `random<k>` (closed Catmull-Rom route through 12 seeded control
points). The reference's YAML occupancy-image maps need a skeletonizer and are
not provided; YAML `route:` files are.
"""
from __future__ import annotations

from pathlib import Path

import numpy as np
from scipy.ndimage import distance_transform_edt, gaussian_filter
from scipy.spatial import cKDTree

SPACING = 0.15  # metres between waypoints in the reference's built-in tracks
HEADER = 12


def arc(route: np.ndarray) -> np.ndarray:
    """Cumulative distance to every point and back around to the first (racer/track.py:133-135)."""
    step = np.linalg.norm(np.roll(route, -1, 0) - route, axis=1)
    return np.concatenate([[0.0], np.cumsum(step)])


def resample(route: np.ndarray, spacing: float = SPACING) -> np.ndarray:
    """Evenly spaced points along the closed polyline (racer/track.py:138-142)."""
    s = arc(route)
    even = np.linspace(0, s[-1], max(2, round(s[-1] / spacing)), endpoint=False)
    closed = np.vstack([route, route[:1]])
    return np.column_stack([np.interp(even, s, closed[:, i]) for i in range(3)])


def headings(route: np.ndarray) -> np.ndarray:
    d = np.roll(route, -1, 0) - route
    return np.arctan2(d[:, 1], d[:, 0])


def centres(shape, x0, y0, cell):
    x, y = np.meshgrid(x0 + cell * np.arange(shape[1]), y0 + cell * np.arange(shape[0]))
    return np.column_stack([x.ravel(), y.ravel()])


def _passes(near: np.ndarray, count: int) -> list[np.ndarray]:
    """Split sorted waypoint indices into runs (gaps > 2 start a new pass); the
    run through the route's start and the run through its end are one pass."""
    passes = np.split(near, np.nonzero(np.diff(near) > 2)[0] + 1)
    if len(passes) > 1 and near[0] == 0 and near[-1] == count - 1:
        passes[0] = np.concatenate([passes.pop(), passes[0]])
    return passes


def _cell_passes(route, xy, radius, chunk=65536, k=128):
    """For every cell: the nearest waypoint of each pass of the route within
    `radius`, sorted by height, reduced to the single nearest pass when all the
    passes are within 0.3 m of height. Returns (cells, waypoints, upper_mask)."""
    tree = cKDTree(route[:, :2])
    count = len(route)
    cells, wps = [], []
    upper = np.zeros(count, bool)
    for start in range(0, len(xy), chunk):
        pts = xy[start : start + chunk]
        dist, idx = tree.query(pts, k=min(k, count), distance_upper_bound=radius)
        if idx.ndim == 1:
            dist, idx = dist[:, None], idx[:, None]
        found = idx < count
        any_near = found.any(1)
        full = found.all(1) & (min(k, count) < count)  # may have missed points: exact query
        srt = np.sort(np.where(found, idx, count + 10), axis=1)
        gaps = (np.diff(srt, axis=1) > 2) & (srt[:, 1:] < count)
        runs = gaps.sum(1) + 1
        wrap = found.any(1) & (srt[:, 0] == 0) & (np.take_along_axis(srt, (found.sum(1) - 1)[:, None].clip(0), 1)[:, 0] == count - 1)
        single = any_near & ~full & ((runs == 1) | ((runs == 2) & wrap))
        rows = np.nonzero(single)[0]
        dmin = np.where(found, dist, np.inf).argmin(1)
        cells.append(start + rows)
        wps.append(idx[rows, dmin[rows]])
        for r in np.nonzero(any_near & ~single)[0]:
            near = np.array(sorted(tree.query_ball_point(pts[r], radius)))
            nearest = [p[np.linalg.norm(route[p, :2] - pts[r], axis=1).argmin()] for p in _passes(near, count)]
            nearest.sort(key=lambda w: route[w, 2])
            if route[nearest[-1], 2] - route[nearest[0], 2] < 0.3:
                nearest = [min(nearest, key=lambda w: np.linalg.norm(route[w, :2] - pts[r]))]
            upper[nearest[1:]] = True
            cells.append(np.full(len(nearest), start + r))
            wps.append(np.array(nearest))
    return np.concatenate(cells).astype(np.int64), np.concatenate(wps).astype(np.int64), upper


def ribbon(route, half_width, bank, cell, ramps=(), spacing=SPACING):
    """racer/track.py:61-105, restated with a chunked nearest-neighbour search."""
    route = resample(np.asarray(route, float), spacing)
    heading = headings(route)
    grade = np.gradient(route[:, 2], arc(route)[:-1])
    lo = route[:, :2].min(0) - 2
    shape = tuple(reversed((np.ceil((route[:, :2].max(0) + 2 - lo) / cell).astype(int) + 1).tolist()))
    xy = centres(shape, *lo, cell)
    cells, wps, upper = _cell_passes(route, xy, half_width + 0.5)
    high = route[:, 2] > route[:, 2].min() + 0.1
    while True:  # the upper layer extends along the route while the road stays raised
        grown = upper | ((np.roll(upper, 1) | np.roll(upper, -1)) & high)
        if (grown == upper).all():
            break
        upper = grown
    flag = upper[wps]
    layers = [(cells[flag == l], wps[flag == l]) for l in (False, True) if (flag == l).any()]
    nl = len(layers)
    clearance = np.full((nl,) + shape, -10.0)
    height = np.zeros(clearance.shape)
    nearest = cKDTree(route[:, :2]).query(xy)[1].reshape(shape)
    waypoint = np.tile(nearest, (nl, 1, 1)).astype(float)
    for layer, (c, w) in enumerate(layers):
        dx, dy = xy[c, 0] - route[w, 0], xy[c, 1] - route[w, 1]
        along = dx * np.cos(heading[w]) + dy * np.sin(heading[w])
        left = -dx * np.sin(heading[w]) + dy * np.cos(heading[w])
        clearance.reshape(nl, -1)[layer, c] = half_width - abs(left)
        height.reshape(nl, -1)[layer, c] = route[w, 2] + grade[w] * along - bank * left
        waypoint.reshape(nl, -1)[layer, c] = w
    return pack(clearance, height, waypoint, route, *lo, cell, float(np.hypot(abs(grade).max(), bank)), ramps)


def pack(clearance, height, waypoint, route, x0, y0, cell, slope, ramps=()):
    """racer/track.py:108-125: ramps, off-road height fill and blur, header."""
    ny, nx = clearance.shape[1:]
    height = np.array(height, float)
    xy = centres((ny, nx), x0, y0, cell)
    for ramp in ramps:  # a box with a sloped top: rise over `rise` metres, flat, then descend
        c, s = np.cos(ramp["heading"]), np.sin(ramp["heading"])
        u = (xy[:, 0] - ramp["center"][0]) * c + (xy[:, 1] - ramp["center"][1]) * s + ramp["length"] / 2
        v = -(xy[:, 0] - ramp["center"][0]) * s + (xy[:, 1] - ramp["center"][1]) * c
        rise = ramp.get("rise", ramp["length"] / 2)
        top = ramp["height"] * np.clip(np.minimum(u, ramp["length"] - u) / rise, 0, 1)
        height += np.where(abs(v) < ramp["width"] / 2, top, 0).reshape(ny, nx)
        slope += ramp["height"] / rise
    for l in range(len(clearance)):
        iy, ix = distance_transform_edt(clearance[l] <= 0, return_indices=True)[1]
        height[l] = gaussian_filter(height[l, iy, ix], 1.5)
    header = [nx, ny, x0, y0, cell, len(route), arc(route)[-1] / len(route), len(clearance), slope, 0, 0, 0]
    grids = np.stack([clearance, height, waypoint], 1).ravel()
    return np.concatenate([header, grids, np.column_stack([route, headings(route)]).ravel()]).astype(np.float32)


def catmull_rom(points: np.ndarray, samples: int = 64) -> np.ndarray:
    """Closed uniform Catmull-Rom spline through `points` (P x 3)."""
    p = np.asarray(points, float)
    out = []
    t = np.linspace(0, 1, samples, endpoint=False)[:, None]
    for i in range(len(p)):
        p0, p1, p2, p3 = p[i - 1], p[i], p[(i + 1) % len(p)], p[(i + 2) % len(p)]
        out.append(0.5 * ((2 * p1) + (-p0 + p2) * t + (2 * p0 - 5 * p1 + 4 * p2 - p3) * t**2 + (-p0 + 3 * p1 - 3 * p2 + p3) * t**3))
    return np.vstack(out)


def random_route(seed: int) -> tuple[np.ndarray, float]:
    """12 control points at evenly spaced angles, radius U(6, 14) m each, heights
    U(0, A) with amplitude A ~ U(0, 0.4) m, joined by a closed Catmull-Rom spline."""
    rng = np.random.default_rng(seed)
    angle = 2 * np.pi * np.arange(12) / 12
    radius = rng.uniform(6, 14, 12)
    amplitude = rng.uniform(0, 0.4)
    z = rng.uniform(0, amplitude, 12)
    return catmull_rom(np.column_stack([radius * np.cos(angle), radius * np.sin(angle), z])), amplitude


def load(name: str, cell: float | None = None) -> np.ndarray:
    """A named track or a YAML route map, as the float32 track array (racer/track.py:37-58)."""
    if name in ("flat", "ramp", "bank", "bridge"):  # the reference's built-in ribbons
        t = np.linspace(0, 2 * np.pi, 160, endpoint=False)
        c = 0.025 if cell is None else cell
        if name == "bridge":
            z = 0.9 * np.clip(np.cos((t - np.pi) * 1.4), 0, 1) ** 2
            return ribbon(np.column_stack([7 * np.sin(t), 3.5 * np.sin(2 * t), z]), 1.1, 0, c)
        z = 0.7 * (1 - np.cos(2 * t)) if name == "ramp" else 0 * t
        return ribbon(np.column_stack([6 * np.cos(t), 6 * np.sin(t), z]), 1.3, 0.25 * (name == "bank"), c)
    c = 0.025 if cell is None else cell
    t = np.linspace(0, 2 * np.pi, 4096, endpoint=False)
    if name in ("ring10", "ramp10", "bank10"):
        z = 0.3 * np.sin(3 * t) if name == "ramp10" else 0 * t
        bank = 0.1 if name == "bank10" else 0.0
        return ribbon(np.column_stack([10 * np.cos(t), 10 * np.sin(t), z]), 1.3, bank, c, spacing=0.1)
    if name == "bridge8":
        z = 0.5 * np.clip(np.cos((t - np.pi) * 1.4), 0, 1) ** 2
        return ribbon(np.column_stack([16 * np.sin(t), 8 * np.sin(2 * t), z]), 1.3, 0, c, spacing=0.1)
    if name.startswith("random"):
        route, _ = random_route(int(name[6:] or 0))
        return ribbon(route, 1.3, 0, c, spacing=0.1)
    import yaml

    meta = yaml.safe_load(Path(name).read_text())
    c = float(meta.get("resolution", 0.025))
    if "route" in meta:
        route = np.loadtxt(Path(name).parent / meta["route"])[:, :3]
        return ribbon(route, meta.get("half_width", 1.3), meta.get("bank", 0), c, meta.get("ramps", []))
    raise ValueError("occupancy-image maps need a skeletonizer, which this package does not provide: " + name)


def header(track: np.ndarray) -> dict:
    d = np.asarray(track)
    return dict(nx=int(d[0]), ny=int(d[1]), x0=float(d[2]), y0=float(d[3]), cell=float(d[4]), waypoints=int(d[5]),
                spacing=float(d[6]), layers=int(d[7]), slope=float(d[8]))


def cached(name: str, directory: str | Path | None = None) -> np.ndarray:
    """load(name), memoised as a compressed .npz next to the package (or in `directory`)."""
    folder = Path(directory) if directory else Path(__file__).resolve().parent / "_tracks"
    folder.mkdir(parents=True, exist_ok=True)
    path = folder / f"{name}.npz"
    if path.exists():
        with np.load(path, allow_pickle=False) as z:
            return z["track"]
    arr = load(name)
    tmp = folder / f"{name}.tmp.npz"
    np.savez_compressed(tmp, track=arr)
    tmp.replace(path)
    return arr


if __name__ == "__main__":
    import sys

    arr = load(sys.argv[1])
    np.save(sys.argv[2], arr)
    print(header(arr), arr.size)
