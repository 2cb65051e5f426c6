"""Track compiler: a closed route (and its surroundings) compiled into the one
float32 array the simulation uploads once.

The array's layout is the reference's format (racer/track.py:15-18, 123-125),
kept as is because the simulation, saved tracks and the viewer read it:

    header[12]  nx, ny, x0, y0, cell, waypoints, spacing, layers,
                steepest slope, 0, 0, 0
    per layer   clearance[ny*nx], height[ny*nx], nearest waypoint[ny*nx]
    route       [waypoints][x, y, z, heading]

Cell (ix, iy) is centred at (x0 + ix cell, y0 + iy cell); rows run south to
north. Clearance is the signed distance (metres) to the nearest wall.

Sources:
  * ribbons -- a road of constant half width around a closed 3-D route,
    banked outward; where the route passes over itself, the raised pass
    becomes layer 1 (racer/track.py:61-105). Built in: the reference's
    `flat`, `ramp`, `bank`, `bridge` (racer/track.py:38-44) and the
    benchmark workloads `ring10`, `ramp10`, `bank10`, `bridge8`,
    `random<k>`;
  * YAML maps (racer/track.py:45-58): a `route:` text file of x y z rows
    with `half_width` and `bank`, or an occupancy `image:` (a ROS map:
    `resolution`, `origin`, pixels >= 230 free) whose free space is thinned
    to its centreline loop;
  * both with optional `ramps:` boxes (sloped tops) added to the heights.

For the built-in ribbons the arrays equal the reference generator's byte for
byte (tests/test_track.py against digests of its output).
"""
from __future__ import annotations

from collections import deque
from dataclasses import dataclass
from pathlib import Path

import numpy as np
from scipy.ndimage import distance_transform_edt, gaussian_filter, label
from scipy.signal import savgol_filter
from scipy.spatial import cKDTree

SPACING = 0.15     # metres between waypoints of the reference's tracks
HEADER = 12        # floats before the grids
MARGIN = 2.0       # metres of grid beyond the route on every side
FREE_LEVEL = 230   # occupancy-image grey level from which a pixel is drivable
LAYER_GAP = 0.3    # metres of height that make two passes over a cell separate roads
RAISED = 0.1       # metres above the route's lowest point that count as raised


# ---------------------------------------------------------------- geometry
class Route:
    """A closed polyline of 3-D points (the last joins the first)."""

    def __init__(self, points):
        self.p = np.asarray(points, float)

    def __len__(self) -> int:
        return len(self.p)

    def _closing_deltas(self) -> np.ndarray:
        return np.diff(np.vstack([self.p, self.p[:1]]), axis=0)

    def arc(self) -> np.ndarray:
        """Distance along the route at every point and, last, around to the start."""
        d = self._closing_deltas()
        return np.concatenate([[0.0], np.cumsum(np.sqrt((d * d).sum(axis=1)))])

    def even(self, spacing: float) -> "Route":
        """Points every `spacing` metres of arc (the count rounded), linearly interpolated."""
        s = self.arc()
        count = max(2, round(s[-1] / spacing))
        at = np.linspace(0.0, s[-1], count, endpoint=False)
        ring = np.vstack([self.p, self.p[:1]])
        return Route(np.stack([np.interp(at, s, ring[:, c]) for c in range(3)], axis=1))

    def headings(self) -> np.ndarray:
        """Direction of travel at every point: towards the next one."""
        d = self._closing_deltas()
        return np.arctan2(d[:, 1], d[:, 0])

    def grades(self) -> np.ndarray:
        """dz / ds at every point (second-order differences over the arc)."""
        return np.gradient(self.p[:, 2], self.arc()[:-1])


@dataclass(frozen=True)
class Grid:
    """Cell centres x0 + ix cell, y0 + iy cell for ix < nx, iy < ny."""

    x0: float
    y0: float
    cell: float
    nx: int
    ny: int

    @classmethod
    def covering(cls, xy: np.ndarray, cell: float, margin: float = MARGIN) -> "Grid":
        lo = xy.min(0) - margin
        hi = xy.max(0) + margin
        n = np.ceil((hi - lo) / cell).astype(int) + 1
        return cls(lo[0], lo[1], cell, int(n[0]), int(n[1]))

    @property
    def shape(self) -> tuple[int, int]:
        return (self.ny, self.nx)

    def centres(self) -> np.ndarray:
        """[ny * nx, 2] centres, x fastest."""
        x = self.x0 + self.cell * np.arange(self.nx)
        y = self.y0 + self.cell * np.arange(self.ny)
        return np.column_stack([np.tile(x, self.ny), np.repeat(y, self.nx)])


# ---------------------------------------------------------------- ribbon layers
def _runs(sorted_idx: np.ndarray, count: int) -> list[np.ndarray]:
    """Consecutive waypoint indices (gaps of at most 2) form one pass of the
    route; a pass through the route's end and start is one pass."""
    cut = np.flatnonzero(np.diff(sorted_idx) > 2) + 1
    runs = np.split(sorted_idx, cut)
    if len(runs) > 1 and sorted_idx[0] == 0 and sorted_idx[-1] == count - 1:
        runs[0] = np.concatenate([runs.pop(), runs[0]])
    return runs


def _cell_roads(route: np.ndarray, xy: np.ndarray, radius: float, chunk: int = 65536, k: int = 128):
    """The roads over every cell: for each pass of the route within `radius`
    of the cell, its waypoint nearest the cell; when all those waypoints lie
    within LAYER_GAP of one height, only the nearest pass (one road). Returns
    (cells, waypoints, seeds): a (cell, waypoint) pair per road, and the
    waypoints of passes that run above another pass over some cell.

    Cells whose neighbourhood holds one pass (almost all) are settled from a
    k-nearest query per chunk; the rest take a full ball query."""
    tree = cKDTree(route[:, :2])
    count = len(route)
    kk = min(k, count)
    cells, wps = [], []
    seeds = np.zeros(count, bool)
    for start in range(0, len(xy), chunk):
        pts = xy[start: start + chunk]
        dist, idx = tree.query(pts, k=kk, distance_upper_bound=radius)
        if idx.ndim == 1:
            dist, idx = dist[:, None], idx[:, None]
        hit = idx < count
        near_any = hit.any(1)
        maybe_more = hit.all(1) & (kk < count)  # the k nearest may not be all of them
        order = np.sort(np.where(hit, idx, count + 10), axis=1)
        breaks = (np.diff(order, axis=1) > 2) & (order[:, 1:] < count)
        last = np.take_along_axis(order, (hit.sum(1) - 1)[:, None].clip(0), 1)[:, 0]
        wraps = near_any & (order[:, 0] == 0) & (last == count - 1)
        passes = breaks.sum(1) + 1
        simple = near_any & ~maybe_more & ((passes == 1) | ((passes == 2) & wraps))
        rows = np.flatnonzero(simple)
        closest = np.where(hit, dist, np.inf).argmin(1)
        cells.append(start + rows)
        wps.append(idx[rows, closest[rows]])
        for r in np.flatnonzero(near_any & ~simple):
            ball = np.array(sorted(tree.query_ball_point(pts[r], radius)))
            here = pts[r]
            best = [p[np.linalg.norm(route[p, :2] - here, axis=1).argmin()] for p in _runs(ball, count)]
            best.sort(key=lambda w: route[w, 2])
            if route[best[-1], 2] - route[best[0], 2] < LAYER_GAP:
                best = [min(best, key=lambda w: np.linalg.norm(route[w, :2] - here))]
            seeds[best[1:]] = True
            cells.append(np.full(len(best), start + r))
            wps.append(np.array(best))
    return np.concatenate(cells).astype(np.int64), np.concatenate(wps).astype(np.int64), seeds


def _raised_layer(seeds: np.ndarray, raised: np.ndarray) -> np.ndarray:
    """Waypoints of layer 1: the seeds and every stretch of consecutive raised
    waypoints (around the loop) that holds a seed or borders one, so the upper
    road runs on along its whole hump."""
    out = seeds.copy()
    if not seeds.any():
        return out
    if raised.all():
        return np.ones_like(seeds)
    shift = int(np.argmin(raised))  # rotate a lowered waypoint to index 0: stretches do not wrap
    r = np.roll(raised, -shift)
    s = np.roll(seeds, -shift)
    n = len(r)
    mark = np.zeros(n, bool)
    edges = np.diff(np.concatenate([[0], r.astype(np.int8), [0]]))
    for a, b in zip(np.flatnonzero(edges == 1), np.flatnonzero(edges == -1)):
        if s[a:b].any() or s[a - 1] or s[b % n]:
            mark[a:b] = True
    return out | np.roll(mark, shift)


def _road_surface(xy, cells, wps, route, heading, grade, half_width, bank):
    """Clearance and height of the ribbon over the given cells: the cell's
    offset from its waypoint, turned into the waypoint's frame (along, left)."""
    dx = xy[cells, 0] - route[wps, 0]
    dy = xy[cells, 1] - route[wps, 1]
    cos, sin = np.cos(heading[wps]), np.sin(heading[wps])
    along = dx * cos + dy * sin
    left = -dx * sin + dy * cos
    return half_width - np.abs(left), route[wps, 2] + grade[wps] * along - bank * left


def ribbon(points, half_width: float, bank: float, cell: float, ramps=(), spacing: float = SPACING) -> np.ndarray:
    """A road `half_width` either side of the closed route through `points`,
    rising `bank` metres per metre outward (racer/track.py:61-105)."""
    route = Route(points).even(spacing)
    p, heading, grade = route.p, route.headings(), route.grades()
    grid = Grid.covering(p[:, :2], cell)
    xy = grid.centres()
    cells, wps, seeds = _cell_roads(p, xy, half_width + 0.5)
    upper = _raised_layer(seeds, p[:, 2] > p[:, 2].min() + RAISED)[wps]
    groups = [(cells[upper == on], wps[upper == on]) for on in (False, True) if (upper == on).any()]
    layers = len(groups)
    plane = grid.nx * grid.ny
    clearance = np.full((layers, plane), -10.0)
    height = np.zeros((layers, plane))
    waypoint = np.repeat(cKDTree(p[:, :2]).query(xy)[1][None, :].astype(float), layers, axis=0)
    for layer, (c, w) in enumerate(groups):
        clearance[layer, c], height[layer, c] = _road_surface(xy, c, w, p, heading, grade, half_width, bank)
        waypoint[layer, c] = w
    slope = float(np.hypot(abs(grade).max(), bank))
    shape = (layers,) + grid.shape
    return assemble(grid, route, clearance.reshape(shape), height.reshape(shape), waypoint.reshape(shape), slope,
                    ramps)


# ---------------------------------------------------------------- assembly
def _ramp_heights(xy: np.ndarray, ramps) -> tuple[np.ndarray, float]:
    """Boxes with sloped tops (racer/track.py:112-119): each `ramp` has a centre,
    heading, length, width, height and an optional `rise` (default half the
    length); its top climbs over `rise` metres, stays level and falls again.
    Returns the added height per cell and the added slope bound."""
    added = np.zeros(len(xy))
    slope = 0.0
    for ramp in ramps:
        length, rise = ramp["length"], ramp.get("rise", ramp["length"] / 2)
        cos, sin = np.cos(ramp["heading"]), np.sin(ramp["heading"])
        ox, oy = xy[:, 0] - ramp["center"][0], xy[:, 1] - ramp["center"][1]
        u = ox * cos + oy * sin + length / 2   # from the box's back edge, along its heading
        v = -ox * sin + oy * cos               # across it
        top = ramp["height"] * np.clip(np.minimum(u, length - u) / rise, 0, 1)
        added += np.where(abs(v) < ramp["width"] / 2, top, 0)
        slope += ramp["height"] / rise
    return added, slope


def _ground(clearance: np.ndarray, height: np.ndarray) -> np.ndarray:
    """Heights off the road: each such cell takes its nearest road cell's height,
    then everything is blurred (sigma 1.5 cells) so rasterisation steps do not
    read as slopes."""
    nearest_road = distance_transform_edt(clearance <= 0, return_distances=False, return_indices=True)
    return gaussian_filter(height[tuple(nearest_road)], 1.5)


def assemble(grid: Grid, route: Route, clearance, height, waypoint, slope: float, ramps=()) -> np.ndarray:
    """The float32 track array from per-layer [layers, ny, nx] grids."""
    height = np.array(height, float)
    added, ramp_slope = _ramp_heights(grid.centres(), ramps)
    height += added.reshape(grid.shape)
    slope = slope + ramp_slope
    for layer in range(len(clearance)):
        height[layer] = _ground(clearance[layer], height[layer])
    p = route.p
    head = [grid.nx, grid.ny, grid.x0, grid.y0, grid.cell, len(p), route.arc()[-1] / len(p), len(clearance), slope,
            0, 0, 0]
    planes = np.stack([clearance, height, waypoint], axis=1)  # [layer][clearance, height, waypoint]
    table = np.concatenate([p, route.headings()[:, None]], axis=1)
    return np.concatenate([np.asarray(head, float), planes.ravel(), table.ravel()]).astype(np.float32)


# ---------------------------------------------------------------- occupancy images
_NEIGHBOURS = [(dr, dc) for dr in (-1, 0, 1) for dc in (-1, 0, 1) if dr or dc]


def _ring(img: np.ndarray) -> list[np.ndarray]:
    """The 8 neighbours P2 .. P9 of every pixel (N, NE, E, SE, S, SW, W, NW),
    zero beyond the border."""
    p = np.pad(img, 1)
    h, w = img.shape
    at = lambda dr, dc: p[1 + dr: 1 + dr + h, 1 + dc: 1 + dc + w]  # noqa: E731
    return [at(-1, 0), at(-1, 1), at(0, 1), at(1, 1), at(1, 0), at(1, -1), at(0, -1), at(-1, -1)]


def thin(free: np.ndarray) -> np.ndarray:
    """Zhang and Suen's parallel thinning (Commun. ACM 27(3), 1984) to a
    one-pixel-wide skeleton: two sub-iterations remove border pixels with 2 to
    6 set neighbours and exactly one 0 -> 1 transition around them, the first
    those with P2 P4 P6 = P4 P6 P8 = 0, the second those with
    P2 P4 P8 = P2 P6 P8 = 0, until nothing changes."""
    img = np.asarray(free, bool).astype(np.uint8)
    while True:
        changed = False
        for first in (True, False):
            ring = _ring(img)
            count = sum(q.astype(np.int32) for q in ring)
            cycle = ring + ring[:1]
            rises = sum(((cycle[i] == 0) & (cycle[i + 1] == 1)).astype(np.int32) for i in range(8))
            p2, p4, p6, p8 = ring[0], ring[2], ring[4], ring[6]
            if first:
                side = (p2 * p4 * p6 == 0) & (p4 * p6 * p8 == 0)
            else:
                side = (p2 * p4 * p8 == 0) & (p2 * p6 * p8 == 0)
            drop = (img == 1) & (count >= 2) & (count <= 6) & (rises == 1) & side
            if drop.any():
                img[drop] = 0
                changed = True
        if not changed:
            return img.astype(bool)


def centreline(free: np.ndarray) -> np.ndarray:
    """The free space's skeleton without dead ends (tips, pixels with at most
    one skeleton neighbour, removed until none is left), largest 8-connected
    piece only (racer/track.py:150-163)."""
    skel = thin(free)
    while True:
        tips = skel & (sum(q.astype(np.int32) for q in _ring(skel.astype(np.uint8))) <= 1)
        if not tips.any():
            break
        skel &= ~tips
    labels, count = label(skel, structure=np.ones((3, 3)))
    if count == 0:
        raise ValueError("no closed centreline loop in the image")
    return labels == 1 + int(np.bincount(labels.ravel())[1:].argmax())


def _pixel_neighbours(skel: np.ndarray, p: tuple[int, int]) -> list[tuple[int, int]]:
    h, w = skel.shape
    return [(p[0] + dr, p[1] + dc) for dr, dc in _NEIGHBOURS
            if 0 <= p[0] + dr < h and 0 <= p[1] + dc < w and skel[p[0] + dr, p[1] + dc]]


def _bfs(skel: np.ndarray, avoid, src, dst):
    """Shortest 8-connected skeleton path src -> dst not through `avoid`, or None."""
    came = {src: src}
    todo = deque([src])
    while todo and dst not in came:
        u = todo.popleft()
        for v in _pixel_neighbours(skel, u):
            if v not in came and v != avoid:
                came[v] = u
                todo.append(v)
    if dst not in came:
        return None
    walk = [dst]
    while walk[-1] != src:
        walk.append(came[walk[-1]])
    return walk[::-1]


def loop_through(skel: np.ndarray, origin, candidates: int = 32) -> np.ndarray:
    """The longest closed pixel loop through one of the `candidates` skeleton
    pixels nearest `origin` (row, col): around each, the shortest path between
    two of its neighbours that avoids it (racer/track.py:186-202). [L, 2] rows
    and columns, in order around the loop."""
    pts = np.argwhere(skel)
    best: list = []
    for k in np.argsort(((pts - np.asarray(origin, float)) ** 2).sum(1))[:candidates]:
        seed = (int(pts[k][0]), int(pts[k][1]))
        nb = _pixel_neighbours(skel, seed)
        for i in range(len(nb)):
            for j in range(i + 1, len(nb)):
                walk = _bfs(skel, seed, nb[i], nb[j])
                if walk and len(walk) + 1 > len(best):
                    best = [seed] + walk
    if not best:
        raise ValueError("could not find a closed centreline loop in the image")
    return np.array(best)


def occupancy(free: np.ndarray, x0: float, y0: float, cell: float, ramps=()) -> np.ndarray:
    """A one-layer flat track from a free-space image (rows south to north):
    the route is the smoothed centreline loop through the world origin's
    neighbourhood (Savitzky-Golay, 51 points, cubic, wrapping), clearance the
    signed distance to the nearest occupied pixel (racer/track.py:50-58)."""
    free = np.asarray(free, bool)
    loop = loop_through(centreline(free), (-y0 / cell, -x0 / cell))
    xy = savgol_filter(loop[:, ::-1] * cell + (x0, y0), 51, 3, axis=0, mode="wrap")
    route = Route(np.column_stack([xy, np.zeros(len(xy))])).even(SPACING)
    grid = Grid(x0, y0, cell, free.shape[1], free.shape[0])
    nearest = cKDTree(route.p[:, :2]).query(grid.centres())[1].reshape(grid.shape)
    clearance = (distance_transform_edt(free) - distance_transform_edt(~free) - 0.5) * cell
    return assemble(grid, route, clearance[None], np.zeros((1,) + grid.shape), nearest[None].astype(float), 0.0,
                    ramps)


# ---------------------------------------------------------------- named tracks
def _circle(radius, z):
    return lambda t: np.column_stack([radius * np.cos(t), radius * np.sin(t), z(t)])


def _eight(a, b, z):
    return lambda t: np.column_stack([a * np.sin(t), b * np.sin(2 * t), z(t)])


def _hump(height):
    """A deck `height` up where the figure eight crosses itself (t near pi)."""
    return lambda t: height * np.clip(np.cos((t - np.pi) * 1.4), 0, 1) ** 2


# name: (route(t), samples of t, half width, bank, waypoint spacing)
BUILTIN = {
    # the reference's own tracks (racer/track.py:38-44)
    "flat": (_circle(6, lambda t: 0 * t), 160, 1.3, 0.0, SPACING),
    "ramp": (_circle(6, lambda t: 0.7 * (1 - np.cos(2 * t))), 160, 1.3, 0.0, SPACING),
    "bank": (_circle(6, lambda t: 0 * t), 160, 1.3, 0.25, SPACING),
    "bridge": (_eight(7, 3.5, _hump(0.9)), 160, 1.1, 0.0, SPACING),
    # the benchmark workloads: R = 10 m rings, the 8 m-lobe figure eight
    "ring10": (_circle(10, lambda t: 0 * t), 4096, 1.3, 0.0, 0.1),
    "ramp10": (_circle(10, lambda t: 0.3 * np.sin(3 * t)), 4096, 1.3, 0.0, 0.1),
    "bank10": (_circle(10, lambda t: 0 * t), 4096, 1.3, 0.1, 0.1),
    "bridge8": (_eight(16, 8, _hump(0.5)), 4096, 1.3, 0.0, 0.1),
}


def catmull_rom(points: np.ndarray, samples: int = 64) -> np.ndarray:
    """Closed uniform Catmull-Rom spline through `points` (P x 3), `samples` per span."""
    p = np.asarray(points, float)
    t = np.linspace(0, 1, samples, endpoint=False)[:, None]
    spans = []
    for i in range(len(p)):
        p0, p1, p2, p3 = p[i - 1], p[i], p[(i + 1) % len(p)], p[(i + 2) % len(p)]
        spans.append(0.5 * ((2 * p1) + (-p0 + p2) * t + (2 * p0 - 5 * p1 + 4 * p2 - p3) * t**2
                            + (-p0 + 3 * p1 - 3 * p2 + p3) * t**3))
    return np.vstack(spans)


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
    """A named track, or a YAML map file, as the float32 track array
    (racer/track.py:37-58)."""
    c = 0.025 if cell is None else cell
    if name in BUILTIN:
        curve, samples, half_width, bank, spacing = BUILTIN[name]
        return ribbon(curve(np.linspace(0, 2 * np.pi, samples, endpoint=False)), half_width, bank, c,
                      spacing=spacing)
    if name.startswith("random") and (name[6:] == "" or name[6:].isdigit()):
        return ribbon(random_route(int(name[6:] or 0))[0], 1.3, 0, c, spacing=0.1)
    path = Path(name)
    if not path.exists():
        raise ValueError(f"unknown track {name!r}: not a built-in name ({', '.join(BUILTIN)}, random<k>) "
                         "and no such YAML file")
    import yaml

    meta = yaml.safe_load(path.read_text())
    c = float(meta.get("resolution", 0.025))
    ramps = meta.get("ramps", [])
    if "route" in meta:
        points = np.loadtxt(path.parent / meta["route"])[:, :3]
        return ribbon(points, meta.get("half_width", 1.3), meta.get("bank", 0), c, ramps)
    if "image" in meta:
        from PIL import Image

        grey = np.asarray(Image.open(path.parent / meta["image"]).convert("L"))
        free = np.flipud(grey >= FREE_LEVEL)  # image rows run from the north
        return occupancy(free, float(meta["origin"][0]), float(meta["origin"][1]), c, ramps)
    raise ValueError(f"{name}: a track map needs a `route` or an `image`")


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
