"""The track compiler (warporacer3d_amd/track.py) against the reference
generator's own output and against closed forms.

* The reference's built-in ribbons (racer/track.py:37-44): byte for byte the
  arrays the reference's load() writes (SHA-256 digests recorded by
  tests/golden/make_track_digests.py from the reference checkout).
* YAML route maps (racer/track.py:53-56): the same array as the ribbon of the
  listed route; ramps (racer/track.py:112-119) add their box tops.
* Occupancy-image maps (racer/track.py:57-58, 150-202): no reference output
  can be produced here (scikit-image, which the reference thins with, is not
  installed), so the byte parity of this path is unpinned; it is pinned by
  closed forms instead: the centreline of an annulus is its middle circle, the
  clearance there is half the road's width, a dead-end spur is pruned away.
"""
import hashlib
import json
from pathlib import Path

import numpy as np
import pytest

from warporacer3d_amd import track as T

DIGESTS = json.loads((Path(__file__).resolve().parent / "golden" / "track_digests.json").read_text())["tracks"]


@pytest.mark.parametrize("name", ["flat", "ramp", "bank", "bridge"])
def test_builtin_ribbons_match_reference_generator(name):
    arr = T.load(name)
    assert arr.dtype == np.float32
    ref = DIGESTS[name]
    assert arr.size == ref["size"] and [float(v) for v in arr[:12]] == ref["header"]
    assert hashlib.sha256(arr.tobytes()).hexdigest() == ref["sha256"]


def _write_route_map(tmp_path, points, **meta):
    np.savetxt(tmp_path / "route.txt", points)
    lines = ["route: route.txt", "resolution: 0.05"] + [f"{k}: {v}" for k, v in meta.items()]
    (tmp_path / "map.yaml").write_text("\n".join(lines) + "\n")
    return str(tmp_path / "map.yaml")


def test_yaml_route_map_is_the_ribbon_of_its_route(tmp_path):
    t = np.linspace(0, 2 * np.pi, 90, endpoint=False)
    pts = np.column_stack([5 * np.cos(t), 3 * np.sin(t), 0.2 * np.sin(t)])
    path = _write_route_map(tmp_path, pts, half_width=1.0, bank=0.05)
    arr = T.load(path)
    assert np.array_equal(arr, T.ribbon(np.loadtxt(tmp_path / "route.txt")[:, :3], 1.0, 0.05, 0.05))
    h = T.header(arr)
    assert h["layers"] == 1 and abs(h["cell"] - 0.05) < 1e-7 and abs(h["spacing"] - 0.15) < 0.01


def test_ramp_adds_a_box_with_a_sloped_top(tmp_path):
    t = np.linspace(0, 2 * np.pi, 90, endpoint=False)
    pts = np.column_stack([6 * np.cos(t), 6 * np.sin(t), 0 * t])
    plain = T.load(_write_route_map(tmp_path, pts))
    ramp = "ramps: [{center: [6.0, 0.0], heading: 1.5707963, length: 2.0, width: 1.0, height: 0.2, rise: 0.5}]"
    (tmp_path / "map.yaml").write_text((tmp_path / "map.yaml").read_text() + ramp + "\n")
    ramped = T.load(str(tmp_path / "map.yaml"))
    h = T.header(plain)
    assert abs(T.header(ramped)["slope"] - (h["slope"] + 0.2 / 0.5)) < 1e-6
    plane = h["nx"] * h["ny"]

    def height_at(arr, x, y):
        ix = int(round((x - h["x0"]) / h["cell"]))
        iy = int(round((y - h["y0"]) / h["cell"]))
        return float(arr[12 + plane + iy * h["nx"] + ix])

    # on the box's level top (after the height fill's blur, well inside it) and far from it
    assert abs(height_at(ramped, 6.0, 0.0) - 0.2) < 0.02 and abs(height_at(plain, 6.0, 0.0)) < 1e-6
    assert abs(height_at(ramped, -6.0, 0.0) - height_at(plain, -6.0, 0.0)) < 1e-6


def _annulus(r_in, r_out, cell, size, spur=False):
    """Free space between two circles about the image centre (world origin),
    optionally with a dead-end corridor leaving the ring outward."""
    c = (np.arange(size) - size / 2 + 0.5) * cell
    x, y = np.meshgrid(c, c)
    r = np.hypot(x, y)
    free = (r >= r_in) & (r <= r_out)
    if spur:
        free |= (np.abs(y) < 0.3) & (x > r_out - 0.2) & (x < r_out + 1.2)
    return free, float(c[0]), float(c[0])


def _image_map(tmp_path, free, x0, y0, cell):
    from PIL import Image

    grey = np.where(np.flipud(free), 254, 0).astype(np.uint8)  # image rows run from the north
    Image.fromarray(grey).save(tmp_path / "map.png")
    (tmp_path / "map.yaml").write_text(f"image: map.png\nresolution: {cell}\norigin: [{x0}, {y0}, 0.0]\n")
    return str(tmp_path / "map.yaml")


@pytest.mark.parametrize("spur", [False, True])
def test_occupancy_image_centreline_of_an_annulus(tmp_path, spur):
    cell, r_in, r_out = 0.05, 3.0, 4.4
    free, x0, y0 = _annulus(r_in, r_out, cell, 220, spur)
    arr = T.load(_image_map(tmp_path, free, x0, y0, cell))
    h = T.header(arr)
    assert h["layers"] == 1 and h["nx"] == 220 and h["ny"] == 220 and h["slope"] == 0
    plane = h["nx"] * h["ny"]
    route = arr[12 + 3 * plane:].reshape(-1, 4).astype(np.float64)
    radius = np.hypot(route[:, 0], route[:, 1])
    mid = (r_in + r_out) / 2
    assert np.all(np.abs(radius - mid) < 2 * cell), (radius.min(), radius.max())  # the spur is pruned
    assert abs(h["waypoints"] * h["spacing"] - 2 * np.pi * mid) < 0.2 and abs(h["spacing"] - 0.15) < 0.005
    # headings follow the loop (tangent to the circle), either way round
    tangent = np.arctan2(route[:, 1], route[:, 0]) + np.pi / 2
    d = np.angle(np.exp(1j * (route[:, 3] - tangent)))
    assert np.all(np.abs(np.abs(d) - np.pi / 2) > np.pi / 2 - 0.15)
    clearance = arr[12: 12 + plane].reshape(h["ny"], h["nx"])
    ix = int(round((mid - x0) / cell))
    iy = int(round((0.0 - y0) / cell))
    assert abs(clearance[iy, ix] - (r_out - r_in) / 2) < 2 * cell
    assert clearance[iy, int(round((0.0 - x0) / cell))] < 0  # the hole is a wall
    waypoint = arr[12 + 2 * plane: 12 + 3 * plane]
    assert waypoint.min() >= 0 and waypoint.max() <= h["waypoints"] - 1


def test_thinning_leaves_a_thin_closed_centreline():
    """An annulus thins to a closed curve along its middle circle: no dead ends,
    about one pixel per pixel of circumference, every pixel within 1.5 pixels
    of the circle, and a loop through it that uses most of its pixels."""
    cell, r_in, r_out = 0.05, 2.0, 3.0
    free, x0, _ = _annulus(r_in, r_out, cell, 140)
    skel = T.centreline(free)
    counts = sum(q.astype(int) for q in T._ring(skel.astype(np.uint8)))[skel]
    assert counts.min() >= 2, "no dead ends"
    circumference = 2 * np.pi * (r_in + r_out) / 2 / cell
    assert 0.7 * circumference < skel.sum() < 1.4 * circumference
    rows, cols = np.nonzero(skel)
    radius = np.hypot(x0 + cols * cell, x0 + rows * cell)
    assert np.all(np.abs(radius - (r_in + r_out) / 2) <= 1.5 * cell)
    loop = T.loop_through(skel, (-x0 / cell, -x0 / cell))
    assert len(loop) > 0.7 * circumference and len({tuple(p) for p in loop.tolist()}) == len(loop)


def test_unknown_names_and_bad_maps_raise(tmp_path):
    with pytest.raises(ValueError, match="unknown track"):
        T.load("no_such_track")
    (tmp_path / "bad.yaml").write_text("resolution: 0.05\n")
    with pytest.raises(ValueError, match="needs a `route` or an `image`"):
        T.load(str(tmp_path / "bad.yaml"))
    free = np.zeros((50, 50), bool)
    with pytest.raises(ValueError, match="no closed centreline loop"):
        T.occupancy(free, -1.25, -1.25, 0.05)
