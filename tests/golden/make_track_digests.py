"""Writes tests/golden/track_digests.json: SHA-256 digests (and per-grid
summaries, to locate a difference) of the reference generator's float32 track
arrays for its built-in ribbons `flat`, `ramp`, `bank` and `bridge`
(racer/track.py:37-44, 61-125).

Runs only in the build container, where the reference checkout is present:
it imports racer/track.py from the reference directory. scikit-image is not
installed here and the ribbon path never calls it, so `skimage.morphology`
is replaced by a stub that raises if the image-map path is taken.

    python tests/golden/make_track_digests.py [reference_root]

The JSON is a fixture (digests and numbers); tests/test_track.py compares
warporacer3d_amd.track's arrays with it.
"""
from __future__ import annotations

import hashlib
import importlib.util
import json
import sys
import types
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
NAMES = ("flat", "ramp", "bank", "bridge")


def summary(arr: np.ndarray) -> dict:
    """Digest of the whole array plus what locates a difference: the header,
    and per layer and grid the sum, the sum of squares and the nonzero count."""
    a = np.ascontiguousarray(arr, np.float32)
    nx, ny, layers, waypoints = int(a[0]), int(a[1]), int(a[7]), int(a[5])
    plane = nx * ny
    grids = {}
    for layer in range(layers):
        for g, name in enumerate(("clearance", "height", "waypoint")):
            v = a[12 + (3 * layer + g) * plane: 12 + (3 * layer + g + 1) * plane].astype(np.float64)
            grids[f"{name}{layer}"] = {"sum": float(v.sum()), "sum_sq": float((v * v).sum()),
                                       "nonzero": int(np.count_nonzero(v)),
                                       "sha256": hashlib.sha256(v.astype(np.float32).tobytes()).hexdigest()}
    route = a[12 + 3 * layers * plane:].astype(np.float64)
    return {"sha256": hashlib.sha256(a.tobytes()).hexdigest(), "size": int(a.size),
            "header": [float(v) for v in a[:12]], "grids": grids,
            "route": {"sum": float(route.sum()), "waypoints": waypoints,
                      "sha256": hashlib.sha256(route.astype(np.float32).tobytes()).hexdigest()}}


def reference_module(root: Path):
    skimage = types.ModuleType("skimage")
    morphology = types.ModuleType("skimage.morphology")

    def skeletonize(*_a, **_k):
        raise RuntimeError("scikit-image is not installed: the image-map path cannot run here")

    morphology.skeletonize = skeletonize
    skimage.morphology = morphology
    sys.modules.setdefault("skimage", skimage)
    sys.modules.setdefault("skimage.morphology", morphology)
    spec = importlib.util.spec_from_file_location("reference_track", root / "racer" / "track.py")
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def main() -> int:
    root = Path(sys.argv[1]) if len(sys.argv) > 1 else Path("/root/reference")
    mod = reference_module(root)
    out = {"generator": "racer/track.py load(name) of the reference checkout", "tracks": {}}
    for name in NAMES:
        arr = np.asarray(mod.load(name), np.float32)
        out["tracks"][name] = summary(arr)
        print(name, arr.size, out["tracks"][name]["sha256"], flush=True)
    (HERE / "track_digests.json").write_text(json.dumps(out, indent=1) + "\n")
    return 0


if __name__ == "__main__":
    sys.exit(main())
