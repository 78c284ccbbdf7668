"""Write the committed golden fixtures of tests/golden/.

- workload_hashes.json: SHA-256 of every packed buffer the seeded generator
  produces for each synthetic workload (seeds and sizes are part of the
  fixture), so a change to the generator or encoder is caught.
- small_frame.json: counts and SHA-256 of the oracle's intermediates and
  image for the 100k / 1280x720 workload with default options; the GPU path
  must reproduce them bit for bit.

Run from the repository root: python tests/golden/make_golden.py
"""
from __future__ import annotations

import hashlib
import importlib
import json
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "oracle"))

FRAME_ARRAYS = ["projected_mean", "tile_counts", "depth_records", "sorted_records",
                "tile_offsets", "image_f16"]


def sha(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def workload_hashes(gs) -> dict:
    out = {}
    for name, w in gs.WORKLOADS.items():
        s = gs.workload_scene(name)
        out[name] = {"count": w.count, "seed": w.seed, "means": sha(s.means),
                     "scales_opacity": sha(s.scales_opacity), "rotations": sha(s.rotations),
                     "sh": sha(s.sh)}
        del s
    return out


def small_frame(gs, oracle) -> dict:
    w = gs.WORKLOADS["small"]
    scene = gs.workload_scene("small")
    cam = w.camera()
    r = oracle.run_frame(scene, gs.frame_for(cam, w.width, w.height), gs.object_frames_for(cam),
                         capacity=1 << 22)
    c = r.counts
    return {"workload": "small", "width": w.width, "height": w.height, "capacity": 1 << 22,
            "visible": int(c.visible), "requested": int(c.total), "keys": int(c.count),
            "chunk_tasks": int(c.chunk_tasks),
            "sha256": {n: sha(r.get(n)) for n in FRAME_ARRAYS}}


def main():
    gs = importlib.import_module("3dgs_tile_webgpu_amd")
    import gs_oracle

    (HERE / "workload_hashes.json").write_text(json.dumps(workload_hashes(gs), indent=1) + "\n")
    (HERE / "small_frame.json").write_text(json.dumps(small_frame(gs, gs_oracle), indent=1) + "\n")


if __name__ == "__main__":
    main()
