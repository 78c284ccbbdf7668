"""Committed golden fixtures (tests/golden/, written by make_golden.py).

The seeded generator must reproduce every workload's packed buffers, and
the oracle and the GPU path must reproduce the small workload's
intermediates and image, bit for bit.
"""
from __future__ import annotations

import hashlib
import json
from pathlib import Path

import numpy as np
import pytest

GOLDEN = Path(__file__).resolve().parent / "golden"


def sha(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def test_workload_buffers_match_fixture(gs):
    want = json.loads((GOLDEN / "workload_hashes.json").read_text())
    assert set(want) == set(gs.WORKLOADS)
    for name, h in want.items():
        w = gs.WORKLOADS[name]
        assert (w.count, w.seed) == (h["count"], h["seed"]), name
        s = gs.workload_scene(name)
        got = {"means": sha(s.means), "scales_opacity": sha(s.scales_opacity),
               "rotations": sha(s.rotations), "sh": sha(s.sh)}
        assert got == {k: h[k] for k in got}, name
        del s


def small_setup(gs):
    g = json.loads((GOLDEN / "small_frame.json").read_text())
    w = gs.WORKLOADS[g["workload"]]
    return g, w, gs.workload_scene(g["workload"]), w.camera()


def test_oracle_small_frame_matches_fixture(gs, oracle):
    g, w, scene, cam = small_setup(gs)
    r = oracle.run_frame(scene, gs.frame_for(cam, g["width"], g["height"]),
                         gs.object_frames_for(cam), capacity=g["capacity"])
    assert (r.counts.visible, r.counts.total, r.counts.count) == (
        g["visible"], g["requested"], g["keys"])
    for name, h in g["sha256"].items():
        assert sha(r.get(name)) == h, name


@pytest.mark.gpu
def test_gpu_small_frame_matches_fixture(gs, gpu_available):
    g, w, scene, cam = small_setup(gs)
    p = gs.gaussian_pass(cam, scene, {"intersectionCapacity": g["capacity"]})
    p.render(g["width"], g["height"])
    st = p.read_state()
    assert (st.visible, st.requested, st.intersections) == (g["visible"], g["requested"], g["keys"])
    res = p.get_resources()
    for name, h in g["sha256"].items():
        arr = p.read_image() if name == "image_f16" else res[name]
        assert sha(arr) == h, name
    p.dispose()
