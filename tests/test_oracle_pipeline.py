"""Whole-frame invariants of the CPU oracle (CPU only): closed forms for the
scans and offsets, lexicographic sort order, overflow semantics, the chunk
composite identity and the FP64 blend bound."""
from __future__ import annotations

import numpy as np
import pytest


@pytest.fixture(scope="module")
def frame(gs, oracle):
    w = gs.WORKLOADS["small"]
    scene = gs.workload_scene("small", count=30_000)
    cam = w.camera()
    f = gs.frame_for(cam, 640, 360)
    obj = gs.object_frames_for(cam)
    return scene, f, obj, oracle.run_frame(scene, f, obj, capacity=1 << 22, threads=4)


def test_scans_and_gathers_closed_form(frame):
    _, _, _, r = frame
    mean = r.get("projected_mean")
    vis = (mean[:, 3] > 0).astype(np.uint64)
    np.testing.assert_array_equal(r.get("visible_offsets"), (np.cumsum(vis) - vis).astype(np.uint32))
    dr = r.get("depth_records")
    assert dr.shape[0] == r.counts.visible == vis.sum()
    counts = r.get("tile_counts")
    ordered = r.get("ordered_counts")
    np.testing.assert_array_equal(ordered, counts[dr[:, 1]])
    off = r.get("intersection_offsets").astype(np.uint64)
    np.testing.assert_array_equal(off, np.cumsum(ordered.astype(np.uint64)) - ordered)
    assert r.counts.total == ordered.sum() == r.counts.count


def test_depth_records_stably_sorted(frame):
    _, _, _, r = frame
    dr = r.get("depth_records")
    mean = r.get("projected_mean")
    vis = np.nonzero(mean[:, 3] > 0)[0]
    keys = mean[vis, 2].view(np.uint32)            # bitcast<u32>(depth), depth > 0
    order = np.lexsort((vis, keys))
    np.testing.assert_array_equal(dr[:, 0], keys[order])
    np.testing.assert_array_equal(dr[:, 1], vis[order])


def test_sorted_records_lexicographic(frame):
    _, _, _, r = frame
    em = r.get("emitted_records")
    so = r.get("sorted_records")
    assert r.counts.sentinels == 0
    # emission is in depth-rank order; the stable tile sort keeps it per tile
    rank = np.empty(r.counts.visible, np.int64)
    rank_of = {int(g): i for i, g in enumerate(r.get("depth_records")[:, 1])}
    rank = np.array([rank_of[int(g)] for g in em[:, 1]])
    order = np.lexsort((np.arange(em.shape[0]), rank, em[:, 0]))
    np.testing.assert_array_equal(so, em[order])
    assert np.all(np.diff(so[:, 0].astype(np.int64)) >= 0)


def test_tile_offsets_closed_form(frame):
    _, _, _, r = frame
    so = r.get("sorted_records")
    off = r.get("tile_offsets")
    tiles = r.counts.tiles_x * r.counts.tiles_y
    expect = np.searchsorted(so[:, 0], np.arange(tiles + 1), side="left")
    expect[tiles] = so.shape[0]
    np.testing.assert_array_equal(off, expect.astype(np.uint32))
    np.testing.assert_array_equal(np.diff(off.astype(np.int64)), np.bincount(so[:, 0], minlength=tiles)[:tiles])


def test_overflow_publishes_nothing(gs, oracle, frame):
    scene, f, obj, r = frame
    small = oracle.run_frame(scene, f, obj, capacity=100, threads=4)
    c = small.counts
    assert c.overflow == 1 and c.count == 0 and c.total == r.counts.total
    assert small.get("sorted_records").shape[0] == 0
    np.testing.assert_array_equal(small.get("tile_offsets"), 0)


def test_chunk_composite_matches_whole_stream(gs, oracle, frame):
    """test/rasterizationWGSL.test.ts:63-101: chunk partials composited in
    order equal the uninterrupted stream (up to binary32 rounding)."""
    scene, f, obj, r = frame
    whole = oracle.run_frame(scene, f, obj, capacity=1 << 22, chunk_size=0, threads=4)
    chunked = oracle.run_frame(scene, f, obj, capacity=1 << 22, chunk_size=256, threads=4)
    assert chunked.counts.chunk_tasks > 0
    a, b = whole.get("image_f32"), chunked.get("image_f32")
    assert np.abs(a - b).max() < 2e-5


def test_composite_identity_reference_samples():
    samples = [((0.8, 0.1, 0.2), 0.25), ((0.2, 0.7, 0.3), 0.4), ((0.1, 0.2, 0.9), 0.15),
               ((0.9, 0.8, 0.1), 0.35)]

    def composite(ss):
        c, t = np.zeros(3), 1.0
        for col, a in ss:
            c += np.array(col) * t * a
            t *= 1 - a
        return c, t

    whole = composite(samples)
    first, second = composite(samples[:2]), composite(samples[2:])
    merged = first[0] + first[1] * second[0]
    np.testing.assert_allclose(merged, whole[0], atol=1e-12)
    assert abs(first[1] * second[1] - whole[1]) < 1e-12


def test_f32_image_within_one_level_of_fp64_blend(frame):
    _, _, _, r = frame
    exact = r.blend_f64()
    img = r.get("image_f32").astype(np.float64)
    diff = np.abs(img - exact)
    assert diff.max() <= 1.0 / 255.0
    mse = float(np.mean(diff[..., :3] ** 2))
    assert mse == 0 or 10 * np.log10(1 / mse) >= 45
    # rgba16float store: the f16 image is the f32 one rounded to nearest even
    np.testing.assert_array_equal(r.get("image_f16"), r.get("image_f32").astype(np.float16).view(np.uint16))


def test_packed16_keys(gs, oracle, frame):
    scene, f, obj, _ = frame
    r = oracle.run_frame(scene, f, obj, capacity=1 << 22, depth_mode=1, threads=4)
    dr = r.get("depth_records")
    mean = r.get("projected_mean")
    d = mean[dr[:, 1], 2].astype(np.float32)
    near, far = np.float32(f.near_plane), np.float32(f.far_plane)
    q = np.clip((d - near) / (far - near), 0, 1).astype(np.float32) * np.float32(65535)
    np.testing.assert_array_equal(dr[:, 0], np.rint(q).astype(np.uint32))
    assert np.all(np.diff(dr[:, 0].astype(np.int64)) >= 0)
