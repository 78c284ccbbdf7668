"""GPU parity: every stage of the frame on the MI355X against the CPU oracle.

Integer outputs (tile counts, visible offsets, depth keys and order, ordered
counts, intersection offsets, sorted (tile, id) records, tile ranges) must be
bit-exact, and so must every projected float and the rgba16float image: both
sides evaluate the same binary32 operations in the reference's order.
"""
from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

STAGES_INT = ["tile_counts", "visible_offsets", "depth_records", "ordered_counts",
              "intersection_offsets", "sorted_records", "tile_offsets"]


def render_both(gs, oracle, scene, camera, width, height, options=None, capacity=1 << 22):
    options = dict(options or {})
    options.setdefault("intersectionCapacity", capacity)
    p = gs.gaussian_pass(camera, scene, options)
    p.render(width, height, retry_overflow=False)
    frame = gs.frame_for(camera, width, height, options)
    objects = p.object_frames()
    o = p.options
    ref = oracle.run_frame(scene, frame, objects, capacity=p.capacity,
                           subpixel=o["subpixelSampleCulling"],
                           compensated=o["antialiasMode"] == "compensated",
                           depth_mode=0 if o["depthSortMode"] == "float32" else 1,
                           chunk_size=o["rasterChunkSize"] or 0,
                           transmittance=o["rasterTransmittanceThreshold"],
                           background=o["background"])
    return p, ref


def assert_frame_matches(p, ref):
    st = p.read_state()
    c = ref.counts
    assert (st.visible, st.requested, st.intersections, st.overflow) == (
        c.visible, c.total, c.count, c.overflow)
    assert c.sentinels == 0
    res = p.get_resources()
    mean_gpu, mean_ref = res["projected_mean"], ref.get("projected_mean")
    np.testing.assert_array_equal(mean_gpu.view(np.uint32), mean_ref.view(np.uint32))
    vis = mean_ref[:, 3] > 0
    for name in ("projected_conic", "projected_color"):
        np.testing.assert_array_equal(res[name][vis].view(np.uint32),
                                      ref.get(name)[vis].view(np.uint32), err_msg=name)
    for name in STAGES_INT:
        np.testing.assert_array_equal(res[name], ref.get(name), err_msg=name)
    img = p.read_image()
    np.testing.assert_array_equal(img.view(np.uint16), ref.get("image_f16"))


def test_small_workload_bit_exact(gs, oracle, gpu_available):
    w = gs.WORKLOADS["small"]
    scene = gs.workload_scene("small")
    p, ref = render_both(gs, oracle, scene, w.camera(), w.width, w.height)
    assert ref.counts.count > 100_000
    assert_frame_matches(p, ref)
    p.dispose()
