"""GPU parity of the raster work counters and of the depth target.

Counters: the reference counts every (pixel, record) pair its raster loop
evaluates and every pair that blends (TileRasterizer.ts:617, 750-768;
GaussianPassStats.rasterWork, types.ts:131-140), each chunk task of a
chunked tile in full. The GPU evaluates only the pairs inside each record's
support rectangle and reconstructs the reference's count from where each
pixel stops; these counts are also the FLOP denominator of the raster's
roofline, so they must equal the oracle's exactly, in total and in the part
that falls in heavy units (tiles above 1024 records or above the chunk size).

Depth target: finalizeGaussianDepth.ts:18-68 with the default depth node
(GaussianContextNodes.ts:125-131), bit-exact against the oracle's restatement,
on the direct path and through the chunk composite.
"""
from __future__ import annotations

import numpy as np
import pytest

from test_gpu_pipeline import assert_frame_matches, render_both

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def small(gs):
    return gs.workload_scene("small")


def assert_counters_match(p, ref):
    st = p.read_state()
    c = ref.counts
    assert st.error_flags == 0 and st.overflow == 0
    print(f"checked={st.raster_checked} blended={st.raster_blended} "
          f"heavy={st.raster_checked_heavy} chunk_tasks={c.chunk_tasks}")
    assert st.raster_checked == c.checked
    assert st.raster_blended == c.blended
    assert st.raster_checked_heavy == c.checked_heavy
    stats = p.read_stats()["rasterWork"]
    assert stats == {"checked": c.checked, "blended": c.blended}


@pytest.mark.parametrize("size,options", [
    ((1280, 720), {}),
    ((640, 360), {"rasterChunkSize": 256}),
    ((640, 360), {"rasterChunkSize": 768}),
    ((640, 360), {"rasterChunkSize": None}),
    ((640, 360), {"heavySplitUnits": 0}),
    ((640, 360), {"heavySplitUnits": 0, "rasterChunkSize": 256}),
    ((640, 360), {"heavySplitUnits": 0, "heavyPairUnits": 0}),
    ((640, 360), {"heavySplitUnits": 0, "heavyPairUnits": 0, "rasterChunkSize": 256}),
    ((640, 360), {"heavySplitUnits": 0, "heavyFineSplitUnits": 0}),
    ((700, 420), {"tileSize": 8, "rasterChunkSize": 256}),
    ((700, 420), {"tileSize": 32, "rasterChunkSize": 2048}),
], ids=lambda v: str(v))
def test_raster_counters_match_oracle_small(gs, oracle, gpu_available, small, size, options):
    w = gs.WORKLOADS["small"]
    opts = {"rasterStats": True, **options}
    p, ref = render_both(gs, oracle, small, w.camera(), *size, opts)
    assert_frame_matches(p, ref, check_f64=False)
    assert_counters_match(p, ref)
    if options.get("rasterChunkSize") in (256, 768):
        assert ref.counts.chunk_tasks > 0
    p.dispose()


def test_raster_counters_match_oracle_1m_1080p(gs, oracle, gpu_available):
    """The bench frame: its counters are the roofline's denominator."""
    w = gs.WORKLOADS["1m_1080p"]
    scene = gs.workload_scene("1m_1080p")
    p, ref = render_both(gs, oracle, scene, w.camera(), w.width, w.height, {"rasterStats": True})
    assert ref.counts.chunk_tasks > 0
    assert_counters_match(p, ref)
    p.dispose()


@pytest.mark.parametrize("options", [
    {"rasterChunkSize": None},
    {"rasterChunkSize": 256},
    {},
    {"heavySplitUnits": 0, "rasterChunkSize": 256},
    {"depthAlphaThreshold": 0.5, "invalidDepthValue": 7.0, "depthRange": (1.0, 40.0)},
    {"background": (0.2, 0.4, 0.6, 0.5)},
], ids=lambda o: ",".join(f"{k}={v}" for k, v in o.items()) or "default")
def test_depth_target_bit_exact(gs, oracle, gpu_available, small, options):
    w = gs.WORKLOADS["small"]
    opts = {"outputDepth": True, **options}
    p, ref = render_both(gs, oracle, small, w.camera(), w.width, w.height, opts)
    assert_frame_matches(p, ref, check_f64=False)
    if options.get("rasterChunkSize") == 256:
        assert ref.counts.chunk_tasks > 0
    got = p.read_depth()
    want = ref.get("depth_target")
    np.testing.assert_array_equal(got.view(np.uint32), want.view(np.uint32))
    inv = np.float32(opts.get("invalidDepthValue", 1.0))
    valid = got != inv
    print(f"valid depth pixels: {valid.sum()} of {valid.size}")
    assert valid.sum() >= 500 and (~valid).sum() >= 500  # both kinds of pixel occur
    assert ((got[valid] >= 0) & (got[valid] <= 1)).all()
    p.dispose()
