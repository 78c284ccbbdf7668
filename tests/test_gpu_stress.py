"""The 6M-Gaussian, 3840x2160 stress frame: strongly anisotropic scales and
near-plane clusters, about 79 M tile-Gaussian keys.

A 200k-Gaussian draw of the same distributions at 960x540 is checked bit for
bit against the oracle; the full frame is checked through properties that do
not depend on size: the key count lies in the 60-100 M band, the sorted
records are ordered by (tile, depth key, ordinal) with no sentinels, the
tile ranges are exactly the boundaries of that order, each Gaussian's keys
equal its tile count, and the image is finite and within [0, 1].
"""
from __future__ import annotations

import numpy as np
import pytest

from test_gpu_pipeline import assert_frame_matches, render_both

pytestmark = pytest.mark.gpu


def test_stress_distribution_reduced_bit_exact(gs, oracle, gpu_available):
    w = gs.WORKLOADS["6m_4k"]
    scene = gs.workload_scene("6m_4k", count=200_000)
    cam = gs.PerspectiveCamera(fov=60.0, aspect=960 / 540, position=(0.0, 0.0, w.camera_distance))
    p, ref = render_both(gs, oracle, scene, cam, 960, 540, {"rasterChunkSize": 1024},
                         capacity=1 << 23)
    assert ref.counts.count > 500_000 and ref.counts.chunk_tasks > 0
    assert_frame_matches(p, ref)
    p.dispose()


def test_stress_full_frame_properties(gs, gpu_available):
    w = gs.WORKLOADS["6m_4k"]
    scene = gs.workload_scene("6m_4k")
    p = gs.gaussian_pass(w.camera(), scene, {"initialIntersectionCapacity": 1 << 26})
    p.render(w.width, w.height)
    st = p.read_state()
    assert not st.overflow and st.error_flags == 0
    k = st.intersections
    print(f"6M 4K: V={st.visible} K={k}")
    assert 60_000_000 <= k <= 100_000_000
    rec = p.read_debug("sorted_records")
    assert rec.shape == (k, 2)
    tiles = st.tiles_x * st.tiles_y
    tile, gid = rec[:, 0].astype(np.int64), rec[:, 1].astype(np.int64)
    assert tile.max() < tiles  # no sentinel records
    mean = p.read_debug("projected_mean")
    dkey = mean[:, 2].view(np.uint32).astype(np.int64)
    assert np.all(mean[gid, 3] > 0)
    # Lexicographic (tile, depth key, ordinal), strictly increasing.
    composite = (tile << 44) | (dkey[gid] << 12)  # tile < 2^19, depth key < 2^31
    step = np.diff(composite)
    assert np.all(step >= 0)
    same = step == 0
    assert np.all(np.diff(gid)[same] > 0)
    del composite, step, same
    off = p.read_debug("tile_offsets").astype(np.int64)
    np.testing.assert_array_equal(off, np.searchsorted(tile, np.arange(tiles + 1), side="left"))
    counts = np.bincount(gid, minlength=mean.shape[0])
    np.testing.assert_array_equal(counts, p.read_debug("tile_counts"))
    img = p.read_image().astype(np.float32)
    assert np.isfinite(img).all() and img.min() >= 0.0 and img.max() <= 1.0
    p.dispose()
