"""GPU parity of the launch shapes and the batched multi-camera launch.

Launch shapes: every tile edge (8, 16, 32) against the oracle run at that
tile edge, with every radix digit width (4, 6, 8) and raster batch depth
(128, 256) -- the last two must not change a single bit. Batches: cameras of
uneven resolution rendered by one gs_render_batch must reproduce, bit for
bit, what one gs_render per camera produces (itself checked against the
oracle), across camera motion that replays the captured graph, and after
the autotuner has picked a launch shape.
"""
from __future__ import annotations

import itertools

import numpy as np
import pytest

from test_gpu_pipeline import assert_frame_matches, render_both

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def small(gs):
    return gs.workload_scene("small", count=60_000)


SHAPES = list(itertools.product((8, 16, 32), (4, 6, 8), (128, 256)))


@pytest.mark.parametrize("tile,bits,batch", SHAPES,
                         ids=[f"tile{t}-bits{b}-batch{r}" for t, b, r in SHAPES])
def test_launch_shapes_bit_exact(gs, oracle, gpu_available, small, tile, bits, batch):
    w = gs.WORKLOADS["small"]
    # Chunk sizes that give every tile edge chunk tasks (and, at 32, whole
    # heavy tiles too): at 700x420 the fullest tile holds 607 / 1554 / 3595
    # records for edges 8 / 16 / 32.
    chunk = {8: 256, 16: 1024, 32: 2048}[tile]
    options = {"tileSize": tile, "radixBits": bits, "rasterBatch": batch, "rasterChunkSize": chunk}
    p, ref = render_both(gs, oracle, small, w.camera(), 700, 420, options)
    assert ref.counts.chunk_tasks > 0
    assert_frame_matches(p, ref, check_f64=(tile == 16 and bits == 8))
    p.dispose()


def batch_cameras(gs, n, seed=11):
    sizes = [(854, 480), (1280, 720), (640, 360), (320, 200), (960, 544)]
    views = gs.orbit_cameras(n, seed=seed, radius=(12.0, 25.0), resolutions=sizes)
    return [c for c, _ in views], [s for _, s in views]


def single_images(gs, scene, cams, sizes, options):
    out = []
    for cam, (w, h) in zip(cams, sizes):
        p = gs.gaussian_pass(cam, scene, dict(options))
        p.render(w, h)
        out.append(p.read_image().copy())
        p.dispose()
    return out


@pytest.mark.parametrize("lanes", [
    1,
    pytest.param(4, marks=pytest.mark.skip(
        reason="multi-lane batch capture is disabled: the four-lane run ended in a host "
               "segmentation fault inside gs_render_batch whose cause is not established")),
])
def test_batch_matches_single_passes(gs, oracle, gpu_available, small, lanes):
    cams, sizes = batch_cameras(gs, 6)
    options = {"batchLanes": lanes, "initialIntersectionCapacity": 4096}
    b = gs.BatchedPass(small, options)
    b.render(cams, sizes)
    states = b.read_states()
    assert all(s.overflow == 0 for s in states)
    want = single_images(gs, small, cams, sizes, {})
    for v in range(len(cams)):
        np.testing.assert_array_equal(b.read_image(v).view(np.uint16), want[v].view(np.uint16),
                                      err_msg=f"view {v}")
    # One view against the oracle directly.
    w, h = sizes[0]
    p, ref = render_both(gs, oracle, small, cams[0], w, h)
    np.testing.assert_array_equal(b.read_image(0).view(np.uint16), ref.get("image_f16"))
    assert states[0].intersections == ref.counts.count
    p.dispose()
    b.dispose()


GROUPS = [
    (4, {}),
    (3, {"tileSize": 8, "radixBits": 4, "rasterChunkSize": 256}),
    (8, {"rasterChunkSize": 1024, "heavySplitUnits": 0}),
    (5, {"heavySplitUnits": 0, "heavyPairUnits": 0}),
]


@pytest.mark.parametrize("group,options", GROUPS,
                         ids=[f"group{g}-" + ",".join(f"{k}={v}" for k, v in o.items())
                              for g, o in GROUPS])
def test_view_groups_match_single_passes(gs, oracle, gpu_available, small, group, options):
    """Views rendered in groups (each stage one launch over a group's views)
    reproduce one gs_render per camera bit for bit: six views of five
    resolutions, growing from a small capacity, with groups of 4 (+2), 3,
    all six and 5 (+1); at tile 8 with 4-bit digits the views' tile sorts
    have 3 and 4 passes (a group runs the larger count); chunk tasks, whole
    tiles with two lanes per pixel and 8-row slabs."""
    cams, sizes = batch_cameras(gs, 6, seed=21)
    b = gs.BatchedPass(small, {"viewGroup": group, "initialIntersectionCapacity": 4096,
                               **options})
    b.render(cams, sizes)
    states = b.read_states()
    assert all(s.overflow == 0 and s.error_flags == 0 for s in states)
    want = single_images(gs, small, cams, sizes, options)
    for v in range(len(cams)):
        np.testing.assert_array_equal(b.read_image(v).view(np.uint16), want[v].view(np.uint16),
                                      err_msg=f"view {v}")
    w, h = sizes[1]
    p, ref = render_both(gs, oracle, small, cams[1], w, h, options)
    np.testing.assert_array_equal(b.read_image(1).view(np.uint16), ref.get("image_f16"))
    assert states[1].intersections == ref.counts.count
    p.dispose()
    b.dispose()


@pytest.mark.parametrize("group", [1, 4])
def test_batch_replay_follows_camera_motion(gs, gpu_available, small, group):
    """The captured batch graph is replayed with new camera uniforms."""
    cams, sizes = batch_cameras(gs, 5, seed=3)
    b = gs.BatchedPass(small, {"initialIntersectionCapacity": 1 << 21, "viewGroup": group})
    b.render(cams, sizes)
    first = [b.read_image(v).copy() for v in range(5)]
    moved, _ = batch_cameras(gs, 5, seed=4)
    for _ in range(3):  # replays of the same graph
        b.render(moved, sizes)
    want = single_images(gs, small, moved, sizes, {})
    for v in range(5):
        got = b.read_image(v)
        np.testing.assert_array_equal(got.view(np.uint16), want[v].view(np.uint16))
        assert not np.array_equal(got.view(np.uint16), first[v].view(np.uint16))
    b.dispose()


def test_autotune_picks_a_shape_and_stays_exact(gs, oracle, gpu_available, small):
    cams, sizes = batch_cameras(gs, 4, seed=9)
    b = gs.BatchedPass(small, {"initialIntersectionCapacity": 1 << 20})
    chosen, rows = b.autotune(cams, sizes, reps=2)
    assert len(rows) == 18 and all(r["ms"] > 0 for r in rows)
    assert min(rows, key=lambda r: r["ms"])["tileSize"] == chosen["tileSize"]
    print("autotune:", chosen, [(r["tileSize"], r["radixBits"], r["rasterBatch"], round(r["ms"], 3))
                                for r in rows])
    b.render(cams, sizes)
    for v in (0, 3):
        w, h = sizes[v]
        p, ref = render_both(gs, oracle, small, cams[v], w, h, {"tileSize": chosen["tileSize"]})
        np.testing.assert_array_equal(b.read_image(v).view(np.uint16), ref.get("image_f16"))
        p.dispose()
    b.dispose()


def test_c4_batch_at_configured_size(gs, oracle, gpu_available):
    """The batched workload at its configured size: 2M Gaussians, 64 seeded
    orbit cameras (radius 12-25) at resolutions drawn from 854x480,
    1280x720, 1920x1080 and 2560x1440, in one gs_render_batch, one view per
    launch and in view groups of 4. Every view ends without overflow or
    wait errors, publishes the same key count as a single-camera pass of it
    and the same image bits; one view of each resolution is checked bit for
    bit against the oracle."""
    w = gs.WORKLOADS["2m_batch"]
    scene = gs.workload_scene("2m_batch")
    views = gs.orbit_cameras(w.views, seed=777, radius=w.orbit, resolutions=gs.BATCH_RESOLUTIONS)
    cams, sizes = [c for c, _ in views], [s for _, s in views]
    assert len(cams) == 64 and set(sizes) == set(gs.BATCH_RESOLUTIONS)
    b = gs.BatchedPass(scene, {"initialIntersectionCapacity": 1 << 22})
    b.render(cams, sizes)
    states = b.read_states()
    assert all(s.overflow == 0 and s.error_flags == 0 for s in states)
    grouped = gs.BatchedPass(scene, {"initialIntersectionCapacity": 1 << 22, "viewGroup": 4})
    grouped.render(cams, sizes)
    gstates = grouped.read_states()
    assert all(s.overflow == 0 and s.error_flags == 0 for s in gstates)
    p = gs.gaussian_pass(cams[0], scene, {"intersectionCapacity": 1 << 23})
    checked = {}
    for v, (cam, (x, y)) in enumerate(views):
        p.camera = cam
        p.render(x, y)
        st = p.read_state()
        assert st.overflow == 0 and st.error_flags == 0
        assert (states[v].visible, states[v].intersections) == (st.visible, st.intersections), v
        np.testing.assert_array_equal(b.read_image(v).view(np.uint16),
                                      p.read_image().view(np.uint16), err_msg=f"view {v}")
        assert (gstates[v].visible, gstates[v].intersections) == (st.visible, st.intersections)
        np.testing.assert_array_equal(grouped.read_image(v).view(np.uint16),
                                      p.read_image().view(np.uint16), err_msg=f"grouped view {v}")
        if (x, y) not in checked:
            checked[(x, y)] = v
            frame = gs.frame_for(cam, x, y, p.options)
            ref = oracle.run_frame(scene, frame, p.object_frames(), capacity=p.capacity)
            assert ref.counts.count == st.intersections
            np.testing.assert_array_equal(p.read_debug("sorted_records"), ref.get("sorted_records"))
            np.testing.assert_array_equal(b.read_image(v).view(np.uint16), ref.get("image_f16"),
                                          err_msg=f"view {v} ({x}x{y}) against the oracle")
    print("keys per view:", [s.intersections for s in states][:8], "...",
          "total", sum(s.intersections for s in states), "oracle views", checked)
    assert len(checked) == 4
    p.dispose()
    b.dispose()
    grouped.dispose()
