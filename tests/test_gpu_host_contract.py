"""The host mirror's frame contract on the GPU path (Python mirror; the
Node facade's twins are in test_node_facade.py).

- redrawStrategy (GaussianPass.ts:159-174, 684-693, 776-790, 1187+; the
  behaviour pinned by test/GaussianPassNodes.test.ts:455-485, 606-611):
  "always" encodes every frame, "auto" reuses the output while the frame's
  inputs are unchanged, "never" only re-encodes when invalidated.
- Capacity feedback (test/IntersectionCapacity.test.ts:67-101, 131-181,
  233-255): a first overflowing frame grows by the policy and the retry
  publishes; a fixed capacity reports budget-exceeded and never grows; a
  failed growth reports allocation-failed and stops encoding; independent
  passes grow from their own demand.
- Viewport resizes (test/ResizeResources.test.ts:17-42, 74-92, 129+): buffers
  are grow-only, a shrunk frame reads back only its own tile ranges plus the
  sentinel, and frames after shrinking and regrowing stay bit-exact.
"""
from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def scene(gs):
    return gs.workload_scene("small", count=20_000)


def test_redraw_always_is_the_default(gs, gpu_available, scene):
    p = gs.gaussian_pass(gs.WORKLOADS["small"].camera(), scene)
    p.render(320, 180)
    p.render(320, 180)
    assert p.redraw_strategy == "always" and p.render_count == 2 and p.cache_hit_count == 0
    # rasterWork is only reported with rasterStats (test/RasterWorkStats.test.ts:27-33)
    assert p.read_stats()["rasterWork"] is None
    p.dispose()


def test_redraw_auto_reuses_clean_frames(gs, gpu_available, scene):
    cam = gs.WORKLOADS["small"].camera()
    p = gs.gaussian_pass(cam, scene, {"redrawStrategy": "auto"})
    p.render(320, 180)
    first = p.read_image().copy()
    p.render(320, 180)
    assert (p.render_count, p.cache_hit_count) == (1, 1)
    np.testing.assert_array_equal(p.read_image().view(np.uint16), first.view(np.uint16))
    p.camera = gs.PerspectiveCamera(fov=60.0, aspect=16 / 9, position=(1.0, 0.5, 14.0))
    p.render(320, 180)      # moved camera: encoded
    p.render(320, 180)      # clean again: cached
    p.render(256, 144)      # new size: encoded
    p.invalidate()
    p.render(256, 144)      # invalidated: encoded
    assert (p.render_count, p.cache_hit_count) == (4, 2)
    p.dispose()


def test_redraw_never_encodes_only_when_invalidated(gs, gpu_available, scene):
    cam = gs.WORKLOADS["small"].camera()
    p = gs.gaussian_pass(cam, scene, {"redrawStrategy": "never"})
    p.render(320, 180)
    p.camera = gs.PerspectiveCamera(fov=60.0, aspect=16 / 9, position=(1.0, 0.5, 14.0))
    p.render(320, 180)      # a moved camera alone does not redraw
    assert (p.render_count, p.cache_hit_count) == (1, 1)
    p.invalidate()
    p.render(320, 180)
    assert (p.render_count, p.cache_hit_count) == (2, 1)
    p.dispose()


@pytest.mark.parametrize("strategy", ["auto", "never"])
def test_first_overflowing_frame_grows_and_retries(gs, gpu_available, scene, strategy):
    cam = gs.WORKLOADS["small"].camera()
    p = gs.gaussian_pass(cam, scene, {"redrawStrategy": strategy,
                                      "initialIntersectionCapacity": 256,
                                      "maxIntersectionCapacity": 1 << 20})
    p.render(320, 180)
    st = p.read_state()
    status = p.get_output_status()
    assert p.growth_count == 1 and status.state == "ready" and status.valid
    assert p.capacity == gs.IntersectionCapacityPolicy({"maxIntersectionCapacity": 1 << 20}).next(
        256, st.requested)
    q = gs.gaussian_pass(cam, scene, {"intersectionCapacity": 1 << 20})
    q.render(320, 180)
    np.testing.assert_array_equal(p.read_image().view(np.uint16), q.read_image().view(np.uint16))
    q.dispose()
    p.dispose()


def test_fixed_capacity_reports_budget_exceeded(gs, gpu_available, scene):
    p = gs.gaussian_pass(gs.WORKLOADS["small"].camera(), scene, {"intersectionCapacity": 256})
    p.render(320, 180)
    status = p.get_output_status()
    assert (status.state, status.valid, p.growth_count) == ("budget-exceeded", False, 0)
    p.render(320, 180)
    assert p.growth_count == 0 and p.capacity == 256
    p.dispose()


def test_failed_growth_reports_allocation_failed(gs, gpu_available, scene, monkeypatch):
    p = gs.gaussian_pass(gs.WORKLOADS["small"].camera(), scene,
                         {"initialIntersectionCapacity": 256})

    def fail(capacity):
        raise gs.GsError(-3, "allocation-failed: records-a")

    monkeypatch.setattr(p, "resize_intersections", fail)
    p.render(320, 180)
    encoded = p.latest_attempt
    p.render(320, 180)
    status = p.get_output_status()
    assert (status.state, status.valid) == ("allocation-failed", False)
    assert p.latest_attempt == encoded == 1  # nothing encoded after the failure
    p.dispose()


def test_independent_passes_grow_from_their_own_demand(gs, gpu_available, scene):
    opts = {"initialIntersectionCapacity": 256}
    a = gs.gaussian_pass(gs.WORKLOADS["small"].camera(), scene, opts)
    b = gs.gaussian_pass(gs.PerspectiveCamera(fov=60.0, aspect=1.0, position=(0, 0, 40.0)),
                         gs.workload_scene("small", count=2_000), opts)
    a.render(320, 180)
    b.render(64, 64)
    policy = gs.IntersectionCapacityPolicy(opts)
    for p in (a, b):
        st = p.read_state()
        assert st.overflow == 0
        assert p.capacity == policy.next(256, st.requested) or (p.growth_count == 0
                                                                 and st.requested <= 256)
    assert a.capacity != b.capacity
    a.dispose()
    b.dispose()


def test_resizes_reuse_buffers_and_read_back_active_tiles(gs, oracle, gpu_available, scene):
    """test/ResizeResources.test.ts:17-42, 74-92, 129+: a pass renders through
    widths 272..329, a 1x1 frame and back to 272 without an error; after each
    shrink the tile ranges it reports cover only the current grid plus the
    sentinel, and the regrown frame is bit-exact against the oracle (buffers
    are grow-only, so nothing stale from the larger frames reaches it)."""
    from test_gpu_pipeline import oracle_for

    cam = gs.WORKLOADS["small"].camera()
    p = gs.gaussian_pass(cam, scene, {"rasterChunkSize": 512, "outputDepth": True})
    objects = gs.object_frames_for(cam)[None]

    def check(w, h):
        p.render(w, h)
        st = p.read_state()
        assert st.overflow == 0 and st.error_flags == 0
        tiles = -(-w // 16) * -(-h // 16)
        assert st.tiles_x * st.tiles_y == tiles
        offsets = p.read_debug("tile_offsets").view(np.uint32)
        assert offsets.size == tiles + 1 and offsets[-1] == st.intersections
        ref = oracle_for(gs, oracle, p, scene, cam, w, h, objects)
        np.testing.assert_array_equal(p.read_image().view(np.uint16), ref.get("image_f16"))
        np.testing.assert_array_equal(p.read_depth().view(np.uint32),
                                      ref.get("depth_target").view(np.uint32))

    check(272, 256)
    for width in range(273, 330, 7):
        p.render(width, 256)
    check(329, 256)
    check(1, 1)
    check(272, 256)
    check(1024, 512)  # growth past the first tile grid
    check(16, 16)
    p.dispose()
