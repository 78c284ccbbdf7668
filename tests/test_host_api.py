"""Host-side API: option validation, capacity policy, camera and object-frame
maths, and the C ABI export list (CPU only; no GPU calls)."""
from __future__ import annotations

import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent


def declared_functions():
    names = []
    for h in (ROOT / "include").glob("*.h"):
        text = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        names += re.findall(r"\b(gs_[a-z0-9_]+)\s*\(", text)
    return sorted(set(names))


def test_library_exports_every_declared_symbol(gs):
    lib = gs.lib()
    names = declared_functions()
    assert len(names) >= 25
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(gs._native.SIGNATURES), "ctypes table out of sync with include/"


def test_capacity_policy_reference_cases(gs):
    # test/IntersectionCapacity.test.ts:26-60
    p = gs.IntersectionCapacityPolicy({"maxIntersectionCapacity": 1000, "initialIntersectionCapacity": 256})
    assert p.next(512, 900) == 1000
    assert p.next(1000, 1001) is None
    assert gs.IntersectionCapacityPolicy({}).initial == 262_144
    with pytest.raises(ValueError, match="combined"):
        gs.IntersectionCapacityPolicy({"initialIntersectionCapacity": 12, "intersectionCapacity": 16})
    with pytest.raises(ValueError, match="limits"):
        gs.IntersectionCapacityPolicy({"maxIntersectionCapacity": (1 << 28) + 1})
    # the feedback test's growth: 256 -> demand 700 -> 1024 (IntersectionCapacity.test.ts:68-103)
    q = gs.IntersectionCapacityPolicy({"initialIntersectionCapacity": 256, "maxIntersectionCapacity": 4096})
    assert q.next(256, 700) == 1024


def test_capacity_rule_c_matches_python(gs):
    lib = gs.lib()
    p = gs.IntersectionCapacityPolicy({"maxIntersectionCapacity": 10_000_000})
    rng = np.random.default_rng(0)
    for _ in range(200):
        cur = int(rng.integers(1, 5_000_000))
        dem = int(rng.integers(1, 12_000_000))
        py = p.next(cur, dem)
        c = lib.gs_capacity_next(cur, dem, 10_000_000, 2.0, 1.25)
        assert (py or 0) == c


@pytest.mark.parametrize("opts,err", [
    ({"depthSortMode": "float16"}, TypeError),
    ({"antialiasMode": "msaa"}, ValueError),  # RangeError, GaussianPass.ts:161-165
    ({"rasterTransmittanceThreshold": 0.0}, ValueError),
    ({"rasterTransmittanceThreshold": 1.0}, ValueError),
    ({"rasterChunkSize": 128}, ValueError),
    ({"rasterChunkSize": 1000}, ValueError),
    ({"filterVariance": -1.0}, ValueError),
    ({"boundaryPolicy": "wrap"}, ValueError),
    ({"background": (0, 0, 0)}, ValueError),
    ({"fixedSize": {"width": 0, "height": 10}}, ValueError),
    ({"depthRange": (5.0, 1.0)}, ValueError),
    ({"unknownOption": 1}, TypeError),
    ({"tileSize": 12}, ValueError),
    ({"radixBits": 5}, ValueError),
    ({"rasterBatch": 512}, ValueError),
    ({"batchLanes": 4}, ValueError),
    ({"viewGroup": 0}, ValueError),
    ({"viewGroup": 9}, ValueError),
    ({"viewGroup": 2.5}, ValueError),
])
def test_option_validation(gs, opts, err):
    with pytest.raises(err):
        gs.validate_options(opts)


def test_defaults_match_reference(gs):
    o = gs.validate_options({})
    assert o["filterVariance"] == 0.3 and o["rasterTransmittanceThreshold"] == 1e-4
    assert o["rasterChunkSize"] == 8192 and o["depthAlphaThreshold"] == 0.95
    assert o["subpixelSampleCulling"] is True and o["background"] == (0, 0, 0, 0)
    cfg = gs._native.GsConfig()
    gs.lib().gs_config_default(C.byref(cfg))
    assert cfg.raster_chunk_size == 8192 and cfg.intersection_capacity == 262_144
    assert abs(cfg.transmittance_threshold - 1e-4) < 1e-9 and cfg.capture_graph == 1


def test_projection_matrix_three_conventions(gs):
    cam = gs.PerspectiveCamera(fov=60, aspect=16 / 9, near=0.1, far=1000, position=(0, 0, 15))
    P = cam.projection_matrix
    f = 1 / np.tan(np.radians(30))
    np.testing.assert_allclose(P[1, 1], f)
    np.testing.assert_allclose(P[0, 0], f / (16 / 9))
    np.testing.assert_allclose(P[3, 2], -1)
    np.testing.assert_allclose(P[2, 2], -(1000.1) / 999.9)
    # a point on the optical axis projects to the image centre
    v = cam.matrix_world_inverse @ np.array([0, 0, 0, 1.0])
    clip = P @ v
    assert abs(clip[0]) < 1e-12 and abs(clip[1]) < 1e-12 and clip[3] == pytest.approx(15)


def test_object_frame_matches_numpy(gs):
    rng = np.random.default_rng(3)
    cam = gs.PerspectiveCamera(position=(3.0, -2.0, 11.0), target=(0.5, 0.2, -0.3))
    world = np.eye(4)
    a = rng.normal(size=(3, 3))
    q, _ = np.linalg.qr(a)
    world[:3, :3] = q * 1.7
    world[:3, 3] = [1.0, 2.0, -3.0]
    out = gs.object_frames_for(cam, world)
    np.testing.assert_allclose(out[:16], gs.elements(world), rtol=1e-6)
    mv = np.linalg.inv(cam.matrix_world) @ world
    np.testing.assert_allclose(out[16:32], gs.elements(mv).astype(np.float32), rtol=1e-5, atol=1e-5)
    cam_local = np.linalg.inv(world) @ np.append(cam.matrix_world[:3, 3], 1.0)
    np.testing.assert_allclose(out[32:35], cam_local[:3], rtol=1e-5, atol=1e-6)
    assert out[35] == 1 and out[36] == 1
    hidden = gs.object_frames_for(cam, world, visible=False)
    assert hidden[36] == 0


def test_gpu_entry_points_fail_cleanly_without_device(gs):
    import torch

    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    cfg = gs._native.GsConfig()
    gs.lib().gs_config_default(C.byref(cfg))
    h = C.c_void_p()
    assert gs.lib().gs_create(C.byref(cfg), C.byref(h)) == gs._native.GS_E_DEVICE
