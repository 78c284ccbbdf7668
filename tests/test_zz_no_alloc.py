"""gs_render_noalloc (the frame without allocating) against gs_render.

The scene, the object table and the output images are device buffers the
caller owns (torch tensors here); every working array of the frame lives in
the library's fixed device scratch. Two calls in a row, the second with a
smaller input (fewer Gaussians, a smaller image, another camera), must each
give gs_render's state and image bits for the same frame.
"""
from __future__ import annotations

import ctypes as C
import importlib

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def device_buffers(torch, scene, active_count):
    t = {
        "means": torch.from_numpy(np.ascontiguousarray(scene.means, np.float32)).cuda(),
        "shape": torch.from_numpy(np.ascontiguousarray(scene.scales_opacity, np.uint32)
                                  .view(np.int32)).cuda(),
        "rot": torch.from_numpy(np.ascontiguousarray(scene.rotations, np.uint32)
                                .view(np.int32)).cuda(),
        "sh": torch.from_numpy(np.ascontiguousarray(scene.sh, np.uint32).view(np.int32)).cuda(),
    }
    d = _desc(scene.count, active_count, scene.sh_degree, t)
    return t, d


def config(gs, options):
    o = gs.validate_options(dict(options))
    gp = importlib.import_module("3dgs_tile_webgpu_amd.gaussian_pass")
    return o, gp.config_for(o, gs.IntersectionCapacityPolicy(o))


def _desc(count, active_count, degree, t):
    N = importlib.import_module("3dgs_tile_webgpu_amd")._native
    d = N.GsDataDesc()
    d.count, d.active_count, d.sh_degree = count, active_count, degree
    d.geometry_format = d.sh_format = 0
    d.memory = N.GS_MEMORY_DEVICE
    d.means, d.scales_opacity = t["means"].data_ptr(), t["shape"].data_ptr()
    d.rotations, d.sh = t["rot"].data_ptr(), t["sh"].data_ptr()
    d.active_slots = None
    return d


def render_noalloc(gs, torch, desc, cam, w, h, options):
    N = gs._native
    o, cfg = config(gs, options)
    objects = torch.from_numpy(gs.object_frames_for(cam)).cuda()
    frame = gs.frame_for(cam, w, h, o)
    frame.object_frames = objects.data_ptr()
    color = torch.zeros((h, w, 4), dtype=torch.float16, device="cuda")
    image = N.GsImage(color.data_ptr(), None, N.GS_MEMORY_DEVICE)
    state = N.GsState()
    torch.cuda.synchronize()
    rc = N.lib().gs_render_noalloc(C.byref(cfg), C.byref(desc), 1, C.byref(frame),
                                   C.byref(image), None, C.byref(state))
    assert rc == N.GS_OK, rc
    return color.cpu().numpy(), state


def render_context(gs, scene, cam, w, h, options):
    p = gs.gaussian_pass(cam, scene, dict(options))
    p.render(w, h, retry_overflow=False)
    img, st = p.read_image().copy(), p.read_state()
    p.dispose()
    return img, st


def test_noalloc_matches_render_twice(gs, gpu_available):
    import torch

    scene = gs.workload_scene("small", count=60_000)
    w = gs.WORKLOADS["small"]
    cases = [
        (60_000, w.camera(), 640, 360, {"intersectionCapacity": 1 << 19}),
        (25_000, gs.PerspectiveCamera(fov=50.0, aspect=1.6, position=(3.0, 2.0, 12.0)), 320, 200,
         {"intersectionCapacity": 1 << 18, "rasterChunkSize": 256}),
    ]
    t, _ = device_buffers(torch, scene, scene.count)
    for n, cam, width, height, options in cases:
        desc = _desc(scene.count, n, scene.sh_degree, t)
        got, st = render_noalloc(gs, torch, desc, cam, width, height, options)
        sub = gs.PackedScene(scene.means, scene.scales_opacity, scene.rotations, scene.sh, 3,
                             np.arange(n, dtype=np.uint32))
        want, ref = render_context(gs, sub, cam, width, height, options)
        print(f"n={n} {width}x{height}: V={st.visible} K={st.intersections}")
        assert st.error_flags == 0 and st.overflow == 0 and st.intersections > 0
        assert (st.visible, st.requested, st.intersections) == (
            ref.visible, ref.requested, ref.intersections)
        np.testing.assert_array_equal(got.view(np.uint16), want.view(np.uint16))


def test_noalloc_rejects_what_does_not_fit(gs, gpu_available):
    """Sizes beyond the scratch and host-memory inputs fail before any launch."""
    import torch

    N = gs._native
    scene = gs.workload_scene("small", count=1000)
    t, desc = device_buffers(torch, scene, scene.count)
    _, cfg = config(gs, {"intersectionCapacity": 1 << 24})
    cam = gs.WORKLOADS["small"].camera()
    objects = torch.from_numpy(gs.object_frames_for(cam)).cuda()
    frame = gs.frame_for(cam, 640, 360)
    frame.object_frames = objects.data_ptr()
    color = torch.zeros((360, 640, 4), dtype=torch.float16, device="cuda")
    image = N.GsImage(color.data_ptr(), None, N.GS_MEMORY_DEVICE)
    lib = N.lib()
    # 2^24 intersections need 256 MiB of records
    assert lib.gs_render_noalloc(C.byref(cfg), C.byref(desc), 1, C.byref(frame), C.byref(image),
                                 None, None) == N.GS_E_RANGE
    desc.memory = N.GS_MEMORY_HOST
    cfg.intersection_capacity = 1 << 16
    assert lib.gs_render_noalloc(C.byref(cfg), C.byref(desc), 1, C.byref(frame), C.byref(image),
                                 None, None) == N.GS_E_TYPE
