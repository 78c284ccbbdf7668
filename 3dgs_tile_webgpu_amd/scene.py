"""Synthetic workloads in the packed-buffer layout the store publishes.

Each workload is generated by the native seeded generator
(gs_generate_scene: xoshiro256** streams, Box-Muller normals, Shoemake
rotations) and encoded by the native producer (compact geometry + rgb8e8
SH), so CPU oracle runs and GPU runs see bit-identical inputs.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _native as N
from .camera import PerspectiveCamera


@dataclass
class PackedScene:
    """GaussianData buffers (GaussianData.ts:10-33), compact + rgb8e8."""

    means: np.ndarray            # (n, 4) f32: xyz, object id
    scales_opacity: np.ndarray   # (n, 2) u32
    rotations: np.ndarray        # (n,) u32
    sh: np.ndarray               # (n, (deg+1)^2) u32
    sh_degree: int
    active_slots: np.ndarray | None = None
    raw: dict = field(default_factory=dict)  # pre-encoding floats, if requested

    @property
    def count(self) -> int:
        return int(self.means.shape[0])

    @property
    def active_count(self) -> int:
        return self.count if self.active_slots is None else int(self.active_slots.shape[0])


@dataclass
class Workload:
    name: str
    count: int
    width: int
    height: int
    camera_distance: float
    extent: float
    seed: int
    anisotropic: bool = False
    near_fraction: float = 0.0
    description: str = ""
    orbit: tuple[float, float] | None = None  # camera orbit radius range (multi-view)
    views: int = 1

    def camera(self) -> PerspectiveCamera:
        return PerspectiveCamera(fov=60.0, aspect=self.width / self.height, near=0.1, far=1000.0,
                                 position=(0.0, 0.0, self.camera_distance))


# The plain-method workloads. Means: 32 isotropic clusters (sigma 0.5) with
# centres uniform in a cube of half-size `extent`; camera on +z looking at
# the origin with a 60 degree vertical field of view.
WORKLOADS = {
    "small": Workload("small", 100_000, 1280, 720, 15.0, 5.0, seed=20241,
                      description="100k Gaussians, 10-unit cube, 1280x720, camera at 15"),
    "1m_1080p": Workload("1m_1080p", 1_000_000, 1920, 1080, 60.0, 20.0, seed=20242,
                         description="1M Gaussians, 40-unit cube, 1920x1080, camera at 60"),
    # Near-plane clusters (8 clusters at view depth 0.5-3, sigma 0.1) hold 17 % of the
    # Gaussians: that puts the frame at about 79 M tile-Gaussian keys.
    "6m_4k": Workload("6m_4k", 6_000_000, 3840, 2160, 60.0, 20.0, seed=20243, anisotropic=True,
                      near_fraction=0.17,
                      description="6M anisotropic Gaussians with near-plane clusters, 3840x2160"),
    # Multi-view workloads: the 10-unit cube of the small scene seen from
    # seeded orbits of radius 12-25 around the origin.
    "2m_batch": Workload("2m_batch", 2_000_000, 1920, 1080, 18.0, 5.0, seed=20244,
                         orbit=(12.0, 25.0), views=64,
                         description="2M Gaussians, 64 orbit cameras of mixed resolution"),
    "3m_cull": Workload("3m_cull", 3_000_000, 1920, 1080, 18.0, 5.0, seed=20245,
                        orbit=(12.0, 25.0), views=64,
                        description="3M Gaussians, 1920x1080 orbit views (512 train, 64 held out)"),
}

# Resolutions of the batched multi-camera workload.
BATCH_RESOLUTIONS = [(854, 480), (1280, 720), (1920, 1080), (2560, 1440)]


def spec_for(w: Workload, count: int | None = None) -> N.GsSceneSpec:
    lib = N.lib()
    spec = N.GsSceneSpec()
    lib.gs_scene_spec_default(C.byref(spec))
    spec.seed = w.seed
    spec.count = count if count is not None else w.count
    spec.extent = w.extent
    spec.anisotropic = 1 if w.anisotropic else 0
    spec.near_fraction = w.near_fraction
    spec.camera_position = (C.c_double * 3)(0.0, 0.0, w.camera_distance)
    spec.camera_target = (C.c_double * 3)(0.0, 0.0, 0.0)
    spec.camera_fov_deg = 60.0
    spec.camera_aspect = w.width / w.height
    return spec


def generate(spec: N.GsSceneSpec, keep_raw: bool = False) -> PackedScene:
    n = int(spec.count)
    deg = int(spec.sh_degree)
    ncoef = (deg + 1) ** 2
    means = np.empty((n, 4), np.float32)
    shape = np.empty((n, 2), np.uint32)
    rot = np.empty((n,), np.uint32)
    sh = np.empty((n, ncoef), np.uint32)
    raw = {}
    ptrs = [None, None, None]
    if keep_raw:
        raw = {"scales": np.empty((n, 3), np.float32), "opacity": np.empty((n,), np.float32),
               "quats": np.empty((n, 4), np.float32)}
        ptrs = [raw["scales"].ctypes.data, raw["opacity"].ctypes.data, raw["quats"].ctypes.data]
    rc = N.lib().gs_generate_scene(C.byref(spec), means.ctypes.data, shape.ctypes.data,
                                   rot.ctypes.data, sh.ctypes.data, *ptrs)
    if rc != 0:
        raise ValueError("invalid scene spec")
    return PackedScene(means, shape, rot, sh, deg, None, raw)


def read_ply(data: bytes) -> dict:
    """A canonical 3DGS PLY (CanonicalGaussianPlyLoader.ts:84-350) as float
    arrays: means (n, 3), scales (n, 3), opacity (n,), quats (n, 4) xyzw,
    sh (n, (deg+1)^2, 3), sh_degree."""
    lib = N.lib()
    buf = np.frombuffer(data, np.uint8)
    info = N.GsPlyInfo()
    if lib.gs_ply_parse(buf.ctypes.data, buf.size, C.byref(info), None, None, None, None, None):
        raise ValueError(lib.gs_ply_last_error().decode())
    n, deg = info.count, info.sh_degree
    out = {"means": np.empty((n, 3), np.float32), "scales": np.empty((n, 3), np.float32),
           "opacity": np.empty(n, np.float32), "quats": np.empty((n, 4), np.float32),
           "sh": np.empty((n, (deg + 1) ** 2, 3), np.float32)}
    if lib.gs_ply_parse(buf.ctypes.data, buf.size, C.byref(info), *(out[k].ctypes.data for k in
                        ("means", "scales", "opacity", "quats", "sh"))):
        raise ValueError(lib.gs_ply_last_error().decode())
    out["sh_degree"] = deg
    return out


def pack_scene(g: dict, capacity: int | None = None, object_id: int = 0) -> PackedScene:
    """The store's slot buffers of one cloud (engine.rs:1801-1949): the
    Gaussians in slots [0, n), holes (means.w = -1) up to `capacity`."""
    n = g["means"].shape[0]
    capacity = n if capacity is None else capacity
    deg = int(g["sh_degree"])
    ncoef = (deg + 1) ** 2
    means = np.empty((capacity, 4), np.float32)
    shape = np.empty((capacity, 2), np.uint32)
    rot = np.empty(capacity, np.uint32)
    sh = np.empty((capacity, ncoef), np.uint32)
    src = [np.ascontiguousarray(g[k], np.float32) for k in ("means", "scales", "opacity", "quats", "sh")]
    if N.lib().gs_pack_scene(n, capacity, object_id, deg, *(a.ctypes.data for a in src),
                             means.ctypes.data, shape.ctypes.data, rot.ctypes.data, sh.ctypes.data):
        raise ValueError("pack_scene: n exceeds capacity or SH degree above 3")
    return PackedScene(means, shape, rot, sh, deg)


def workload_scene(name: str, count: int | None = None, keep_raw: bool = False) -> PackedScene:
    return generate(spec_for(WORKLOADS[name], count), keep_raw)
