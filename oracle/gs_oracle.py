"""ctypes wrapper of the CPU oracle (oracle/_build/libgs_oracle.so).

TEST INFRASTRUCTURE ONLY: loaded by tests, the smoke check and the
benchmark's CPU-baseline leg as the checker, never by the product.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "_build" / "libgs_oracle.so"

u32 = C.c_uint32
f32 = C.c_float
P = C.c_void_p


class OracleScene(C.Structure):
    _fields_ = [
        ("count", u32), ("active_count", u32), ("sh_degree", u32), ("object_count", u32),
        ("means", P), ("scales_opacity", P), ("rotations", P), ("sh", P), ("active_slots", P),
        ("objects", P), ("projection", f32 * 16),
        ("width", f32), ("height", f32), ("near_plane", f32), ("far_plane", f32),
        ("filter_variance", f32), ("boundary_axis", u32), ("subpixel_culling", u32),
        ("antialias_compensated", u32), ("depth_mode", u32), ("raster_chunk_size", u32),
        ("capacity", u32), ("transmittance_threshold", f32), ("background", f32 * 4),
        ("threads", u32), ("tile_size", u32), ("cull_outputs", u32),
        ("output_depth", u32), ("depth_alpha_threshold", f32), ("invalid_depth_value", f32),
    ]


class OracleCounts(C.Structure):
    _fields_ = [
        ("visible", u32), ("total", u32), ("count", u32), ("overflow", u32),
        ("tiles_x", u32), ("tiles_y", u32), ("sentinels", u32), ("chunk_tasks", u32),
        ("checked", C.c_uint64), ("blended", C.c_uint64), ("stage_seconds", C.c_double * 8),
        ("checked_heavy", C.c_uint64),
    ]


_LIB = None


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB_PATH


def lib():
    global _LIB
    if _LIB is None:
        if not LIB_PATH.exists():
            build()
        h = C.CDLL(str(LIB_PATH))
        h.gs_oracle_run.restype = P
        h.gs_oracle_run.argtypes = [C.POINTER(OracleScene)]
        h.gs_oracle_counts_of.argtypes = [P, C.POINTER(OracleCounts)]
        h.gs_oracle_get.restype = C.c_size_t
        h.gs_oracle_get.argtypes = [P, C.c_char_p, P, C.c_size_t]
        h.gs_oracle_free.argtypes = [P]
        h.gs_oracle_blend_f64.argtypes = [P, P]
        for name in ("gs_oracle_perspective_projection", "gs_oracle_projection_result",
                     "gs_oracle_view_covariance", "gs_oracle_boundary_delta",
                     "gs_oracle_decode_scale_opacity", "gs_oracle_decode_rotation",
                     "gs_oracle_decode_sh_rgb8e8"):
            getattr(h, name).restype = None
        h.gs_oracle_projection_result.argtypes = [P, f32, P, f32, P]
        h.gs_oracle_perspective_projection.argtypes = [P, P, P, P, f32, P]
        h.gs_oracle_projection_is_valid.argtypes = [P]
        h.gs_oracle_view_covariance.argtypes = [P, P, P, P]
        h.gs_oracle_tile_hit.argtypes = [P, P, f32, P, C.c_int, C.c_int, P, u32]
        h.gs_oracle_count_tiles.restype = u32
        h.gs_oracle_count_tiles.argtypes = [P, P, f32, P, C.c_int, C.c_int, C.c_int, C.c_int, P, u32]
        h.gs_oracle_subpixel_has_sample.argtypes = [P, P, f32, P, u32, u32, u32]
        h.gs_oracle_boundary_delta.argtypes = [P, P, P, P, u32, P]
        h.gs_oracle_decode_scale_opacity.argtypes = [u32, u32, P]
        h.gs_oracle_decode_rotation.argtypes = [u32, P]
        h.gs_oracle_decode_sh_rgb8e8.argtypes = [u32, P]
        h.gs_oracle_evaluate_sh.restype = None
        h.gs_oracle_evaluate_sh.argtypes = [P, u32, P, P]
        h.gs_oracle_finalize_depth.restype = f32
        h.gs_oracle_finalize_depth.argtypes = [f32, f32, f32, f32, f32, f32, u32]
        _LIB = h
    return _LIB


def fptr(a):
    return np.ascontiguousarray(a, np.float32).ctypes.data


class OracleResult:
    """One frame of the oracle; arrays are copied out on demand."""

    SHAPES = {
        "projected_mean": (np.float32, 4), "projected_conic": (np.float32, 4),
        "projected_color": (np.float32, 4), "tile_counts": (np.uint32, 0),
        "visible_offsets": (np.uint32, 0), "ordered_counts": (np.uint32, 0),
        "intersection_offsets": (np.uint32, 0), "tile_offsets": (np.uint32, 0),
        "depth_records": (np.uint32, 2), "emitted_records": (np.uint32, 2),
        "sorted_records": (np.uint32, 2), "image_f32": (np.float32, 4),
        "image_f16": (np.uint16, 4), "depth_f32": (np.float32, 0), "final_t": (np.float32, 0),
        "depth_target": (np.float32, 0),
        "cull_features": (np.float32, 16), "max_contribution": (np.uint32, 0),
    }

    def __init__(self, handle, width, height, keep):
        self._h = handle
        self.width, self.height = width, height
        self._keep = keep
        c = OracleCounts()
        lib().gs_oracle_counts_of(handle, C.byref(c))
        self.counts = c

    def get(self, name: str) -> np.ndarray:
        dtype, inner = self.SHAPES[name]
        size = lib().gs_oracle_get(self._h, name.encode(), None, 0)
        item = np.dtype(dtype).itemsize * max(inner, 1)
        out = np.empty((size // item, inner) if inner else (size // item,), dtype)
        if size:
            got = lib().gs_oracle_get(self._h, name.encode(), out.ctypes.data, size)
            assert got == size
        if name.startswith("image") :
            out = out.reshape(self.height, self.width, 4)
        elif name in ("depth_f32", "final_t", "depth_target"):
            out = out.reshape(self.height, self.width)
        return out

    def blend_f64(self) -> np.ndarray:
        out = np.empty((self.height, self.width, 4), np.float64)
        lib().gs_oracle_blend_f64(self._h, out.ctypes.data)
        return out

    def __del__(self):
        if getattr(self, "_h", None):
            lib().gs_oracle_free(self._h)
            self._h = None


def run_frame(scene, frame, objects, *, capacity, subpixel=True, compensated=True,
              depth_mode=0, chunk_size=8192, transmittance=1e-4, background=(0, 0, 0, 0),
              threads=0, tile_size=16, cull_outputs=False, output_depth=False,
              depth_alpha_threshold=0.95, invalid_depth_value=1.0) -> OracleResult:
    """Run the oracle on a PackedScene with the exact f32 uniforms of `frame`
    (a GsFrame) and `objects` ((k, 40) f32 object table)."""
    keep = [np.ascontiguousarray(scene.means, np.float32),
            np.ascontiguousarray(scene.scales_opacity, np.uint32),
            np.ascontiguousarray(scene.rotations, np.uint32),
            np.ascontiguousarray(scene.sh, np.uint32),
            np.ascontiguousarray(objects, np.float32).reshape(-1, 40)]
    active = None
    if scene.active_slots is not None:
        active = np.ascontiguousarray(scene.active_slots, np.uint32)
        keep.append(active)
    s = OracleScene()
    s.count = scene.count
    s.active_count = scene.active_count
    s.sh_degree = scene.sh_degree
    s.object_count = keep[4].shape[0]
    s.means, s.scales_opacity, s.rotations, s.sh = (a.ctypes.data for a in keep[:4])
    s.active_slots = None if active is None else active.ctypes.data
    s.objects = keep[4].ctypes.data
    s.projection = (f32 * 16)(*list(frame.projection))
    s.width, s.height = float(frame.width), float(frame.height)
    s.near_plane, s.far_plane = frame.near_plane, frame.far_plane
    s.filter_variance = frame.filter_variance
    s.boundary_axis = frame.boundary_axis
    s.subpixel_culling = 1 if subpixel else 0
    s.antialias_compensated = 1 if compensated else 0
    s.depth_mode = depth_mode
    s.raster_chunk_size = chunk_size or 0
    s.capacity = capacity
    s.transmittance_threshold = transmittance
    s.background = (f32 * 4)(*background)
    s.threads = threads
    s.tile_size = tile_size
    s.cull_outputs = 1 if cull_outputs else 0
    s.output_depth = 1 if output_depth else 0
    s.depth_alpha_threshold = depth_alpha_threshold
    s.invalid_depth_value = invalid_depth_value
    h = lib().gs_oracle_run(C.byref(s))
    if not h:
        raise ValueError("oracle rejected the scene")
    return OracleResult(h, frame.width, frame.height, keep)


def evaluate_sh(words, degree: int, direction) -> np.ndarray:
    """evaluate_gaussian_sh_rgb8e8 of one Gaussian: clamp(sum + 0.5, 0, 1)."""
    w = np.ascontiguousarray(words, np.uint32)
    d = np.ascontiguousarray(direction, np.float32)
    out = np.empty(3, np.float32)
    lib().gs_oracle_evaluate_sh(w.ctypes.data, degree, d.ctypes.data, out.ctypes.data)
    return out


def finalize_depth(weighted, transmittance, *, alpha_threshold=0.95, invalid=1.0, near=0.1,
                   far=1000.0, encoding=0) -> float:
    """finalizeGaussianDepth for one pixel (encodings: see gs_oracle.h)."""
    return float(lib().gs_oracle_finalize_depth(weighted, transmittance, alpha_threshold, invalid,
                                                near, far, encoding))
