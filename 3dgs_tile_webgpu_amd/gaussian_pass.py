"""Host-side mirror of the reference's pass API over the C ABI.

``gaussian_pass(camera, data, options)`` mirrors ``gaussianPass(renderer,
camera, store, options)`` (createGaussianPass.ts:7-14, GaussianPass.ts:144-307)
for the frame path this package accelerates: options are validated with the
reference's rules and error types (RangeError -> GsRangeError(ValueError),
TypeError -> GsTypeError(TypeError)), ``render()`` runs one frame on the GPU,
``read_stats()`` / ``get_output_status()`` / ``get_resources()`` mirror the
reference's introspection, and intersection overflow follows the reference's
capacity-feedback contract (docs/dynamic-intersections.md): the overflowing
frame publishes nothing, the capacity grows by IntersectionCapacityPolicy's
rule and the next frame retries.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass

import numpy as np

from . import _native as N
from .active_slots import ActiveSlots, check_data
from .camera import PerspectiveCamera, elements
from .scene import PackedScene

DEFAULT_OPTIONS = {
    "depthSortMode": "float32",
    "antialiasMode": "compensated",
    "filterVariance": 0.3,
    "boundaryPolicy": "clip",
    "rasterTransmittanceThreshold": 1e-4,
    "depthAlphaThreshold": 0.95,
    "invalidDepthValue": 1.0,
    "background": (0.0, 0.0, 0.0, 0.0),
    "subpixelSampleCulling": True,
    "rasterChunkSize": 8192,
    "outputDepth": False,
    "rasterStats": False,
    "intersectionGrowthFactor": 2.0,
    "intersectionHeadroom": 1.25,
    # Launch shape (no reference counterpart): tile edge, radix digit width,
    # raster LDS batch; only tileSize 16 gives keys comparable with the
    # reference, the other two never change any output.
    "tileSize": 16,
    "radixBits": 8,
    "rasterBatch": 256,
}
# Options of the reference that take no part in this path's arithmetic:
# accepted (and validated where the reference validates them) for drop-in
# use. radixBackend picks a WebGPU sort kernel (the sort here is always the
# library's own, with identical order); colorSpace only tags the output
# texture; profileKernels asks for timestamp queries (gs_profile_begin/end
# is the counterpart).
ACCEPTED_OPTIONS = {"redrawStrategy", "radixBackend", "dataMode", "colorSpace", "profileKernels",
                    "intersectionReadbackInterval"}
KNOWN_OPTIONS = set(DEFAULT_OPTIONS) | ACCEPTED_OPTIONS | {
    "fixedSize", "depthRange", "intersectionCapacity", "initialIntersectionCapacity",
    "maxIntersectionCapacity", "device", "batchLanes", "heavySplitUnits",
    "heavyFineSplitUnits", "heavyPairUnits", "viewGroup",
}
MAX_CAPACITY = 1 << 28


def _f32(v: float) -> float:
    """Math.fround: the nearest binary32 value (inf beyond its range)."""
    with np.errstate(over="ignore"):
        return float(np.float32(v))


class IntersectionCapacityPolicy:
    """IntersectionCapacityPolicy.ts:13-101 for the options this API takes."""

    def __init__(self, options: dict, device_limit: int = MAX_CAPACITY):
        fixed = options.get("intersectionCapacity")
        for name in ("intersectionCapacity", "initialIntersectionCapacity", "maxIntersectionCapacity"):
            v = options.get(name)
            if v is not None and (not isinstance(v, int) or isinstance(v, bool) or v < 1 or v > 2**53 - 1):
                raise N.GsRangeError(N.GS_E_RANGE, f"{name} must be a positive safe integer")
        if fixed is not None and (options.get("initialIntersectionCapacity") is not None
                                  or options.get("maxIntersectionCapacity") is not None):
            raise N.GsRangeError(N.GS_E_RANGE,
                                 "intersectionCapacity cannot be combined with dynamic capacity options")
        self.maximum = fixed if fixed is not None else options.get("maxIntersectionCapacity", device_limit)
        if self.maximum > device_limit:
            raise N.GsRangeError(N.GS_E_RANGE, "Intersection capacity exceeds device buffer or dispatch limits")
        initial = fixed if fixed is not None else options.get("initialIntersectionCapacity")
        self.initial = initial if initial is not None else min(262_144, self.maximum)
        if self.initial > self.maximum:
            raise N.GsRangeError(N.GS_E_RANGE, "initialIntersectionCapacity exceeds maximum capacity")
        self.growth = options.get("intersectionGrowthFactor", 2.0)
        self.headroom = options.get("intersectionHeadroom", 1.25)
        if not math.isfinite(self.growth) or self.growth <= 1:
            raise N.GsRangeError(N.GS_E_RANGE, "intersectionGrowthFactor must exceed 1")
        if not math.isfinite(self.headroom) or self.headroom < 1:
            raise N.GsRangeError(N.GS_E_RANGE, "intersectionHeadroom must be at least 1")
        self.interval = options.get("intersectionReadbackInterval", 100)
        if not math.isfinite(self.interval) or self.interval < 0:
            raise N.GsRangeError(N.GS_E_RANGE, "intersectionReadbackInterval must be nonnegative")

    def next(self, current: int, demand: int) -> int | None:
        """Next capacity after an overflow with `demand` records, or None."""
        if demand > self.maximum:
            return None
        desired = max(math.ceil(demand * self.headroom), math.ceil(current * self.growth))
        return min(self.maximum, math.ceil(desired / 256) * 256)


def validate_options(options: dict) -> dict:
    """GaussianPass.ts:157-260 and the render-domain setters (:386-466) for
    the supported option subset, with the reference's error classes: every
    value check is a RangeError there, and so a GsRangeError (ValueError)
    here. depthSortMode has no run-time check in the reference (its type
    system rejects other values); an unsupported variant or an unknown
    option is a GsTypeError (TypeError) here."""
    unknown = set(options) - KNOWN_OPTIONS
    if unknown:
        raise N.GsTypeError(N.GS_E_TYPE, f"unsupported options: {sorted(unknown)}")
    o = {**DEFAULT_OPTIONS, **options}
    if o["depthSortMode"] not in ("float32", "packed16"):
        raise N.GsTypeError(N.GS_E_TYPE, "depthSortMode must be float32 or packed16")
    if o["antialiasMode"] not in ("compensated", "classic"):  # GaussianPass.ts:161-165
        raise N.GsRangeError(N.GS_E_RANGE, 'antialiasMode must be either "compensated" or "classic"')
    if o.get("redrawStrategy", "always") not in ("always", "auto", "never"):  # :166-174
        raise N.GsRangeError(N.GS_E_RANGE, 'redrawStrategy must be "always", "auto", or "never"')
    if o.get("radixBackend", "auto") not in ("auto", "subgroup", "workgroup"):  # radixBackend.ts:7-15
        raise N.GsRangeError(N.GS_E_RANGE, 'radixBackend must be "auto", "subgroup", or "workgroup"')
    if o.get("dataMode", "managed") not in ("managed", "external"):  # :212-213
        raise N.GsRangeError(N.GS_E_RANGE, 'dataMode must be "managed" or "external"')
    if o["boundaryPolicy"] not in N.GS_BOUNDARY:  # :462-466
        raise N.GsRangeError(N.GS_E_RANGE, 'boundaryPolicy must be "clip", "repeat-x", or "repeat-y"')
    t = o["rasterTransmittanceThreshold"]  # :246-255
    if not (isinstance(t, (int, float)) and math.isfinite(t) and 0 < t < 1):
        raise N.GsRangeError(N.GS_E_RANGE, "rasterTransmittanceThreshold must be finite and in (0, 1)")
    fv = o["filterVariance"]  # :436-441
    if not (isinstance(fv, (int, float)) and math.isfinite(fv) and fv >= 0
            and math.isfinite(_f32(fv))):
        raise N.GsRangeError(N.GS_E_RANGE, "filterVariance must be finite and nonnegative")
    iv = o["invalidDepthValue"]  # :449-455 (Infinity is allowed)
    if not isinstance(iv, (int, float)) or math.isnan(iv) or (math.isfinite(iv)
                                                              and not math.isfinite(_f32(iv))):
        raise N.GsRangeError(N.GS_E_RANGE,
                             "invalidDepthValue must be representable as float32 and cannot be NaN")
    bg = o["background"]
    if len(bg) != 4 or not all(math.isfinite(v) for v in bg):
        raise N.GsRangeError(N.GS_E_RANGE, "background must be four finite numbers")
    cs = o["rasterChunkSize"]  # rasterChunks.ts:80-94
    if cs is not None and (not isinstance(cs, int) or isinstance(cs, bool) or cs < 256
                           or cs % 256 != 0):
        raise N.GsRangeError(N.GS_E_RANGE, "rasterChunkSize must be a multiple of 256 and at least 256")
    da = o["depthAlphaThreshold"]  # :236-242
    if not (isinstance(da, (int, float)) and 0 <= da <= 1):
        raise N.GsRangeError(N.GS_E_RANGE, "depthAlphaThreshold must be finite and in [0, 1]")
    fs = o.get("fixedSize")  # :386-395
    if fs is not None:
        w, h = fs["width"], fs["height"]
        if not all(isinstance(v, int) and not isinstance(v, bool) and 1 <= v <= 2**53 - 1
                   for v in (w, h)):
            raise N.GsRangeError(N.GS_E_RANGE, "fixedSize width/height must be positive safe integers")
    dr = o.get("depthRange")  # :410-423
    if dr is not None:
        n, f = float(dr[0]), float(dr[1])
        if not (math.isfinite(n) and math.isfinite(f) and n > 0 and _f32(n) > 0 and f > n
                and math.isfinite(_f32(f)) and _f32(n) < _f32(f)):
            raise N.GsRangeError(N.GS_E_RANGE,
                                 "depthRange must contain finite positive increasing float32 values")
    if (o["tileSize"] not in (8, 16, 32) or o["radixBits"] not in (4, 6, 8)
            or o["rasterBatch"] not in (128, 256)):
        raise N.GsRangeError(N.GS_E_RANGE,
                             "tileSize must be 8, 16 or 32, radixBits 4, 6 or 8, rasterBatch 128 or 256")
    lanes = o.get("batchLanes", 1)
    if lanes != 1:
        raise N.GsRangeError(N.GS_E_RANGE, "batch lanes above 1 are disabled")
    group = o.get("viewGroup", 1)
    if not isinstance(group, int) or isinstance(group, bool) or not 1 <= group <= 8:
        raise N.GsRangeError(N.GS_E_RANGE, "viewGroup must be an integer in [1, 8]")
    for name in ("heavySplitUnits", "heavyFineSplitUnits", "heavyPairUnits"):
        split = o.get(name)
        if split is not None and (not isinstance(split, int) or isinstance(split, bool)
                                  or not 0 <= split < 0xFFFFFFFF):
            raise N.GsRangeError(N.GS_E_RANGE, f"{name} must be a non-negative integer")
    return o


def config_for(o: dict, policy: IntersectionCapacityPolicy) -> N.GsConfig:
    """The gs_config of validated options."""
    cfg = N.GsConfig()
    N.lib().gs_config_default(C.byref(cfg))
    cfg.device = int(o.get("device", 0))
    cfg.depth_sort_mode = N.GS_DEPTH_FLOAT32 if o["depthSortMode"] == "float32" else N.GS_DEPTH_PACKED16
    cfg.antialias_mode = N.GS_AA_COMPENSATED if o["antialiasMode"] == "compensated" else N.GS_AA_CLASSIC
    cfg.subpixel_sample_culling = 1 if o["subpixelSampleCulling"] else 0
    cfg.raster_chunk_size = 0 if o["rasterChunkSize"] is None else int(o["rasterChunkSize"])
    cfg.intersection_capacity = policy.initial
    cfg.max_intersection_capacity = policy.maximum
    cfg.output_depth = 1 if o["outputDepth"] else 0
    cfg.transmittance_threshold = float(o["rasterTransmittanceThreshold"])
    cfg.depth_alpha_threshold = float(o["depthAlphaThreshold"])
    cfg.invalid_depth_value = float(o["invalidDepthValue"])
    cfg.background = (C.c_float * 4)(*o["background"])
    cfg.raster_stats = 1 if o["rasterStats"] else 0
    cfg.tile_size = int(o["tileSize"])
    cfg.radix_bits = int(o["radixBits"])
    cfg.raster_batch = int(o["rasterBatch"])
    cfg.batch_lanes = int(o.get("batchLanes", 1))
    split = o.get("heavySplitUnits")
    cfg.heavy_split_units = 0xFFFFFFFF if split is None else int(split)
    fine = o.get("heavyFineSplitUnits")
    cfg.heavy_fine_split_units = 0xFFFFFFFF if fine is None else int(fine)
    cfg.view_group = int(o.get("viewGroup", 1))
    pair = o.get("heavyPairUnits")
    cfg.heavy_pair_units = 0xFFFFFFFF if pair is None else int(pair)
    return cfg


def upload_scene(ctx, data: PackedScene) -> list:
    """gs_set_data from a PackedScene (after GaussianData's checks); returns
    the arrays the call read."""
    check_data(data)
    keep = [np.ascontiguousarray(data.means, np.float32),
            np.ascontiguousarray(data.scales_opacity, np.uint32),
            np.ascontiguousarray(data.rotations, np.uint32),
            np.ascontiguousarray(data.sh, np.uint32)]
    active = None if data.active_slots is None else np.ascontiguousarray(data.active_slots, np.uint32)
    d = N.GsDataDesc()
    d.count = data.count
    d.active_count = data.active_count
    d.geometry_format = 0
    d.sh_format = 0
    d.sh_degree = data.sh_degree
    d.memory = N.GS_MEMORY_HOST
    d.means, d.scales_opacity, d.rotations, d.sh = (a.ctypes.data for a in keep)
    d.active_slots = None if active is None else active.ctypes.data
    N.check(N.lib().gs_set_data(ctx, C.byref(d)), ctx)
    return keep


def frame_for(camera: PerspectiveCamera, width: int, height: int, options: dict | None = None,
              attempt: int = 1, reset_valid: bool = False) -> N.GsFrame:
    """The per-frame uniforms (FrameUniforms.update, FrameUniforms.ts:41-58)."""
    o = {**DEFAULT_OPTIONS, **(options or {})}
    f = N.GsFrame()
    f.projection = (C.c_float * 16)(*elements(camera.projection_matrix).astype(np.float32))
    f.width, f.height = width, height
    near, far = o.get("depthRange") or (camera.near, camera.far)
    f.near_plane, f.far_plane = near, far
    f.filter_variance = float(o["filterVariance"])
    f.boundary_axis = N.GS_BOUNDARY[o["boundaryPolicy"]]
    f.attempt = attempt
    f.reset_valid = 1 if reset_valid else 0
    return f


def object_frames_for(camera: PerspectiveCamera, world: np.ndarray | None = None,
                      visible: bool = True) -> np.ndarray:
    """ObjectFrameState.writeCloud for one cloud (ObjectFrameState.ts:97-120)."""
    out = np.zeros(40, np.float32)
    cam = elements(camera.matrix_world)
    w = elements(np.eye(4) if world is None else world)
    N.lib().gs_object_frame(cam.ctypes.data, w.ctypes.data, 1 if visible else 0, out.ctypes.data)
    return out


@dataclass
class OutputStatus:
    state: str
    valid: bool
    version: int
    confirmed_attempt: int
    latest_attempt: int
    requested_intersections: int
    intersection_capacity: int
    max_intersection_capacity: int
    growth_count: int


class GaussianPass:
    """One camera's tiled Gaussian-splatting pass on one GPU."""

    def __init__(self, camera: PerspectiveCamera, data: PackedScene, options: dict | None = None):
        self.options = validate_options(dict(options or {}))
        self.policy = IntersectionCapacityPolicy(self.options)
        cfg = config_for(self.options, self.policy)
        lib = N.lib()
        handle = C.c_void_p()
        N.check(lib.gs_create(C.byref(cfg), C.byref(handle)))
        self._ctx = handle
        self.camera = camera
        self.capacity = self.policy.initial
        self.growth_count = 0
        self.latest_attempt = 0
        self.width = self.height = 0
        self.data = None
        # redrawStrategy (GaussianPass.ts:159-174, 684-693, 776-790, 1187+):
        # "always" encodes every frame; "auto" only when the frame's inputs
        # (size, depth range, projection, camera pose, object table, data)
        # changed or the pass was invalidated; "never" only when invalidated.
        self.redraw_strategy = self.options.get("redrawStrategy", "always")
        self.render_count = 0       # frames encoded (renderCount)
        self.cache_hit_count = 0    # frames that reused the last output (cacheHitCount)
        self._dirty = True
        self._snapshot = None
        self.allocation_error: str | None = None  # set when growth failed (allocation-failed)
        self._tracked: list = []
        self.update_data(data)

    # -- data --------------------------------------------------------------
    def update_data(self, data: PackedScene) -> None:
        """TiledGaussianPipeline.updateData (TiledGaussianPipeline.ts:203-218);
        the buffers are checked as GaussianData's constructor checks them."""
        self._keep = upload_scene(self._ctx, data)
        self.data = data
        self._dirty = True

    def update_active(self, start: int, slots: np.ndarray, active_count: int) -> None:
        s = np.ascontiguousarray(slots, np.uint32)
        N.check(N.lib().gs_update_active(self._ctx, start, s.size, s.ctypes.data if s.size else None,
                                         active_count), self._ctx)
        self._dirty = True

    def track_active(self, active: ActiveSlots) -> None:
        """Upload every committed change of ``active`` as it is committed
        (GaussianData.subscribeActiveSlots and ProjectionStage.updateActiveSlots,
        ProjectionStage.ts:298-329): the changed ordinal range and the new
        active count."""
        if active.capacity != (self.data.count if self.data is not None else active.capacity):
            raise N.GsRangeError(N.GS_E_RANGE, "Active GPU slot exceeds capacity")

        def on_commit(rng: dict) -> None:
            start, n = rng["start"], rng["count"]
            self.update_active(start, active.slots[start:start + n], active.count)

        active.subscribe(on_commit)
        self._tracked.append((active, on_commit))

    def invalidate(self) -> None:
        """Force the next render to encode a frame (GaussianPass.ts:604-606)."""
        self._dirty = True

    # -- frames ------------------------------------------------------------
    def frame_struct(self, width: int, height: int, attempt: int, reset_valid: bool = False) -> N.GsFrame:
        return frame_for(self.camera, width, height, self.options, attempt, reset_valid)

    def object_frames(self) -> np.ndarray:
        return object_frames_for(self.camera)

    def set_objects(self, frames: np.ndarray) -> None:
        f = np.ascontiguousarray(frames, np.float32).reshape(-1, 40)
        N.check(N.lib().gs_set_objects(self._ctx, f.shape[0], f.ctypes.data), self._ctx)

    def render(self, width: int | None = None, height: int | None = None, objects=None,
               retry_overflow: bool = True) -> None:
        """prepareFrame + render; grows capacity and retries on overflow."""
        fs = self.options.get("fixedSize")
        if width is None:
            width, height = (fs["width"], fs["height"]) if fs else (self.width, self.height)
        if not width or not height:
            raise N.GsRangeError(N.GS_E_RANGE, "render size is required")
        if self.allocation_error is not None:
            return  # GaussianPass.ts:683: nothing is encoded after a failed allocation
        table = np.ascontiguousarray(self.object_frames() if objects is None else objects,
                                     np.float32).reshape(-1, 40)
        if (width, height) != (self.width, self.height):
            self._dirty = True  # a resized target holds no valid frame
        snapshot = None
        if self.redraw_strategy == "auto":
            f0 = self.frame_struct(width, height, 0)
            snapshot = (width, height, f0.near_plane, f0.far_plane, bytes(f0.projection),
                        table.tobytes())
            if snapshot != self._snapshot:
                self._dirty = True
        if self.redraw_strategy != "always" and not self._dirty:
            self.cache_hit_count += 1
            return
        self.width, self.height = width, height
        self.set_objects(table)
        self.latest_attempt += 1
        f = self.frame_struct(width, height, self.latest_attempt)
        N.check(N.lib().gs_render(self._ctx, C.byref(f), None), self._ctx)
        self.render_count += 1
        self._dirty = False
        self._snapshot = snapshot
        if retry_overflow:
            st = self.read_state()
            if st.overflow:
                nxt = self.policy.next(self.capacity, st.requested)
                if nxt is not None and nxt > self.capacity:
                    try:
                        self.resize_intersections(nxt)
                    except N.GsError as e:  # GaussianPass.ts:1010-1017 (failAllocation)
                        self.allocation_error = str(e)
                        return
                    self.latest_attempt += 1
                    f = self.frame_struct(width, height, self.latest_attempt)
                    N.check(N.lib().gs_render(self._ctx, C.byref(f), None), self._ctx)

    def render_frame(self, frame: N.GsFrame) -> None:
        N.check(N.lib().gs_render(self._ctx, C.byref(frame), None), self._ctx)

    def resize_intersections(self, capacity: int) -> None:
        N.check(N.lib().gs_resize_intersections(self._ctx, capacity), self._ctx)
        self.capacity = capacity
        self.growth_count += 1

    # -- introspection -----------------------------------------------------
    def read_state(self) -> N.GsState:
        st = N.GsState()
        N.check(N.lib().gs_read_state(self._ctx, C.byref(st)), self._ctx)
        return st

    def read_stats(self) -> dict:
        """GaussianPassStats (types.ts:94-103)."""
        st = self.read_state()
        return {
            "visibleGaussianCount": st.visible,
            "intersectionCount": st.intersections,
            "requestedIntersections": st.requested,
            "intersectionCapacity": st.capacity,
            "overflow": bool(st.overflow),
            "rasterWork": {"checked": st.raster_checked, "blended": st.raster_blended}
            if self.options["rasterStats"] else None,
        }

    def get_output_status(self) -> OutputStatus:
        """getOutputStatus (GaussianPass.ts:989-1008), synchronous form."""
        st = self.read_state()
        if self.allocation_error is not None:
            state = "allocation-failed"
        elif st.attempt == 0:
            state = "awaiting-first-result"
        elif st.overflow:
            state = "budget-exceeded" if self.policy.next(self.capacity, st.requested) is None else "overflow-pending"
        else:
            state = "ready"
        valid = st.version != 0 and not st.overflow and self.allocation_error is None
        return OutputStatus(state, valid, st.version, st.attempt,
                            self.latest_attempt, st.requested, self.capacity, self.policy.maximum,
                            self.growth_count)

    def read_image(self) -> np.ndarray:
        """The rgba16float colour target as (H, W, 4) float16."""
        img = np.empty((self.height, self.width, 4), np.float16)
        N.check(N.lib().gs_read_image(self._ctx, img.ctypes.data, img.nbytes, None, 0), self._ctx)
        return img

    def read_depth(self) -> np.ndarray:
        d = np.empty((self.height, self.width), np.float32)
        N.check(N.lib().gs_read_image(self._ctx, None, 0, d.ctypes.data, d.nbytes), self._ctx)
        return d

    def read_debug(self, name: str) -> np.ndarray:
        size = C.c_size_t()
        lib = N.lib()
        N.check(lib.gs_read_debug(self._ctx, name.encode(), None, 0, C.byref(size)), self._ctx)
        if name in ("projected_mean", "projected_conic", "projected_color"):
            out = np.empty((size.value // 16, 4), np.float32)
        elif name in ("depth_records", "sorted_records"):
            out = np.empty((size.value // 8, 2), np.uint32)
        else:
            out = np.empty((size.value // 4,), np.uint32)
        N.check(lib.gs_read_debug(self._ctx, name.encode(), out.ctypes.data if out.nbytes else None,
                                  out.nbytes, None), self._ctx)
        return out

    def get_resources(self) -> dict:
        """getResources as host copies (TiledGaussianPipeline.ts:391-406)."""
        names = ["projected_mean", "projected_conic", "projected_color", "visible_offsets",
                 "depth_records", "tile_counts", "ordered_counts", "intersection_offsets",
                 "sorted_records", "tile_offsets"]
        return {n: self.read_debug(n) for n in names}

    @property
    def stream(self) -> int:
        return N.lib().gs_stream(self._ctx) or 0

    def dispose(self) -> None:
        for active, listener in getattr(self, "_tracked", []):
            active.unsubscribe(listener)
        self._tracked = []
        if getattr(self, "_ctx", None):
            N.lib().gs_destroy(self._ctx)
            self._ctx = None

    def __del__(self):
        try:
            self.dispose()
        except Exception:
            pass


def gaussian_pass(camera: PerspectiveCamera, data: PackedScene, options: dict | None = None) -> GaussianPass:
    """gaussianPass(renderer, camera, store, options) (createGaussianPass.ts:7-14)."""
    return GaussianPass(camera, data, options)
