"""Many cameras over one scene in one captured launch.

The reference renders one camera per pass and lets several passes share a
store (``dataMode: "external"``, docs/render-extensions.md:5-47). Here the
cameras of a batch -- of uneven resolutions -- are rendered by one
``gs_render_batch`` call: one hipGraph holds every camera's frame and is
replayed while resolutions, capacities and buffers stay the same (camera
matrices and object frames are uploaded per launch and never force a
re-capture). Each view keeps its own intersection
capacity, grown by the reference's IntersectionCapacityPolicy on overflow.

``autotune`` runs the library's autotuner over tile size {8, 16, 32}, radix
digit width {4, 6, 8} and raster LDS batch {128, 256} and keeps the fastest.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _native as N
from .camera import PerspectiveCamera
from .gaussian_pass import (IntersectionCapacityPolicy, config_for, frame_for,
                            object_frames_for, upload_scene, validate_options)
from .scene import PackedScene


def check_object_tables(tables: list[np.ndarray]) -> None:
    """Every camera of a batch brings its own object table (40 floats per
    object); the library reads as many rows from each as the first has, so
    they must all have that many."""
    rows = {t.shape[0] for t in tables}
    if len(rows) > 1:
        raise N.GsRangeError(N.GS_E_RANGE, "every camera's object table must have the same rows")


class BatchedPass:
    """A batch of cameras rendered together on one GPU."""

    def __init__(self, data: PackedScene, options: dict | None = None):
        self.options = validate_options(dict(options or {}))
        self.policy = IntersectionCapacityPolicy(self.options)
        handle = C.c_void_p()
        N.check(N.lib().gs_create(C.byref(config_for(self.options, self.policy)), C.byref(handle)))
        self._ctx = handle
        self._keep_scene = upload_scene(self._ctx, data)
        self.data = data
        self.attempt = 0
        self.sizes: list[tuple[int, int]] = []
        self.capacities: list[int] = []
        self._frames = None
        self._objects: list[np.ndarray] = []

    # -- frames ------------------------------------------------------------
    def build_frames(self, cameras: list[PerspectiveCamera], sizes: list[tuple[int, int]],
                     objects: list[np.ndarray] | None = None):
        """The gs_frame array for the batch (one object table per camera)."""
        if len(cameras) != len(sizes) or not cameras:
            raise N.GsRangeError(N.GS_E_RANGE, "one size per camera is required")
        tables = [np.ascontiguousarray(object_frames_for(cam) if objects is None else objects[i],
                                       np.float32).reshape(-1, 40)
                  for i, cam in enumerate(cameras)]
        check_object_tables(tables)
        self.attempt += 1
        arr = (N.GsFrame * len(cameras))()
        self._objects = tables
        for i, (cam, (w, h)) in enumerate(zip(cameras, sizes)):
            f = frame_for(cam, w, h, self.options, self.attempt)
            f.object_frames = tables[i].ctypes.data
            arr[i] = f
        if len(self._objects):
            N.check(N.lib().gs_set_objects(self._ctx, self._objects[0].shape[0],
                                           self._objects[0].ctypes.data), self._ctx)
        return arr

    def launch(self, frames) -> None:
        """One gs_render_batch of prepared frames (asynchronous)."""
        N.check(N.lib().gs_render_batch(self._ctx, len(frames), frames, None), self._ctx)

    def render(self, cameras: list[PerspectiveCamera], sizes: list[tuple[int, int]],
               retry_overflow: bool = True, objects: list[np.ndarray] | None = None) -> None:
        """Render every camera; overflowing views grow and the batch re-renders."""
        frames = self.build_frames(cameras, sizes, objects)
        self._frames = frames
        self.sizes = list(sizes)
        n = len(frames)
        if len(self.capacities) < n:
            self.capacities += [self.policy.initial] * (n - len(self.capacities))
        for _ in range(8):
            self.launch(frames)
            if not retry_overflow:
                return
            grown = False
            for v, st in enumerate(self.read_states()):
                if st.overflow:
                    nxt = self.policy.next(self.capacities[v], st.requested)
                    if nxt is None:
                        raise N.GsRangeError(N.GS_E_RANGE, "intersection demand exceeds maximum")
                    self.resize_view(v, nxt)
                    grown = True
            if not grown:
                return
        raise N.GsError(N.GS_E_STATE, "batch capacity did not settle")

    def resize_view(self, view: int, capacity: int) -> None:
        N.check(N.lib().gs_resize_batch_intersections(self._ctx, view, capacity), self._ctx)
        self.capacities[view] = capacity

    # -- results -----------------------------------------------------------
    def read_states(self) -> list[N.GsState]:
        n = len(self.sizes)
        out = (N.GsState * n)()
        N.check(N.lib().gs_read_batch_state(self._ctx, 0, n, out), self._ctx)
        return list(out)

    def read_image(self, view: int) -> np.ndarray:
        w, h = self.sizes[view]
        img = np.empty((h, w, 4), np.float16)
        N.check(N.lib().gs_read_batch_image(self._ctx, view, img.ctypes.data, img.nbytes, None, 0),
                self._ctx)
        return img

    def read_depth(self, view: int) -> np.ndarray:
        w, h = self.sizes[view]
        d = np.empty((h, w), np.float32)
        N.check(N.lib().gs_read_batch_image(self._ctx, view, None, 0, d.ctypes.data, d.nbytes),
                self._ctx)
        return d

    # -- launch shape ------------------------------------------------------
    def tuning(self) -> dict:
        t = N.GsTuning()
        N.check(N.lib().gs_get_tuning(self._ctx, C.byref(t)), self._ctx)
        return {"tileSize": t.tile_size, "radixBits": t.radix_bits, "rasterBatch": t.raster_batch}

    def set_tuning(self, tile_size: int, radix_bits: int, raster_batch: int) -> None:
        t = N.GsTuning(tile_size, radix_bits, raster_batch)
        N.check(N.lib().gs_set_tuning(self._ctx, C.byref(t)), self._ctx)
        self.options.update(tileSize=tile_size, radixBits=radix_bits, rasterBatch=raster_batch)

    def autotune(self, cameras: list[PerspectiveCamera], sizes: list[tuple[int, int]],
                 reps: int = 3) -> tuple[dict, list[dict]]:
        """Time every launch shape on this batch; keep and return the fastest."""
        frames = self.build_frames(cameras, sizes)
        self._frames = frames
        self.sizes = list(sizes)
        best = N.GsTuning()
        table = (N.GsTuneEntry * 18)()
        count = C.c_uint32()
        N.check(N.lib().gs_autotune(self._ctx, len(frames), frames, reps, C.byref(best), table, 18,
                                    C.byref(count)), self._ctx)
        rows = [{"tileSize": e.tuning.tile_size, "radixBits": e.tuning.radix_bits,
                 "rasterBatch": e.tuning.raster_batch, "ms": e.ms} for e in table[:count.value]]
        chosen = {"tileSize": best.tile_size, "radixBits": best.radix_bits,
                  "rasterBatch": best.raster_batch}
        self.options.update(chosen)
        n = len(frames)
        st = (N.GsState * n)()
        N.check(N.lib().gs_read_batch_state(self._ctx, 0, n, st), self._ctx)
        self.capacities = [s.capacity for s in st]
        return chosen, rows

    @property
    def stream(self) -> int:
        return N.lib().gs_stream(self._ctx) or 0

    def dispose(self) -> None:
        if getattr(self, "_ctx", None):
            N.lib().gs_destroy(self._ctx)
            self._ctx = None

    def __del__(self):
        try:
            self.dispose()
        except Exception:
            pass
