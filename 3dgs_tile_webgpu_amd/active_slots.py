"""The store's active slot list and the GaussianData checks, on the host.

``ActiveSlots`` restates GaussianData's active list (src/renderer/
GaussianData.ts:170-232): a compact list of active store slots with each
slot's ordinal in it, changed through staged batches of removals and
additions. A removal moves the list's last slot into the removed slot's
place (the order is not kept); an addition appends. Staged changes go to a
back list; a commit swaps the lists, bumps the version and tells every
subscriber the ordinal range that changed, which is what
ProjectionStage.updateActiveSlots (ProjectionStage.ts:298-329) uploads;
here ``GaussianPass.track_active`` subscribes and calls gs_update_active
with that range.

``check_data`` applies GaussianData's constructor checks
(GaussianData.ts:76-166, 265-318) to a PackedScene, with the reference's
error classes and messages.
"""
from __future__ import annotations

from typing import Callable

import numpy as np

from ._native import GS_E_RANGE, GS_E_TYPE, GsRangeError, GsTypeError
from .scene import PackedScene


class ActiveSlots:
    """A compact active list over ``capacity`` store slots (GaussianData.ts:
    86-97, 170-232)."""

    def __init__(self, capacity: int, slots=None):
        if not (isinstance(capacity, (int, np.integer)) and capacity > 0):
            raise GsRangeError(GS_E_RANGE, "GaussianData count must be a positive integer")
        self.capacity = int(capacity)
        self._values = np.zeros(self.capacity, np.uint32)
        self._ordinals = np.full(self.capacity, -1, np.int32)
        self._staged_values = np.zeros(self.capacity, np.uint32)
        self._staged_ordinals = np.zeros(self.capacity, np.int32)
        self.count = 0
        self.version = 0
        self._staged_count = 0
        self._staging = False
        self._first = 0
        self._last = -1
        self._listeners: list[Callable[[dict], None]] = []
        if slots is not None:
            self.stage_activation(np.asarray(slots, np.uint32), np.zeros(0, np.uint32), True)

    @property
    def slots(self) -> np.ndarray:
        """The committed list (a view of its first ``count`` entries)."""
        return self._values[: self.count]

    def subscribe(self, listener: Callable[[dict], None]) -> None:
        self._listeners.append(listener)

    def unsubscribe(self, listener: Callable[[dict], None]) -> None:
        self._listeners.remove(listener)

    def stage_activation(self, added, removed, commit: bool) -> dict | None:
        """Stage removals, then additions; with ``commit`` publish them.
        Returns the committed range {"start", "count"} (None when nothing was
        committed)."""
        added = np.asarray(added, np.int64).ravel()
        removed = np.asarray(removed, np.int64).ravel()
        if not self._staging and added.size == 0 and removed.size == 0:
            return None
        if added.size and (added.min() < 0 or added.max() >= self.capacity):
            raise GsRangeError(GS_E_RANGE, "Active GPU slot exceeds capacity")
        if removed.size and (removed.min() < 0 or removed.max() >= self.capacity):
            raise GsRangeError(GS_E_RANGE, "Removed GPU slot exceeds capacity")
        if not self._staging:
            self._staged_values[: self.count] = self._values[: self.count]
            self._staged_ordinals[:] = self._ordinals
            self._staged_count = self.count
            self._first, self._last = self.capacity, -1  # "infinity", -1
            self._staging = True
        # Removals come before additions across delta batches.
        for slot in removed.tolist():
            ordinal = int(self._staged_ordinals[slot])
            if ordinal < 0:
                continue
            self._staged_count -= 1
            last = int(self._staged_values[self._staged_count])
            self._staged_values[ordinal] = last
            self._staged_ordinals[last] = ordinal
            self._staged_ordinals[slot] = -1
            self._first = min(self._first, ordinal)
            self._last = max(self._last, ordinal)
        for slot in added.tolist():
            if self._staged_ordinals[slot] >= 0:
                continue
            ordinal = self._staged_count
            self._staged_count += 1
            self._staged_values[ordinal] = slot
            self._staged_ordinals[slot] = ordinal
            self._first = min(self._first, ordinal)
            self._last = max(self._last, ordinal)
        if not commit:
            return None
        self._values, self._staged_values = self._staged_values, self._values
        self._ordinals, self._staged_ordinals = self._staged_ordinals, self._ordinals
        self.count = self._staged_count
        self._staging = False
        self.version += 1
        start = min(self._first, self.count)
        end = min(self._last + 1, self.count)
        rng = {"start": start, "count": max(0, end - start)}
        for listener in list(self._listeners):
            listener(rng)
        return rng


def check_data(data: PackedScene) -> None:
    """GaussianData's checks (GaussianData.ts:76-166, 265-318) for the
    compact + rgb8e8 contract this path renders."""
    count = data.count
    if count <= 0:
        raise GsRangeError(GS_E_RANGE, "GaussianData count must be a positive integer")
    if data.sh_degree not in (0, 1, 2, 3):
        raise GsRangeError(GS_E_RANGE, "GaussianData shDegree must be 0, 1, 2, or 3")
    means = np.asarray(data.means)
    if means.dtype != np.float32:
        raise GsTypeError(GS_E_TYPE, "GaussianData means must use Float32Array storage")
    if means.ndim != 2 or means.shape[1] != 4:
        width = means.shape[1] if means.ndim == 2 else 1
        raise GsRangeError(GS_E_RANGE, f"GaussianData means itemSize is {width}; vec4 data requires itemSize 4")
    for arr, width in ((np.asarray(data.scales_opacity), 2), (np.asarray(data.rotations), 1)):
        shape_ok = arr.ndim == (1 if width == 1 else 2) and (width == 1 or arr.shape[1] == width)
        if arr.dtype != np.uint32 or not shape_ok or arr.shape[0] < count:
            raise GsTypeError(GS_E_TYPE, "Invalid compact Gaussian shape attribute")
    sh = np.asarray(data.sh)
    if sh.dtype != np.uint32:
        raise GsTypeError(GS_E_TYPE, "GaussianData rgb8e8 shCoefficients use the wrong typed array")
    need = count * (data.sh_degree + 1) ** 2
    if sh.size < need:
        raise GsRangeError(GS_E_RANGE, f"GaussianData shCoefficients has {sh.size} items; at least {need} are required")
    if data.active_slots is not None:
        slots = np.asarray(data.active_slots)
        if slots.size and int(slots.max()) >= count:
            raise GsRangeError(GS_E_RANGE, "Active GPU slot exceeds capacity")
