"""The learned cull: training on the GPU and use in the frame pipeline.

A 16-64-64-1 MLP (bf16 weights on the matrix cores, fp32 accumulation),
fused into projection, predicts each visible Gaussian's largest alpha * T;
with the cull enabled, Gaussians predicted below the threshold (1e-4 by
default) are dropped before duplication and sort. Its training targets are
recorded by the plain rasterizer itself (an instrumented frame raises a
per-Gaussian record to every alpha * T it blends), over seeded orbit views;
training runs the library's own forward and backward kernels with Adam.
The plain pipeline is both the fallback (cull disabled) and the yardstick:
the image loss of the cull is measured as PSNR against the plain frame.

No counterpart in the reference; the nearest mechanisms there are the
opacity and subpixel culls (ProjectionStage.ts:533, 572-587).
"""
from __future__ import annotations

import ctypes as C
import math
from pathlib import Path

import numpy as np

from . import _native as N
from .camera import PerspectiveCamera
from .gaussian_pass import frame_for, object_frames_for

PARAMS = 5313
FEATURES = 16
# gs_cull_save / gs_cull_load blob: 32-byte header, params, Adam m and v, mu, inv_sigma.
CHECKPOINT_BYTES = 32 + (3 * PARAMS + 2 * FEATURES) * 4
HIDDEN = 64
# Feature indices (the order of include/gs.h).
FEATURE_LOG_RADIUS = 0
FEATURE_OPACITY = 4


class LearnedCull:
    """The learned cull of one context (a GaussianPass or BatchedPass)."""

    def __init__(self, owner, seed: int = 0):
        self.owner = owner
        self._ctx = owner._ctx
        N.check(N.lib().gs_cull_init(self._ctx, seed), self._ctx)
        self.threshold = 1e-4
        self.enabled = False

    # -- training set ------------------------------------------------------
    def reserve(self, capacity: int) -> None:
        N.check(N.lib().gs_cull_reserve(self._ctx, capacity), self._ctx)

    def collect(self, camera: PerspectiveCamera, width: int, height: int, rate: float,
                seed: int, max_retries: int = 6) -> N.GsState:
        """Record one plain frame and append a sample of its visible Gaussians;
        an overflowing frame grows the owner's capacity and is recorded again."""
        table = np.ascontiguousarray(object_frames_for(camera), np.float32).reshape(-1, 40)
        N.check(N.lib().gs_set_objects(self._ctx, table.shape[0], table.ctypes.data), self._ctx)
        for _ in range(max_retries):
            f = frame_for(camera, width, height, self.owner.options)
            f.object_frames = table.ctypes.data
            st = N.GsState()
            N.check(N.lib().gs_cull_collect(self._ctx, C.byref(f), rate, seed, C.byref(st)),
                    self._ctx)
            if not st.overflow:
                return st
            nxt = self.owner.policy.next(st.capacity, st.requested)
            if nxt is None:
                raise N.GsRangeError(N.GS_E_RANGE, "intersection demand exceeds maximum")
            N.check(N.lib().gs_resize_intersections(self._ctx, nxt), self._ctx)
            if hasattr(self.owner, "capacity"):
                self.owner.capacity = nxt
        raise N.GsError(N.GS_E_STATE, "collect frame kept overflowing")

    def dataset_size(self) -> int:
        n = C.c_uint32()
        N.check(N.lib().gs_cull_dataset_size(self._ctx, C.byref(n)), self._ctx)
        return n.value

    def read_dataset(self, first: int = 0, n: int | None = None) -> tuple[np.ndarray, np.ndarray]:
        n = self.dataset_size() - first if n is None else n
        f = np.empty((n, FEATURES), np.float32)
        t = np.empty((n,), np.float32)
        N.check(N.lib().gs_cull_read_dataset(self._ctx, first, n, f.ctypes.data, t.ctypes.data),
                self._ctx)
        return f, t

    # -- training ----------------------------------------------------------
    def fit_normalization(self) -> None:
        N.check(N.lib().gs_cull_fit_normalization(self._ctx), self._ctx)

    def train(self, steps: int, batch: int = 65536, lr: float = 2e-3, under_weight: float = 4.0,
              seed: int = 0) -> float:
        loss = C.c_float()
        N.check(N.lib().gs_cull_train(self._ctx, steps, batch, lr, under_weight, seed,
                                      C.byref(loss)), self._ctx)
        return loss.value

    # -- parameters --------------------------------------------------------
    def params(self) -> np.ndarray:
        p = np.empty(PARAMS, np.float32)
        N.check(N.lib().gs_cull_get_params(self._ctx, p.ctypes.data), self._ctx)
        return p

    def set_params(self, p: np.ndarray) -> None:
        p = np.ascontiguousarray(p, np.float32)
        assert p.shape == (PARAMS,)
        N.check(N.lib().gs_cull_set_params(self._ctx, p.ctypes.data), self._ctx)

    def normalization(self) -> tuple[np.ndarray, np.ndarray]:
        mu = np.empty(FEATURES, np.float32)
        inv = np.empty(FEATURES, np.float32)
        N.check(N.lib().gs_cull_get_normalization(self._ctx, mu.ctypes.data, inv.ctypes.data),
                self._ctx)
        return mu, inv

    def set_normalization(self, mu: np.ndarray, inv_sigma: np.ndarray) -> None:
        mu = np.ascontiguousarray(mu, np.float32)
        inv = np.ascontiguousarray(inv_sigma, np.float32)
        N.check(N.lib().gs_cull_set_normalization(self._ctx, mu.ctypes.data, inv.ctypes.data),
                self._ctx)

    # -- checkpoint --------------------------------------------------------
    def checkpoint(self) -> bytes:
        """The training state (parameters, Adam moments and step count,
        standardisation, threshold) as the library's checkpoint blob of
        CHECKPOINT_BYTES bytes (gs_cull_save)."""
        buf = (C.c_ubyte * CHECKPOINT_BYTES)()
        N.check(N.lib().gs_cull_save(self._ctx, buf, CHECKPOINT_BYTES), self._ctx)
        return bytes(buf)

    def restore(self, blob: bytes) -> None:
        """Resume from a checkpoint blob (gs_cull_load); the library rejects
        a blob of another size, version or network, or with non-finite
        values, with GsRangeError and keeps its state."""
        blob = bytes(blob)
        N.check(N.lib().gs_cull_load(self._ctx, blob, len(blob)), self._ctx)
        self.threshold = float(np.frombuffer(blob, np.float32, 1, 24)[0])

    def save(self, path) -> None:
        Path(path).write_bytes(self.checkpoint())

    def load(self, path) -> None:
        self.restore(Path(path).read_bytes())

    # -- use ---------------------------------------------------------------
    def enable(self, threshold: float | None = None) -> None:
        if threshold is not None:
            self.threshold = threshold
        N.check(N.lib().gs_cull_enable(self._ctx, 1, self.threshold), self._ctx)
        self.enabled = True

    def disable(self) -> None:
        N.check(N.lib().gs_cull_enable(self._ctx, 0, self.threshold), self._ctx)
        self.enabled = False

    def set_threshold(self, threshold: float) -> None:
        self.threshold = threshold
        N.check(N.lib().gs_cull_enable(self._ctx, 1 if self.enabled else 0, threshold), self._ctx)

    def calibrate(self, cameras: list[PerspectiveCamera], sizes: list[tuple[int, int]],
                  target_psnr: float = 46.0,
                  shifts=(-3.0, -2.5, -2.0, -1.875, -1.75, -1.625, -1.5, -1.375, -1.25, -1.125,
                          -1.0, -0.75, -0.5, -0.25, 0.0, 0.5, 1.0, 1.5, 2.0, 3.0, 4.0, 6.0,
                          8.0)) -> tuple[float, list]:
        """Move the network's output bias by the lowest shift (most culling)
        whose culled frames of these cameras (a BatchedPass owner) keep a
        mean PSNR of at least target_psnr against their plain frames. Returns
        the shift and (shift, key fraction kept, mean PSNR, min PSNR) rows.
        When no shift reaches the target the shift is None, the parameters
        are left as they were and the cull is disabled."""
        b = self.owner
        was = self.enabled
        self.disable()
        b.render(cameras, sizes)
        plain = [b.read_image(v).copy() for v in range(len(cameras))]
        keys = sum(s.intersections for s in b.read_states())
        p0 = self.params()
        rows = []
        chosen = None
        for shift in shifts:
            p = p0.copy()
            p[PARAMS - 1] += shift
            self.set_params(p)
            self.enable()
            b.render(cameras, sizes)
            kept = sum(s.intersections for s in b.read_states()) / max(1, keys)
            ps = [psnr(b.read_image(v), plain[v]) for v in range(len(cameras))]
            rows.append((shift, kept, float(np.mean(ps)), float(min(ps))))
            if np.mean(ps) >= target_psnr:
                chosen = shift
                break
        p = p0.copy()
        if chosen is not None:
            p[PARAMS - 1] += chosen
        self.set_params(p)
        if not was or chosen is None:
            self.disable()
        return chosen, rows

    def set_linear_score(self, feature: int, scale: float, bias: float) -> None:
        """A non-learned cull on one feature (see linear_score_params), with
        the current normalisation."""
        mu, inv = self.normalization()
        self.set_params(linear_score_params(feature, scale, bias, mu, inv))

    def predict(self, features: np.ndarray) -> np.ndarray:
        f = np.ascontiguousarray(features, np.float32).reshape(-1, FEATURES)
        out = np.empty(f.shape[0], np.float32)
        if f.shape[0]:
            N.check(N.lib().gs_cull_eval(self._ctx, f.shape[0], f.ctypes.data, out.ctypes.data),
                    self._ctx)
        return out

    def gradient(self, features: np.ndarray, targets: np.ndarray,
                 under_weight: float = 4.0) -> tuple[np.ndarray, float]:
        f = np.ascontiguousarray(features, np.float32).reshape(-1, FEATURES)
        t = np.ascontiguousarray(targets, np.float32)
        g = np.empty(PARAMS, np.float32)
        loss = C.c_float()
        N.check(N.lib().gs_cull_gradient(self._ctx, f.shape[0], f.ctypes.data, t.ctypes.data,
                                         under_weight, g.ctypes.data, C.byref(loss)), self._ctx)
        return g, loss.value


def linear_score_params(feature: int, scale: float, bias: float, mu: np.ndarray,
                        inv_sigma: np.ndarray) -> np.ndarray:
    """Parameters under which the network outputs scale * f[feature] + bias
    (f the raw feature; the bf16 rounding of the standardised input and of
    the output weights aside). With scale > 0 the cull then drops the
    Gaussians whose feature lies below (log(threshold) - bias) / scale: a
    non-learned cull (an opacity or footprint bound) to compare the trained
    network with. Two hidden units carry x and -x through the ReLUs."""
    off_w2 = HIDDEN * FEATURES + HIDDEN
    off_w3 = off_w2 + HIDDEN * HIDDEN + HIDDEN
    p = np.zeros(PARAMS, np.float32)
    p[0 * FEATURES + feature] = 1.0        # W1: hidden 0 = x, hidden 1 = -x
    p[1 * FEATURES + feature] = -1.0
    p[off_w2 + 0 * HIDDEN + 0] = 1.0       # W2: identity on the two units
    p[off_w2 + 1 * HIDDEN + 1] = 1.0
    p[off_w3 + 0] = scale / float(inv_sigma[feature])
    p[off_w3 + 1] = -scale / float(inv_sigma[feature])
    p[PARAMS - 1] = scale * float(mu[feature]) + bias
    return p


def psnr(a: np.ndarray, b: np.ndarray) -> float:
    """PSNR of two rgba16float images over their rgb channels (peak 1)."""
    d = a[..., :3].astype(np.float64) - b[..., :3].astype(np.float64)
    mse = float(np.mean(d * d))
    return float("inf") if mse == 0 else 10.0 * math.log10(1.0 / mse)
