"""Known-answer tests of the projection helpers, ported from the reference's
executed-WGSL tests (test/ProjectionBoundary.test.ts), run on the oracle's
restatement of the same helpers."""
from __future__ import annotations

import math

import numpy as np
import pytest

F = np.float32


def arr(*v):
    return np.ascontiguousarray(v, np.float32)


def persp(oracle, view, cov, proj, viewport, fv):
    out = np.empty(16, np.float32)
    v, vp = arr(*view), arr(*viewport)
    oracle.lib().gs_oracle_perspective_projection(v.ctypes.data, cov.ctypes.data,
                                                  proj.ctypes.data, vp.ctypes.data,
                                                  fv, out.ctypes.data)
    return out.reshape(4, 4)  # [column][row]


def result(oracle, center, depth, cov, fv):
    out = np.empty(16, np.float32)
    c = arr(*center)
    k = arr(*cov)
    oracle.lib().gs_oracle_projection_result(c.ctypes.data, depth, k.ctypes.data, fv, out.ctypes.data)
    return out.reshape(4, 4)


def valid(oracle, p):
    p = np.ascontiguousarray(p, np.float32)
    return oracle.lib().gs_oracle_projection_is_valid(p.ctypes.data)


def test_default_perspective_projection(oracle):
    # ProjectionBoundary.test.ts:30-54
    cov = np.empty(9, np.float32)
    scale, rot, ident = arr(1, 1, 1), arr(0, 0, 0, 1), np.eye(4, dtype=np.float32).ravel()
    oracle.lib().gs_oracle_view_covariance(scale.ctypes.data, rot.ctypes.data, ident.ctypes.data,
                                           cov.ctypes.data)
    np.testing.assert_array_equal(cov.reshape(3, 3), np.eye(3))
    proj = np.eye(4, dtype=np.float32).ravel()
    p = persp(oracle, (0, 0, -2), cov, proj, (32, 32, 1, 10), 0.3)
    assert list(p[0, :3]) == [16, 16, 2]
    assert p[2, 0] == pytest.approx(64.3, abs=5e-5)
    assert p[1, 0] == pytest.approx(1 / 64.3, abs=5e-7)
    assert p[2, 3] == 4096
    assert valid(oracle, p) == 1
    clipped = persp(oracle, (0, 0, 2), cov, proj, (32, 32, 1, 10), 0.3)
    assert clipped[0, 3] == 0


def test_custom_scalar_projection_and_rejections(oracle):
    # ProjectionBoundary.test.ts:56-72
    p = result(oracle, (-3, 7), 42, (4, 1, 9), 0)
    assert p[0, 2] == 42
    # The reference pins p[2].x == 4 under WgslExec's evaluation; the binary32
    # eigen reconstruction here lands within 2 ulp of it.
    assert abs(float(p[2, 0]) - 4.0) <= 2 * np.spacing(F(4))
    assert p[2, 2] == pytest.approx(9, abs=5e-5)
    assert valid(oracle, p) == 1
    neg = p.copy()
    neg[0] = [p[0, 0], p[0, 1], -1, 1]
    assert valid(oracle, neg) == 0
    inf = p.copy()
    inf[0, 2] = np.inf
    assert valid(oracle, inf) == 0
    bad = p.copy()
    bad[1, 0] = 99
    assert valid(oracle, bad) == 0


@pytest.mark.parametrize("axis", [1, 2])
def test_periodic_delta_picks_strongest_image(oracle, axis):
    # ProjectionBoundary.test.ts:74-104
    conic = arr(0.1, 0.08, 0.12)
    viewport = arr(7, 9)
    center = arr(-15.25, 26.5)
    for pixel in ([0.5, 0.5], [6.5, 8.5], [3.5, 4.5]):
        out = np.empty(2, np.float32)
        px = arr(*pixel)
        oracle.lib().gs_oracle_boundary_delta(px.ctypes.data, center.ctypes.data,
                                              conic.ctypes.data, viewport.ctypes.data, axis,
                                              out.ctypes.data)

        def power(x, y):
            return 0.1 * x * x + 2 * 0.08 * x * y + 0.12 * y * y

        best = min(power(pixel[0] - center[0] - (i * 7 if axis == 1 else 0),
                         pixel[1] - center[1] - (i * 9 if axis == 2 else 0))
                   for i in range(-100, 101))
        assert power(float(out[0]), float(out[1])) == pytest.approx(best, abs=5e-5)


FIXTURES = [
    dict(center=[-0.25, 8.5], cov=[4, 0, 4], width=33, height=17, axis=1),
    dict(center=[16.5, 18.25], cov=[4, 0, 4], width=33, height=17, axis=2),
    dict(center=[-35.5, 8.5], cov=[100, 24, 9], width=7, height=17, axis=1),
    dict(center=[8.5, -41.25], cov=[9, 24, 100], width=17, height=7, axis=2),
]


@pytest.mark.parametrize("fx", FIXTURES)
def test_tile_hits_cover_every_sampled_pixel(oracle, fx):
    # ProjectionBoundary.test.ts:106-163 (brute force over periodic images)
    cov = fx["cov"]
    det = cov[0] * cov[2] - cov[1] ** 2
    conic = arr(cov[2] / det, -cov[1] / det, cov[0] / det)
    thr = math.log(255)
    radius = arr(math.ceil(math.sqrt(2 * thr * cov[0])), math.ceil(math.sqrt(2 * thr * cov[2])))
    w, h, axis = fx["width"], fx["height"], fx["axis"]
    nx, ny = math.ceil(w / 16), math.ceil(h / 16)
    center = arr(*fx["center"])
    vp = arr(w, h)
    L = oracle.lib()
    hits = np.array([[L.gs_oracle_tile_hit(center.ctypes.data, conic.ctypes.data, thr,
                                           radius.ctypes.data, tx, ty, vp.ctypes.data, axis)
                      for tx in range(nx)] for ty in range(ny)])
    count = L.gs_oracle_count_tiles(center.ctypes.data, conic.ctypes.data, thr, radius.ctypes.data,
                                    0, 0, nx - 1, ny - 1, vp.ctypes.data, axis)
    assert count == hits.sum() <= nx * ny
    sampled = 0
    c = [float(v) for v in conic]
    for y in range(h):
        for x in range(w):
            best = min(0.5 * (c[0] * dx * dx + 2 * c[1] * dx * dy + c[2] * dy * dy)
                       for i in range(-100, 101)
                       for dx, dy in [(x + 0.5 - fx["center"][0] - (i * w if axis == 1 else 0),
                                       y + 0.5 - fx["center"][1] - (i * h if axis == 2 else 0))])
            if best <= thr:
                sampled += 1
                assert hits[y // 16, x // 16] == 1
    assert sampled > 0


def test_subpixel_periodic_sample(oracle):
    # ProjectionBoundary.test.ts:165-175
    L = oracle.lib()
    center, conic, extent = arr(33.5, 0.5), arr(100, 0, 100), arr(0.1, 0.1)
    args = (center.ctypes.data, conic.ctypes.data, 1.0, extent.ctypes.data, 33, 17)
    assert L.gs_oracle_subpixel_has_sample(*args, 1) == 1
    assert L.gs_oracle_subpixel_has_sample(*args, 0) == 0


def test_tile_test_is_conservative_brute_force(oracle):
    """Invariant for clip mode: a tile that contains a pixel centre inside the
    power threshold is always counted (random Gaussians, brute force)."""
    rng = np.random.default_rng(11)
    L = oracle.lib()
    vp = arr(64, 48)
    for _ in range(150):
        a, b = rng.uniform(0.5, 40, 2)
        rho = rng.uniform(-0.9, 0.9)
        cov = np.array([a, rho * math.sqrt(a * b), b])
        det = cov[0] * cov[2] - cov[1] ** 2
        conic = arr(cov[2] / det, -cov[1] / det, cov[0] / det)
        thr = float(np.log(255 * rng.uniform(0.01, 1)))
        if thr <= 0:
            continue
        radius = arr(math.ceil(math.sqrt(2 * thr * cov[0])), math.ceil(math.sqrt(2 * thr * cov[2])))
        center = arr(*rng.uniform([-10, -10], [74, 58]))
        hits = np.array([[L.gs_oracle_tile_hit(center.ctypes.data, conic.ctypes.data, thr,
                                               radius.ctypes.data, tx, ty, vp.ctypes.data, 0)
                          for tx in range(4)] for ty in range(3)])
        ys, xs = np.mgrid[0:48, 0:64]
        dx = xs + 0.5 - center[0]
        dy = ys + 0.5 - center[1]
        c = conic.astype(np.float64)
        sigma = 0.5 * (c[0] * dx * dx + 2 * c[1] * dx * dy + c[2] * dy * dy)
        inside = sigma <= thr * (1 - 1e-5)
        assert np.all(hits[ys[inside] // 16, xs[inside] // 16] == 1)
