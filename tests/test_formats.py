"""Packed-buffer formats: the producer (libgs host code) against the
in-kernel decoders restated in the oracle, with the reference's own
tolerances (rust/gaussian-backend/tests/backend.rs:1085-1174,
test/GaussianSh.test.ts:19-40)."""
from __future__ import annotations

import ctypes as C
import math

import numpy as np
import pytest


def decode_scale_opacity(oracle, x, y):
    out = np.empty(4, np.float32)
    oracle.lib().gs_oracle_decode_scale_opacity(int(x), int(y), out.ctypes.data)
    return out


def decode_rotation(oracle, packed):
    out = np.empty(4, np.float32)
    oracle.lib().gs_oracle_decode_rotation(int(packed), out.ctypes.data)
    return out


def decode_sh(oracle, packed):
    out = np.empty(3, np.float32)
    oracle.lib().gs_oracle_decode_sh_rgb8e8(int(packed), out.ctypes.data)
    return out


def encode(gs, scales, opacity, quats):
    scales = np.ascontiguousarray(scales, np.float32)
    opacity = np.ascontiguousarray(opacity, np.float32)
    quats = np.ascontiguousarray(quats, np.float32)
    n = opacity.shape[0]
    shape = np.empty((n, 2), np.uint32)
    rot = np.empty(n, np.uint32)
    gs.lib().gs_encode_compact(n, scales.ctypes.data, opacity.ctypes.data, quats.ctypes.data,
                               shape.ctypes.data, rot.ctypes.data)
    return shape, rot


def test_backend_compact_precision_fixture(gs, oracle):
    # backend.rs:1088-1095: two Gaussians, opacity 0.73, extreme scales.
    half = 1.2
    scales = [[1e-8, 1.234, 1e5], [0.2, 0.2, 0.2]]
    quats = [[0.0, math.sin(half), 0.0, math.cos(half)], [0.0, 0.0, 0.0, 1.0]]
    shape, rot = encode(gs, scales, [0.73, 0.73], quats)
    for i in range(2):
        d = decode_scale_opacity(oracle, shape[i, 0], shape[i, 1])
        for axis in range(3):
            assert abs(d[axis] / scales[i][axis] - 1.0) < 0.01   # backend.rs:1145
        assert abs(d[3] - 0.73) < 0.001                           # backend.rs:1149
        q = decode_rotation(oracle, rot[i])
        assert abs(float(np.dot(q, quats[i]))) > 0.99999          # backend.rs:1156


def test_compact_round_trip_random(gs, oracle):
    rng = np.random.default_rng(3)
    n = 2000
    scales = np.exp(rng.uniform(-9, 0, (n, 3))).astype(np.float32)
    opacity = rng.uniform(0, 1, n).astype(np.float32)
    q = rng.normal(size=(n, 4)).astype(np.float32)
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    shape, rot = encode(gs, scales, opacity, q)
    for i in range(0, n, 7):
        d = decode_scale_opacity(oracle, shape[i, 0], shape[i, 1])
        np.testing.assert_allclose(d[:3], scales[i], rtol=0.01)
        assert abs(d[3] - opacity[i]) < 0.001
        assert abs(float(np.dot(decode_rotation(oracle, rot[i]), q[i]))) > 0.9999


def test_rotation_decode_closed_form(oracle):
    # Identity: r = 0 half-angle -> (0, 0, 0, 1) whatever the axis bits.
    np.testing.assert_allclose(decode_rotation(oracle, (512 << 10) | 512), [0, 0, 0, 1], atol=1e-7)
    # Axis +z (u = v = 511.5 -> centre), r = 4095 -> half angle pi/2.
    q = decode_rotation(oracle, (4095 << 20) | (511 << 10) | 511)
    assert abs(q[3]) < 1e-6 and abs(np.linalg.norm(q) - 1) < 1e-6


@pytest.mark.parametrize("rgb", [[0, 0, 0], [1, -0.5, 0.25], [-3.75, 2.125, 0.03125],
                                 [1e-4, -2e-4, 3e-4]])
def test_rgb8e8_round_trip(gs, oracle, rgb):
    # GaussianSh.test.ts:19-40: |decoded - value| <= 2^ceil(log2 max) / 127.
    a = np.ascontiguousarray(rgb, np.float32)
    packed = gs.lib().gs_pack_sh_rgb8e8(a.ctypes.data)
    d = decode_sh(oracle, packed)
    m = max(abs(v) for v in rgb)
    tol = 0.0 if m == 0 else 2.0 ** math.ceil(math.log2(m)) / 127
    assert np.all(np.abs(d - a) <= tol + 1e-12)


def test_rgb8e8_closed_form(oracle):
    # bytes (127, -127, 0), exponent byte 127 + 2 -> (4, -4, 0)
    packed = 127 | (0x81 << 8) | (0 << 16) | (129 << 24)
    np.testing.assert_array_equal(decode_sh(oracle, packed), [4.0, -4.0, 0.0])
    # -128 clamps to -1 (unpack4x8snorm), exponent byte 0 -> 2^-127 (subnormal)
    packed = 0x80 | (0 << 24)
    d = decode_sh(oracle, packed)
    assert d[0] == -np.float32(2.0) ** -127


def test_f16_encoder_matches_numpy(gs):
    x = np.random.default_rng(0).normal(scale=30, size=5000).astype(np.float32)
    got = np.array([gs.lib().gs_f32_to_f16(float(v)) for v in x], np.uint16)
    np.testing.assert_array_equal(got, x.astype(np.float16).view(np.uint16))


def test_generator_is_deterministic(gs):
    a = gs.workload_scene("small", count=5000)
    b = gs.workload_scene("small", count=5000)
    for f in ("means", "scales_opacity", "rotations", "sh"):
        np.testing.assert_array_equal(getattr(a, f), getattr(b, f))
    # distributions: log-scale mean, opacity median, DC coefficient mean
    c = gs.workload_scene("small", count=20000, keep_raw=True)
    ls = np.log(c.raw["scales"])
    assert abs(ls.mean() - (-4.5)) < 0.05 and abs(ls.std() - 0.6) < 0.05
    assert abs(np.median(c.raw["opacity"]) - 0.5) < 0.02
    assert np.all(np.abs(c.means[:, :3]) < 5.0 + 6 * 0.5)
