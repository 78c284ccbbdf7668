"""Accuracy of the shared binary32 elementary functions (gs_math.h) against
libm in double precision, plus their special values. The GPU kernels use
the same header, so these bounds hold for both sides."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

EXP, LOG, SIN, COS, ATAN, ATAN2, H2F, F2H, LOGPOS, EXP_NONPOS = range(10)


@pytest.fixture(scope="module")
def mathfn(oracle):
    h = oracle.lib()
    h.gs_oracle_math.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t]

    def run(fn, a, b=None):
        a = np.ascontiguousarray(a, np.float32)
        b = np.ascontiguousarray(a if b is None else b, np.float32)
        out = np.empty_like(a)
        h.gs_oracle_math(fn, a.ctypes.data, b.ctypes.data, out.ctypes.data, a.size)
        return out

    return run


def ulps(got, exact):
    ref = exact.astype(np.float32)
    spacing = np.spacing(np.abs(ref)).astype(np.float64)
    return np.abs(got.astype(np.float64) - exact) / spacing


RNG = np.random.default_rng(7)


@pytest.mark.parametrize("lo,hi", [(-87.0, 88.0), (-5.55, 0.0)])
def test_exp(mathfn, lo, hi):
    x = RNG.uniform(lo, hi, 100_000).astype(np.float32)
    assert ulps(mathfn(EXP, x), np.exp(x.astype(np.float64))).max() <= 1.0


def test_log(mathfn):
    x = np.exp(RNG.uniform(-85, 85, 100_000)).astype(np.float32)
    assert ulps(mathfn(LOG, x), np.log(x.astype(np.float64))).max() <= 1.0
    x = RNG.uniform(1.0, 255.0, 100_000).astype(np.float32)  # ln(255 * opacity)
    assert ulps(mathfn(LOG, x), np.log(x.astype(np.float64))).max() <= 1.0


def test_log_positive_normal(mathfn):
    """The learned-cull features' log (fused-multiply-add Cephes, no special
    values) over positive normal arguments: within 1 ulp of libm, and within
    1 ulp of gs_logf."""
    x = np.exp(RNG.uniform(-85, 85, 100_000)).astype(np.float32)
    x = x[x >= np.finfo(np.float32).tiny]
    got = mathfn(LOGPOS, x)
    assert ulps(got, np.log(x.astype(np.float64))).max() <= 1.0
    assert ulps(got, mathfn(LOG, x).astype(np.float64)).max() <= 1.0
    x = np.arange(1, 512, dtype=np.float32)  # pixel radii
    assert ulps(mathfn(LOGPOS, x), np.log(x.astype(np.float64))).max() <= 1.0


def test_exp_of_nonpositive_arguments(mathfn):
    """The rasterizer's exp (no range branches, one scaling product) gives
    gs_expf's bits on its whole domain [-87, 0]: every candidate power in
    [-17, 0] on a fine grid, random arguments over the domain, its end points,
    the arguments nearest each scaling step, and NaN."""
    grid = np.arange(-17.0, 0.0, 1.0 / 4096, dtype=np.float64).astype(np.float32)
    steps = (np.arange(-126, 1) - 0.5) * np.log(2.0)  # where n = floor(x log2 e + 1/2) steps
    near = np.concatenate([np.nextafter(steps.astype(np.float32), -np.inf),
                           steps.astype(np.float32), np.nextafter(steps.astype(np.float32), np.inf)])
    near = near[(near >= -87.0) & (near <= 0.0)]
    x = np.concatenate([grid, RNG.uniform(-87.0, 0.0, 200_000).astype(np.float32), near,
                        np.array([-87.0, 0.0, -0.0, np.finfo(np.float32).tiny * -1], np.float32)])
    np.testing.assert_array_equal(mathfn(EXP_NONPOS, x).view(np.uint32),
                                  mathfn(EXP, x).view(np.uint32))
    assert np.isnan(mathfn(EXP_NONPOS, np.array([np.nan], np.float32))).all()


@pytest.mark.parametrize("fn,ref", [(SIN, np.sin), (COS, np.cos)])
def test_sin_cos(mathfn, fn, ref):
    x = RNG.uniform(-3.3, 3.3, 100_000).astype(np.float32)
    got = mathfn(fn, x)
    exact = ref(x.astype(np.float64))
    assert np.abs(got - exact).max() <= 1.2e-7
    big = np.abs(exact) > 1e-3
    assert ulps(got[big], exact[big]).max() <= 2.0


def test_atan_atan2(mathfn):
    x = RNG.uniform(-100, 100, 100_000).astype(np.float32)
    assert ulps(mathfn(ATAN, x), np.arctan(x.astype(np.float64))).max() <= 3.0
    y = RNG.normal(size=100_000).astype(np.float32)
    x = RNG.normal(size=100_000).astype(np.float32)
    assert ulps(mathfn(ATAN2, y, x), np.arctan2(y.astype(np.float64), x.astype(np.float64))).max() <= 3.5


def test_special_values(mathfn):
    nan, inf = np.float32("nan"), np.float32("inf")
    assert np.isnan(mathfn(EXP, [nan]))[0] and mathfn(EXP, [100.0])[0] == inf
    assert mathfn(EXP, [-200.0])[0] == 0.0 and mathfn(EXP, [0.0])[0] == 1.0
    assert mathfn(LOG, [1.0])[0] == 0.0 and mathfn(LOG, [0.0])[0] == -inf
    assert np.isnan(mathfn(LOG, [-1.0]))[0] and mathfn(LOG, [inf])[0] == inf
    sub = np.float32(1e-40)
    assert abs(mathfn(LOG, [sub])[0] - np.log(1e-40)) < 1e-4
    # atan2 signed-zero conventions (C99), used by the eigen rotation angle
    y = np.array([0.0, -0.0, 0.0, -0.0, 1.0, -1.0], np.float32)
    x = np.array([0.0, 0.0, -0.0, -0.0, 0.0, 0.0], np.float32)
    np.testing.assert_array_equal(mathfn(ATAN2, y, x), np.arctan2(y, x))


def test_half_conversions_exhaustive(mathfn):
    bits = np.arange(65536, dtype=np.uint32)
    got = mathfn(H2F, bits.view(np.float32))
    ref = bits.astype(np.uint16).view(np.float16).astype(np.float32)
    same = (got.view(np.uint32) == ref.view(np.uint32)) | (np.isnan(got) & np.isnan(ref))
    assert same.all()
    # float -> half round-to-nearest-even agrees with numpy on a wide sweep
    x = np.concatenate([RNG.normal(scale=s, size=50_000) for s in (1e-6, 1e-3, 1.0, 100.0, 7e4)]).astype(np.float32)
    x = np.concatenate([x, ref[np.isfinite(ref)], ref[np.isfinite(ref)] * np.float32(1.0004883)])
    got = mathfn(F2H, x).view(np.uint32) & 0xFFFF
    with np.errstate(over="ignore"):
        expect = x.astype(np.float16).view(np.uint16)
    np.testing.assert_array_equal(got.astype(np.uint16), expect)


def test_small_integer_division_matches_ieee(oracle):
    """gs_div_small (the kernels' b / d for the packed formats' integer
    numerators: rgb8e8 bytes / 127, oct101012 fields / 1023 and / 4095) is
    the IEEE quotient for every numerator in range. The oracle divides
    directly, so this is what keeps the two bit-identical."""
    h = oracle.lib()
    h.gs_oracle_div_small_mismatches.restype = C.c_uint32
    assert h.gs_oracle_div_small_mismatches() == 0


def test_small_integer_division_rational_proof():
    """The same identity from exact rationals, independent of any float
    library: q1 = RN(b/d'), r = RN(b - q1 d), q = RN(r/d' + q1) with
    d' = RN(1/d) equals RN(b/d)."""
    from fractions import Fraction as F

    def rn32(x):
        c = np.float32(float(x))
        best = None
        for v in (np.nextafter(c, np.float32(-np.inf)), c, np.nextafter(c, np.float32(np.inf))):
            d = abs(F(float(v)) - x)
            even = (int(np.float32(v).view(np.uint32)) & 1) == 0
            if best is None or d < best[0] or (d == best[0] and even):
                best = (d, v)
        return F(float(best[1]))

    for d, lo, hi in ((127, -128, 127), (1023, 0, 1023), (4095, 0, 4095)):
        inv = rn32(F(1, d))
        for b in range(lo, hi + 1, 7 if d == 4095 else 1):
            q1 = rn32(F(b) * inv)
            r = rn32(F(b) - q1 * d)
            assert rn32(r * inv + q1) == rn32(F(b, d)), (d, b)
