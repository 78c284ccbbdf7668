"""Independent pins of four oracle functions that the bit-exact GPU tests
cannot pin by themselves (the kernels were written from the same WGSL):

- SH colour (projectionHelpers.ts:118-179): against a float64 real
  spherical-harmonic basis built from scipy's associated Legendre functions
  (Condon-Shortley phase) and the normalisation
  sqrt((2l+1)/(4 pi) (l-m)!/(l+m)!), not from the WGSL constants; rgb8e8
  words decoded by their closed form (projectionHelpers.ts:123-127).
- The multi-Gaussian composite (TileRasterizer.ts:608-709, storeFinalPixel
  :930-964): a float64 numpy renderer (closed-form decode, EWA projection,
  SH, one whole-image front-to-back blend per pixel in depth order, no tiles,
  no lists) against the oracle's binary32 frame on tiny scenes.
- Depth finalisation (finalizeGaussianDepth.ts:18-68 with the default depth
  node, GaussianContextNodes.ts:125-131): the known answers of
  test/ScalarOutput.test.ts:61-96 and the closed form of the encoding.
- The raster work counters (TileRasterizer.ts:617, :750-768): restated per
  pixel and chunk task over the oracle's own sorted records, vectorised in
  numpy, and shown to tell apart the plausible miscounts.
"""
from __future__ import annotations

import math

import numpy as np
import pytest
from scipy.special import lpmv


# ------------------------------------------------------------------ SH ----

def real_sh_basis(degree: int, dirs: np.ndarray) -> np.ndarray:
    """Real SH Y_lm (index l*l + l + m) at unit directions, float64."""
    x, y, z = (dirs[:, i].astype(np.float64) for i in range(3))
    cos_t = np.clip(z, -1.0, 1.0)
    phi = np.arctan2(y, x)
    cols = []
    for l in range(degree + 1):
        for m in range(-l, l + 1):
            a = abs(m)
            norm = math.sqrt((2 * l + 1) / (4 * math.pi) * math.factorial(l - a) / math.factorial(l + a))
            p = lpmv(a, l, cos_t)  # includes the Condon-Shortley phase (-1)^m
            if m == 0:
                cols.append(norm * p)
            elif m > 0:
                cols.append(math.sqrt(2.0) * norm * p * np.cos(a * phi))
            else:
                cols.append(math.sqrt(2.0) * norm * p * np.sin(a * phi))
    return np.stack(cols, axis=1)


def decode_rgb8e8(words: np.ndarray) -> np.ndarray:
    """Closed form: snorm8 mantissas (/127, -128 clamps to -1) times 2^(e - 127)."""
    w = np.asarray(words, np.uint32)
    b = np.stack([(w >> s) & 0xFF for s in (0, 8, 16)], axis=-1).astype(np.uint8).view(np.int8)
    mant = np.maximum(b.astype(np.float64) / 127.0, -1.0)
    return mant * np.exp2(((w >> 24) & 0xFF).astype(np.float64) - 127.0)[..., None]


def random_sh_words(rng, n: int, degree: int) -> np.ndarray:
    ncoef = (degree + 1) ** 2
    mant = rng.integers(-128, 128, size=(n, ncoef, 3)).astype(np.int8).view(np.uint8).astype(np.uint32)
    exp = np.empty((n, ncoef), np.uint32)
    exp[:, 0] = rng.integers(127 - 6, 127 - 1, size=n)          # DC up to 1/2
    if ncoef > 1:
        exp[:, 1:] = rng.integers(127 - 12, 127 - 5, size=(n, ncoef - 1))  # bands up to 1/32
    return mant[..., 0] | (mant[..., 1] << 8) | (mant[..., 2] << 16) | (exp << 24)


def test_real_sh_basis_is_orthonormal():
    """The float64 basis is orthonormal on the sphere (Gauss-Legendre in cos
    theta times a uniform azimuth grid, exact for these degrees)."""
    nodes, weights = np.polynomial.legendre.leggauss(16)
    phis = np.arange(32) * (2 * math.pi / 32)
    ct, ph = np.meshgrid(nodes, phis, indexing="ij")
    st = np.sqrt(1 - ct ** 2)
    dirs = np.stack([(st * np.cos(ph)).ravel(), (st * np.sin(ph)).ravel(), ct.ravel()], axis=1)
    w = (weights[:, None] * np.full(ph.shape[1], 2 * math.pi / 32)[None, :]).ravel()
    y = real_sh_basis(3, dirs)
    gram = (y * w[:, None]).T @ y
    np.testing.assert_allclose(gram, np.eye(16), atol=1e-12)


@pytest.mark.parametrize("degree", [0, 1, 2, 3])
def test_sh_matches_legendre_basis(oracle, degree):
    rng = np.random.default_rng(100 + degree)
    n = 400
    dirs = rng.normal(size=(n, 3))
    dirs = (dirs / np.linalg.norm(dirs, axis=1, keepdims=True)).astype(np.float32)
    words = random_sh_words(rng, n, degree)
    coef = decode_rgb8e8(words)                                  # (n, ncoef, 3)
    basis = real_sh_basis(degree, dirs.astype(np.float64))      # (n, ncoef)
    want = np.einsum("nk,nkc->nc", basis, coef)
    assert np.abs(want).max() < 0.49, "the test's coefficients must not clamp"
    got = np.stack([oracle.evaluate_sh(words[i], degree, dirs[i]) for i in range(n)])
    err = np.abs(got.astype(np.float64) - 0.5 - want)
    assert err.max() <= 1e-6, f"max |error| {err.max():.3g}"
    # The check has the power to see a one-term slip: flipping the sign of
    # any single basis function (a sign or band-constant mistake) moves the
    # result far outside the tolerance.
    for k in range(basis.shape[1]):
        flipped = basis.copy()
        flipped[:, k] = -flipped[:, k]
        alt = np.einsum("nk,nkc->nc", flipped, coef)
        assert np.abs(got.astype(np.float64) - 0.5 - alt).max() > 1e-3, f"term {k}"


def test_sh_degree0_is_the_dc_constant(oracle):
    """Y_00 = 1 / (2 sqrt(pi)) in every direction."""
    words = random_sh_words(np.random.default_rng(5), 1, 0)[0]
    c = decode_rgb8e8(words)[0]
    for d in ([0, 0, 1], [1, 0, 0], [0, -1, 0], [0.6, 0.0, -0.8]):
        got = oracle.evaluate_sh(words, 0, np.array(d, np.float32))
        np.testing.assert_allclose(got, c / (2 * math.sqrt(math.pi)) + 0.5, atol=1e-7)


# ------------------------------------------------------- composite ----

def decode_compact(shape: np.ndarray, rot: np.ndarray):
    """compactGaussians.ts:2-19 by closed form, float64."""
    halves = np.stack([shape[:, 0] & 0xFFFF, shape[:, 0] >> 16, shape[:, 1] & 0xFFFF,
                       shape[:, 1] >> 16], axis=1).astype(np.uint16).view(np.float16)
    h = halves.astype(np.float64)
    scales, opacity = np.exp(h[:, :3]), h[:, 3]
    u = (rot & 1023).astype(np.float64)
    v = ((rot >> 10) & 1023).astype(np.float64)
    px, py = u / 1023.0 * 2.0 - 1.0, v / 1023.0 * 2.0 - 1.0
    z = 1.0 - np.abs(px) - np.abs(py)
    t = np.maximum(-z, 0.0)
    x = px + np.where(px >= 0, -t, t)
    y = py + np.where(py >= 0, -t, t)
    axis = np.stack([x, y, z], axis=1)
    axis /= np.linalg.norm(axis, axis=1, keepdims=True)
    half = (rot >> 20).astype(np.float64) / 4095.0 * (math.pi / 2)
    quats = np.concatenate([axis * np.sin(half)[:, None], np.cos(half)[:, None]], axis=1)
    return scales, opacity, quats


def quat_matrix(q: np.ndarray) -> np.ndarray:
    x, y, z, w = q / np.linalg.norm(q)
    return np.array([
        [1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
        [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
        [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)],
    ])


def brute_force_frame(scene, frame, objects, transmittance=1e-4, background=(0, 0, 0, 0),
                      slip=None):
    """Float64 render: every Gaussian against every pixel, in depth order.

    Returns (image (H, W, 4), margin (H, W)): margin is the smallest relative
    distance of any of the pixel's discrete decisions (alpha against 1/255,
    T against the stop threshold) from its threshold. `slip` plants one
    plausible mistake ("cross_sign", "no_compensation", "back_to_front",
    "sh_band") to show that the comparison sees it."""
    W, H = frame.width, frame.height
    proj = np.array(list(frame.projection), np.float32).astype(np.float64).reshape(4, 4).T
    obj = np.asarray(objects, np.float32).astype(np.float64).reshape(-1, 40)[0]
    model_view = obj[16:32].reshape(4, 4).T
    cam_local = obj[32:35]
    scales, opac, quats = decode_compact(scene.scales_opacity, scene.rotations)
    coef = decode_rgb8e8(scene.sh)
    fv = float(np.float32(frame.filter_variance))
    fx, fy = 0.5 * W * proj[0, 0], 0.5 * H * proj[1, 1]
    splats = []
    for g in range(scene.count):
        pos = scene.means[g, :3].astype(np.float64)
        view = model_view @ np.append(pos, 1.0)
        depth = -view[2]
        assert frame.near_plane < depth < frame.far_plane
        clip = proj @ np.append(view[:3], 1.0)
        ndc = clip[:2] / clip[3]
        center = np.array([(ndc[0] * 0.5 + 0.5) * W, (0.5 - ndc[1] * 0.5) * H])
        r = quat_matrix(quats[g])
        cov_local = r @ np.diag(scales[g] ** 2) @ r.T
        w3 = model_view[:3, :3]
        cov_view = w3 @ cov_local @ w3.T
        j = np.array([[fx / depth, 0.0, fx * view[0] / depth ** 2],
                      [0.0, -fy / depth, -fy * view[1] / depth ** 2]])
        cov2 = j @ cov_view @ j.T
        det0 = max(cov2[0, 0] * cov2[1, 1] - cov2[0, 1] ** 2, 0.0)
        s = cov2 + fv * np.eye(2)
        lam = np.linalg.eigvalsh(s)
        assert 1e-6 < lam[0] and lam[1] < 1e4, "outside the eigen clamp: pick another scene"
        det = s[0, 0] * s[1, 1] - s[0, 1] ** 2
        conic = np.array([s[1, 1], -s[0, 1], s[0, 0]]) / det
        if slip == "cross_sign":
            conic[1] = -conic[1]
        o = min(max(opac[g], 0.0), 1.0)
        if slip != "no_compensation":
            o *= math.sqrt(min(max(det0 / det, 0.0), 1.0))
        assert o >= 1 / 255
        d = pos - cam_local
        d = (d / np.linalg.norm(d)).astype(np.float32).astype(np.float64)
        basis = real_sh_basis(scene.sh_degree, d[None])[0]
        if slip == "sh_band":
            basis[1:4] = -basis[1:4]
        color = np.clip(basis @ coef[g] + 0.5, 0.0, 1.0)
        splats.append((depth, g, center, conic, o, color))
    splats.sort(key=lambda s: (s[0], s[1]), reverse=slip == "back_to_front")
    img = np.zeros((H, W, 4))
    margin = np.full((H, W), np.inf)
    bg = np.asarray(background, np.float64)
    bg_a = min(max(bg[3], 0.0), 1.0)
    ys, xs = np.mgrid[0:H, 0:W]
    pxc, pyc = xs + 0.5, ys + 0.5
    c = np.zeros((H, W, 3))
    t = np.ones((H, W))
    live = np.ones((H, W), bool)
    for _, _, center, conic, o, color in splats:
        dx, dy = pxc - center[0], pyc - center[1]
        power = -0.5 * (conic[0] * dx * dx + 2 * conic[1] * dx * dy + conic[2] * dy * dy)
        raw = o * np.exp(power)
        alpha = np.minimum(raw, 0.99)
        use = live & (alpha >= 1 / 255)
        margin = np.where(live, np.minimum(margin, np.abs(raw * 255 - 1)), margin)
        contrib = np.where(use, t * alpha, 0.0)
        c += contrib[..., None] * color
        t_new = np.where(use, t * (1 - alpha), t)
        margin = np.where(use, np.minimum(margin, np.abs(t_new / transmittance - 1)), margin)
        t = t_new
        live &= ~(use & (t < transmittance))
    img[..., :3] = c + bg[:3] * (t * bg_a)[..., None]
    img[..., 3] = 1 - t * (1 - bg_a)
    return img, margin


def tiny_scene(gs, seed: int, n: int = 16):
    """n Gaussians in a 2-unit cube seen from distance 3 on a 32 x 32 image."""
    rng = np.random.default_rng(seed)
    means = np.zeros((n, 4), np.float32)
    means[:, :3] = rng.uniform(-1.0, 1.0, size=(n, 3))
    scales = np.exp(rng.normal(-1.3, 0.4, size=(n, 3))).astype(np.float32)
    opacity = (1 / (1 + np.exp(-rng.normal(1.0, 1.2, size=n)))).astype(np.float32)
    q = rng.normal(size=(n, 4)).astype(np.float32)
    shape = np.empty((n, 2), np.uint32)
    rot = np.empty(n, np.uint32)
    gs.lib().gs_encode_compact(n, scales.ctypes.data, opacity.ctypes.data, q.ctypes.data,
                               shape.ctypes.data, rot.ctypes.data)
    sh = random_sh_words(rng, n, 3)
    scene = gs.PackedScene(means, shape, rot, sh, 3)
    cam = gs.PerspectiveCamera(fov=60.0, aspect=1.0, near=0.1, far=100.0,
                               position=tuple(rng.normal(size=3) * 0.3 + np.array([0, 0, 3.0])))
    return scene, cam


@pytest.mark.parametrize("seed", range(8))
def test_frame_matches_bruteforce_composite(gs, oracle, seed):
    scene, cam = tiny_scene(gs, seed)
    frame = gs.frame_for(cam, 32, 32)
    objects = gs.object_frames_for(cam)[None]
    background = (0.1, 0.2, 0.3, 0.5) if seed % 2 else (0, 0, 0, 0)
    ref = oracle.run_frame(scene, frame, objects, capacity=1 << 16, threads=1,
                           background=background)
    assert ref.counts.visible == scene.count and ref.counts.overflow == 0
    want, margin = brute_force_frame(scene, frame, objects, background=background)
    got = ref.get("image_f32").astype(np.float64)
    decided = margin > 1e-4  # no decision within 1e-4 (relative) of its threshold
    assert decided.mean() > 0.97
    err = np.abs(got - want)[decided]
    assert err.max() <= 2e-6, f"max |error| {err.max():.3g}"
    # The frame exercises overlaps: many pixels blend several Gaussians.
    assert (want[..., 3] > 0.5).mean() > 0.2


@pytest.mark.parametrize("slip", ["cross_sign", "no_compensation", "back_to_front", "sh_band"])
def test_bruteforce_comparison_sees_a_slip(gs, oracle, slip):
    """Each planted mistake moves the float64 frame far outside 2e-6."""
    worst = 0.0
    for seed in range(4):
        scene, cam = tiny_scene(gs, seed)
        frame = gs.frame_for(cam, 32, 32)
        objects = gs.object_frames_for(cam)[None]
        ref = oracle.run_frame(scene, frame, objects, capacity=1 << 16, threads=1)
        want, _ = brute_force_frame(scene, frame, objects, slip=slip)
        worst = max(worst, float(np.abs(ref.get("image_f32") - want).max()))
    assert worst > 1e-3


# ------------------------------------------------------------ depth ----

INF = float("inf")


def test_scalar_finalisation_known_answers(oracle):
    """test/ScalarOutput.test.ts:61-96 (raw scalar encoding, invalid value
    Infinity, coverage threshold 0.5)."""
    fin = lambda w, t, enc=1: oracle.finalize_depth(w, t, alpha_threshold=0.5, invalid=INF,
                                                    encoding=enc)
    # empty pixels, insufficient coverage, non-finite accumulation
    for w, t in [(0, 1), (0.6, 0.8), (INF, 0.1), (float("nan"), 0.1), (0, 0.1)]:
        assert fin(w, t) == INF
    # the alpha-weighted scalar: 8.8 / 0.8 and 880 / 0.8
    assert fin(8.8, 0.2) == pytest.approx(11, abs=5e-6)
    assert fin(880, 0.2) == pytest.approx(1100, abs=5e-4)
    # non-finite custom encodings never divide by zero into the target
    assert fin(8.8, 0.2, 2) == INF and fin(0, 1, 2) == INF
    # coverage is not conflated with background alpha
    assert fin(0.1, 0.9, 3) == INF


def test_default_depth_encoding_closed_form(oracle):
    """clamp((near + z) far / ((far - near) z), 0, 1) with z = -depth: 0 at
    the near plane, 1 at the far plane, monotone between, invalid below the
    coverage threshold (GaussianPass.ts:243 default 0.95)."""
    near, far = 0.5, 50.0
    enc = lambda depth, t=0.0: oracle.finalize_depth(depth * (1 - t), t, near=near, far=far)
    assert enc(near) == 0.0
    assert enc(far) == pytest.approx(1.0, abs=1e-6)
    ds = np.geomspace(near, far, 50)
    vals = [enc(float(d)) for d in ds]
    assert all(b >= a for a, b in zip(vals, vals[1:]))
    for d in (1.0, 3.0, 20.0):
        exact = (near - d) * far / ((far - near) * -d)
        assert enc(d) == pytest.approx(exact, rel=1e-6)
    assert enc(0.2) == 0.0                       # in front of near: clamped
    assert oracle.finalize_depth(3.0 * 0.9, 0.1, near=near, far=far) == 1.0  # coverage 0.9 < 0.95
    assert oracle.finalize_depth(3.0 * 0.9, 0.1, near=near, far=far, alpha_threshold=0.5) \
        == pytest.approx((near - 3.0) * far / ((far - near) * -3.0), rel=1e-6)


# ------------------------------------------------------- raster counters ----

def counted_pairs(ref, width, height, chunk, thr_t=1e-4, edge=16, slip=None):
    """The reference's raster counters restated over the oracle's sorted
    records and projected values (TileRasterizer.ts:547-709): per in-image
    pixel and per chunk task (the whole tile when it is not chunked), every
    record up to and including the one that drops T below the threshold is
    checked; those with power in range and alpha >= 1/255 are blended. Every
    chunk task of a chunked tile runs from T = 1 and is counted in full.
    Vectorised over a tile's pixels in binary32 (numpy's exp and log)."""
    f = np.float32
    off = ref.get("tile_offsets").astype(np.int64)
    ids = ref.get("sorted_records")[:, 1]
    pm, pc = ref.get("projected_mean"), ref.get("projected_conic")
    tiles_x = -(-width // edge)
    checked = blended = heavy = tasks = 0
    for t in range(len(off) - 1):
        b, n = off[t], off[t + 1] - off[t]
        if n == 0:
            continue
        x0, y0 = (t % tiles_x) * edge, (t // tiles_x) * edge
        xs, ys = np.meshgrid(np.arange(edge) + x0, np.arange(edge) + y0)
        inside = (xs < width) & (ys < height)
        if slip == "all_pixels":
            inside[:] = True
        px, py = (xs[inside] + f(0.5)).astype(f), (ys[inside] + f(0.5)).astype(f)
        chunked = chunk and n > chunk
        spans = [(c, min(n, c + chunk)) for c in range(0, n, chunk)] if chunked else [(0, n)]
        tasks += len(spans) if chunked else 0
        tile_checked = 0
        t_comp = np.ones(px.shape, f)
        for lo, hi in spans:
            g = ids[b + lo: b + hi]
            m, k = pm[g], pc[g]
            dx = px[None, :] - m[:, 0:1]
            dy = py[None, :] - m[:, 1:2]
            power = (k[:, 0:1] * (dx * dx) + k[:, 1:2] * f(2) * dx * dy
                     + k[:, 2:3] * (dy * dy)) * f(-0.5)
            cut = np.log(m[:, 3:4] * f(255))
            alpha = np.clip(m[:, 3:4] * np.exp(power), f(0), f(0.99)).astype(f)
            ok = ~((power > 0) | (power < -cut)) & (alpha >= f(1 / 255))
            trans = np.cumprod(np.where(ok, f(1) - alpha, f(1)).astype(f), axis=0, dtype=f)
            below = trans < f(thr_t)
            stop = np.where(below.any(0), below.argmax(0) + 1, hi - lo)
            if slip == "exclude_stop":
                stop = np.where(below.any(0), stop - 1, stop)
            if slip == "stop_after_composite" and chunked:
                stop = np.where(t_comp < f(thr_t), 0, stop)
            t_comp = (t_comp * trans[-1]).astype(f)
            rows = np.arange(hi - lo)[:, None]
            tile_checked += int(stop.sum())
            blended += int((ok & (rows < stop[None, :])).sum())
        checked += tile_checked
        if chunked or n > 1024:
            heavy += tile_checked
    return {"checked": checked, "blended": blended, "checked_heavy": heavy, "chunk_tasks": tasks}


@pytest.mark.parametrize("chunk", [256, 0, 8192])
def test_raster_counters_restated(gs, oracle, chunk):
    """The oracle's checked / blended / heavy counters equal the restatement
    on a dense 72 x 40 frame (ragged edge tiles; with 256-record chunks, most
    tiles are chunked and many pixels saturate inside an early chunk)."""
    scene, cam = tiny_scene(gs, 7, 4000)
    w, h = 72, 40
    frame = gs.frame_for(cam, w, h)
    objects = gs.object_frames_for(cam)[None]
    ref = oracle.run_frame(scene, frame, objects, capacity=1 << 22, threads=1, chunk_size=chunk)
    assert ref.counts.overflow == 0
    c = ref.counts
    got = {"checked": c.checked, "blended": c.blended, "checked_heavy": c.checked_heavy,
           "chunk_tasks": c.chunk_tasks}
    want = counted_pairs(ref, w, h, chunk)
    print(chunk, got)
    assert got == want
    if chunk == 256:
        assert c.chunk_tasks > 20 and 0 < c.checked_heavy
        # The restatement tells the plausible miscounts apart.
        for slip in ("all_pixels", "exclude_stop", "stop_after_composite"):
            assert counted_pairs(ref, w, h, chunk, slip=slip)["checked"] != c.checked, slip
