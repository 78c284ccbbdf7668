"""CPU checks of the learned cull's checkers.

The MLP reference (oracle/gs_cull_ref.py) is pinned by finite differences
of its own loss (exact float64 path) and by known bfloat16 roundings; the
oracle's training data (features and per-Gaussian max alpha * T) is pinned
by a single-Gaussian closed form and by invariants against the oracle's own
projected outputs.
"""
from __future__ import annotations

import importlib
import math

import numpy as np
import pytest

import gs_cull_ref as R


def test_bf16_round_to_nearest_even():
    one = 1.0
    cases = {
        one: 1.0,
        1.0 + 2.0**-8: 1.0,                       # tie -> even (down)
        1.0 + 3 * 2.0**-8: 1.0 + 2 * 2.0**-7,     # tie -> even (up)
        1.0 + 2.0**-8 + 2.0**-20: 1.0 + 2.0**-7,  # above the tie
        -2.5: -2.5,
        65504.0: 65536.0,
        3.0e-39: float(np.array([0x00210000], np.uint32).view(np.float32)[0]),  # 0x0020aac8 rounds up
    }
    for v, want in cases.items():
        assert R.bf16(np.array([v], np.float32))[0] == pytest.approx(want, rel=0, abs=0), v


def test_loss_gradient_matches_finite_differences():
    rng = np.random.default_rng(3)
    p = R.random_params(rng).astype(np.float64)
    feats = rng.normal(0, 1, (96, 16)).astype(np.float32)
    mu = np.zeros(16, np.float32)
    inv = np.ones(16, np.float32)
    targets = rng.uniform(-16, 0, 96)
    log_thr = math.log(1e-4)
    loss, g = R.loss_and_grad(p, mu, inv, feats, targets, 4.0, log_thr, rounding=False)
    idx = np.concatenate([rng.choice(np.arange(R.OFF_W1, R.OFF_B1), 12),
                          rng.choice(np.arange(R.OFF_B1, R.OFF_W2), 6),
                          rng.choice(np.arange(R.OFF_W2, R.OFF_B2), 12),
                          rng.choice(np.arange(R.OFF_B2, R.OFF_W3), 6),
                          rng.choice(np.arange(R.OFF_W3, R.OFF_B3), 6), [R.OFF_B3]])
    h = 1e-6
    for k in idx:
        q = p.copy()
        q[k] += h
        lp, _ = R.loss_and_grad(q, mu, inv, feats, targets, 4.0, log_thr, rounding=False)
        q[k] -= 2 * h
        lm, _ = R.loss_and_grad(q, mu, inv, feats, targets, 4.0, log_thr, rounding=False)
        fd = (lp - lm) / (2 * h)
        assert fd == pytest.approx(g[k], rel=1e-4, abs=1e-7), k


def test_rounded_gradient_stays_close_to_exact():
    """The bf16 rounding points perturb the gradient by bf16 precision only."""
    rng = np.random.default_rng(4)
    p = R.random_params(rng)
    feats = rng.normal(0, 1, (256, 16)).astype(np.float32)
    targets = rng.uniform(-16, 0, 256)
    args = (p, np.zeros(16, np.float32), np.ones(16, np.float32), feats, targets, 4.0,
            math.log(1e-4))
    _, g_exact = R.loss_and_grad(*args, rounding=False)
    _, g_bf16 = R.loss_and_grad(*args, rounding=True)
    assert np.linalg.norm(g_bf16 - g_exact) <= 0.05 * np.linalg.norm(g_exact)


def test_oracle_record_single_gaussian_closed_form(gs, oracle):
    """With one Gaussian, T = 1 at every pixel: its max alpha * T is the
    largest clamp(o e^power, 0, 0.99) over pixel centres (float64 here)."""
    scene = gs.workload_scene("small", count=1)
    cam = gs.PerspectiveCamera(fov=60.0, aspect=64 / 48, position=(0.0, 0.0, 3.0))
    means = scene.means.copy()
    means[0, :3] = [0.05, -0.03, 0.0]
    scene = gs.PackedScene(means, scene.scales_opacity, scene.rotations, scene.sh, 3, None)
    w, h = 64, 48
    f = gs.frame_for(cam, w, h, {})
    r = oracle.run_frame(scene, f, gs.object_frames_for(cam), capacity=1 << 16, chunk_size=0,
                         cull_outputs=True)
    pm, pc = r.get("projected_mean"), r.get("projected_conic")
    rec = r.get("max_contribution").view(np.float32)
    if pm[0, 3] <= 0:
        pytest.skip("Gaussian not visible in this setup")
    cx, cy, o = pm[0, 0], pm[0, 1], pm[0, 3]
    a, b, c = (float(v) for v in pc[0, :3])
    ys, xs = np.mgrid[0:h, 0:w]
    dx, dy = xs + 0.5 - cx, ys + 0.5 - cy
    power = -0.5 * (a * dx * dx + 2 * b * dx * dy + c * dy * dy)
    alpha = np.clip(o * np.exp(power), 0, 0.99)
    alpha[(alpha < 1 / 255) | (power > 0) | (power < -math.log(o * 255))] = 0
    assert rec[0] == pytest.approx(alpha.max(), rel=1e-5)


def test_oracle_features_consistent_with_projection(gs, oracle):
    scene = gs.workload_scene("small", count=20_000)
    wl = gs.WORKLOADS["small"]
    cam = wl.camera()
    f = gs.frame_for(cam, 320, 180, {})
    r = oracle.run_frame(scene, f, gs.object_frames_for(cam), capacity=1 << 20, chunk_size=0,
                         cull_outputs=True)
    pm, pc = r.get("projected_mean"), r.get("projected_conic")
    feats = r.get("cull_features")
    rec = r.get("max_contribution").view(np.float32)
    vis = pm[:, 3] > 0
    assert vis.sum() > 1000
    fv = feats[vis]
    np.testing.assert_array_equal(fv[:, 4], pm[vis, 3])                 # opacity
    np.testing.assert_allclose(fv[:, 5], np.sum(fv[:, 10:13] * fv[:, 7:10], axis=1),
                               rtol=1e-5, atol=1e-5)                     # position . direction
    np.testing.assert_allclose(np.exp(fv[:, 6]), pm[vis, 2], rtol=1e-6)  # log depth
    np.testing.assert_allclose(fv[:, 14] * 320, pm[vis, 0], rtol=1e-6)
    np.testing.assert_allclose(fv[:, 15] * 180, pm[vis, 1], rtol=1e-6)
    np.testing.assert_allclose(np.linalg.norm(fv[:, 7:10], axis=1), 1.0, atol=1e-6)
    np.testing.assert_allclose(fv[:, 10:13], scene.means[:20_000][vis, :3], rtol=0, atol=0)
    radius = np.maximum(pc[vis, 3] % 512, pc[vis, 3] // 512)
    np.testing.assert_allclose(np.exp(fv[:, 0]), radius, rtol=1e-6)
    # A Gaussian's largest contribution is at most its alpha cap, and only
    # visible Gaussians contribute.
    assert np.all(rec[vis] <= np.minimum(pm[vis, 3], 0.99) + 1e-7)
    assert np.all(rec[~vis] == 0)
    assert (rec[vis] > 0).mean() > 0.3


@pytest.mark.parametrize("feature,scale,bias", [(4, 10.0, -12.0), (0, 3.0, -5.0)])
def test_linear_score_params_compute_the_linear_score(gs, feature, scale, bias):
    """The non-learned baseline cull's parameters make the network (the
    numpy reference at the kernels' bf16 rounding points) output
    scale * feature + bias, to bf16 precision of the standardised input."""
    lc = importlib.import_module("3dgs_tile_webgpu_amd.learned_cull")
    rng = np.random.default_rng(8)
    mu = rng.normal(0, 1, 16).astype(np.float32)
    inv = np.exp(rng.normal(0, 0.5, 16)).astype(np.float32)
    feats = (mu + rng.normal(0, 1, (500, 16)) / inv).astype(np.float32)
    p = lc.linear_score_params(feature, scale, bias, mu, inv)
    y = R.forward(p, mu, inv, feats)["y"]
    want = scale * feats[:, feature].astype(np.float64) + bias
    x = np.abs((feats[:, feature] - mu[feature]) * inv[feature])
    # two bf16 roundings (input, output weight), unit roundoff 2^-8 each
    tol = scale / inv[feature] * (x * 2.0**-7 + 1e-6) + 1e-4 * np.abs(want) + 1e-5
    assert (np.abs(y - want) <= tol).all()
