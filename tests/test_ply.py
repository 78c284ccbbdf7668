"""The packed-buffer producer for real scenes: canonical 3DGS PLY files
(src/streaming-backend-impl/CanonicalGaussianPlyLoader.ts:84-350) into the
store's slot buffers with hole rows (rust/gaussian-backend/src/engine.rs
:1801-1949, :1917-1923).

The committed fixture tests/golden/scene_64.ply (made by
tests/golden/make_ply_fixture.py) is read back independently with a numpy
structured dtype; the encoded buffers are held to the reference's
round-trip tolerances (rust/gaussian-backend/tests/backend.rs:1141-1156:
scale ratio within 1 %, opacity within 0.001, |dot| > 0.99999 for
quaternions; test/GaussianSh.test.ts:19-40 for rgb8e8: 2^ceil(log2 max)/127).
"""
from __future__ import annotations

import math
from pathlib import Path

import numpy as np
import pytest

import make_ply_fixture as F  # tests/golden, on the path through conftest
from test_formats import decode_rotation, decode_scale_opacity, decode_sh

FIXTURE = Path(__file__).resolve().parent / "golden" / "scene_64.ply"


def raw_vertices() -> np.ndarray:
    data = FIXTURE.read_bytes()
    end = data.index(b"end_header\n") + len(b"end_header\n")
    return np.frombuffer(data[end:], dtype=[(k, "<f4") for k in F.NAMES])


def test_fixture_reads_as_the_loader_computes(gs):
    v = raw_vertices()
    g = gs.read_ply(FIXTURE.read_bytes())
    assert g["sh_degree"] == 3 and g["means"].shape == (64, 3)
    np.testing.assert_array_equal(g["means"], np.stack([v["x"], v["y"], v["z"]], 1))
    s = np.stack([v[f"scale_{i}"] for i in range(3)], 1).astype(np.float64)
    np.testing.assert_allclose(g["scales"], np.maximum(np.exp(s), 1e-6).astype(np.float32), rtol=1e-7)
    o = v["opacity"].astype(np.float64)
    np.testing.assert_allclose(g["opacity"], (1 / (1 + np.exp(-o))).astype(np.float32), rtol=1e-7)
    q = np.stack([v["rot_1"], v["rot_2"], v["rot_3"], v["rot_0"]], 1).astype(np.float64)
    np.testing.assert_allclose(g["quats"], q / np.linalg.norm(q, axis=1, keepdims=True), rtol=1e-6,
                               atol=1e-7)
    dc = np.stack([v[f"f_dc_{i}"] for i in range(3)], 1)
    np.testing.assert_array_equal(g["sh"][:, 0], dc)
    rest = np.stack([v[f"f_rest_{i}"] for i in range(45)], 1).reshape(64, 3, 15)  # channel-major
    np.testing.assert_array_equal(g["sh"][:, 1:], rest.transpose(0, 2, 1))


def test_ascii_and_big_endian_read_the_same(gs, tmp_path):
    v = raw_vertices()
    want = gs.read_ply(FIXTURE.read_bytes())
    ascii_rows = "\n".join(" ".join(repr(float(x)) for x in row) for row in v.tolist())
    for data in (F.header(64, "ascii") + ascii_rows.encode() + b"\n",
                 F.header(64, "binary_big_endian") + v.astype([(k, ">f4") for k in F.NAMES]).tobytes(),
                 F.header(64, "binary_little_endian", "double")
                 + v.astype([(k, "<f8") for k in F.NAMES]).tobytes()):
        got = gs.read_ply(data)
        for k in ("means", "scales", "opacity", "quats", "sh"):
            np.testing.assert_array_equal(got[k], want[k], err_msg=k)


@pytest.mark.parametrize("mutate,message", [
    (lambda h, b: (h.replace(b"property float opacity\n", b""), b), "missing property opacity"),
    (lambda h, b: (h.replace(b"f_rest_3\n", b"f_rest_99\n"), b), "contiguous from f_rest_0"),
    (lambda h, b: (h.replace(b"property float f_rest_44\n", b""), b[:-4]), "divisible by three"),
    (lambda h, b: (h.replace(b"end_header", b"end_hdr"), b), "end_header is missing"),
    (lambda h, b: (h.replace(b"ply\n", b"plx\n", 1), b), "Invalid PLY signature"),
    (lambda h, b: (h, b[:100]), "ends before the vertex data is complete"),
    (lambda h, b: (h.replace(b"binary_little_endian", b"binary_middle_endian"), b),
     "Unsupported PLY format"),
    (lambda h, b: (h.replace(b"element vertex 64\n", b"element face 1\nelement vertex 64\n"), b),
     "vertex element must be first"),
    (lambda h, b: (h.replace(b"property float nx\n", b"property list uchar int nx\n"), b),
     "List properties are not supported"),
])
def test_invalid_files_raise_the_loaders_errors(gs, mutate, message):
    data = FIXTURE.read_bytes()
    end = data.index(b"end_header\n") + len(b"end_header\n")
    h, b = mutate(data[:end], data[end:])
    with pytest.raises(ValueError, match=message):
        gs.read_ply(h + b)


def test_sh_degree_follows_the_rest_count(gs):
    """Files with 0, 3 and 8 higher-band coefficients per channel read as SH
    degrees 0, 1 and 2, the f_rest block channel-major."""
    v = raw_vertices()
    base = [k for k in F.NAMES if not k.startswith("f_rest_")]
    for deg in (0, 1, 2):
        per = (deg + 1) ** 2 - 1
        src = [f"f_rest_{c * 15 + j}" for c in range(3) for j in range(per)]  # kept, in order
        names = base + [f"f_rest_{i}" for i in range(len(src))]
        sub = np.zeros(len(v), dtype=[(k, "<f4") for k in names])
        for k in base:
            sub[k] = v[k]
        for i, k in enumerate(src):
            sub[f"f_rest_{i}"] = v[k]
        lines = ["ply", "format binary_little_endian 1.0", f"element vertex {len(v)}"]
        lines += [f"property float {k}" for k in names] + ["end_header"]
        g = gs.read_ply(("\n".join(lines) + "\n").encode() + sub.tobytes())
        assert g["sh_degree"] == deg and g["sh"].shape == (len(v), per + 1, 3)
        for c in range(3):
            for j in range(per):
                np.testing.assert_array_equal(g["sh"][:, 1 + j, c], v[f"f_rest_{c * 15 + j}"])


def test_pack_scene_slots_and_holes(gs, oracle):
    g = gs.read_ply(FIXTURE.read_bytes())
    s = gs.pack_scene(g, capacity=80, object_id=2)
    assert s.count == 80 and s.sh_degree == 3
    np.testing.assert_array_equal(s.means[:64, :3], g["means"])
    assert (s.means[:64, 3] == 2.0).all() and (s.means[64:, 3] == -1.0).all()
    assert not s.means[64:, :3].any() and not s.scales_opacity[64:].any()
    assert not s.rotations[64:].any() and not s.sh[64:].any()
    for i in range(64):
        d = decode_scale_opacity(oracle, *s.scales_opacity[i])
        for axis in range(3):
            assert abs(d[axis] / g["scales"][i, axis] - 1.0) < 0.01      # backend.rs:1145
        assert abs(d[3] - g["opacity"][i]) < 0.001                       # backend.rs:1149
        q = decode_rotation(oracle, s.rotations[i])
        assert abs(float(np.dot(q, g["quats"][i]))) > 0.99999            # backend.rs:1156
        for k in range(16):
            c = g["sh"][i, k]
            tol = 2.0 ** math.ceil(math.log2(max(np.abs(c).max(), 1e-30))) / 127  # GaussianSh.test.ts
            np.testing.assert_allclose(decode_sh(oracle, s.sh[i, k]), c, atol=tol * 0.5 + 1e-12)


def test_ply_scene_frame_on_the_oracle_skips_holes(gs, oracle):
    """The holes are skipped: the frame of the packed scene equals the frame
    of the same Gaussians without hole rows."""
    g = gs.read_ply(FIXTURE.read_bytes())
    full = gs.pack_scene(g, capacity=96)
    tight = gs.pack_scene(g)
    cam = gs.PerspectiveCamera(fov=60.0, aspect=1.0, near=0.1, far=100.0, position=(0.0, 0.0, 4.0))
    frame = gs.frame_for(cam, 64, 64)
    objects = gs.object_frames_for(cam)[None]
    a = oracle.run_frame(full, frame, objects, capacity=1 << 16, threads=1)
    b = oracle.run_frame(tight, frame, objects, capacity=1 << 16, threads=1)
    assert a.counts.visible == b.counts.visible > 32
    np.testing.assert_array_equal(a.get("image_f16"), b.get("image_f16"))
