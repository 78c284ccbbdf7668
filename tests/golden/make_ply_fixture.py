"""Writes tests/golden/scene_64.ply: a canonical 3DGS PLY (binary little
endian, SH degree 3) of 64 seeded Gaussians in a unit cluster, with the
property set and order of the canonical exporter (x y z nx ny nz f_dc_0..2
f_rest_0..44 opacity scale_0..2 rot_0..3). Run from the repository root:
python tests/golden/make_ply_fixture.py"""
from __future__ import annotations

from pathlib import Path

import numpy as np

NAMES = (["x", "y", "z", "nx", "ny", "nz", "f_dc_0", "f_dc_1", "f_dc_2"]
         + [f"f_rest_{i}" for i in range(45)] + ["opacity", "scale_0", "scale_1", "scale_2"]
         + [f"rot_{i}" for i in range(4)])


def vertices(n: int = 64, seed: int = 64) -> np.ndarray:
    rng = np.random.default_rng(seed)
    v = np.zeros(n, dtype=[(k, "<f4") for k in NAMES])
    for k, col in zip("xyz", rng.normal(0.0, 0.6, (3, n))):
        v[k] = col
    for i in range(3):
        v[f"f_dc_{i}"] = rng.normal(0.0, 0.5, n)
    for i in range(45):
        v[f"f_rest_{i}"] = rng.normal(0.0, 0.05, n)
    v["opacity"] = rng.normal(0.5, 1.5, n)
    for i in range(3):
        v[f"scale_{i}"] = rng.normal(-3.0, 0.5, n)
    q = rng.normal(size=(n, 4))
    for i in range(4):
        v[f"rot_{i}"] = q[:, i]
    return v


def header(n: int, fmt: str = "binary_little_endian", kind: str = "float") -> bytes:
    lines = ["ply", f"format {fmt} 1.0", "comment seeded test scene", f"element vertex {n}"]
    lines += [f"property {kind} {k}" for k in NAMES]
    lines += ["end_header"]
    return ("\n".join(lines) + "\n").encode()


def main() -> None:
    v = vertices()
    out = Path(__file__).resolve().parent / "scene_64.ply"
    out.write_bytes(header(len(v)) + v.tobytes())
    print(out, out.stat().st_size)


if __name__ == "__main__":
    main()
