"""Camera matrices with Three.js conventions, computed in float64.

The reference takes ``camera.projectionMatrix`` and ``camera.matrixWorld``
from Three.js (FrameUniforms.ts:41-58, ObjectFrameState.ts:97-120). This
module reproduces PerspectiveCamera.updateProjectionMatrix and
Matrix4.lookAt / compose so callers without Three.js get the same uniforms.
Matrices are 4x4 numpy float64 arrays; ``.T.ravel()`` gives Three.js's
column-major ``elements``.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np


def perspective_matrix(fov_deg: float, aspect: float, near: float, far: float,
                       zoom: float = 1.0) -> np.ndarray:
    """PerspectiveCamera.updateProjectionMatrix -> Matrix4.makePerspective."""
    top = near * math.tan(math.radians(0.5 * fov_deg)) / zoom
    height = 2.0 * top
    width = aspect * height
    left = -0.5 * width
    right = left + width
    bottom = top - height
    x = 2.0 * near / (right - left)
    y = 2.0 * near / (top - bottom)
    a = (right + left) / (right - left)
    b = (top + bottom) / (top - bottom)
    c = -(far + near) / (far - near)
    d = -2.0 * far * near / (far - near)
    return np.array([[x, 0, a, 0], [0, y, b, 0], [0, 0, c, d], [0, 0, -1, 0]], dtype=np.float64)


def look_at_world(position, target, up=(0.0, 1.0, 0.0)) -> np.ndarray:
    """matrixWorld of a camera at `position` looking at `target` (Matrix4.lookAt)."""
    eye = np.asarray(position, dtype=np.float64)
    z = eye - np.asarray(target, dtype=np.float64)
    if np.dot(z, z) == 0.0:
        z = np.array([0.0, 0.0, 1.0])
    z = z / np.linalg.norm(z)
    upv = np.asarray(up, dtype=np.float64)
    x = np.cross(upv, z)
    if np.dot(x, x) == 0.0:
        z = z + np.array([1e-4, 0.0, 0.0]) if abs(upv[2]) == 1.0 else z + np.array([0.0, 0.0, 1e-4])
        z = z / np.linalg.norm(z)
        x = np.cross(upv, z)
    x = x / np.linalg.norm(x)
    y = np.cross(z, x)
    m = np.eye(4)
    m[:3, 0] = x
    m[:3, 1] = y
    m[:3, 2] = z
    m[:3, 3] = eye
    return m


@dataclass
class PerspectiveCamera:
    """The subset of THREE.PerspectiveCamera the pass reads."""

    fov: float = 60.0
    aspect: float = 16.0 / 9.0
    near: float = 0.1
    far: float = 1000.0
    position: tuple = (0.0, 0.0, 15.0)
    target: tuple = (0.0, 0.0, 0.0)
    up: tuple = (0.0, 1.0, 0.0)
    zoom: float = 1.0
    matrix_world: np.ndarray = field(default=None, repr=False)

    def __post_init__(self):
        if self.matrix_world is None:
            self.look_at(self.target)

    def look_at(self, target) -> None:
        self.target = tuple(target)
        self.matrix_world = look_at_world(self.position, self.target, self.up)

    @property
    def projection_matrix(self) -> np.ndarray:
        return perspective_matrix(self.fov, self.aspect, self.near, self.far, self.zoom)

    @property
    def matrix_world_inverse(self) -> np.ndarray:
        return np.linalg.inv(self.matrix_world)


def elements(m: np.ndarray) -> np.ndarray:
    """Three.js Matrix4.elements order (column-major) as float64."""
    return np.ascontiguousarray(np.asarray(m, dtype=np.float64).T.ravel())


def orbit_cameras(n: int, seed: int, radius=(12.0, 25.0), resolutions=None, fov=60.0,
                  near=0.1, far=1000.0):
    """Seeded orbit views around the origin: radius ~ U(radius), random
    azimuth, elevation ~ U(-40, 40) degrees, resolution drawn uniformly
    from `resolutions` ([(w, h), ...]). Returns [(camera, (w, h))]."""
    rng = np.random.default_rng(seed)
    resolutions = resolutions or [(1920, 1080)]
    out = []
    for _ in range(n):
        r = rng.uniform(*radius)
        az = rng.uniform(0.0, 2.0 * math.pi)
        el = math.radians(rng.uniform(-40.0, 40.0))
        pos = (r * math.cos(el) * math.sin(az), r * math.sin(el), r * math.cos(el) * math.cos(az))
        w, h = resolutions[int(rng.integers(len(resolutions)))]
        cam = PerspectiveCamera(fov=fov, aspect=w / h, near=near, far=far, position=pos)
        out.append((cam, (w, h)))
    return out
