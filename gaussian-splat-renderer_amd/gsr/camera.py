"""Viewer-camera conventions of the reference (app/camera.cpp:29-52), used to build the
SplatCamera of a frame on the host: glm::lookAt (right-handed, -z forward, +y up),
glm::perspective (zero-to-one depth) with the Vulkan y flip, yaw/pitch orbit. Matrices are
column-major float32, flattened m[col * 4 + row] (glm's memory order).
"""
from __future__ import annotations

import math

import numpy as np

WORLD_UP = np.array([0.0, 1.0, 0.0])


def _normalize(v):
    return v / np.sqrt(np.dot(v, v))


def look_at(eye, target, up=WORLD_UP) -> np.ndarray:
    """glm::lookAtRH as a flattened column-major 4x4."""
    eye, target, up = (np.asarray(x, np.float64) for x in (eye, target, up))
    f = _normalize(target - eye)
    s = _normalize(np.cross(f, up))
    u = np.cross(s, f)
    m = np.eye(4)  # m[col][row]
    m[0][0], m[1][0], m[2][0] = s
    m[0][1], m[1][1], m[2][1] = u
    m[0][2], m[1][2], m[2][2] = -f
    m[3][0] = -np.dot(s, eye)
    m[3][1] = -np.dot(u, eye)
    m[3][2] = np.dot(f, eye)
    return m.astype(np.float32).reshape(16)


def perspective(fov_y_deg: float, aspect: float, near: float = 0.01, far: float = 1000.0) -> np.ndarray:
    """glm::perspectiveRH_ZO with proj[1][1] negated (Camera::projection, camera.cpp:38-45)."""
    t = math.tan(math.radians(fov_y_deg) / 2.0)
    m = np.zeros((4, 4))
    m[0][0] = 1.0 / (aspect * t)
    m[1][1] = -1.0 / t
    m[2][2] = far / (near - far)
    m[2][3] = -1.0
    m[3][2] = -(far * near) / (far - near)
    return m.astype(np.float32).reshape(16)


IDENTITY = np.eye(4, dtype=np.float32).reshape(16)


def forward(yaw_deg: float, pitch_deg: float) -> np.ndarray:
    """Camera::forward (camera.cpp:29-33)."""
    y, p = math.radians(yaw_deg), math.radians(pitch_deg)
    return np.array([math.cos(p) * math.sin(y), math.sin(p), -math.cos(p) * math.cos(y)])


def orbit_view(radius: float, elevation_deg: float, yaw_deg: float, target=(0.0, 0.0, 0.0)) -> np.ndarray:
    """The view of an orbit camera at `radius` around `target`, `elevation_deg` above the
    horizontal plane (looking down at it: pitch = -elevation), after yaw_deg of orbit
    (Camera::orbit, camera.cpp:47-52: position = target - forward * distance)."""
    f = forward(yaw_deg, -elevation_deg)
    pos = np.asarray(target, np.float64) - f * radius
    return look_at(pos, pos + f)


def orbit_frames(frames: int, radius: float, elevation_deg: float):
    """The benchmark orbit: frame i at yaw 360 i / frames degrees (application.cpp:1057)."""
    for i in range(frames):
        yield orbit_view(radius, elevation_deg, 360.0 * i / frames)


def warmup_frames(frames: int) -> int:
    """Frames the reference benchmark excludes: frames / 10 + 5 (application.cpp:1042)."""
    return frames // 10 + 5
