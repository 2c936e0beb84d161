"""Host matrix helpers with glm's left-handed, depth-zero-to-one conventions.

The reference configures glm with GLM_FORCE_LEFT_HANDED and
GLM_FORCE_DEPTH_ZERO_TO_ONE (VFS/pch.h:33-35). These restate glm 0.9.9.9's
lookAtLH, orthoLH_ZO and perspectiveLH_ZO from their published formulas.
Matrices are numpy (4, 4) arrays in math (row, col) order; `flat()` gives
glm's column-major 16 floats.
"""
from __future__ import annotations

import numpy as np


def normalize(v) -> np.ndarray:
    v = np.asarray(v, dtype=np.float64)
    return v / np.linalg.norm(v)


def look_at_lh(eye, center, up) -> np.ndarray:
    eye = np.asarray(eye, dtype=np.float64)
    f = normalize(np.asarray(center, dtype=np.float64) - eye)
    s = normalize(np.cross(up, f))
    u = np.cross(f, s)
    m = np.eye(4)
    m[0, :3] = s
    m[1, :3] = u
    m[2, :3] = f
    m[0, 3] = -np.dot(s, eye)
    m[1, 3] = -np.dot(u, eye)
    m[2, 3] = -np.dot(f, eye)
    return m


def ortho_lh_zo(left, right, bottom, top, z_near, z_far) -> np.ndarray:
    m = np.eye(4)
    m[0, 0] = 2.0 / (right - left)
    m[1, 1] = 2.0 / (top - bottom)
    m[2, 2] = 1.0 / (z_far - z_near)
    m[0, 3] = -(right + left) / (right - left)
    m[1, 3] = -(top + bottom) / (top - bottom)
    m[2, 3] = -z_near / (z_far - z_near)
    return m


def perspective_lh_zo(fovy, aspect, z_near, z_far) -> np.ndarray:
    t = np.tan(fovy / 2.0)
    m = np.zeros((4, 4))
    m[0, 0] = 1.0 / (aspect * t)
    m[1, 1] = 1.0 / t
    m[2, 2] = z_far / (z_far - z_near)
    m[3, 2] = 1.0
    m[2, 3] = -(z_far * z_near) / (z_far - z_near)
    return m


def flat(m: np.ndarray) -> np.ndarray:
    """glm column-major storage: flat[col * 4 + row] = m[row, col]."""
    return np.ascontiguousarray(np.asarray(m, dtype=np.float64).T.reshape(16)).astype(np.float32)
