"""Enumerations and small host-side math helpers shared by the host package.

The integer values of the enums are the reference's own values, because they travel through the
C-ABI and are compared inside the kernels:

* ``GeoType``      -- newton/_src/geometry/types.py:38-78
* ``ShapeFlags``   -- newton/_src/geometry/flags.py:16-33
* ``BodyFlags``    -- newton/_src/sim/enums.py:8-22
* ``JointType``    -- newton/_src/sim/enums.py:29-56

Transforms are ``(px, py, pz, qx, qy, qz, qw)`` and spatial vectors are ``(linear, angular)``,
the reference's layouts (newton/_src/sim/state.py; Warp ``transform`` / ``spatial_vector``).
"""

from __future__ import annotations

import enum

import numpy as np


class GeoType(enum.IntEnum):
    NONE = 0
    PLANE = 1
    HFIELD = 2
    SPHERE = 3
    CAPSULE = 4
    ELLIPSOID = 5
    CYLINDER = 6
    BOX = 7
    MESH = 8
    CONE = 9
    CONVEX_MESH = 10


class ShapeFlags(enum.IntFlag):
    VISIBLE = 1 << 0
    COLLIDE_SHAPES = 1 << 1
    COLLIDE_PARTICLES = 1 << 2
    SITE = 1 << 3
    HYDROELASTIC = 1 << 4


class BodyFlags(enum.IntFlag):
    DYNAMIC = 1 << 0
    KINEMATIC = 1 << 1


class JointType(enum.IntEnum):
    PRISMATIC = 0
    REVOLUTE = 1
    BALL = 2
    FIXED = 3
    FREE = 4
    DISTANCE = 5
    D6 = 6
    CABLE = 7


class Axis(enum.IntEnum):
    X = 0
    Y = 1
    Z = 2

    @staticmethod
    def from_any(v) -> "Axis":
        if isinstance(v, Axis):
            return v
        if isinstance(v, str):
            return Axis("XYZ".index(v.upper()))
        return Axis(int(v))


# ---------------------------------------------------------------------------------------------
# host math (float64 internally, results stored as float32 by the builder)
# ---------------------------------------------------------------------------------------------


def quat_identity() -> np.ndarray:
    return np.array([0.0, 0.0, 0.0, 1.0])


def quat_mul(a, b) -> np.ndarray:
    ax, ay, az, aw = a
    bx, by, bz, bw = b
    return np.array(
        [
            aw * bx + bw * ax + ay * bz - by * az,
            aw * by + bw * ay + az * bx - bz * ax,
            aw * bz + bw * az + ax * by - bx * ay,
            aw * bw - ax * bx - ay * by - az * bz,
        ]
    )


def quat_rotate(q, v) -> np.ndarray:
    """v (2 w^2 - 1) + u (2 u.v) + (u x v) 2 w for q = (u, w), in float64 scalar arithmetic (3-vectors are too
    small for numpy's per-call cost)."""
    x, y, z, w = (float(c) for c in q)
    vx, vy, vz = (float(c) for c in v)
    a = 2.0 * w * w - 1.0
    d = 2.0 * (x * vx + y * vy + z * vz)
    w2 = 2.0 * w
    return np.array([vx * a + x * d + (y * vz - z * vy) * w2,
                     vy * a + y * d + (z * vx - x * vz) * w2,
                     vz * a + z * d + (x * vy - y * vx) * w2])


def quat_from_axis_angle(axis, angle) -> np.ndarray:
    axis = np.asarray(axis, dtype=np.float64)
    axis = axis / np.linalg.norm(axis)
    s = np.sin(0.5 * angle)
    return np.array([axis[0] * s, axis[1] * s, axis[2] * s, np.cos(0.5 * angle)])


def quat_between_vectors(a, b) -> np.ndarray:
    """Shortest-arc rotation taking unit vector ``a`` onto unit vector ``b``."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    a = a / np.linalg.norm(a)
    b = b / np.linalg.norm(b)
    d = float(np.dot(a, b))
    if d > 1.0 - 1e-12:
        return quat_identity()
    if d < -1.0 + 1e-12:
        axis = np.cross(a, [1.0, 0.0, 0.0])
        if np.linalg.norm(axis) < 1e-6:
            axis = np.cross(a, [0.0, 1.0, 0.0])
        return quat_from_axis_angle(axis, np.pi)
    c = np.cross(a, b)
    q = np.array([c[0], c[1], c[2], 1.0 + d])
    return q / np.linalg.norm(q)


def quat_to_matrix(q) -> np.ndarray:
    """Columns: the rotated unit vectors."""
    x, y, z, w = (float(c) for c in q)
    a = 2.0 * w * w - 1.0
    w2 = 2.0 * w
    return np.array([[a + 2.0 * x * x, 2.0 * x * y - z * w2, 2.0 * x * z + y * w2],
                     [2.0 * y * x + z * w2, a + 2.0 * y * y, 2.0 * y * z - x * w2],
                     [2.0 * z * x - y * w2, 2.0 * z * y + x * w2, a + 2.0 * z * z]])


def transform_identity() -> np.ndarray:
    return np.array([0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 1.0])


def transform(p=(0.0, 0.0, 0.0), q=(0.0, 0.0, 0.0, 1.0)) -> np.ndarray:
    return np.concatenate([np.asarray(p, dtype=np.float64), np.asarray(q, dtype=np.float64)])


def as_transform(x) -> np.ndarray:
    if x is None:
        return transform_identity()
    if isinstance(x, (tuple, list)) and len(x) == 2:
        return transform(x[0], x[1])
    arr = np.asarray(x, dtype=np.float64).reshape(-1)
    if arr.shape[0] != 7:
        raise ValueError(f"transform must have 7 entries (p, q), got {arr.shape[0]}")
    return arr.copy()


def transform_multiply(a, b) -> np.ndarray:
    a = as_transform(a)
    b = as_transform(b)
    return transform(quat_rotate(a[3:], b[:3]) + a[:3], quat_mul(a[3:], b[3:]))


def transform_point(t, p) -> np.ndarray:
    t = as_transform(t)
    return quat_rotate(t[3:], p) + t[:3]


def transform_inertia(mass: float, inertia: np.ndarray, offset: np.ndarray, q) -> np.ndarray:
    """Rotate an inertia tensor by ``q`` and shift it by ``offset`` (parallel axis theorem)."""
    R = quat_to_matrix(q)
    I = R @ np.asarray(inertia, dtype=np.float64) @ R.T
    p = np.asarray(offset, dtype=np.float64)
    return I + mass * (np.dot(p, p) * np.eye(3) - np.outer(p, p))
