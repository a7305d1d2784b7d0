"""CPU oracle pinned against the reference's known-answer tables and closed forms.

The reference holds no golden vectors for this path; its tests pin the colliders with analytic
distance tables (newton/tests/test_collision_primitives.py) and the narrow phase / solver with
invariants (test_narrow_phase.py, test_solver_xpbd.py, test_rigid_contact.py).  The same tables are
restated here as data and run through the oracle's full collide path (AABB -> SAP -> narrow phase ->
write_contact), reading the signed distance back the way the XPBD solver does:
    d = dot(n, X_b * point1 - X_a * point0) - (margin0 + margin1)
"""

from __future__ import annotations

import math

import numpy as np
import pytest

from newton_custom_contact_amd._types import quat_from_axis_angle, transform
from newton_custom_contact_amd.builder import ModelBuilder
from oracle import Oracle

QX = quat_from_axis_angle((0.0, 1.0, 0.0), 0.5 * math.pi)  # local +Z -> world +X
QY = quat_from_axis_angle((1.0, 0.0, 0.0), -0.5 * math.pi)  # local +Z -> world +Y
QI = (0.0, 0.0, 0.0, 1.0)


def _qrot(q, v):
    q = np.asarray(q, np.float64)
    u, w = q[:3], q[3]
    v = np.asarray(v, np.float64)
    return v * (2 * w * w - 1) + u * (2 * np.dot(u, v)) + np.cross(u, v) * (2 * w)


def contact_distances(model, c, q=None):
    """Signed distance of every contact, computed from the stored contact record."""
    q = model.host["body_q"] if q is None else q
    sb = model.host["shape_body"]
    out = []
    for k in range(len(c["shape0"])):
        ba, bb = sb[c["shape0"][k]], sb[c["shape1"][k]]
        pa = c["point0"][k] if ba < 0 else _qrot(q[ba][3:], c["point0"][k]) + q[ba][:3]
        pb = c["point1"][k] if bb < 0 else _qrot(q[bb][3:], c["point1"][k]) + q[bb][:3]
        out.append(float(np.dot(c["normal"][k], pb - pa) - (c["margin0"][k] + c["margin1"][k])))
    return np.asarray(out)


def make_scene(shapes, gap: float = 2.0):
    """shapes: list of (kind, dims, pos, quat, static). Planes are infinite, pos/quat define them."""
    b = ModelBuilder()
    cfg = ModelBuilder.ShapeConfig(gap=gap)
    for kind, dims, pos, quat, static in shapes:
        if kind == "plane":
            b.add_shape_plane(xform=transform(pos, quat), width=0.0, length=0.0, cfg=cfg)
            continue
        body = -1 if static else b.add_body(xform=transform(pos, quat))
        xf = transform(pos, quat) if static else None
        if kind == "sphere":
            b.add_shape_sphere(body, xform=xf, radius=dims[0], cfg=cfg)
        elif kind == "box":
            b.add_shape_box(body, xform=xf, hx=dims[0], hy=dims[1], hz=dims[2], cfg=cfg)
        elif kind == "capsule":
            b.add_shape_capsule(body, xform=xf, radius=dims[0], half_height=dims[1], cfg=cfg)
    return b.finalize("cpu")


def collide(shapes, gap=2.0):
    m = make_scene(shapes, gap)
    o = Oracle(m)
    c = o.collide(m.host["body_q"])
    return m, c, contact_distances(m, c)


# ----------------------------------------------------------------- test_collision_primitives.py:429-504
PLANE_SPHERE = [
    ((0, 0, 1), (0, 0, 0), (0, 0, 2.0), 1.0, 1.0),
    ((0, 0, 1), (0, 0, 0), (0, 0, 1.5), 1.0, 0.5),
    ((0, 0, 1), (0, 0, 0), (0, 0, 1.0), 1.0, 0.0),
    ((0, 0, 1), (0, 0, 0), (0, 0, 0.8), 1.0, -0.2),
    ((0, 0, 1), (0, 0, 0), (0, 0, 0.5), 1.0, -0.5),
    ((0, 0, 1), (0, 0, 0), (0, 0, 0.2), 1.0, -0.8),
    ((1, 0, 0), (1, 0, 0), (2.0, 0, 0), 0.5, 0.5),
    ((1, 0, 0), (1, 0, 0), (1.5, 0, 0), 0.5, 0.0),
    ((1, 0, 0), (1, 0, 0), (1.3, 0, 0), 0.5, -0.2),
]


@pytest.mark.parametrize("n,pp,sp,r,expected", PLANE_SPHERE)
def test_plane_sphere(n, pp, sp, r, expected):
    from newton_custom_contact_amd._types import quat_between_vectors

    q = quat_between_vectors((0, 0, 1), n)
    _m, c, d = collide([("plane", None, pp, q, True), ("sphere", (r,), sp, QI, False)])
    assert len(d) == 1
    assert d[0] == pytest.approx(expected, abs=1e-5)
    np.testing.assert_allclose(c["normal"][0], n, atol=1e-6)


# ----------------------------------------------------------------- test_collision_primitives.py:505-586
SPHERE_SPHERE = [
    (1.0, 3.5, 1.0, 1.5), (1.0, 3.0, 1.0, 1.0), (1.0, 2.5, 1.0, 0.5), (1.0, 2.0, 1.0, 0.0),
    (1.0, 1.8, 1.0, -0.2), (1.0, 1.5, 1.0, -0.5), (1.0, 1.2, 1.0, -0.8),
    (0.5, 2.0, 1.0, 0.5), (0.5, 1.5, 1.0, 0.0), (0.5, 1.2, 1.0, -0.3),
]


@pytest.mark.parametrize("r1,x2,r2,expected", SPHERE_SPHERE)
def test_sphere_sphere(r1, x2, r2, expected):
    _m, c, d = collide([("sphere", (r1,), (0, 0, 0), QI, False), ("sphere", (r2,), (x2, 0, 0), QI, False)])
    assert len(d) == 1
    assert d[0] == pytest.approx(expected, abs=1e-5)
    np.testing.assert_allclose(c["normal"][0], (1, 0, 0), atol=1e-6)


# ----------------------------------------------------------------- test_collision_primitives.py:696-890
CAPSULE_CAPSULE = [((0, 2.0, 0), QX, 1.0), ((0, 1.5, 0), QX, 0.5), ((0, 1.0, 0), QX, 0.0),
                   ((0, 0.9, 0), QX, -0.1), ((0, 0.8, 0), QX, -0.2), ((0, 0.6, 0), QX, -0.4),
                   ((0, 0, 0), QY, -1.0)]


@pytest.mark.parametrize("p2,q2,expected", CAPSULE_CAPSULE)
def test_capsule_capsule(p2, q2, expected):
    _m, _c, d = collide([("capsule", (0.5, 1.0), (0, 0, 0), QX, False), ("capsule", (0.5, 1.0), p2, q2, False)])
    assert 1 <= len(d) <= 2  # parallel axes give two contacts (collision_primitive.py:180-275)
    assert d.min() == pytest.approx(expected, abs=0.01)


# ----------------------------------------------------------------- test_collision_primitives.py:1136-1249
SPHERE_BOX = [
    ((2.5, 0, 0), 0.5, 1.0), ((2.0, 0, 0), 0.5, 0.5), ((1.5, 0, 0), 0.5, 0.0), ((1.4, 0, 0), 0.5, -0.1),
    ((1.3, 0, 0), 0.5, -0.2), ((1.2, 0, 0), 0.5, -0.3), ((0, 0, 2.0), 0.5, 0.5), ((0, 0, 1.5), 0.5, 0.0),
    ((0, 0, 1.3), 0.5, -0.2), ((0, 0, 0.4), 0.3, -0.9),
]


@pytest.mark.parametrize("sp,r,expected", SPHERE_BOX)
def test_sphere_box(sp, r, expected):
    _m, _c, d = collide([("sphere", (r,), sp, QI, False), ("box", (1.0, 1.0, 1.0), (0, 0, 0), QI, True)])
    assert len(d) == 1
    assert d[0] == pytest.approx(expected, abs=0.01)


# ----------------------------------------------------------------- test_collision_primitives.py:1251-1334
PLANE_CAPSULE = [(2.0, 1.5), (1.5, 1.0), (1.0, 0.5), (0.5, 0.0), (0.4, -0.1), (0.3, -0.2), (0.2, -0.3)]


@pytest.mark.parametrize("z,expected", PLANE_CAPSULE)
def test_plane_capsule(z, expected):
    _m, _c, d = collide([("plane", None, (0, 0, 0), QI, True), ("capsule", (0.5, 1.0), (0, 0, z), QX, False)])
    assert len(d) == 2  # both end spheres
    np.testing.assert_allclose(d, expected, atol=1e-5)


# ----------------------------------------------------------------- test_collision_primitives.py:1335-1428
PLANE_BOX = [(2.0, 0, 1.0), (1.0, 4, 0.0), (0.9, 4, -0.1), (0.8, 4, -0.2), (0.7, 4, -0.3), (0.5, 4, -0.5)]


@pytest.mark.parametrize("z,count,expected", PLANE_BOX)
def test_plane_box(z, count, expected):
    # gap 0: only corners at or below the plane count, as in the reference table (margin 0)
    _m, c, d = collide([("plane", None, (0, 0, 0), QI, True), ("box", (1.0, 1.0, 1.0), (0, 0, z), QI, False)], gap=0.0)
    assert len(d) == count
    if count:
        np.testing.assert_allclose(d, expected, atol=0.01)
        assert np.all(c["normal"] @ np.array([0, 0, 1.0]) > 0.99)


# ----------------------------------------------------------------- test_collision_primitives.py:1749-1880
CAPSULE_BOX = [(3.0, 0.5), (2.5, 0.0), (2.4, -0.1), (2.3, -0.2), (2.2, -0.3)]


@pytest.mark.parametrize("z,expected", CAPSULE_BOX)
def test_capsule_box(z, expected):
    _m, _c, d = collide([("capsule", (0.5, 1.0), (0, 0, z), QI, False), ("box", (1.0, 1.0, 1.0), (0, 0, 0), QI, True)])
    assert len(d) >= 1
    assert d.min() == pytest.approx(expected, abs=0.01)


def test_box_box_separated_and_overlapping():
    """test_collision_primitives.py:1529-1612: unit normals pointing from box 1 to box 2."""
    _m, c, d = collide([("box", (1, 1, 1), (0, 0, 0), QI, False), ("box", (1, 1, 1), (3.0, 0, 0), QI, False)], gap=0.0)
    assert len(d) == 0
    _m, c, d = collide([("box", (1, 1, 1), (0, 0, 0), QI, False), ("box", (1, 1, 1), (1.5, 0, 0), QI, False)], gap=0.0)
    assert 1 <= len(d) <= 5
    np.testing.assert_allclose(np.linalg.norm(c["normal"], axis=1), 1.0, atol=1e-5)
    assert np.all(c["normal"] @ np.array([1.0, 0, 0]) > -0.1)
    np.testing.assert_allclose(d, -0.5, atol=0.01)  # SAT depth along x


@pytest.mark.parametrize("overlap", [0.05, 0.1, 0.2])
def test_box_box_face_manifold(overlap):
    """test_narrow_phase.py:635-706: stacked boxes overlapping by `overlap` -> 4-point face manifold at
    depth -overlap, contact points on the two touching faces (surface reconstruction)."""
    m, c, d = collide([("box", (0.5, 0.5, 0.5), (0, 0, 0), QI, False),
                       ("box", (0.3, 0.3, 0.3), (0.05, -0.03, 0.8 - overlap), QI, False)], gap=0.0)
    assert len(d) in (4, 5)
    np.testing.assert_allclose(d, -overlap, atol=2e-3)
    np.testing.assert_allclose(c["normal"], np.tile([0, 0, 1.0], (len(d), 1)), atol=1e-4)
    # point0 on the lower box's top face, point1 on the upper box's bottom face
    np.testing.assert_allclose(c["point0"][:, 2], 0.5, atol=2e-3)
    np.testing.assert_allclose(c["point1"][:, 2], -0.3, atol=2e-3)
    # manifold corners span the upper box's footprint (0.6 x 0.6)
    span = c["point1"][:, :2].max(axis=0) - c["point1"][:, :2].min(axis=0)
    np.testing.assert_allclose(span, 0.6, atol=2e-3)


def test_sort_keys_and_sub_keys():
    """contact_data.py:59-83 key layout; slots 0-3 for analytic pairs (narrow_phase.py:569-593)."""
    _m, c, _d = collide([("plane", None, (0, 0, 0), QI, True), ("box", (1, 1, 1), (0, 0, 0.9), QI, False)], gap=0.0)
    k = c["sort_key"]
    assert np.all(np.diff(k) > 0)
    assert np.all((k >> 43) & 0xFFFFF == c["shape0"])
    assert np.all((k >> 23) & 0xFFFFF == c["shape1"])
    assert sorted((k & 0x7FFFFF).tolist()) == [0, 1, 2, 3]
    # plane is the lower GeoType: shape0 is the plane, normal points plane -> box
    assert np.all(c["shape0"] == 0) and np.all(c["normal"][:, 2] > 0.99)
