"""CPU oracle of the ellipsoid, cylinder and cone colliders (analytic pairs and the MPR/GJK route with their support
maps and the axial rolling post-process), pinned by the reference's tables and by closed forms.

* test_collision_primitives.py:891-963 (plane-ellipsoid), :965-1090 (sphere-cylinder, side and cap),
  :1430-1525 (plane-cylinder, upright): the signed distance read back from the written contact within 0.01 and
  the normal along the expected direction.
* Closed forms for the MPR/GJK route, whose result the support map decides: the depth of a tilted cone, cylinder
  or ellipsoid resting on a box (lowest point of the shape from its support along -z, written out by hand for
  each shape), the distance between two separated ellipsoids along their common axis, and the sphere-cylinder
  rim case.
* Mass properties: mass, centre of mass and inertia of each new solid against a voxel integration of the shape.
* The rolling post-process (collision_core.py:172-275): a cylinder lying on a box, or a cone lying on its side,
  gets every contact point projected onto the plane through its centre spanned by its axis and the normal, so the
  points sit on the centre's line across the axis; the same shapes standing upright keep their rim points.
"""

from __future__ import annotations

import math

import numpy as np
import pytest

from newton_custom_contact_amd._types import quat_from_axis_angle, transform
from newton_custom_contact_amd.builder import ModelBuilder, compute_inertia_shape
from newton_custom_contact_amd._types import GeoType
from oracle import Oracle
from test_oracle_known_answers import contact_distances

QI = (0.0, 0.0, 0.0, 1.0)


def _scene(shapes, gap=0.5, up_plane=True):
    """shapes: (kind, dims, pos, quat, static).  A ground plane at z = 0 first when up_plane."""
    b = ModelBuilder()
    cfg = ModelBuilder.ShapeConfig(gap=gap, mu=0.5)
    if up_plane:
        b.add_shape_plane(xform=transform((0.0, 0.0, 0.0), QI), width=0.0, length=0.0, cfg=cfg)
    for kind, dims, pos, quat, static in shapes:
        body = -1 if static else b.add_body(xform=transform(pos, quat))
        xf = transform(pos, quat) if static else None
        if kind == "sphere":
            b.add_shape_sphere(body, xform=xf, radius=dims[0], cfg=cfg)
        elif kind == "box":
            b.add_shape_box(body, xform=xf, hx=dims[0], hy=dims[1], hz=dims[2], cfg=cfg)
        elif kind == "ellipsoid":
            b.add_shape_ellipsoid(body, xform=xf, rx=dims[0], ry=dims[1], rz=dims[2], cfg=cfg)
        elif kind == "cylinder":
            b.add_shape_cylinder(body, xform=xf, radius=dims[0], half_height=dims[1], cfg=cfg)
        elif kind == "cone":
            b.add_shape_cone(body, xform=xf, radius=dims[0], half_height=dims[1], cfg=cfg)
        elif kind == "capsule":
            b.add_shape_capsule(body, xform=xf, radius=dims[0], half_height=dims[1], cfg=cfg)
    return b.finalize("cpu")


def _collide(shapes, gap=0.5, up_plane=True):
    m = _scene(shapes, gap, up_plane)
    c = Oracle(m).collide(m.host["body_q"])
    return m, c, contact_distances(m, c)


# ------------------------------------------------------------------ test_collision_primitives.py:891-963
@pytest.mark.parametrize("z,expected", [(2.5, 1.0), (2.0, 0.5), (1.5, 0.0), (1.4, -0.1), (1.3, -0.2), (1.0, -0.5)])
def test_plane_ellipsoid_table(z, expected):
    _m, c, d = _collide([("ellipsoid", (1.0, 1.0, 1.5), (0.0, 0.0, z), QI, False)], gap=2.0)
    assert c["count"] == 1
    assert abs(d[0] - expected) <= 0.01
    assert np.dot(c["normal"][0], (0.0, 0.0, 1.0)) > 0.99


# ------------------------------------------------------------------ test_collision_primitives.py:965-1090
@pytest.mark.parametrize("pos,expected", [((2.0, 0.0, 0.0), 0.5), ((1.5, 0.0, 0.0), 0.0), ((1.4, 0.0, 0.0), -0.1),
                                          ((1.3, 0.0, 0.0), -0.2), ((0.0, 0.0, 2.0), 0.5), ((0.0, 0.0, 1.5), 0.0),
                                          ((0.0, 0.0, 1.4), -0.1), ((0.0, 0.0, 1.3), -0.2)])
def test_sphere_cylinder_table(pos, expected):
    _m, c, d = _collide([("cylinder", (1.0, 1.0), (0.0, 0.0, 0.0), QI, True),
                         ("sphere", (0.5,), pos, QI, False)], gap=1.0, up_plane=False)
    assert c["count"] == 1
    assert abs(d[0] - expected) <= 0.01
    # the normal points from the sphere (shape 0 by GeoType) toward the cylinder
    assert np.dot(c["normal"][0], -np.asarray(pos) / np.linalg.norm(pos)) > 0.99


def test_sphere_cylinder_rim_closed_form():
    """Centre beyond the rim and above the cap: distance to the rim circle minus the radius."""
    pos = np.array([1.3, 0.2, 1.25])
    _m, c, d = _collide([("cylinder", (1.0, 1.0), (0.0, 0.0, 0.0), QI, True),
                         ("sphere", (0.5,), tuple(pos), QI, False)], gap=1.0, up_plane=False)
    rho = math.hypot(pos[0], pos[1])
    rim = np.array([pos[0] / rho, pos[1] / rho, 1.0])
    assert c["count"] == 1
    assert abs(d[0] - (np.linalg.norm(pos - rim) - 0.5)) <= 1e-5


# ------------------------------------------------------------------ test_collision_primitives.py:1430-1525
@pytest.mark.parametrize("z,expected", [(5.0, 4.0), (1.0, 0.0), (0.9, -0.1), (0.8, -0.2), (0.7, -0.3), (0.5, -0.5)])
def test_plane_cylinder_table(z, expected):
    _m, c, d = _collide([("cylinder", (1.0, 1.0), (0.0, 0.0, z), QI, False)], gap=5.0)
    assert c["count"] >= 1
    assert abs(min(d) - expected) <= 0.01
    assert all(np.dot(n, (0.0, 0.0, 1.0)) > 0.99 for n in c["normal"])
    if expected <= 0.0:
        # upright: the flat-mode tripod on the bottom cap, all at the cap's depth
        assert c["count"] == 3 and np.allclose(d, expected, atol=1e-5)


def test_plane_cylinder_rolling_generators():
    """A cylinder lying on the plane: its two side generators at the deepest line (rolling mode)."""
    q = quat_from_axis_angle(np.array([1.0, 0.0, 0.0]), 0.5 * math.pi)
    _m, c, d = _collide([("cylinder", (0.3, 0.5), (0.0, 0.0, 0.29), q, False)], gap=0.05)
    assert c["count"] >= 2
    assert np.allclose(sorted(d)[:2], -0.01, atol=1e-5)


# ------------------------------------------------------------------ MPR/GJK route: closed-form depths
def _lowest(kind, dims, q):
    """Extent of the shape below its centre along -z (q rotates the local frame into the world)."""
    R = _mat(q)
    if kind == "ellipsoid":
        return math.sqrt(sum((R[2, k] * dims[k]) ** 2 for k in range(3)))
    az = R[2, 2]  # world z component of the local axis
    if kind == "cylinder":
        return dims[1] * abs(az) + dims[0] * math.sqrt(max(0.0, 1.0 - az * az))
    if kind == "cone":  # apex (0, 0, h) or the base rim at z = -h
        h, r = dims[1], dims[0]
        return max(-h * az, h * az + r * math.sqrt(max(0.0, 1.0 - az * az)))
    raise ValueError(kind)


def _mat(q):
    x, y, z, w = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


@pytest.mark.parametrize("kind,dims", [("cone", (0.3, 0.25)), ("cylinder", (0.2, 0.3)), ("ellipsoid", (0.4, 0.3, 0.2))])
@pytest.mark.parametrize("angle", [0.3, 1.1, 2.4])
def test_shape_on_box_depth_closed_form(kind, dims, angle):
    """A tilted shape 5 mm into a static box's top face: the deepest contact's distance is -5 mm."""
    q = quat_from_axis_angle(np.array([1.0, 0.4, 0.0]) / math.hypot(1.0, 0.4), angle)
    low = _lowest(kind, dims, q)
    pos = (0.1, -0.05, 0.5 + low - 0.005)
    _m, c, d = _collide([("box", (1.0, 1.0, 0.5), (0.0, 0.0, 0.0), QI, True), (kind, dims, pos, q, False)],
                        gap=0.02, up_plane=False)
    assert c["count"] >= 1
    assert abs(min(d) - (-0.005)) <= 2e-4, (kind, angle, d)


def test_ellipsoids_separated_along_axis():
    """test_narrow_phase.py:1544-1608: separated along x the gap is 3 - 2 a; overlapping, -(2 a - 1.8)."""
    for x, expected in ((3.0, 1.0), (1.8, -0.2)):
        _m, c, d = _collide([("ellipsoid", (1.0, 0.5, 0.3), (0.0, 0.0, 0.0), QI, False),
                             ("ellipsoid", (1.0, 0.5, 0.3), (x, 0.0, 0.0), QI, False)], gap=1.5, up_plane=False)
        assert c["count"] == 1  # an ellipsoid pair keeps a single contact
        assert abs(d[0] - expected) <= 1e-4
        assert abs(np.linalg.norm(c["normal"][0]) - 1.0) <= 1e-5 and c["normal"][0][0] > 0.99


# ------------------------------------------------------------------ mass properties
def _voxel_props(kind, dims, n=120):
    lo = -np.array([dims[0], dims[0], dims[1]]) if kind in ("cylinder", "cone") else -np.array(dims)
    hi = -lo
    axes = [np.linspace(lo[k], hi[k], n, endpoint=False) + (hi[k] - lo[k]) / (2 * n) for k in range(3)]
    X, Y, Z = np.meshgrid(*axes, indexing="ij")
    if kind == "ellipsoid":
        inside = (X / dims[0]) ** 2 + (Y / dims[1]) ** 2 + (Z / dims[2]) ** 2 <= 1.0
    elif kind == "cylinder":
        inside = X * X + Y * Y <= dims[0] ** 2
    else:  # cone: radius shrinks linearly from the base (z = -h) to the apex (z = +h)
        rz = dims[0] * (dims[1] - Z) / (2 * dims[1])
        inside = X * X + Y * Y <= rz * rz
    dv = np.prod([(hi[k] - lo[k]) / n for k in range(3)])
    m = inside.sum() * dv
    P = np.stack([X[inside], Y[inside], Z[inside]], 1)
    com = P.mean(0)
    P = P - com
    r2 = (P * P).sum(1)
    I = (np.eye(3)[None] * r2[:, None, None] - P[:, :, None] * P[:, None, :]).sum(0) * dv
    return m, com, I


@pytest.mark.parametrize("kind,gt,dims,scale", [("ellipsoid", GeoType.ELLIPSOID, (0.4, 0.3, 0.2), (0.4, 0.3, 0.2)),
                                                ("cylinder", GeoType.CYLINDER, (0.2, 0.35), (0.2, 0.35, 0.0)),
                                                ("cone", GeoType.CONE, (0.3, 0.25), (0.3, 0.25, 0.0))])
def test_mass_properties_against_voxels(kind, gt, dims, scale):
    m, com, I = compute_inertia_shape(gt, scale, 1.0)
    mv, cv, Iv = _voxel_props(kind, dims)
    assert abs(m - mv) <= 0.01 * m
    assert np.allclose(com, cv, atol=0.01 * max(dims))
    assert np.allclose(I, Iv, rtol=0.02, atol=0.02 * np.abs(I).max())


# ------------------------------------------------------------------ axial rolling post-process
@pytest.mark.parametrize("kind,dims", [("cylinder", (0.2, 0.4)), ("cone", (0.25, 0.3))])
def test_rolling_projection(kind, dims):
    r, h = dims
    if kind == "cylinder":
        q = quat_from_axis_angle(np.array([0.0, 1.0, 0.0]), 0.5 * math.pi)  # axis along x: lying on its side
        z = r - 0.004
    else:
        half = math.atan2(r, 2 * h)  # lying on a generator: the axis tilted by (90 deg - half angle) from z
        q = quat_from_axis_angle(np.array([0.0, 1.0, 0.0]), 0.5 * math.pi + half)
        z = _lowest("cone", dims, q) - 0.004
    pos = (0.05, 0.12, 0.5 + z)
    _m, c, d = _collide([("box", (1.0, 1.0, 0.5), (0.0, 0.0, 0.0), QI, True), (kind, dims, pos, q, False)],
                        gap=0.02, up_plane=False)
    assert c["count"] >= 2
    # every contact point (the midpoint between the written surface points) lies on the plane through the
    # centre spanned by the axis (in the x-z plane) and the normal (z): its y is the centre's
    qb = _m.host["body_q"][0]
    mid = []
    for k in range(c["count"]):
        p1 = _mat(qb[3:]) @ c["point1"][k] + qb[:3]
        mid.append(0.5 * (c["point0"][k] + p1))
    assert np.allclose(np.array(mid)[:, 1], 0.12, atol=1e-5), np.array(mid)


def test_upright_cylinder_keeps_rim_points():
    """Standing on its cap the cylinder is not rolling: its manifold spreads across the cap (y differs)."""
    pos = (0.05, 0.12, 0.5 + 0.4 - 0.004)
    _m, c, _d = _collide([("box", (1.0, 1.0, 0.5), (0.0, 0.0, 0.0), QI, True), ("cylinder", (0.2, 0.4), pos, QI, False)],
                         gap=0.02, up_plane=False)
    assert c["count"] >= 3
    assert np.ptp(c["point0"][:, 1]) > 0.1


def test_sphere_inside_cylinder_nearer_the_cap():
    """Centre inside the cylinder, 0.1 below the top cap and 0.7 from the side: the cap is the nearer surface, so
    the cap collider answers (distance (0.9 - 1) - 0.5), not the side one (0.3 - 1.5)."""
    _m, c, d = _collide([("cylinder", (1.0, 1.0), (0.0, 0.0, 0.0), QI, True),
                         ("sphere", (0.5,), (0.3, 0.0, 0.9), QI, False)], gap=1.0, up_plane=False)
    assert c["count"] == 1
    assert abs(d[0] - (-0.6)) <= 1e-5


def test_plane_cylinder_tilted_rolling_mode_depths():
    """Axis 30 degrees from the normal (beyond the 22.5-degree flat mode): the deepest rim point, the near-cap rim
    points across it (raised by r sin 30) and the far end of the deepest generator (raised by 2 hh cos 30)."""
    r, hh, delta, tilt = 0.2, 0.4, 0.01, math.radians(30.0)
    q = quat_from_axis_angle(np.array([1.0, 0.0, 0.0]), tilt)
    z = hh * math.cos(tilt) + r * math.sin(tilt) - delta
    _m, c, d = _collide([("cylinder", (r, hh), (0.0, 0.0, z), q, False)], gap=2.0)
    assert c["count"] == 3
    expected = sorted([-delta, -delta + r * math.sin(tilt), -delta + 2 * hh * math.cos(tilt)])
    assert np.allclose(sorted(d), expected, atol=1e-5), (sorted(d), expected)


@pytest.mark.parametrize("angle", [0.4, 1.2])
def test_plane_ellipsoid_rotated_depth(angle):
    """A tilted ellipsoid 5 mm into the plane: depth from its extent below the centre, |A R^T n|."""
    dims = (0.5, 0.3, 0.2)
    q = quat_from_axis_angle(np.array([1.0, 0.7, 0.2]) / np.linalg.norm([1.0, 0.7, 0.2]), angle)
    z = _lowest("ellipsoid", dims, q) - 0.005
    _m, c, d = _collide([("ellipsoid", dims, (0.0, 0.0, z), q, False)], gap=0.1)
    assert c["count"] == 1
    assert abs(d[0] - (-0.005)) <= 1e-5
