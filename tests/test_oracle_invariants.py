"""CPU oracle pinned by brute force, closed forms and physical invariants.

* broad phase: the SAP pair set equals an independent O(n^2) AABB-overlap + world/group/exclusion
  filter in numpy (test_broad_phase.py:90-147 strategy);
* inertia: closed forms of sphere/box/capsule (inertia.py:42-244);
* XPBD: semi-implicit free fall in closed form, resting contact, no bounce above the drop height,
  drop speed bounded by free fall (test_rigid_contact.py:766-819), stacks stay put;
* FP32 vs FP64 restatement agree to float tolerance (the FP64 rerun the parity statement cites).
"""

from __future__ import annotations

import math

import numpy as np
import pytest

from newton_custom_contact_amd import scenes
from newton_custom_contact_amd._types import GeoType, quat_to_matrix, transform
from newton_custom_contact_amd.builder import ModelBuilder
from oracle import Oracle

G = 9.81


def numpy_aabbs(model, q):
    """Closed-form world AABBs (expanded by margin + gap) of plane/sphere/box/capsule shapes."""
    h = model.host
    S = model.shape_count
    lo = np.zeros((S, 3))
    hi = np.zeros((S, 3))
    for s in range(S):
        b = h["shape_body"][s]
        xs = h["shape_transform"][s].astype(np.float64)
        if b >= 0:
            xb = q[b].astype(np.float64)
            Rb = quat_to_matrix(xb[3:])
            p = Rb @ xs[:3] + xb[:3]
            R = Rb @ quat_to_matrix(xs[3:])
        else:
            p = xs[:3]
            R = quat_to_matrix(xs[3:])
        t = h["shape_type"][s]
        sc = h["shape_scale"][s].astype(np.float64)
        e = h["shape_margin"][s] + h["shape_gap"][s]
        if t == GeoType.PLANE:
            half = np.full(3, h["shape_collision_radius"][s], np.float64)
        elif t == GeoType.SPHERE:
            half = np.full(3, sc[0])
        elif t == GeoType.BOX:
            half = np.abs(R) @ sc
        elif t == GeoType.CAPSULE:
            half = np.abs(R[:, 2]) * sc[1] + sc[0]
        lo[s] = p - half - e
        hi[s] = p + half + e
    return lo, hi


def brute_pairs(model, q, tol=1e-5):
    h = model.host
    lo, hi = numpy_aabbs(model, q)
    sw, grp = h["shape_world"], h["shape_collision_group"]
    excl = {tuple(p) for p in h["shape_filter_pairs"].tolist()}
    sure, ambiguous = set(), set()
    S = model.shape_count
    for a in range(S):
        for b in range(a + 1, S):
            wa, wb = sw[a], sw[b]
            if wa != -1 and wb != -1 and wa != wb:
                continue
            ga, gb = grp[a], grp[b]
            if ga == 0 or gb == 0:
                continue
            if ga > 0 and not (ga == gb or gb < 0):
                continue
            if ga < 0 and ga == gb:
                continue
            if (a, b) in excl:
                continue
            sep = np.maximum(lo[a] - hi[b], lo[b] - hi[a]).max()  # > 0: separated on some axis
            if sep < -tol:
                sure.add((a, b))
            elif sep <= tol:
                ambiguous.add((a, b))
    return sure, ambiguous


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_sap_matches_brute_force(seed):
    model = scenes.mixed_worlds(6, seed=seed, device="cpu")
    o = Oracle(model)
    q = model.host["body_q"].copy()
    qd = model.host["body_qd"].copy()
    for i in range(3):
        got = {tuple(p) for p in o.broadphase(q).tolist()}
        assert len(got) == len(o.broadphase(q)), "duplicate pairs"
        sure, amb = brute_pairs(model, q)
        assert sure <= got
        assert got - sure <= amb
        for _ in range(80):  # let the pile fall to get dense contact states
            c = o.collide(q)
            q, qd = o.step(q, qd, np.zeros_like(qd), c, 1 / 240, iterations=4)


def test_sap_filters_worlds_groups_exclusions():
    b = ModelBuilder()
    b.add_ground_plane()
    for w in range(2):
        b.begin_world()
        b1 = b.add_body(xform=transform((0, 0, 0.3)))
        b.add_shape_box(b1, hx=0.2, hy=0.2, hz=0.2)
        b2 = b.add_body(xform=transform((0.1, 0, 0.5)))
        b.add_shape_sphere(b2, radius=0.2, cfg=ModelBuilder.ShapeConfig(collision_group=-1 if w else 2))
        b3 = b.add_body(xform=transform((0, 0.1, 0.4)))
        s3 = b.add_shape_sphere(b3, radius=0.2)
        b.add_shape_collision_filter_pair(s3 - 2, s3)  # exclude box <-> third sphere
        b.end_world()
    m = b.finalize("cpu")
    o = Oracle(m)
    got = {tuple(p) for p in o.broadphase(m.host["body_q"]).tolist()}
    sure, amb = brute_pairs(m, m.host["body_q"])
    assert got == sure and not amb
    # no cross-world pairs, group 2 only meets group 2 / negative groups, exclusion honoured
    sw = m.host["shape_world"]
    assert all(sw[a] == -1 or sw[b] == -1 or sw[a] == sw[b] for a, b in got)
    assert (1, 3) not in got and (5, 7) not in got


def test_inertia_closed_forms():
    b = ModelBuilder()
    rho = 1000.0
    bs = b.add_body()
    b.add_shape_sphere(bs, radius=0.2)
    bb = b.add_body()
    b.add_shape_box(bb, hx=0.1, hy=0.2, hz=0.3)
    bc = b.add_body()
    b.add_shape_capsule(bc, radius=0.1, half_height=0.2)
    m = b.finalize("cpu")
    h = m.host
    ms = rho * 4 / 3 * math.pi * 0.2**3
    assert h["body_mass"][bs] == pytest.approx(ms, rel=1e-6)
    np.testing.assert_allclose(np.diag(h["body_inertia"][bs]), 0.4 * ms * 0.04, rtol=1e-6)
    mb = rho * 8 * 0.1 * 0.2 * 0.3
    assert h["body_mass"][bb] == pytest.approx(mb, rel=1e-6)
    np.testing.assert_allclose(np.diag(h["body_inertia"][bb]),
                               [mb / 3 * (0.04 + 0.09), mb / 3 * (0.01 + 0.09), mb / 3 * (0.01 + 0.04)], rtol=1e-6)
    # capsule: cylinder (h = 0.4) + two hemispheres
    r, hh = 0.1, 0.4
    mc = rho * math.pi * r * r * hh
    msph = rho * 4 / 3 * math.pi * r**3
    assert h["body_mass"][bc] == pytest.approx(mc + msph, rel=1e-6)
    iz = (mc * 0.5 + msph * 0.4) * r * r
    assert h["body_inertia"][bc][2, 2] == pytest.approx(iz, rel=1e-6)
    for k in (bs, bb, bc):
        np.testing.assert_allclose(h["body_inv_inertia"][k] @ h["body_inertia"][k], np.eye(3), atol=1e-5)
        assert h["body_inv_mass"][k] == pytest.approx(1.0 / h["body_mass"][k], rel=1e-6)


def _single(kind, z, **kw):
    b = ModelBuilder()
    b.add_ground_plane()
    body = b.add_body(xform=transform((0, 0, z), kw.pop("q", (0, 0, 0, 1))))
    if kind == "sphere":
        b.add_shape_sphere(body, radius=kw["r"])
    else:
        b.add_shape_box(body, hx=kw["h"], hy=kw["h"], hz=kw["h"])
    return b.finalize("cpu")


def test_free_fall_closed_form():
    """Semi-implicit Euler (solver.py:50-95): v_n = -g n dt, z_n = z0 - g dt^2 n (n + 1) / 2."""
    m = _single("sphere", 50.0, r=0.1)
    o = Oracle(m)
    q, qd = m.host["body_q"].copy(), m.host["body_qd"].copy()
    dt = 1 / 240
    n = 120
    for _ in range(n):
        c = o.collide(q)
        assert c["count"] == 0
        q, qd = o.step(q, qd, np.zeros_like(qd), c, dt, iterations=10, precision="f32")
    assert qd[0, 2] == pytest.approx(-G * n * dt, rel=1e-5)
    assert q[0, 2] == pytest.approx(50.0 - G * dt * dt * n * (n + 1) / 2, rel=1e-6)
    np.testing.assert_allclose(q[0, 3:], [0, 0, 0, 1], atol=1e-7)


@pytest.mark.parametrize("kind", ["sphere", "box"])
def test_resting_on_ground(kind):
    """test_rigid_contact.py:28-229: a shape resting on the plane stays there."""
    m = _single(kind, 0.2, r=0.2, h=0.2)
    o = Oracle(m)
    q, qd = m.host["body_q"].copy(), m.host["body_qd"].copy()
    for _ in range(480):
        c = o.collide(q)
        q, qd = o.step(q, qd, np.zeros_like(qd), c, 1 / 240, iterations=10)
    assert abs(q[0, 2] - 0.2) < 2e-3
    assert np.abs(qd).max() < 0.05


def test_drop_speed_bounded_and_no_bounce():
    """test_rigid_contact.py:766-819 and test_shapes_no_bounce.py: |v| <= sqrt(2 g h), never above z0."""
    z0 = 1.5
    m = _single("box", z0, h=0.15, q=(0.2, 0.1, 0.0, math.sqrt(1 - 0.05)))
    o = Oracle(m)
    q, qd = m.host["body_q"].copy(), m.host["body_qd"].copy()
    vmax = math.sqrt(2 * G * z0) * 1.05
    for _ in range(600):
        c = o.collide(q)
        q, qd = o.step(q, qd, np.zeros_like(qd), c, 1 / 240, iterations=10)
        assert np.linalg.norm(qd[0, :3]) <= vmax
        assert q[0, 2] <= z0 + 1e-4
    assert q[0, 2] < 0.35  # came to rest on the ground (half-diagonal 0.26)


def test_box_stack_stays():
    b = ModelBuilder()
    b.add_ground_plane()
    for i in range(3):
        body = b.add_body(xform=transform((0, 0, 0.2 + 0.4 * i)))
        b.add_shape_box(body, hx=0.2, hy=0.2, hz=0.2)
    m = b.finalize("cpu")
    o = Oracle(m)
    q, qd = m.host["body_q"].copy(), m.host["body_qd"].copy()
    for _ in range(480):
        c = o.collide(q)
        q, qd = o.step(q, qd, np.zeros_like(qd), c, 1 / 240, iterations=10)
    np.testing.assert_allclose(q[:, 2], [0.2, 0.6, 1.0], atol=5e-3)
    np.testing.assert_allclose(q[:, :2], 0.0, atol=5e-3)


def test_fp32_vs_fp64_restatement():
    """One collide + step in FP32 and FP64 from the same state: the FP32 error statement."""
    model = scenes.mixed_worlds(4, seed=5, device="cpu")
    o = Oracle(model)
    q, qd = model.host["body_q"].copy(), model.host["body_qd"].copy()
    for _ in range(200):
        c = o.collide(q)
        q, qd = o.step(q, qd, np.zeros_like(qd), c, 1 / 240, iterations=10)
    c32 = o.collide(q)
    c64 = o.collide(q.astype(np.float64), precision="f64")
    assert c32["count"] == c64["count"]
    assert np.array_equal(c32["sort_key"], c64["sort_key"])
    np.testing.assert_allclose(c32["normal"], c64["normal"], atol=1e-5)
    q32, qd32 = o.step(q, qd, np.zeros_like(qd), c32, 1 / 240, iterations=10)
    q64, qd64 = o.step(q.astype(np.float64), qd.astype(np.float64), np.zeros_like(qd, np.float64), c64, 1 / 240,
                       iterations=10, precision="f64")
    np.testing.assert_allclose(q32, q64, atol=2e-5)
    # XPBD velocities are position corrections divided by dt (kernels.py:855-868): an FP32 rounding of
    # 1e-7 m in a resolved penetration becomes ~1e-5 m/s per correction at dt = 1/240 s, and ten
    # iterations of coupled contacts amplify it; 1e-2 m/s bounds it on this scene
    np.testing.assert_allclose(qd32, qd64, atol=1e-2)


def test_fp32_asin_acos_error_bound():
    """The revolute twist/swing angles use an FP32 asin/acos (Cephes method) that the HIP kernels
    share bit for bit; it is pinned here against float64 arcsin/arccos: asin within 3 ulp, acos
    within 2 ulp, exact at the special points, odd/even symmetry."""
    import oracle

    x = np.linspace(-1, 1, 2_000_001, dtype=np.float32)
    extra = np.float32([-1, -0.5, 0.5, 1, 0, 1e-8, -1e-8, np.nextafter(np.float32(0.5), np.float32(1)),
                        np.nextafter(np.float32(-0.5), np.float32(-1))])
    x = np.unique(np.concatenate([x, extra]))
    for f, ref, bound in ((oracle.asin_f32, np.arcsin, 3.0), (oracle.acos_f32, np.arccos, 2.0)):
        y = f(x).astype(np.float64)
        r = ref(x.astype(np.float64))
        ulp = np.spacing(np.abs(r).astype(np.float32)).astype(np.float64)
        assert np.max(np.abs(y - r) / ulp) <= bound
    a = oracle.asin_f32(np.float32([0.0, 1.0, -1.0, 0.5]))
    assert a[0] == 0.0 and a[1] == np.float32(np.pi / 2) and a[2] == -np.float32(np.pi / 2)
    assert abs(float(a[3]) - np.pi / 6) < 1e-7
    assert np.array_equal(oracle.asin_f32(-x), -oracle.asin_f32(x))
    c = oracle.acos_f32(np.float32([1.0, -1.0, 0.0]))
    assert c[0] == 0.0 and c[1] == np.float32(np.pi) and c[2] == np.float32(np.pi / 2)
