"""CPU tests of the mesh / SDF restatement and the mesh host logic.

Pins (none of them retypes the restatement's formula):
  * Mesh.edges on hand-built meshes (order of first occurrence, geometric de-duplication) and the Euler
    count E = 3V - 6 of convex hulls;
  * solid mass properties against the reference's own known answers (tests/test_inertia.py:25-140:
    unit cube, the same cube with a split edge, a fine sphere);
  * closest point on a triangle against brute-force minimisation; solid angle against closed forms
    (octant = pi/2, winding number 1 inside / 0 outside a closed mesh);
  * the SDF cook of a box mesh against the exact box SDF at every grid vertex (face, edge and corner
    regions, inside and out);
  * trilinear sampling and gradient reproduce a linear field exactly, extrapolate by the distance to the
    grid box outside it, and match finite differences inside a cell;
  * the Brent edge search against a brute-force minimum along the segment;
  * mesh-plane contacts of a box resting on the ground (closed form, reduced and unreduced); reduced contacts
    are a subset of the unreduced ones that keeps each pair's deepest contact (the reduction's own pins are in
    test_oracle_reduction.py).
"""

from __future__ import annotations

import math

import numpy as np
import pytest
import torch

import oracle
from newton_custom_contact_amd.builder import ModelBuilder
from newton_custom_contact_amd.mesh import Mesh, random_convex_hull
from tests.mesh_helpers import box_mesh, box_sdf, cook_cpu, icosphere


# ---------------------------------------------------------------- topology and mass
def test_edges_first_occurrence_order():
    # tetrahedron: triangle edges (0,1) (1,2) (0,2) | (0,3) (3,1) (0,1)* | (1,3)* (3,2) (1,2)* | (2,3)* (3,0)* (2,0)*
    m = Mesh(np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], np.float32),
             np.array([0, 1, 2, 0, 3, 1, 1, 3, 2, 2, 3, 0], np.int32), compute_inertia=False)
    assert m.edges.tolist() == [[0, 1], [1, 2], [0, 2], [0, 3], [3, 1], [3, 2]]


def test_edges_geometric_dedup():
    # two triangles sharing an edge through duplicated (coincident) vertices: 5 unique edges, not 6
    v = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [1, 0, 0], [0, 1, 0], [1, 1, 0]], np.float32)
    m = Mesh(v, np.array([0, 1, 2, 3, 5, 4], np.int32), compute_inertia=False)
    assert len(m.edges) == 5
    assert m.edges.tolist()[:3] == [[0, 1], [1, 2], [0, 2]]


def test_convex_hull_euler_count():
    rng = np.random.Generator(np.random.PCG64(3))
    for _ in range(5):
        h = random_convex_hull(rng)
        assert len(h.edges) == 3 * len(h.vertices) - 6
        assert len(h.indices) // 3 == 2 * len(h.vertices) - 4


_CUBE_V = [[1.0, 0.0, 0.0], [1.0, 0.0, 1.0], [0.0, 0.0, 1.0], [0.0, 0.0, 0.0], [1.0, 1.0, 0.0], [1.0, 1.0, 1.0],
           [0.0, 1.0, 1.0], [0.0, 1.0, 0.0]]
_CUBE_T = [[1, 2, 3], [7, 6, 5], [4, 5, 1], [5, 6, 2], [2, 6, 7], [0, 3, 7], [0, 1, 3], [4, 7, 5], [0, 4, 1],
           [1, 5, 2], [3, 2, 7], [4, 0, 7]]


def test_cube_mass_properties_known_answer():
    m, com, inertia, vol = Mesh(_CUBE_V, _CUBE_T, compute_inertia=False).mass_properties(1000.0)
    assert abs(m - 1000.0) < 1e-6 and abs(vol - 1.0) < 1e-6
    np.testing.assert_allclose(com, [0.5, 0.5, 0.5], atol=1e-6)
    np.testing.assert_allclose(inertia, np.diag([1000.0 / 6.0] * 3), atol=1e-4)
    # the same cube with a vertex inserted on one edge (reference: identical mass properties)
    v = _CUBE_V + [[0.5, 0.0, 0.0]]
    t = [list(x) for x in _CUBE_T]
    t[5] = [0, 8, 7]
    t.append([8, 3, 7])
    t[6] = [0, 1, 8]
    t.append([8, 1, 3])
    m1, com1, inertia1, vol1 = Mesh(v, t, compute_inertia=False).mass_properties(1000.0)
    assert abs(m1 - m) < 1e-6 and abs(vol1 - vol) < 1e-6
    np.testing.assert_allclose(com1, com, atol=1e-6)
    np.testing.assert_allclose(inertia1, inertia, atol=1e-4)


def test_sphere_mass_properties_known_answer():
    s = icosphere(2.5, 6)  # 81920 triangles (the reference uses a 500 x 500 latitude-longitude sphere)
    off = np.array([1.0, 2.0, 3.0], np.float32)
    m, com, inertia, vol = Mesh(s.vertices + off, s.indices, compute_inertia=False).mass_properties(1000.0)
    ms = 1000.0 * 4.0 / 3.0 * math.pi * 2.5 ** 3
    assert abs(m - ms) < 1e2
    np.testing.assert_allclose(com, off, atol=2e-3)
    np.testing.assert_allclose(inertia, np.diag([0.4 * ms * 2.5 ** 2] * 3), atol=4e2)
    assert abs(vol - 4.0 / 3.0 * math.pi * 2.5 ** 3) < 3e-2


def test_builder_mesh_shape_mass_and_radius():
    mesh = box_mesh(0.1, 0.2, 0.3)
    b = ModelBuilder()
    body = b.add_body()
    cfg = ModelBuilder.ShapeConfig()
    cfg.density = 500.0
    b.add_shape_mesh(body, mesh=mesh, cfg=cfg)
    m = b.finalize("cpu")
    mass = 500.0 * 0.2 * 0.4 * 0.6
    assert abs(m.host["body_mass"][0] - mass) < 1e-4
    np.testing.assert_allclose(m.host["body_inertia"][0].reshape(3, 3).diagonal(),
                               [mass / 12 * (0.4 ** 2 + 0.6 ** 2), mass / 12 * (0.2 ** 2 + 0.6 ** 2),
                                mass / 12 * (0.2 ** 2 + 0.4 ** 2)], rtol=1e-5)
    assert abs(m.host["shape_collision_radius"][0] - math.sqrt(0.1 ** 2 + 0.2 ** 2 + 0.3 ** 2)) < 1e-6
    np.testing.assert_allclose(m.host["shape_aabb_lower"][0], [-0.1, -0.2, -0.3], atol=1e-7)
    np.testing.assert_allclose(m.host["shape_aabb_upper"][0], [0.1, 0.2, 0.3], atol=1e-7)
    assert m.host["shape_mesh"][0] == 0 and m.host["mesh_edge_start"][-1] == 18


# ---------------------------------------------------------------- cook building blocks
def _brute_closest(p, a, b, c, n=401):
    """Minimise |p - (a + u (b - a) + v (c - a))| over the triangle by grid search + local refinement."""
    best, arg = np.inf, None
    for k in range(3):
        lo_u, lo_v, span = (0.0, 0.0, 1.0) if arg is None else (arg[0] - span / 20, arg[1] - span / 20, span / 10)
        uu, vv = np.meshgrid(np.linspace(lo_u, lo_u + span, n), np.linspace(lo_v, lo_v + span, n))
        ok = (uu >= 0) & (vv >= 0) & (uu + vv <= 1)
        q = a + uu[ok, None] * (b - a) + vv[ok, None] * (c - a)
        d = np.linalg.norm(q - p, axis=1)
        i = int(np.argmin(d))
        if d[i] < best:
            best, arg = d[i], (uu[ok][i], vv[ok][i])
    return best


def test_closest_on_triangle_vs_brute_force():
    rng = np.random.Generator(np.random.PCG64(7))
    P, A, Bv, C = (rng.normal(size=(60, 3)) for _ in range(4))
    # points aimed at every Voronoi region: vertices, edges, the face
    A[:6] = [0, 0, 0]
    Bv[:6] = [1, 0, 0]
    C[:6] = [0, 1, 0]
    P[:6] = [[-1, -1, 0.5], [2, -0.5, 0.2], [-0.3, 2, -0.1], [0.5, -1, 0.3], [1, 1, 0.7], [0.2, 0.2, 1.0]]
    q = oracle.closest_on_triangle_f64(P, A, Bv, C)
    for i in range(len(P)):
        d = np.linalg.norm(q[i] - P[i])
        assert d <= _brute_closest(P[i], A[i], Bv[i], C[i]) + 1e-9
        # the returned point lies on the triangle
        n = np.cross(Bv[i] - A[i], C[i] - A[i])
        assert abs(np.dot(q[i] - A[i], n)) < 1e-9 * max(1.0, np.linalg.norm(n))
    np.testing.assert_allclose(q[:6], [[0, 0, 0], [1, 0, 0], [0, 1, 0], [0.5, 0, 0], [0.5, 0.5, 0], [0.2, 0.2, 0]],
                               atol=1e-12)


def test_solid_angle_closed_forms():
    om = oracle.solid_angle_f64([[1, 0, 0]], [[0, 1, 0]], [[0, 0, 1]])
    assert abs(om[0] - math.pi / 2) < 1e-12
    # winding number of a closed mesh: 1 inside, 0 outside
    s = icosphere(1.0, 2)
    t = s.indices.reshape(-1, 3)
    v = s.vertices.astype(np.float64)
    for p, wn in (([0.1, -0.2, 0.3], 1.0), ([1.5, 0.0, 0.0], 0.0), ([0.0, 0.0, -3.0], 0.0)):
        om = oracle.solid_angle_f64(v[t[:, 0]] - p, v[t[:, 1]] - p, v[t[:, 2]] - p)
        assert abs(om.sum() / (4 * math.pi) - wn) < 1e-9


def _fine_points(geo, dtype=np.float64):
    """Positions of the padded fine grid's vertices, (nz, ny, nx, 3), as the cook computes them in FP32."""
    w, h, d = geo["coarse_dims"]
    lo = np.asarray(geo["min_extents"], np.float32)
    ce = np.asarray(geo["cell_size"], np.float64).astype(np.float32)
    iz, iy, ix = np.meshgrid(np.arange(8 * d + 1), np.arange(8 * h + 1), np.arange(8 * w + 1), indexing="ij")
    p = np.stack([lo[0] + ix.astype(np.float32) * ce[0], lo[1] + iy.astype(np.float32) * ce[1],
                  lo[2] + iz.astype(np.float32) * ce[2]], -1)
    return p.astype(dtype)


def test_sdf_cook_box_exact():
    """The mesh query at every vertex of the padded fine grid equals the closed-form box SDF (FP64 cook to
    1e-12, FP32 to 2e-7), beyond the mesh's padded box too."""
    h = (0.3, 0.2, 0.1)
    m = box_mesh(*h)
    geo = oracle.texture_geometry(m.vertices, 0.05, (-0.1, 0.1), 16)
    w, hh, d = geo["coarse_dims"]
    n = (8 * w + 1, 8 * hh + 1, 8 * d + 1)
    ce = np.asarray(geo["cell_size"]).astype(np.float32)
    g = oracle.sdf_fine_grid(m.vertices, m.indices, geo["min_extents"], ce, n)
    g64 = oracle.sdf_fine_grid(m.vertices, m.indices, geo["min_extents"], ce, n, precision="f64")
    hf = np.array(h, np.float32).astype(np.float64)  # the mesh's vertices are FP32
    p = _fine_points(geo)
    lo64 = np.asarray(geo["min_extents"], np.float64)
    iz, iy, ix = np.meshgrid(*[np.arange(k) for k in g.shape], indexing="ij")
    p64 = np.stack([lo64[0] + ce[0] * ix, lo64[1] + ce[1] * iy, lo64[2] + ce[2] * iz], -1)
    np.testing.assert_allclose(g64, box_sdf(p64, hf), rtol=0, atol=1e-12)
    np.testing.assert_allclose(g, box_sdf(p, hf), rtol=0, atol=2e-7)
    assert (g < 0).sum() > 0 and (g > 0).sum() > 0


def _field_sdf(fine, geo, fmt="float32", linear_threshold=None):
    """A model with one box mesh whose texture SDF is built from the given fine-grid field."""
    g = dict(geo)
    if linear_threshold is not None:
        g["linear_threshold"] = linear_threshold
    sparse = oracle.sdf_build_sparse(None, None, g, fmt, fine=fine)
    m = box_mesh(0.2, 0.2, 0.1)
    m.attach_sdf(sparse, "cpu")
    b = ModelBuilder()
    b.add_shape_mesh(b.add_body(), mesh=m)
    return b.finalize("cpu"), sparse


@pytest.mark.parametrize("linear_threshold", [None, 0.0])
def test_sdf_sample_linear_field_and_extrapolation(linear_threshold):
    """A linear field is reproduced exactly, value and gradient, whether every subgrid is demoted to the
    coarse grid (the linearity test passes a linear field) or keeps its texture (threshold 0); outside the
    SDF box the value gains the distance to the box and the gradient is the unit offset."""
    a = np.array([0.03, -0.07, 0.05])
    geo = oracle.texture_geometry(box_mesh(0.2, 0.2, 0.1).vertices, 0.05, (-1.0, 1.0), 16)
    p = _fine_points(geo)
    model, sparse = _field_sdf((p @ a + 0.005).astype(np.float32), geo, linear_threshold=linear_threshold)
    if linear_threshold is None:
        assert sparse["num_subgrids"] == 0 and sparse["subgrid_linear"].all()
        assert sparse["subgrid_data"].shape == (1, 1, 1) and sparse["subgrid_data"].dtype == np.float32
    else:
        assert sparse["num_subgrids"] == len(sparse["subgrid_required"]) and not sparse["subgrid_linear"].any()
    o = oracle.Oracle(model)
    lo = np.asarray(sparse["min_extents"], np.float64)
    hi = np.asarray(sparse["max_extents"], np.float32).astype(np.float64)
    rng = np.random.Generator(np.random.PCG64(1))
    q = rng.uniform(lo, hi, size=(500, 3))
    v, gr = o.sdf_sample(0, q)
    np.testing.assert_allclose(v, q @ a + 0.005, atol=2e-7)
    inv = (1.0 / np.asarray(sparse["cell_size"])).astype(np.float32)
    scale = 1.0 if linear_threshold is not None else 8.0  # coarse cells: gradient of coarse weights x fine 1/dx
    np.testing.assert_allclose(gr, np.tile(a * scale, (len(q), 1)), rtol=2e-4, atol=1e-6)
    assert inv.min() > 0
    q2 = q.copy()
    q2[:, 0] = hi[0] + rng.uniform(0.1, 1.0, len(q2))
    v2, g2 = o.sdf_sample(0, q2)
    cl = np.clip(q2, lo, hi)
    np.testing.assert_allclose(v2, cl @ a + 0.005 + np.linalg.norm(q2 - cl, axis=1), atol=3e-6)
    np.testing.assert_allclose(g2, (q2 - cl) / np.linalg.norm(q2 - cl, axis=1)[:, None], atol=1e-6)


def test_sdf_gradient_matches_finite_differences():
    m = cook_cpu(icosphere(0.25, 2), 16)
    b = ModelBuilder()
    b.add_shape_mesh(b.add_body(), mesh=m)
    o = oracle.Oracle(b.finalize("cpu"))
    sp = m.sdf.sparse_data()
    rng = np.random.Generator(np.random.PCG64(2))
    lo = np.asarray(sp["min_extents"], np.float64)
    ce = np.asarray(sp["cell_size"], np.float64)
    w, h, d = sp["coarse_dims"]
    # cell centres +- a quarter cell inside subgrids that keep their texture, so +-h stays inside one cell
    req = np.flatnonzero(sp["subgrid_required"])
    s = req[rng.integers(0, len(req), 200)]
    sub = np.stack([s % w, (s // w) % h, s // (w * h)], -1)
    cells = sub * 8 + rng.integers(0, 8, size=(200, 3))
    p = lo + (cells + 0.5 + rng.uniform(-0.25, 0.25, size=(200, 3))) * ce
    _, g = o.sdf_sample(0, p)
    hstep = 0.05 * ce.min()
    fd = np.zeros_like(p)
    for k in range(3):
        e = np.zeros(3)
        e[k] = hstep
        fd[:, k] = (o.sdf_sample(0, p + e)[0].astype(np.float64) - o.sdf_sample(0, p - e)[0]) / (2 * hstep)
    np.testing.assert_allclose(g, fd, atol=2e-3 * np.abs(fd).max())


def test_edge_search_vs_brute_force():
    m = cook_cpu(icosphere(0.25, 2), 24)
    b = ModelBuilder()
    b.add_shape_mesh(b.add_body(), mesh=m)
    o = oracle.Oracle(b.finalize("cpu"))
    rng = np.random.Generator(np.random.PCG64(4))
    v0 = rng.uniform(-0.3, 0.3, size=(40, 3))
    v1 = v0 + rng.normal(size=(40, 3)) * 0.15
    val, pt = o.edge_search(0, v0, v1)
    t = np.linspace(0.0, 1.0, 2001)
    for i in range(len(v0)):
        seg = v0[i] + t[:, None] * (v1[i] - v0[i])
        vals, _ = o.sdf_sample(0, seg)
        # the search reports the SDF at the point it returns, on the segment, and within a few mm of the
        # segment's true minimum (5 Brent iterations + endpoint check)
        sv, _ = o.sdf_sample(0, pt[i:i + 1])
        assert abs(sv[0] - val[i]) < 1e-6
        u = np.dot(pt[i] - v0[i], v1[i] - v0[i]) / np.dot(v1[i] - v0[i], v1[i] - v0[i])
        assert -1e-6 <= u <= 1 + 1e-6
        assert val[i] <= vals.min() + 5e-3
        assert val[i] >= vals.min() - 1e-6


# ---------------------------------------------------------------- mesh contacts
def _box_on_plane(z, gap=0.1, reduce_contacts=True):
    mesh = cook_cpu(box_mesh(0.1, 0.1, 0.1), 8)
    b = ModelBuilder()
    cfg = ModelBuilder.ShapeConfig()
    cfg.gap = gap
    b.add_ground_plane(cfg=cfg)
    body = b.add_body(xform=(0.0, 0.0, z, 0.0, 0.0, 0.0, 1.0))
    b.add_shape_mesh(body, mesh=mesh, cfg=cfg)
    model = b.finalize("cpu")
    return model, oracle.Oracle(model, reduce_contacts=reduce_contacts)


def test_mesh_plane_contacts_closed_form():
    for reduce in (False, True):
        model, o = _box_on_plane(0.12, reduce_contacts=reduce)
        c = o.collide(model.host["body_q"])
        # the 4 bottom vertices (0..3) at height 0.02; the top ones at 0.22 exceed gap_sum = 0.2.  Reduced, each
        # bottom corner is alone in its voxel (a 0.2 m cube is split 4 x 5 x 5), so all four stay; the octahedral
        # round trip of (0, 0, -1) is exact
        assert c["count"] == 4
        np.testing.assert_array_equal(c["shape0"], [1] * 4)  # (mesh, plane)
        np.testing.assert_array_equal(c["shape1"], [0] * 4)
        np.testing.assert_array_equal(c["sort_key"] & 0x7FFFFF, [0, 1, 2, 3])
        np.testing.assert_allclose(c["normal"], np.tile([0, 0, -1], (4, 1)), atol=0)
        # point0 = vertex (mesh frame), point1 = its projection on the plane (world frame)
        np.testing.assert_allclose(c["point0"], model.host["mesh_vertex"][:4], atol=1e-6)
        np.testing.assert_allclose(c["point1"][:, 2], 0.0, atol=1e-6)


def _contact_depth(model, q, c):
    """Signed gap of each contact from its body-frame points: dot(p1 - p0, n) - margin0 - margin1."""
    from newton_custom_contact_amd._types import quat_rotate  # noqa: PLC0415

    sb = model.host["shape_body"]

    def world(pts, shapes):
        out = np.array(pts, np.float64)
        for i, s_ in enumerate(shapes):
            bd = sb[s_]
            if bd >= 0:
                out[i] = quat_rotate(q[bd, 3:], pts[i]) + q[bd, :3]
        return out

    p0 = world(c["point0"], c["shape0"])
    p1 = world(c["point1"], c["shape1"])
    return np.einsum("ij,ij->i", p1 - p0, c["normal"]) - c["margin0"] - c["margin1"]


def test_reduced_mesh_contacts_subset_of_unreduced():
    """Reduced contacts are a subset (by sort key) of the unreduced ones, in sub-key order, at most 240 per pair;
    each pair keeps its deepest contact (the max-depth slot of that contact's normal bin)."""
    rng = np.random.Generator(np.random.PCG64(5))
    lib = [cook_cpu(random_convex_hull(rng, 60, 90), 12) for _ in range(2)]
    b = ModelBuilder()
    b.add_ground_plane()
    for k, z in enumerate((0.12, 0.3)):
        body = b.add_body(xform=(0.02 * k, 0.0, z, 0.0, 0.0, 0.0, 1.0))
        b.add_shape_mesh(body, mesh=lib[k])
    model = b.finalize("cpu")
    q = model.host["body_q"].astype(np.float64)
    full = oracle.Oracle(model, reduce_contacts=False).collide(model.host["body_q"])
    pairs = sorted({(int(a), int(b_)) for a, b_ in zip(full["shape0"], full["shape1"])})
    assert (1, 0) in pairs and (1, 2) in pairs  # a mesh-plane pair and the mesh-mesh pair
    d_full = _contact_depth(model, q, full)
    for det in (True, False):
        red = oracle.Oracle(model, reduce_contacts=True, deterministic=det).collide(model.host["body_q"])
        for a, b_ in pairs:
            sel = (full["shape0"] == a) & (full["shape1"] == b_)
            sel_r = (red["shape0"] == a) & (red["shape1"] == b_)
            assert 0 < sel_r.sum() <= min(240, sel.sum())
            assert set(red["sort_key"][sel_r].tolist()).issubset(set(full["sort_key"][sel].tolist()))
            assert np.all(np.diff(red["sort_key"][sel_r]) > 0)  # sub-key order
            deepest = full["sort_key"][sel][np.argmin(d_full[sel])]
            assert deepest in set(red["sort_key"][sel_r].tolist())
    # mesh-mesh sub keys carry (edge << 2) | (mode << 1)
    mm = full["shape1"] != 0
    assert mm.any()
    sub = full["sort_key"][mm] & 0x7FFFFF
    assert np.all(sub & 1 == 0)


# ---------------------------------------------------------------- winding threshold and baked scale
def test_inward_winding_threshold():
    """An inward-wound mesh has negative signed volume and cooks, with threshold -0.5
    (sdf_utils.py:506-510, 640-651), to the outward mesh's field; with +0.5 it would be outside everywhere."""
    m = box_mesh(0.3, 0.2, 0.1)
    inv = Mesh(m.vertices, m.indices.reshape(-1, 3)[:, [0, 2, 1]].reshape(-1))
    np.testing.assert_allclose(m.signed_volume(), 0.6 * 0.4 * 0.2, rtol=1e-6)
    np.testing.assert_allclose(inv.signed_volume(), -0.6 * 0.4 * 0.2, rtol=1e-6)
    geo = oracle.texture_geometry(m.vertices, 0.05, (-0.1, 0.1), 16)
    w, h, d = geo["coarse_dims"]
    n, ce = (8 * w + 1, 8 * h + 1, 8 * d + 1), np.asarray(geo["cell_size"]).astype(np.float32)
    g = oracle.sdf_fine_grid(m.vertices, m.indices, geo["min_extents"], ce, n)
    gi = oracle.sdf_fine_grid(inv.vertices, inv.indices, geo["min_extents"], ce, n, winding_threshold=-0.5)
    wrong = oracle.sdf_fine_grid(inv.vertices, inv.indices, geo["min_extents"], ce, n, winding_threshold=0.5)
    np.testing.assert_allclose(gi, g, rtol=0, atol=1e-7)
    assert np.array_equal(np.sign(gi[np.abs(g) > 1e-6]), np.sign(g[np.abs(g) > 1e-6]))
    assert (g < 0).sum() > 0 and np.all(wrong >= 0)
    cook_cpu(inv, 16)
    cook_cpu(m, 16)
    np.testing.assert_allclose(inv.sdf.sparse_data()["coarse_sdf"], m.sdf.sparse_data()["coarse_sdf"], atol=1e-7)
    np.testing.assert_allclose(inv.sdf.sparse_data()["subgrid_data"], m.sdf.sparse_data()["subgrid_data"], atol=1e-7)


def _stacked_boxes(baked: bool):
    """Box A (0.2, 0.2, 0.05) at shape scale (1, 1, 2) -> half extents (0.2, 0.2, 0.1), static at the
    origin; unit-scale box B (0.1) on a body at z = 0.19, 1 cm into A's top face."""
    a = box_mesh(0.2, 0.2, 0.05)
    cook_cpu(a, 16, scale=(1.0, 1.0, 2.0) if baked else None)
    bm = cook_cpu(box_mesh(0.1, 0.1, 0.1), 16)
    b = ModelBuilder()
    cfg = ModelBuilder.ShapeConfig()
    cfg.gap = 0.005
    b.add_shape_mesh(-1, mesh=a, scale=(1.0, 1.0, 2.0), cfg=cfg)
    body = b.add_body(xform=(0.0, 0.0, 0.19, 0.0, 0.0, 0.0, 1.0))
    b.add_shape_mesh(body, mesh=bm, cfg=cfg)
    model = b.finalize("cpu")
    assert model.host["mesh_sdf_flags"][0] == (1 if baked else 0)
    return model


def test_baked_scale_sdf_depth():
    """Edges of B in the SDF of A (mode 1): with the scale baked into A's grid the depth is the true 1 cm;
    sampled through the unit grid it is the conservative min-scale estimate, 0.5 cm (sdf_contact.py:51-82,
    918-930)."""
    for baked, depth in ((True, -0.01), (False, -0.005)):
        model = _stacked_boxes(baked)
        o = oracle.Oracle(model)
        q = model.host["body_q"]
        c = o.collide(q)
        sub = c["sort_key"] & 0x7FFFFF
        mode1 = (sub >> 1) & 1 == 1
        assert mode1.any()
        d = _contact_depth(model, q.astype(np.float64), c)
        np.testing.assert_allclose(d[mode1].min(), depth, atol=2e-6)
        np.testing.assert_allclose(c["normal"][mode1][np.argmin(d[mode1])], [0.0, 0.0, 1.0], atol=1e-5)
