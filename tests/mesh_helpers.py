"""Mesh fixtures shared by the mesh / SDF tests: closed-form meshes and CPU-cooked SDFs."""

from __future__ import annotations

import numpy as np
import torch

from newton_custom_contact_amd.mesh import Mesh


def box_mesh(hx=0.5, hy=0.5, hz=0.5) -> Mesh:
    """Axis-aligned box, 8 vertices, 12 outward-wound triangles."""
    v = np.array([[-hx, -hy, -hz], [hx, -hy, -hz], [hx, hy, -hz], [-hx, hy, -hz],
                  [-hx, -hy, hz], [hx, -hy, hz], [hx, hy, hz], [-hx, hy, hz]], np.float32)
    t = np.array([0, 2, 1, 0, 3, 2,  # bottom (-z)
                  4, 5, 6, 4, 6, 7,  # top (+z)
                  0, 1, 5, 0, 5, 4,  # -y
                  2, 3, 7, 2, 7, 6,  # +y
                  1, 2, 6, 1, 6, 5,  # +x
                  3, 0, 4, 3, 4, 7], np.int32)  # -x
    return Mesh(v, t)


def icosphere(radius=0.5, subdivisions=3) -> Mesh:
    """Subdivided icosahedron projected on a sphere (outward winding)."""
    p = (1.0 + 5.0 ** 0.5) / 2.0
    v = [(-1, p, 0), (1, p, 0), (-1, -p, 0), (1, -p, 0), (0, -1, p), (0, 1, p), (0, -1, -p), (0, 1, -p),
         (p, 0, -1), (p, 0, 1), (-p, 0, -1), (-p, 0, 1)]
    v = [np.array(x, np.float64) / np.linalg.norm(x) for x in v]
    f = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2),
         (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11),
         (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    for _ in range(subdivisions):
        cache = {}

        def mid(a, b):
            k = (min(a, b), max(a, b))
            if k not in cache:
                m = v[a] + v[b]
                v.append(m / np.linalg.norm(m))
                cache[k] = len(v) - 1
            return cache[k]

        nf = []
        for a, b, c in f:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        f = nf
    return Mesh((np.array(v) * radius).astype(np.float32), np.array(f, np.int32).reshape(-1))


def cook_cpu(mesh: Mesh, max_resolution: int = 16, margin: float = 0.05, scale=None, texture_format="float32",
             narrow_band_range=(-0.1, 0.1)):
    """Attach a texture SDF cooked by the oracle (CPU tests have no GPU cook), with Mesh.build_sdf's geometry,
    winding threshold and scale baking."""
    import oracle

    sc = (1.0, 1.0, 1.0) if scale is None else scale
    verts = (mesh.vertices * np.array(sc)[None, :]).astype(np.float32)
    geo = oracle.texture_geometry(verts, margin, narrow_band_range, max_resolution)
    thr = 0.5 if mesh.signed_volume(sc) >= 0.0 else -0.5
    sparse = oracle.sdf_build_sparse(verts, mesh.indices, geo, texture_format, winding_threshold=thr)
    mesh.attach_sdf(sparse, "cpu", scale_baked=scale is not None)
    return mesh


def box_sdf(p: np.ndarray, h) -> np.ndarray:
    """Exact signed distance to an axis-aligned box of half extents h."""
    q = np.abs(p) - np.asarray(h, np.float64)
    outside = np.linalg.norm(np.maximum(q, 0.0), axis=-1)
    inside = np.minimum(np.max(q, axis=-1), 0.0)
    return outside + inside
