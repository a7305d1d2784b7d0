"""Triangle meshes and their signed distance fields (newton/_src/geometry/types.py ``Mesh``, ``SDF``).

``Mesh`` keeps the reference's vertex / index layout, its unique-edge list (the contact sort key of a
mesh-mesh contact encodes the edge index, so the edge order follows ``Mesh.edges``,
types.py:831-866: canonical vertex ids from coordinates quantized at 1e-7, edges (0,1), (1,2), (0,2) of
each triangle interleaved, first occurrence of each canonical pair kept in order) and solid mass
properties (inertia.py:361-440: signed tetrahedra against the origin).

``Mesh.build_sdf`` cooks a dense FP32 signed distance grid on the GPU (``nmc_sdf_build``).  Grid
geometry follows the reference's cook (sdf_texture.py:959-1370): the mesh AABB padded by ``margin``
(0.05 m), ``max_resolution`` cells along the longest axis, the other axes with the same voxel size;
values are the distance to the nearest triangle, negative inside (generalized winding number above 0.5,
or below -0.5 for an inward-wound mesh).  The reference stores the same field as a sparse narrow-band
texture (coarse grid + 8^3 subgrids, optionally quantized); here every grid vertex holds its exact FP32
value (64^3 fp32 = 1 MiB per mesh).  Cooked grids can be kept on disk (``cache_dir``, ``_sdf_cache``).
"""

from __future__ import annotations

import ctypes

import numpy as np
import torch


class SDF:
    """Dense signed distance grid of a mesh in its local frame (unscaled unless ``scale_baked``)."""

    def __init__(self, grid: torch.Tensor, lower, voxel_size: float, cells, *, scale_baked: bool = False,
                 shape_margin: float = 0.0, narrow_band_range=(-0.1, 0.1), texture_format: str = "float32"):
        self.grid = grid                    # (nz + 1, ny + 1, nx + 1) float32, x fastest
        self.lower = np.asarray(lower, np.float32)
        self.voxel_size = float(voxel_size)
        self.cells = np.asarray(cells, np.int32)
        self.upper = (self.lower.astype(np.float64) + self.cells * self.voxel_size).astype(np.float32)
        self.scale_baked = bool(scale_baked)
        self.shape_margin = float(shape_margin)
        self.narrow_band_range = (float(narrow_band_range[0]), float(narrow_band_range[1]))
        self.texture_format = texture_format

    @property
    def shape(self):
        return tuple(int(c) + 1 for c in self.cells[::-1])


class Mesh:
    def __init__(self, vertices, indices, compute_inertia: bool = True, is_solid: bool = True):
        self.vertices = np.ascontiguousarray(np.asarray(vertices, np.float32).reshape(-1, 3))
        self.indices = np.ascontiguousarray(np.asarray(indices, np.int32).reshape(-1))
        if self.indices.size % 3:
            raise ValueError("indices must hold whole triangles")
        self.is_solid = is_solid
        self._edges = None
        self.sdf: SDF | None = None
        self.has_inertia = False
        if compute_inertia:
            self.mass, self.com, self.inertia, self.volume = self.mass_properties(1.0)
            self.has_inertia = True

    # ------------------------------------------------------------------ topology
    @property
    def edges(self) -> np.ndarray:
        """Unique edges (N, 2) in the reference's order (see module docstring)."""
        if self._edges is None:
            tris = self.indices.reshape(-1, 3)
            if len(tris) == 0:
                self._edges = np.zeros((0, 2), np.int32)
                return self._edges
            q = np.round(self.vertices * np.float32(1e7)).astype(np.int64)  # FP32 product, as the reference
            _, canon = np.unique(q, axis=0, return_inverse=True)
            canon = canon.reshape(-1)
            c = canon[tris]
            n = len(tris)
            key = np.empty((3 * n, 2), np.int64)
            orig = np.empty((3 * n, 2), np.int32)
            for k, (a, b) in enumerate(((0, 1), (1, 2), (0, 2))):
                key[k::3, 0] = np.minimum(c[:, a], c[:, b])
                key[k::3, 1] = np.maximum(c[:, a], c[:, b])
                orig[k::3, 0] = tris[:, a]
                orig[k::3, 1] = tris[:, b]
            _, first = np.unique(key, axis=0, return_index=True)
            first.sort()
            self._edges = np.ascontiguousarray(orig[first])
        return self._edges

    def aabb(self, scale=(1.0, 1.0, 1.0)):
        """AABB of the scaled vertices (the bounding-sphere input, geometry/utils.py:86-97)."""
        v = self.vertices.astype(np.float64) * np.asarray(scale, np.float64)
        return v.min(0), v.max(0)

    def local_aabb(self, scale=(1.0, 1.0, 1.0)):
        """Shape-local collision AABB with the scale baked in (builder.py:9978-9985)."""
        sc = np.asarray(scale, np.float64)
        lo = self.vertices.min(axis=0) * sc
        hi = self.vertices.max(axis=0) * sc
        return np.minimum(lo, hi), np.maximum(lo, hi)

    def shape_mass_properties(self, density: float, scale=(1.0, 1.0, 1.0)):
        """(mass, com, inertia) of a shape made of this mesh (inertia.py:595-622): the unit-density
        properties rescaled when they are known, else computed from the scaled vertices."""
        if self.has_inertia and self.mass > 0.0:
            sx, sy, sz = (float(x) for x in scale)
            ratio = sx * sy * sz * density
            I = self.inertia
            m = self.mass * ratio
            c = np.asarray(self.com, np.float64) * np.array([sx, sy, sz])
            Ixx = I[0, 0] * (sy ** 2 + sz ** 2) / 2 * ratio
            Iyy = I[1, 1] * (sx ** 2 + sz ** 2) / 2 * ratio
            Izz = I[2, 2] * (sx ** 2 + sy ** 2) / 2 * ratio
            Ixy = I[0, 1] * sx * sy * ratio
            Ixz = I[0, 2] * sx * sz * ratio
            Iyz = I[1, 2] * sy * sz * ratio
            return m, c, np.array([[Ixx, Ixy, Ixz], [Ixy, Iyy, Iyz], [Ixz, Iyz, Izz]])
        m, c, inertia, _vol = self.mass_properties(density, scale)
        return m, c, inertia

    # ------------------------------------------------------------------ mass properties
    def mass_properties(self, density: float, scale=(1.0, 1.0, 1.0)):
        """(mass, com, inertia about the com, volume) of the solid mesh (inertia.py:361-440)."""
        v = self.vertices.astype(np.float64) * np.asarray(scale, np.float64)
        t = self.indices.reshape(-1, 3)
        a, b, c = v[t[:, 0]], v[t[:, 1]], v[t[:, 2]]
        det = np.einsum("ij,ij->i", a, np.cross(b, c))  # 6 x signed tetra volume
        vol = det.sum() / 6.0
        first = (det[:, None] * (a + b + c)).sum(0) / 24.0
        # second moment of each tetra (0, a, b, c): det/120 * (sum_i x_i x_i^T + (sum_i x_i)(sum_i x_i)^T)
        s = a + b + c
        S = (det[:, None, None] / 120.0 * (np.einsum("ni,nj->nij", a, a) + np.einsum("ni,nj->nij", b, b) +
                                            np.einsum("ni,nj->nij", c, c) + np.einsum("ni,nj->nij", s, s))).sum(0)
        if vol < 0:
            vol, first, S = -vol, -first, -S
        mass = density * vol
        com = first / vol if vol > 0 else first
        S = S * density
        I_origin = np.trace(S) * np.eye(3) - S
        I_com = I_origin - mass * (np.dot(com, com) * np.eye(3) - np.outer(com, com))
        return mass, com, I_com, vol

    # ------------------------------------------------------------------ SDF
    def signed_volume(self, scale=(1.0, 1.0, 1.0)) -> float:
        """Signed volume of the closed surface: positive for outward winding (sdf_utils.py:618-636)."""
        v = self.vertices.astype(np.float64) * np.asarray(scale, np.float64)
        t = self.indices.reshape(-1, 3)
        return float(np.einsum("ij,ij->i", v[t[:, 0]], np.cross(v[t[:, 1]], v[t[:, 2]])).sum() / 6.0)

    def sdf_grid_geometry(self, max_resolution: int = 64, margin: float = 0.05, scale=(1.0, 1.0, 1.0)):
        """(lower, voxel size, cells) of the grid: the AABB of the (scaled) vertices padded by ``margin``,
        ``max_resolution`` cells along its longest axis and the same voxel size along the others."""
        lo, hi = self.aabb(scale)
        lo = lo - margin
        hi = hi + margin
        ext = hi - lo
        voxel = float(ext.max()) / float(max_resolution)
        cells = np.maximum(1, np.ceil(ext / voxel - 1e-9)).astype(np.int32)
        return lo.astype(np.float32), np.float32(voxel), cells

    def sdf_resolution(self, max_resolution: int | None = None, target_voxel_size: float | None = None,
                       margin: float = 0.05, scale=(1.0, 1.0, 1.0)) -> int:
        """Effective resolution along the longest padded axis: ``target_voxel_size`` takes precedence
        (ceil(extent / size) rounded up to a multiple of 8, at least 8, sdf_texture.py:1448-1456), else
        ``max_resolution`` (64 when both are None)."""
        if target_voxel_size is not None:
            if not target_voxel_size > 0.0:
                raise ValueError("target_voxel_size must be > 0")
            lo, hi = self.aabb(scale)
            ext = float(((hi + margin) - (lo - margin)).max())
            res = int(np.ceil(ext / float(target_voxel_size)))
            return max(8, ((res + 7) // 8) * 8)
        res = 64 if max_resolution is None else int(max_resolution)
        if res <= 0:
            raise ValueError("max_resolution must be > 0")
        if res % 8:
            raise ValueError("max_resolution must be divisible by 8")
        return res

    def build_sdf(self, *, device=None, narrow_band_range: tuple[float, float] | None = None,
                  target_voxel_size: float | None = None, max_resolution: int | None = None,
                  margin: float | None = None, shape_margin: float = 0.0,
                  scale: tuple[float, float, float] | None = None, texture_format: str = "uint16",
                  cache_dir=None) -> SDF:
        """Cook and attach the dense signed distance grid on the GPU (``nmc_sdf_build``).

        Mirrors ``Mesh.build_sdf`` (types.py:719-797) and ``SDF.create_from_mesh`` (sdf_utils.py:346-545):

        * ``max_resolution`` (default 64, divisible by 8) cells along the longest axis of the AABB padded by
          ``margin`` (default 0.05 m), or the resolution ``target_voxel_size`` implies (it takes precedence);
        * ``scale`` is multiplied into the vertices before the cook and the SDF is marked ``scale_baked``, so
          the narrow phase samples it with unit scale;
        * inside is decided by the generalized winding number against +0.5, or -0.5 when the mesh's signed
          volume is negative (inward winding);
        * ``cache_dir`` keeps cooked grids on disk as ``{hash}.sdf.npz`` (``_sdf_cache``): a hit loads the
          grid instead of cooking, a miss cooks and publishes it;
        * ``shape_margin``, ``narrow_band_range`` and ``texture_format`` are validated and recorded.  They
          select the reference's sparse, quantized texture storage (and a sample-time offset its contact
          path does not read, sdf_utils.py:236-256); every vertex of the dense grid holds its FP32 distance.

        Raises ``RuntimeError`` when the mesh already has an SDF or ``device`` is not a GPU.
        """
        from . import _sdf_cache
        from ._native import check, load, ptr, stream_handle

        if self.sdf is not None:
            raise RuntimeError("Mesh already has an SDF. Call clear_sdf() before rebuilding.")
        valid_formats = ("float32", "uint16", "uint8")
        if texture_format not in valid_formats:
            raise ValueError(f"Unknown texture_format {texture_format!r}. Expected one of {list(valid_formats)}.")
        band = (-0.1, 0.1) if narrow_band_range is None else tuple(float(x) for x in narrow_band_range)
        if len(band) != 2 or not band[0] < band[1]:
            raise ValueError("narrow_band_range must be (inner, outer) with inner < outer")
        margin = 0.05 if margin is None else float(margin)
        dev = torch.device("cuda", torch.cuda.current_device()) if device is None else torch.device(device)
        if dev.type != "cuda":
            raise RuntimeError("Mesh.build_sdf runs on the GPU only")
        if len(self.indices) == 0:
            raise ValueError("Mesh.build_sdf needs at least one triangle")
        bake = scale is not None
        sc = (1.0, 1.0, 1.0) if scale is None else tuple(float(x) for x in scale)
        res = self.sdf_resolution(max_resolution, target_voxel_size, margin, sc)
        lower, voxel, cells = self.sdf_grid_geometry(res, margin, sc)
        nx, ny, nz = (int(c) + 1 for c in cells)
        if nx * ny * nz >= 2**31:
            raise ValueError(f"SDF grid of {nx} x {ny} x {nz} vertices is too large")
        meta = {"scale_baked": bake, "shape_margin": shape_margin, "narrow_band_range": band,
                "texture_format": texture_format}
        verts_f32 = (self.vertices.astype(np.float64) * np.asarray(sc, np.float64)).astype(np.float32)
        key = None
        if cache_dir is not None:
            key = _sdf_cache.hash_inputs(vertices=verts_f32, indices=self.indices, is_solid=self.is_solid,
                                         max_resolution=res, margin=margin, scale=scale)
            hit = _sdf_cache.try_load_sdf_data(cache_dir, key)
            if hit is not None and bool(hit["scale_baked"]) == bake:
                grid = torch.from_numpy(hit["grid"]).to(dev)
                self.sdf = SDF(grid, hit["lower"], float(hit["voxel_size"]), hit["cells"], **meta)
                return self.sdf
        threshold = 0.5 if self.signed_volume(sc) >= 0.0 else -0.5
        verts = torch.from_numpy(verts_f32).to(dev)
        tris = torch.from_numpy(self.indices).to(dev)
        grid = torch.empty((nz, ny, nx), dtype=torch.float32, device=dev)
        L = load()
        lo = (ctypes.c_float * 3)(*[float(x) for x in lower])
        ce = (ctypes.c_int * 3)(*[int(x) for x in cells])
        check(L.nmc_sdf_build(ctypes.c_void_p(ptr(verts)), int(len(verts_f32)), ctypes.c_void_p(ptr(tris)),
                              int(self.indices.size // 3), lo, ctypes.c_float(float(voxel)), ce,
                              ctypes.c_float(threshold), ctypes.c_void_p(ptr(grid)),
                              ctypes.c_void_p(stream_handle(dev))), "nmc_sdf_build")
        self.sdf = SDF(grid, lower, voxel, cells, **meta)
        if key is not None:
            _sdf_cache.write(cache_dir, key, {"grid": grid.cpu().numpy(), "lower": lower, "voxel_size": voxel,
                                              "cells": cells, "scale_baked": bake})
        return self.sdf

    def clear_sdf(self) -> None:
        """Detach the attached SDF (types.py:799-805)."""
        self.sdf = None

    def attach_sdf(self, grid, lower, voxel_size, cells, scale_baked: bool = False):
        """Attach a grid cooked elsewhere (for example by the CPU restatement in the tests)."""
        self.sdf = SDF(grid, lower, voxel_size, cells, scale_baked=scale_baked)
        return self.sdf


def convex_hull_mesh(points) -> Mesh:
    """Triangulated convex hull of a point set, outward winding (scipy.spatial.ConvexHull)."""
    from scipy.spatial import ConvexHull

    pts = np.asarray(points, np.float64)
    hull = ConvexHull(pts)
    used = np.unique(hull.simplices)
    remap = -np.ones(len(pts), np.int64)
    remap[used] = np.arange(len(used))
    verts = pts[used]
    tris = remap[hull.simplices]
    center = verts.mean(0)
    for t in tris:  # orient every face outward
        a, b, c = verts[t]
        if np.dot(np.cross(b - a, c - a), a - center) < 0:
            t[1], t[2] = t[2], t[1]
    return Mesh(verts.astype(np.float32), tris.astype(np.int32).reshape(-1))


def random_convex_hull(rng: np.random.Generator, n_min: int = 64, n_max: int = 256, r_min: float = 0.1,
                       r_max: float = 0.3) -> Mesh:
    """Seeded convex hull of n ~ U[n_min, n_max] points on a sphere of radius U[r_min, r_max]."""
    n = int(rng.integers(n_min, n_max + 1))
    r = float(rng.uniform(r_min, r_max))
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return convex_hull_mesh(d * r)


__all__ = ["Mesh", "SDF", "convex_hull_mesh", "random_convex_hull"]
