"""Triangle meshes and their signed distance fields (newton/_src/geometry/types.py ``Mesh``, ``SDF``).

``Mesh`` keeps the reference's vertex / index layout, its unique-edge list (the contact sort key of a
mesh-mesh contact encodes the edge index, so the edge order follows ``Mesh.edges``,
types.py:831-866: canonical vertex ids from coordinates quantized at 1e-7, edges (0,1), (1,2), (0,2) of
each triangle interleaved, first occurrence of each canonical pair kept in order), its watertight test
(types.py:868-907) and solid mass properties (inertia.py:361-440: signed tetrahedra against the origin).

``Mesh.build_sdf`` cooks the reference's sparse narrow-band SDF texture on the GPU
(``SDF.create_from_mesh`` / ``create_texture_sdf_from_mesh`` / ``build_sparse_sdf_from_mesh``,
sdf_utils.py:346-545, sdf_texture.py:959-1505): the mesh AABB padded by ``margin``, ``max_resolution`` cells
along its longest axis and per-axis cell sizes that make the grid span the padded box exactly; a coarse grid
at the corners of 8^3-cell subgrids; the subgrids whose centre lies within the narrow band (widened by the
subgrid's half diagonal) and that the coarse grid does not already reproduce within 1e-6 of the box diagonal
keep their 9^3 samples in a texture of float32, unorm16 or unorm8 texels.  The host arithmetic of the grid
geometry is the reference's numpy arithmetic, operation for operation (float32 extents, float64 cell sizes),
so the same mesh gives the same grid.  Cooked textures can be kept on disk in the reference's cache format
(``cache_dir``, ``_sdf_cache``).
"""

from __future__ import annotations

import ctypes

import numpy as np
import torch

SUBGRID_SIZE = 8
"""Cells per subgrid edge (the reference's default ``subgrid_size``, the only one ``Mesh.build_sdf`` uses)."""
SLOT_EMPTY = 0xFFFFFFFF
"""Indirection slot of a far-field subgrid (sampled from the coarse grid)."""
SLOT_LINEAR = 0xFFFFFFFE
"""Indirection slot of a subgrid demoted to the coarse grid by the linearity test."""
QUANTIZATION_MODES = {"float32": 4, "uint16": 2, "uint8": 1}
"""``texture_format`` -> the reference's ``QuantizationMode`` value (= bytes per texel)."""


def texture_sdf_geometry(points, *, margin: float = 0.05, narrow_band_range=(-0.1, 0.1),
                         max_resolution: int | None = None, target_voxel_size: float | None = None) -> dict | None:
    """Grid geometry of a texture SDF cook, or ``None`` for a degenerate (point-sized) mesh.

    The host arithmetic of ``create_texture_sdf_from_mesh`` (sdf_texture.py:1436-1480) and
    ``build_sparse_sdf_from_mesh`` (:1017-1160, :1240-1244) on float32 ``points``: extents in float32,
    ``cells + 1`` vertices per axis with float64 cell sizes ``extent / cells``, ``w, h, d`` subgrids (ceiling
    division), the narrow band widened by the subgrid half diagonal, the linearity threshold 1e-6 times the
    padded box diagonal, the unorm quantisation range and the padded upper corner.
    """
    pts = np.asarray(points, np.float32).reshape(-1, 3)
    margin = float(margin)
    min_ext = pts.min(axis=0) - margin
    max_ext = pts.max(axis=0) + margin
    ext = max_ext - min_ext
    max_ext_scalar = np.max(ext)
    if max_ext_scalar < 1e-10:
        return None
    if target_voxel_size is not None:
        if target_voxel_size <= 0.0:
            raise ValueError("target_voxel_size must be > 0")
        derived = int(np.ceil(max_ext_scalar / float(target_voxel_size)))
        max_resolution = max(8, ((derived + 7) // 8) * 8)
    elif max_resolution is None:
        max_resolution = 64
    max_resolution = int(max_resolution)
    if max_resolution <= 0:
        raise ValueError("max_resolution must be > 0")
    if max_resolution >= (1 << 16):
        raise ValueError(f"max_resolution must be less than {1 << 16}")
    cell_size_scalar = max_ext_scalar / max_resolution
    dims = np.ceil(ext / cell_size_scalar).astype(int) + 1
    cell_size = ext / (dims - 1)
    nb = max(abs(float(narrow_band_range[0])), abs(float(narrow_band_range[1])))
    w, h, d = ((int(n) - 1 + SUBGRID_SIZE - 1) // SUBGRID_SIZE for n in dims)
    subgrid_radius = float(np.linalg.norm(SUBGRID_SIZE * 0.5 * cell_size))
    extents = max_ext - min_ext
    sdf_min = -nb - subgrid_radius
    sdf_range = (nb + subgrid_radius) - sdf_min
    if sdf_range < 1e-10:
        sdf_range = 1.0
    return {
        "min_extents": min_ext, "max_extents_unpadded": max_ext, "cell_size": cell_size, "dims": dims,
        "max_resolution": max_resolution, "coarse_dims": (w, h, d), "narrow_band_thickness": nb,
        "band": (np.float32(-nb - subgrid_radius), np.float32(nb + subgrid_radius)),
        "linear_threshold": float(1e-6 * np.linalg.norm(extents)),
        "sdf_min": sdf_min, "sdf_range": sdf_range,
        "padded_max": min_ext + np.array([w, h, d], dtype=float) * SUBGRID_SIZE * cell_size,
    }


def texture_blocks_per_axis(num_required: int) -> int:
    """Texture blocks per axis for ``num_required`` subgrids (sdf_texture.py:1181-1184: the ceiling of the
    float cube root, raised until its cube holds them all)."""
    b = max(1, int(np.ceil(num_required ** (1.0 / 3.0))))
    while b ** 3 < num_required:
        b += 1
    return b


class SDF:
    """The reference's texture SDF of a mesh in its local frame (``TextureSDFData`` + the cooked arrays).

    Device tensors: ``coarse`` (float32, (d+1, h+1, w+1)), ``slots`` (int32 holding the uint32 slots,
    (w, h, d)), ``texels`` (uint8 bytes of the (ts, ts, ts) texture, float32 / uint16 / uint8 samples).  The
    sampling parameters are those ``create_sparse_sdf_textures`` (sdf_texture.py:1311-1367) derives.
    """

    def __init__(self, sparse: dict, device, *, scale_baked: bool = False, shape_margin: float = 0.0,
                 narrow_band_range=(-0.1, 0.1), texture_format: str = "uint16", device_arrays=None,
                 device_flags=None):
        self.scale_baked = bool(scale_baked)
        self.shape_margin = float(shape_margin)
        self.narrow_band_range = (float(narrow_band_range[0]), float(narrow_band_range[1]))
        self.texture_format = texture_format
        self._sparse = sparse
        self._flags = device_flags  # occupied | linear << 1 | required << 2 per subgrid, until sparse_data()
        w, h, d = (int(x) for x in sparse["coarse_dims"])
        self.coarse_dims = (w, h, d)
        self.subgrid_size = int(sparse["subgrid_size"])
        if self.subgrid_size != SUBGRID_SIZE:
            raise ValueError(f"only subgrid_size {SUBGRID_SIZE} is supported, got {self.subgrid_size}")
        self.subgrid_tex_size = int(sparse["subgrid_tex_size"])
        self.num_subgrids = int(sparse["num_subgrids"])
        self.quantization_mode = int(sparse["quantization_mode"])
        cell = np.asarray(sparse["cell_size"], np.float64).reshape(3)
        self.sdf_box_lower = np.asarray(sparse["min_extents"], np.float64).reshape(3).astype(np.float32)
        self.sdf_box_upper = np.asarray(sparse["max_extents"], np.float64).reshape(3).astype(np.float32)
        self.inv_sdf_dx = np.array([1.0 / float(c) for c in cell], np.float32)
        self.voxel_size = cell.astype(np.float32)
        self.voxel_radius = np.float32(0.5 * np.linalg.norm(cell))
        self.subgrids_min_sdf_value = np.float32(sparse["subgrids_min_sdf_value"])
        self.subgrids_sdf_value_range = np.float32(sparse["subgrids_sdf_value_range"])
        if device_arrays is not None:
            self.coarse, self.slots, self.texels, self.texel_bytes = device_arrays
        else:
            data = np.ascontiguousarray(sparse["subgrid_data"])
            self.texel_bytes = int(data.dtype.itemsize)
            if data.dtype not in (np.float32, np.uint16, np.uint8):
                raise ValueError(f"subgrid_data must be float32, uint16 or uint8, got {data.dtype}")
            self.coarse = torch.from_numpy(np.ascontiguousarray(sparse["coarse_sdf"], np.float32)).to(device)
            slots = np.ascontiguousarray(sparse["subgrid_start_slots"], np.uint32).view(np.int32)
            self.slots = torch.from_numpy(slots.copy()).to(device)
            self.texels = torch.from_numpy(data.reshape(-1).view(np.uint8).copy()).to(device)

    def sparse_data(self) -> dict:
        """The cooked arrays as the reference's ``build_sparse_sdf_from_mesh`` dictionary (host numpy)."""
        if self._sparse.get("subgrid_data") is None:
            ts = self.subgrid_tex_size
            dt = {4: np.float32, 2: np.uint16, 1: np.uint8}[self.texel_bytes]
            sp = dict(self._sparse)
            sp["coarse_sdf"] = self.coarse.cpu().numpy()
            sp["subgrid_start_slots"] = self.slots.cpu().numpy().view(np.uint32)
            sp["subgrid_data"] = self.texels.cpu().numpy().view(dt).reshape(ts, ts, ts)
            if self._flags is not None:
                fl = self._flags.cpu().numpy()
                sp["subgrid_required"] = ((fl >> 2) & 1).astype(np.int32)
                sp["subgrid_occupied"] = (fl & 1).astype(np.int32)
                self._flags = None
            self._sparse = sp
        return self._sparse

    @property
    def params(self) -> dict:
        """Per-mesh table entries of the narrow phase (box, inverse cell size, quantisation, dimensions)."""
        w, h, d = self.coarse_dims
        return {
            "box": np.concatenate([self.sdf_box_lower, self.sdf_box_upper, self.inv_sdf_dx]).astype(np.float32),
            "quant": np.array([self.subgrids_min_sdf_value, self.subgrids_sdf_value_range], np.float32),
            "dims": np.array([w, h, d, self.subgrid_tex_size, self.texel_bytes], np.int32),
        }


class Mesh:
    def __init__(self, vertices, indices, compute_inertia: bool = True, is_solid: bool = True):
        self.vertices = np.ascontiguousarray(np.asarray(vertices, np.float32).reshape(-1, 3))
        self.indices = np.ascontiguousarray(np.asarray(indices, np.int32).reshape(-1))
        if self.indices.size % 3:
            raise ValueError("indices must hold whole triangles")
        self.is_solid = is_solid
        self._edges = None
        self.sdf: SDF | None = None
        self._is_watertight = None
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

    @property
    def is_watertight(self) -> bool:
        """Every geometric edge shared by exactly two triangles (types.py:868-907: vertices identified by
        their coordinates rounded at 1e-7)."""
        if self._is_watertight is None:
            tris = self.indices.reshape(-1, 3)
            if len(tris) == 0 or len(self.vertices) == 0:
                self._is_watertight = False
                return False
            q = np.round(self.vertices * np.float32(1e7)).astype(np.int64)
            _, canon = np.unique(q, axis=0, return_inverse=True)
            c = canon.reshape(-1)[tris]
            pairs = np.empty((3 * len(tris), 2), np.int64)
            for k, (a, b) in enumerate(((0, 1), (1, 2), (0, 2))):
                pairs[k::3, 0] = np.minimum(c[:, a], c[:, b])
                pairs[k::3, 1] = np.maximum(c[:, a], c[:, b])
            _, counts = np.unique(pairs, axis=0, return_counts=True)
            self._is_watertight = bool(np.all(counts == 2))
        return self._is_watertight

    def build_sdf(self, *, device=None, narrow_band_range: tuple[float, float] | None = None,
                  target_voxel_size: float | None = None, max_resolution: int | None = None,
                  margin: float | None = None, shape_margin: float = 0.0,
                  scale: tuple[float, float, float] | None = None, texture_format: str = "uint16",
                  cache_dir=None) -> SDF:
        """Cook and attach the reference's sparse texture SDF on the GPU.

        Mirrors ``Mesh.build_sdf`` (types.py:719-797) and ``SDF.create_from_mesh`` (sdf_utils.py:346-545):

        * ``max_resolution`` (default 64) cells along the longest axis of the AABB padded by ``margin``
          (default 0.05 m), or the resolution ``target_voxel_size`` implies (it takes precedence);
        * ``narrow_band_range`` (default (-0.1, 0.1) m) selects the subgrids that keep their samples;
        * ``texture_format`` ``"uint16"`` (default), ``"float32"`` or ``"uint8"`` stores them;
        * ``scale`` is multiplied into the vertices before the cook and the SDF is marked ``scale_baked``, so
          the narrow phase samples it with unit scale;
        * inside is decided by the generalized winding number against +0.5, or -0.5 when the mesh's signed
          volume is negative (inward winding).  The reference resolves ``sign_method="auto"`` to ray parity for
          watertight meshes, which classifies every point of a watertight mesh the same way; the resolved
          name enters the cache key as there;
        * ``shape_margin`` is recorded (the reference applies it at sample time, outside the contact path);
        * ``cache_dir`` keeps cooked textures on disk as ``{hash}.sdf.npz`` in the reference's format and
          key (``_sdf_cache``): a hit loads the arrays instead of cooking, a miss cooks and publishes them.

        Raises ``RuntimeError`` when the mesh already has an SDF or ``device`` is not a GPU.
        """
        from . import _sdf_cache

        if self.sdf is not None:
            raise RuntimeError("Mesh already has an SDF. Call clear_sdf() before rebuilding.")
        if texture_format not in QUANTIZATION_MODES:
            raise ValueError(f"Unknown texture_format {texture_format!r}. "
                             f"Expected one of {list(QUANTIZATION_MODES)}.")
        band = (-0.1, 0.1) if narrow_band_range is None else tuple(float(x) for x in narrow_band_range)
        margin = 0.05 if margin is None else float(margin)
        dev = torch.device("cuda", torch.cuda.current_device()) if device is None else torch.device(device)
        if dev.type != "cuda":
            raise RuntimeError("Mesh.build_sdf runs on the GPU only")
        if len(self.indices) == 0:
            raise ValueError("Mesh.build_sdf needs at least one triangle")
        bake = scale is not None
        sc = (1.0, 1.0, 1.0) if scale is None else tuple(float(x) for x in scale)
        eff_res = 64 if max_resolution is None and target_voxel_size is None else max_resolution
        meta = {"scale_baked": bake, "shape_margin": shape_margin, "narrow_band_range": band,
                "texture_format": texture_format}
        key = None
        if cache_dir is not None:
            key = _sdf_cache.hash_inputs(
                vertices=np.asarray(self.vertices, np.float32) * np.array(sc, np.float32),
                indices=np.asarray(self.indices, np.int32).reshape(-1), is_solid=bool(self.is_solid),
                narrow_band_range=band, target_voxel_size=target_voxel_size, max_resolution=eff_res,
                margin=margin, texture_format=texture_format,
                sign_method_resolved="parity" if self.is_watertight else "winding", winding_threshold=0.5,
                scale=scale)
            hit = _sdf_cache.try_load_sparse_data(cache_dir, key)
            if hit is not None:
                self.sdf = SDF(hit, dev, **meta)
                return self.sdf
        verts = (self.vertices * np.array(sc)[None, :]).astype(np.float32)
        geo = texture_sdf_geometry(verts, margin=margin, narrow_band_range=band,
                                   max_resolution=64 if eff_res is None else eff_res,
                                   target_voxel_size=target_voxel_size)
        if geo is None:
            raise ValueError("Mesh.build_sdf: the mesh has no extent")
        threshold = 0.5 if self.signed_volume(sc) >= 0.0 else -0.5
        self.sdf = cook_texture_sdf(verts, self.indices, geo, texture_format, threshold, dev, scale_baked=bake,
                                    shape_margin=shape_margin, narrow_band_range=band)
        if key is not None:
            _sdf_cache.write(cache_dir, key, self.sdf.sparse_data())
        return self.sdf

    def clear_sdf(self) -> None:
        """Detach the attached SDF (types.py:799-805)."""
        self.sdf = None

    def attach_sdf(self, sparse: dict, device="cpu", scale_baked: bool = False) -> SDF:
        """Attach a texture SDF cooked elsewhere (a ``build_sparse_sdf_from_mesh`` dictionary, for example one
        the CPU restatement in the tests produced)."""
        self.sdf = SDF(sparse, device, scale_baked=scale_baked)
        return self.sdf


def _cook_start(L, vertices, indices, geo: dict, winding_threshold: float, dev, st, count_slot: torch.Tensor) -> dict:
    """Queue the first two cook calls of one mesh (no host synchronisation): the fine grid, then the coarse
    grid, flags and addresses; the required count lands in ``count_slot`` (one device int)."""
    from ._native import check, ptr

    w, h, d = geo["coarse_dims"]
    nx, ny, nz = (SUBGRID_SIZE * n + 1 for n in (w, h, d))
    if nx * ny * nz >= 2 ** 31:
        raise ValueError(f"SDF grid of {nx} x {ny} x {nz} vertices is too large")
    lower = np.asarray(geo["min_extents"], np.float32)
    cell = np.asarray(geo["cell_size"], np.float64).astype(np.float32)
    vt = torch.from_numpy(np.ascontiguousarray(vertices, np.float32).reshape(-1, 3)).to(dev, non_blocking=True)
    tr = torch.from_numpy(np.ascontiguousarray(indices, np.int32).reshape(-1)).to(dev, non_blocking=True)
    fine = torch.empty((nz, ny, nx), dtype=torch.float32, device=dev)
    check(L.nmc_sdf_build(ctypes.c_void_p(ptr(vt)), int(vt.shape[0]), ctypes.c_void_p(ptr(tr)), int(tr.numel() // 3),
                          (ctypes.c_float * 3)(*lower.tolist()), (ctypes.c_float * 3)(*cell.tolist()),
                          (ctypes.c_int * 3)(nx, ny, nz), ctypes.c_float(winding_threshold),
                          ctypes.c_void_p(ptr(fine)), st), "nmc_sdf_build")
    n = w * h * d
    job = {"geo": geo, "lower": lower, "fine": fine, "vt": vt, "tr": tr,
           "coarse": torch.empty((d + 1, h + 1, w + 1), dtype=torch.float32, device=dev),
           "flags": torch.empty(n, dtype=torch.int32, device=dev),
           "address": torch.empty(n, dtype=torch.int32, device=dev)}
    lo, hi = geo["band"]
    check(L.nmc_sdf_subgrids(ctypes.c_void_p(ptr(fine)), (ctypes.c_int * 3)(w, h, d), ctypes.c_float(float(lo)),
                             ctypes.c_float(float(hi)), ctypes.c_float(geo["linear_threshold"]),
                             ctypes.c_void_p(ptr(job["coarse"])), ctypes.c_void_p(ptr(job["flags"])),
                             ctypes.c_void_p(ptr(job["address"])), ctypes.c_void_p(ptr(count_slot)), st),
          "nmc_sdf_subgrids")
    return job


def _cook_finish(L, job: dict, num_required: int, texture_format: str, dev, st, **meta) -> SDF:
    """Size the texture from the required count and queue the pack call (sdf_texture.py:1163-1288)."""
    from ._native import check, ptr

    geo = job["geo"]
    w, h, d = geo["coarse_dims"]
    qmode = QUANTIZATION_MODES[texture_format]
    slots = torch.empty((w, h, d), dtype=torch.int32, device=dev)
    if num_required == 0:
        # no subgrid keeps samples: a 1-texel float32 texture, unit quantisation (sdf_texture.py:1174-1179)
        blocks, tbytes, ts, qmin, qrange = 1, 4, 1, 0.0, 1.0
        texels = torch.zeros(729 * 4, dtype=torch.uint8, device=dev)
    else:
        blocks = texture_blocks_per_axis(num_required)
        ts = 9 * blocks
        tbytes = qmode
        qmin, qrange = (geo["sdf_min"], geo["sdf_range"]) if qmode != 4 else (0.0, 1.0)
        texels = torch.zeros(ts ** 3 * tbytes, dtype=torch.uint8, device=dev)
    check(L.nmc_sdf_pack(ctypes.c_void_p(ptr(job["fine"])), (ctypes.c_int * 3)(w, h, d),
                         ctypes.c_void_p(ptr(job["flags"])), ctypes.c_void_p(ptr(job["address"])), int(blocks),
                         int(tbytes), ctypes.c_float(qmin if tbytes != 4 else 0.0),
                         ctypes.c_float(1.0 / geo["sdf_range"]), ctypes.c_void_p(ptr(slots)),
                         ctypes.c_void_p(ptr(texels)), st), "nmc_sdf_pack")
    if num_required == 0:
        texels = texels[:4]
    sparse = {
        "coarse_sdf": None, "subgrid_data": None, "subgrid_start_slots": None,
        "coarse_dims": (w, h, d), "subgrid_tex_size": ts, "num_subgrids": num_required,
        "min_extents": job["lower"], "max_extents": geo["padded_max"],
        "cell_size": np.asarray(geo["cell_size"], np.float64), "subgrid_size": SUBGRID_SIZE,
        "quantization_mode": qmode, "subgrids_min_sdf_value": qmin, "subgrids_sdf_value_range": qrange,
        "subgrid_required": None, "subgrid_occupied": None,
    }
    return SDF(sparse, dev, texture_format=texture_format, device_arrays=(job["coarse"], slots, texels, tbytes),
               device_flags=job["flags"], **meta)


def cook_texture_sdf(vertices, indices, geo: dict, texture_format: str, winding_threshold: float, device, *,
                     scale_baked: bool = False, shape_margin: float = 0.0, narrow_band_range=(-0.1, 0.1)) -> SDF:
    """The GPU cook of ``build_sparse_sdf_from_mesh`` (sdf_texture.py:959-1308) for float32 ``vertices`` and
    the grid geometry ``geo`` of :func:`texture_sdf_geometry`: ``nmc_sdf_build`` (padded fine grid),
    ``nmc_sdf_subgrids`` (coarse grid, occupancy, linearity, addresses), the count of required subgrids read
    back to size the texture, ``nmc_sdf_pack`` (slots and texels)."""
    from ._native import load, stream_handle

    L = load()
    dev = torch.device(device)
    st = ctypes.c_void_p(stream_handle(dev))
    count = torch.zeros(1, dtype=torch.int32, device=dev)
    job = _cook_start(L, vertices, indices, geo, winding_threshold, dev, st, count)
    return _cook_finish(L, job, int(count.item()), texture_format, dev, st, scale_baked=scale_baked,
                        shape_margin=shape_margin, narrow_band_range=narrow_band_range)


def cook_texture_sdfs(meshes, device, *, max_resolution: int = 64, texture_format: str = "uint16",
                      margin: float = 0.05, narrow_band_range=(-0.1, 0.1), chunk: int = 512) -> None:
    """``Mesh.build_sdf`` for many meshes at once (no scale, no cache): the first two cook calls of up to
    ``chunk`` meshes are queued back to back, their required counts read back in one transfer, then the pack
    calls queued; the same arrays as cooking them one by one."""
    from ._native import load, stream_handle

    L = load()
    dev = torch.device(device)
    if dev.type != "cuda":
        raise RuntimeError("cook_texture_sdfs runs on the GPU only")
    st = ctypes.c_void_p(stream_handle(dev))
    band = tuple(float(x) for x in narrow_band_range)
    todo = [m for m in meshes if m.sdf is None]
    for c0 in range(0, len(todo), chunk):
        part = todo[c0:c0 + chunk]
        counts = torch.zeros(len(part), dtype=torch.int32, device=dev)
        jobs = []
        for k, m in enumerate(part):
            verts = np.ascontiguousarray(m.vertices, np.float32)
            geo = texture_sdf_geometry(verts, margin=margin, narrow_band_range=band, max_resolution=max_resolution)
            if geo is None:
                raise ValueError("cook_texture_sdfs: a mesh has no extent")
            thr = 0.5 if m.signed_volume() >= 0.0 else -0.5
            jobs.append(_cook_start(L, verts, m.indices, geo, thr, dev, st, counts[k:k + 1]))
        nreq = counts.cpu().numpy()
        for m, job, n in zip(part, jobs, nreq):
            m.sdf = _cook_finish(L, job, int(n), texture_format, dev, st, narrow_band_range=band)
            job.clear()


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
    a, b, c = verts[tris[:, 0]], verts[tris[:, 1]], verts[tris[:, 2]]
    inward = np.einsum("ij,ij->i", np.cross(b - a, c - a), a - center) < 0
    tris[inward] = tris[inward][:, [0, 2, 1]]  # orient every face outward
    return Mesh(verts.astype(np.float32), tris.astype(np.int32).reshape(-1))


def random_convex_hull(rng: np.random.Generator, n_min: int = 64, n_max: int = 256, r_min: float = 0.1,
                       r_max: float = 0.3) -> Mesh:
    """Seeded convex hull of n ~ U[n_min, n_max] points on a sphere of radius U[r_min, r_max]."""
    n = int(rng.integers(n_min, n_max + 1))
    r = float(rng.uniform(r_min, r_max))
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return convex_hull_mesh(d * r)


__all__ = ["Mesh", "SDF", "convex_hull_mesh", "cook_texture_sdf", "cook_texture_sdfs", "random_convex_hull",
           "texture_sdf_geometry"]
