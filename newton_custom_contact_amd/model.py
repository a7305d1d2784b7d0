"""``Model`` / ``State`` / ``Control``: structure-of-arrays device storage for the contact step.

Mirrors newton/_src/sim/model.py (``Model``, line 25; ``state``/``control``/``contacts``/``collide``
at 827-1022), newton/_src/sim/state.py (``State``, line 47; ``clear_forces`` 156, ``assign`` 169) and
newton/_src/sim/control.py (``Control``, line 9).

All arrays are torch tensors on ``model.device`` with the reference's field names and shapes:
transforms are ``(N, 7)`` float32 ``(p, q_xyzw)``, spatial vectors ``(N, 6)`` float32
``(linear, angular)``, 3x3 matrices ``(N, 3, 3)`` float32 row-major.  ``model.host`` keeps the same
arrays as numpy for host-side consumers.

Besides the reference's fields the model carries the per-world CSR tables the per-world kernels
use (``world_body_start``, ``world_shape_start``, ``world_joint_start``, ``segment_start`` /
``segment_shape``), computed once at finalize.
"""

from __future__ import annotations

import numpy as np
import torch

from ._types import GeoType, JointType, ShapeFlags


def _dev(device) -> torch.device:
    if device is None:
        return torch.device("cuda:0") if torch.cuda.is_available() else torch.device("cpu")
    return torch.device(device)


class State:
    """Time-varying body state (state.py:47): ``body_q`` (B,7), ``body_qd`` (B,6), ``body_f`` (B,6)."""

    def __init__(self):
        self.body_q: torch.Tensor | None = None
        self.body_qd: torch.Tensor | None = None
        self.body_f: torch.Tensor | None = None
        self.body_parent_f: torch.Tensor | None = None
        self.joint_q: torch.Tensor | None = None
        self.joint_qd: torch.Tensor | None = None
        self.requires_grad = False

    @property
    def body_count(self) -> int:
        return 0 if self.body_q is None else int(self.body_q.shape[0])

    def clear_forces(self):
        """state.py:156."""
        if self.body_f is not None:
            self.body_f.zero_()

    def assign(self, other: State):
        """state.py:169: copy every array of ``other`` into this state."""
        for name in ("body_q", "body_qd", "body_f", "joint_q", "joint_qd"):
            a = getattr(self, name)
            b = getattr(other, name)
            if a is not None and b is not None:
                a.copy_(b)


class Control:
    """Control inputs (control.py:9): ``joint_f`` and joint targets."""

    def __init__(self):
        self.joint_f: torch.Tensor | None = None
        self.joint_target_pos: torch.Tensor | None = None
        self.joint_target_vel: torch.Tensor | None = None

    def clear(self):
        if self.joint_f is not None:
            self.joint_f.zero_()


_F32 = np.float32
_I32 = np.int32


class Model:
    """Static scene description on the device (model.py:25)."""

    def __init__(self, device=None):
        self.device = _dev(device)
        self.host: dict[str, np.ndarray] = {}
        self.requires_grad = False
        self.rigid_contact_max = 0
        self.body_count = 0
        self.shape_count = 0
        self.joint_count = 0
        self.joint_dof_count = 0
        self.joint_coord_count = 0
        self.world_count = 0
        self.particle_count = 0
        self.shape_collision_filter_pairs: set[tuple[int, int]] = set()
        self._collision_pipeline = None
        self._shape_contact_pairs = None
        self._requested_state_attributes: set[str] = set()
        self._requested_contact_attributes: set[str] = set()

    # ------------------------------------------------------------------ construction
    @classmethod
    def _from_builder(cls, b, device) -> Model:
        m = cls(device)
        h = m.host
        B, S, J = b.body_count, b.shape_count, b.joint_count
        m.body_count, m.shape_count, m.joint_count = B, S, J
        m.joint_dof_count = b.joint_dof_count
        m.joint_coord_count = b.joint_coord_count
        m.world_count = b.world_count
        m.up_axis = b.up_axis

        def arr(x, dtype, shape):
            a = np.asarray(x, dtype=dtype)
            return a.reshape(shape) if a.size else np.zeros(shape, dtype=dtype)

        h["body_q"] = arr(b.body_q, _F32, (B, 7))
        h["body_qd"] = arr(b.body_qd, _F32, (B, 6))
        h["body_com"] = arr(b.body_com, _F32, (B, 3))
        h["body_mass"] = arr(b.body_mass, _F32, (B,))
        h["body_inv_mass"] = arr(b.body_inv_mass, _F32, (B,))
        h["body_inertia"] = arr(b.body_inertia, _F32, (B, 3, 3))
        h["body_inv_inertia"] = arr(b.body_inv_inertia, _F32, (B, 3, 3))
        h["body_flags"] = arr(b.body_flags, _I32, (B,))
        h["body_world"] = arr(b.body_world, _I32, (B,))

        h["shape_transform"] = arr(b.shape_transform, _F32, (S, 7))
        h["shape_body"] = arr(b.shape_body, _I32, (S,))
        h["shape_type"] = arr(b.shape_type, _I32, (S,))
        h["shape_scale"] = arr(b.shape_scale, _F32, (S, 3))
        h["shape_flags"] = arr(b.shape_flags, _I32, (S,))
        h["shape_margin"] = arr(b.shape_margin, _F32, (S,))
        h["shape_gap"] = arr(b.shape_gap, _F32, (S,))
        h["shape_collision_radius"] = arr(b.shape_collision_radius, _F32, (S,))
        h["shape_collision_group"] = arr(b.shape_collision_group, _I32, (S,))
        h["shape_world"] = arr(b.shape_world, _I32, (S,))
        h["shape_material_mu"] = arr(b.shape_material_mu, _F32, (S,))
        h["shape_material_mu_torsional"] = arr(b.shape_material_mu_torsional, _F32, (S,))
        h["shape_material_mu_rolling"] = arr(b.shape_material_mu_rolling, _F32, (S,))
        h["shape_material_restitution"] = arr(b.shape_material_restitution, _F32, (S,))

        D = b.joint_dof_count
        h["joint_type"] = arr(b.joint_type, _I32, (J,))
        h["joint_enabled"] = arr(b.joint_enabled, _I32, (J,))
        h["joint_parent"] = arr(b.joint_parent, _I32, (J,))
        h["joint_child"] = arr(b.joint_child, _I32, (J,))
        h["joint_X_p"] = arr(b.joint_X_p, _F32, (J, 7))
        h["joint_X_c"] = arr(b.joint_X_c, _F32, (J, 7))
        h["joint_q_start"] = arr(b.joint_q_start, _I32, (J,))
        h["joint_qd_start"] = arr(b.joint_qd_start, _I32, (J,))
        h["joint_dof_dim"] = arr(b.joint_dof_dim, _I32, (J, 2))
        h["joint_world"] = arr(b.joint_world, _I32, (J,))
        h["joint_axis"] = arr(b.joint_axis, _F32, (len(b.joint_axis), 3))
        h["joint_limit_lower"] = arr(b.joint_limit_lower, _F32, (len(b.joint_limit_lower),))
        h["joint_limit_upper"] = arr(b.joint_limit_upper, _F32, (len(b.joint_limit_upper),))
        h["joint_target_ke"] = arr(b.joint_target_ke, _F32, (len(b.joint_target_ke),))
        h["joint_target_kd"] = arr(b.joint_target_kd, _F32, (len(b.joint_target_kd),))
        h["joint_target_pos"] = arr(b.joint_target_pos, _F32, (len(b.joint_target_pos),))
        h["joint_target_vel"] = arr(b.joint_target_vel, _F32, (len(b.joint_target_vel),))
        h["joint_q"] = arr(b.joint_q, _F32, (b.joint_coord_count,))
        h["joint_qd"] = arr(b.joint_qd, _F32, (D,))
        h["joint_f"] = arr(b.joint_f, _F32, (D,))

        gv = list(b.world_gravity)
        if not gv:
            u = b.up_vector
            gv = [(u[0] * b.gravity, u[1] * b.gravity, u[2] * b.gravity)]
        h["gravity"] = arr(gv, _F32, (len(gv), 3))

        m.shape_collision_filter_pairs = set(b.shape_collision_filter_pairs)
        fp = sorted(m.shape_collision_filter_pairs)
        h["shape_filter_pairs"] = arr(fp, _I32, (len(fp), 2))
        m._mesh_tables(list(getattr(b, "shape_source", [None] * S)), [tuple(x) for x in b.shape_scale])
        m._finish()
        return m

    _COUNTS = ("body_count", "shape_count", "joint_count", "joint_dof_count", "joint_coord_count", "world_count")

    def save(self, path: str) -> None:
        """The host arrays and counts of a model without mesh shapes, as one .npz (plain arrays; a model built on
        the host of one process can be handed to another)."""
        if getattr(self, "meshes", []):
            raise ValueError("Model.save: models with mesh shapes are not supported")
        meta = np.asarray([getattr(self, k) for k in self._COUNTS], np.int64)
        up = np.asarray([str(getattr(self, "up_axis", "Z"))])
        np.savez(path, _counts=meta, _up_axis=up, **self.host)

    @classmethod
    def load(cls, path: str, device=None) -> Model:
        """Inverse of Model.save: the arrays are read without unpickling and the device tables rebuilt."""
        m = cls(device)
        with np.load(path, allow_pickle=False) as z:
            for k, v in zip(cls._COUNTS, z["_counts"].tolist()):
                setattr(m, k, int(v))
            m.up_axis = str(z["_up_axis"][0])
            m.host = {k: z[k] for k in z.files if not k.startswith("_")}
        fp = m.host["shape_filter_pairs"].astype(np.int64)
        m.shape_collision_filter_pairs = {(int(a), int(b)) for a, b in fp}
        m.meshes = []
        m._finish()
        return m

    @staticmethod
    def voxel_resolution(lower, upper, budget: int = 100) -> tuple[int, int, int]:
        """Voxel grid of a shape's contact reduction (compute_voxel_resolution_from_aabb, builder.py:9936-9961):
        about cubic voxels, at most ``budget`` (NUM_VOXEL_DEPTH_SLOTS), shrinking the largest axis first."""
        size = np.maximum(np.asarray(upper, np.float64) - np.asarray(lower, np.float64), 1e-6)
        volume = size[0] * size[1] * size[2]
        v = max((volume / budget) ** (1.0 / 3.0), 1e-6)
        nx, ny, nz = (max(1, round(float(x) / v)) for x in size)
        while nx * ny * nz > budget:
            if nx >= ny and nx >= nz and nx > 1:
                nx -= 1
            elif ny >= nz and ny > 1:
                ny -= 1
            elif nz > 1:
                nz -= 1
            else:
                break
        return nx, ny, nz

    def _voxel_table(self, mesh_aabb: dict):
        """Per-shape voxel resolution of the contact reduction (builder.py:9924-10065) for the mesh shapes, from
        their local AABB (scaled vertices, in float64 as the builder computes it).  Only mesh pairs are reduced
        in this build, so the other shapes keep 1 x 1 x 1."""
        res = np.ones((self.shape_count, 3), _I32)
        cache = {}
        for s, (lo, hi) in mesh_aabb.items():
            key = (tuple(np.asarray(lo, np.float64)), tuple(np.asarray(hi, np.float64)))
            if key not in cache:
                cache[key] = self.voxel_resolution(lo, hi)
            res[s] = cache[key]
        self.host["shape_voxel_resolution"] = res

    def _mesh_tables(self, sources: list, scales: list | None = None):
        """Per-shape mesh index and local AABB (scale baked, collide.py:215-245), and the concatenated
        mesh vertex / edge tables (builder.py:10316-10351)."""
        h = self.host
        S = self.shape_count
        meshes, index, mesh_aabb = [], {}, {}
        shape_mesh = np.full(S, -1, _I32)
        lo = np.zeros((S, 3), _F32)
        hi = np.zeros((S, 3), _F32)
        for s, src in enumerate(sources):
            if src is None:
                continue
            if id(src) not in index:
                index[id(src)] = len(meshes)
                meshes.append(src)
            shape_mesh[s] = index[id(src)]
            a, b_ = src.local_aabb(h["shape_scale"][s] if scales is None else scales[s])
            lo[s], hi[s] = a, b_
            mesh_aabb[s] = (a, b_)
        self.meshes = meshes
        h["shape_mesh"] = shape_mesh
        h["shape_aabb_lower"] = lo
        h["shape_aabb_upper"] = hi
        vs = [np.zeros((0, 3), _F32)] + [mm.vertices for mm in meshes]
        es = [np.zeros((0, 2), _I32)] + [mm.edges.astype(_I32) for mm in meshes]
        h["mesh_vertex"] = np.concatenate(vs).astype(_F32) if meshes else np.zeros((1, 3), _F32)
        h["mesh_vertex_start"] = np.concatenate([[0], np.cumsum([len(mm.vertices) for mm in meshes])]).astype(_I32)
        h["mesh_edge"] = np.concatenate(es).astype(_I32) if meshes else np.zeros((1, 2), _I32)
        h["mesh_edge_start"] = np.concatenate([[0], np.cumsum([len(mm.edges) for mm in meshes])]).astype(_I32)
        self._voxel_table(mesh_aabb)

    def _mesh_sdf_tables(self):
        """Concatenate the meshes' texture SDFs on the device (cooked on the GPU when missing): coarse grids,
        slot tables and texel bytes, with per-mesh offsets and sampling parameters."""
        meshes = getattr(self, "meshes", [])
        M = len(meshes)
        if M and self.device.type == "cuda":
            for mm in meshes:
                if mm.sdf is None:
                    mm.build_sdf(device=self.device)
        have = M > 0 and all(mm.sdf is not None for mm in meshes)
        if have:
            sdfs = [mm.sdf for mm in meshes]
            coarse = [sd.coarse.reshape(-1).to(self.device) for sd in sdfs]
            slots = [sd.slots.reshape(-1).to(self.device) for sd in sdfs]
            texels = [sd.texels.reshape(-1).to(self.device) for sd in sdfs]
            self.mesh_sdf_coarse = torch.cat(coarse)
            self.mesh_sdf_slots = torch.cat(slots)
            # texel offsets stay 4-byte aligned so float32 textures are read as whole words
            pad = [(-t.numel()) % 4 for t in texels]
            parts = []
            for t, p in zip(texels, pad):
                parts.append(t)
                if p:
                    parts.append(torch.zeros(p, dtype=torch.uint8, device=self.device))
            self.mesh_sdf_texels = torch.cat(parts)
            starts = np.zeros((M, 3), np.int64)
            starts[:, 0] = np.concatenate([[0], np.cumsum([c.numel() for c in coarse])[:-1]])
            starts[:, 1] = np.concatenate([[0], np.cumsum([c.numel() for c in slots])[:-1]])
            starts[:, 2] = np.concatenate([[0], np.cumsum([t.numel() + p for t, p in zip(texels, pad)])[:-1]])
            box = np.stack([sd.params["box"] for sd in sdfs]).astype(_F32)
            quant = np.stack([sd.params["quant"] for sd in sdfs]).astype(_F32)
            dims = np.stack([sd.params["dims"] for sd in sdfs]).astype(_I32)
            flags = np.asarray([1 if sd.scale_baked else 0 for sd in sdfs], _I32)
        else:
            self.mesh_sdf_coarse = torch.zeros(1, dtype=torch.float32, device=self.device)
            self.mesh_sdf_slots = torch.zeros(1, dtype=torch.int32, device=self.device)
            self.mesh_sdf_texels = torch.zeros(4, dtype=torch.uint8, device=self.device)
            k = max(M, 1)
            starts = np.zeros((k, 3), np.int64)
            box = np.zeros((k, 9), _F32)
            quant = np.zeros((k, 2), _F32)
            dims = np.ones((k, 5), _I32)
            flags = np.zeros(k, _I32)
        self.mesh_has_sdf = have
        for k, v in (("mesh_sdf_start", starts), ("mesh_sdf_box", box), ("mesh_sdf_quant", quant),
                     ("mesh_sdf_dims", dims), ("mesh_sdf_flags", flags)):
            self.host[k] = v
            setattr(self, k, torch.from_numpy(np.ascontiguousarray(v)).to(self.device))

    def _finish(self):
        """Derive the per-world tables and upload every host array to the device."""
        self._build_world_tables()
        if "shape_mesh" not in self.host:
            self._mesh_tables([None] * self.shape_count)
        for name, a in self.host.items():
            setattr(self, name, torch.from_numpy(np.ascontiguousarray(a)).to(self.device))
        self._mesh_sdf_tables()
        self._shape_contact_pairs = None

    @property
    def shape_contact_pairs(self) -> torch.Tensor:
        """Every shape pair the collision pipeline may test, ``(K, 2)`` int32 canonical (min, max), sorted
        (find_shape_contact_pairs, builder.py:10855-10913: colliding shapes, the world and group rules,
        minus the filter pairs).  Computed on first use; the explicit broad phase tests exactly these."""
        if self._shape_contact_pairs is None:
            self._shape_contact_pairs = torch.from_numpy(self._find_shape_contact_pairs()).to(self.device)
        return self._shape_contact_pairs

    @property
    def shape_contact_pair_count(self) -> int:
        return int(self.shape_contact_pairs.shape[0])

    def _find_shape_contact_pairs(self) -> np.ndarray:
        h = self.host
        ss, sh = h["segment_start"], h["segment_shape"]
        sw, grp = h["shape_world"], h["shape_collision_group"]
        nseg = len(ss) - 1
        out = []
        for seg in range(nseg):
            ids = sh[ss[seg]:ss[seg + 1]]
            if len(ids) < 2:
                continue
            i, j = np.triu_indices(len(ids), 1)
            a, b = np.minimum(ids[i], ids[j]), np.maximum(ids[i], ids[j])
            wa, wb = sw[a], sw[b]
            ga, gb = grp[a], grp[b]
            keep = (wa == -1) & (wb == -1) if seg == nseg - 1 else ~((wa == -1) & (wb == -1))
            keep &= ~((wa != -1) & (wb != -1) & (wa != wb))
            # test_group_pair (broad_phase_common.py:132-150)
            keep &= (ga != 0) & (gb != 0) & np.where(ga > 0, (ga == gb) | (gb < 0), ga != gb)
            out.append(np.stack([a[keep], b[keep]], 1))
        pairs = np.concatenate(out).astype(np.int64) if out else np.zeros((0, 2), np.int64)
        fp = h["shape_filter_pairs"].astype(np.int64)
        if len(pairs) and len(fp):
            key = pairs[:, 0] * (1 << 32) + pairs[:, 1]
            pairs = pairs[~np.isin(key, fp[:, 0] * (1 << 32) + fp[:, 1])]
        pairs = np.unique(pairs, axis=0) if len(pairs) else pairs
        return np.ascontiguousarray(pairs, np.int32).reshape(-1, 2)

    @classmethod
    def _replicate(cls, template: Model, world_count: int, body_q: np.ndarray, device) -> Model:
        """Tile a one-world template ``world_count`` times (ModelBuilder.replicate without per-entity
        Python work).

        ``template`` holds global entities (world -1, ordered first) and exactly one world (world 0);
        ``body_q`` is ``(world_count, bodies_per_world, 7)``.  The result is identical to building the
        same worlds one by one with ``begin_world``/``add_body``/``add_shape_*``/``end_world``.
        """
        th = template.host
        W = int(world_count)
        if getattr(template, "meshes", []):
            raise ValueError("templates with mesh shapes are replicated with the ModelBuilder")
        if template.world_count != 1:
            raise ValueError("template must contain exactly one world")
        bw, sw, jw = th["body_world"], th["shape_world"], th["joint_world"]
        gb, gs, gj = np.nonzero(bw == -1)[0], np.nonzero(sw == -1)[0], np.nonzero(jw == -1)[0]
        wb, ws, wj = np.nonzero(bw == 0)[0], np.nonzero(sw == 0)[0], np.nonzero(jw == 0)[0]
        for g, w in ((gb, wb), (gs, ws), (gj, wj)):
            if g.size and w.size and g.max() > w.min():
                raise ValueError("template: global entities must precede the world's entities")
        nb, ns, nj = wb.size, ws.size, wj.size
        if body_q.shape != (W, nb, 7):
            raise ValueError(f"body_q must be ({W}, {nb}, 7)")
        m = cls(device)
        m.world_count = W
        m.up_axis = template.up_axis
        m.body_count = gb.size + W * nb
        m.shape_count = gs.size + W * ns
        m.joint_count = gj.size + W * nj
        D = len(th["joint_qd"])
        C = len(th["joint_q"])
        ndof_g = int(th["joint_qd_start"][wj[0]]) if nj else D
        ndof_w = D - ndof_g
        ncoord_g = int(th["joint_q_start"][wj[0]]) if nj else C
        ncoord_w = C - ncoord_g
        m.joint_dof_count = ndof_g + W * ndof_w
        m.joint_coord_count = ncoord_g + W * ncoord_w
        h = m.host
        wid_b = np.repeat(np.arange(W, dtype=_I32), nb)
        wid_s = np.repeat(np.arange(W, dtype=_I32), ns)
        wid_j = np.repeat(np.arange(W, dtype=_I32), nj)
        b_off = gb.size + np.arange(W, dtype=np.int64) * nb  # first body of each world

        def tile(name, sel_g, sel_w):
            a = th[name]
            return np.concatenate([a[sel_g], np.tile(a[sel_w], (W,) + (1,) * (a.ndim - 1))])

        for name in ("body_qd", "body_com", "body_mass", "body_inv_mass", "body_inertia", "body_inv_inertia",
                     "body_flags"):
            h[name] = tile(name, gb, wb)
        h["body_q"] = np.concatenate([th["body_q"][gb], body_q.reshape(W * nb, 7).astype(_F32)])
        h["body_world"] = np.concatenate([bw[gb], wid_b]).astype(_I32)
        for name in ("shape_transform", "shape_type", "shape_scale", "shape_flags", "shape_margin", "shape_gap",
                     "shape_collision_radius", "shape_collision_group", "shape_material_mu",
                     "shape_material_mu_torsional", "shape_material_mu_rolling", "shape_material_restitution"):
            h[name] = tile(name, gs, ws)
        sb = th["shape_body"]
        tb = sb[ws].astype(np.int64)
        shape_body_w = np.where(tb >= 0, (tb - wb[0] if nb else 0)[None, :] + b_off[:, None], -1) if ns else \
            np.zeros((W, 0), np.int64)
        h["shape_body"] = np.concatenate([sb[gs], shape_body_w.reshape(-1)]).astype(_I32)
        h["shape_world"] = np.concatenate([sw[gs], wid_s]).astype(_I32)
        s_off = gs.size + np.arange(W, dtype=np.int64) * ns
        # joints
        for name in ("joint_type", "joint_enabled", "joint_X_p", "joint_X_c", "joint_dof_dim"):
            h[name] = tile(name, gj, wj)
        jp, jc = th["joint_parent"], th["joint_child"]
        par_w = jp[wj].astype(np.int64)
        chi_w = jc[wj].astype(np.int64)
        h["joint_parent"] = np.concatenate(
            [jp[gj], np.where(par_w >= 0, (par_w - (wb[0] if nb else 0))[None, :] + b_off[:, None], -1).reshape(-1)]
        ).astype(_I32)
        h["joint_child"] = np.concatenate(
            [jc[gj], ((chi_w - (wb[0] if nb else 0))[None, :] + b_off[:, None]).reshape(-1)]).astype(_I32)
        h["joint_world"] = np.concatenate([jw[gj], wid_j]).astype(_I32)
        qds = th["joint_qd_start"].astype(np.int64)
        qs = th["joint_q_start"].astype(np.int64)
        dof_off = ndof_g + np.arange(W, dtype=np.int64) * ndof_w
        crd_off = ncoord_g + np.arange(W, dtype=np.int64) * ncoord_w
        h["joint_qd_start"] = np.concatenate(
            [qds[gj], ((qds[wj] - ndof_g)[None, :] + dof_off[:, None]).reshape(-1)]).astype(_I32)
        h["joint_q_start"] = np.concatenate(
            [qs[gj], ((qs[wj] - ncoord_g)[None, :] + crd_off[:, None]).reshape(-1)]).astype(_I32)
        for name in ("joint_axis", "joint_limit_lower", "joint_limit_upper", "joint_target_ke", "joint_target_kd",
                     "joint_target_pos", "joint_target_vel", "joint_qd", "joint_f"):
            a = th[name]
            h[name] = np.concatenate([a[:ndof_g], np.tile(a[ndof_g:], (W,) + (1,) * (a.ndim - 1))])
        jq = np.concatenate([th["joint_q"][:ncoord_g], np.tile(th["joint_q"][ncoord_g:], W)]).astype(_F32)
        # FREE joints carry the child body pose as coordinates
        free = np.nonzero(h["joint_type"] == int(JointType.FREE))[0]
        if free.size:
            st = h["joint_q_start"][free].astype(np.int64)
            for k in range(7):
                jq[st + k] = h["body_q"][h["joint_child"][free], k]
        h["joint_q"] = jq
        h["gravity"] = np.tile(th["gravity"][0:1], (W, 1)).astype(_F32)
        fp = th["shape_filter_pairs"].astype(np.int64)
        if fp.size:
            is_w = sw[fp] >= 0
            g_pairs = fp[~is_w.any(axis=1)]
            w_pairs = fp[is_w.all(axis=1)] - ws[0]
            mixed = fp[is_w.any(axis=1) & ~is_w.all(axis=1)]
            parts = [g_pairs]
            if w_pairs.size:
                parts.append((w_pairs[None, :, :] + s_off[:, None, None]).reshape(-1, 2))
            if mixed.size:
                mw = np.where(sw[mixed] >= 0, mixed - ws[0], -1)
                rep = np.where(mw[None] >= 0, mw[None] + s_off[:, None, None], mixed[None])
                parts.append(np.sort(rep, axis=2).reshape(-1, 2))
            allp = np.concatenate(parts).astype(_I32)
            allp = np.unique(allp, axis=0)
            h["shape_filter_pairs"] = allp
        else:
            h["shape_filter_pairs"] = np.zeros((0, 2), _I32)
        m.shape_collision_filter_pairs = {tuple(map(int, p)) for p in h["shape_filter_pairs"]} \
            if len(h["shape_filter_pairs"]) < 100000 else set()
        m._finish()
        return m

    def _build_world_tables(self):
        """Per-world CSR tables and the broad-phase segment map.

        The segment map follows ``precompute_world_map`` (broad_phase_common.py:183-300): segment ``w``
        holds world ``w``'s colliding shapes plus every colliding world ``-1`` shape, and a final
        dedicated segment holds only the world ``-1`` shapes.  Inside a segment shapes are listed in
        increasing global index, so ranks inside a segment order exactly like global shape ids.
        """
        h = self.host
        W = self.world_count
        bw = h["body_world"]
        sw = h["shape_world"]
        jw = h["joint_world"]

        def ranges(ids, what):
            """first index and count of each world's entities (entities of a world are contiguous)."""
            first = np.zeros(W + 1, dtype=_I32)
            count = np.zeros(W + 1, dtype=_I32)
            idx = np.nonzero(ids >= 0)[0]
            if idx.size:
                wid = ids[idx]
                if np.any(np.diff(wid) < 0):
                    raise ValueError(f"{what} of each world must be contiguous and in world order")
                cnt = np.bincount(wid, minlength=W)[:W]
                starts = np.searchsorted(wid, np.arange(W))
                first[:W] = np.where(cnt > 0, idx[np.minimum(starts, idx.size - 1)], 0)
                count[:W] = cnt
                # contiguity inside a world: last - first + 1 == count
                ends = np.where(cnt > 0, idx[np.minimum(starts + cnt - 1, idx.size - 1)], -1)
                if np.any((cnt > 0) & (ends - first[:W] + 1 != cnt)):
                    raise ValueError(f"{what} of a world must be contiguous")
            return first, count

        h["world_body_first"], h["world_body_count"] = ranges(bw, "bodies")
        h["world_shape_first"], h["world_shape_count"] = ranges(sw, "shapes")
        h["world_joint_first"], h["world_joint_count"] = ranges(jw, "joints")

        colliding = (h["shape_flags"] & int(ShapeFlags.COLLIDE_SHAPES)) != 0
        glob = np.nonzero(colliding & (sw == -1))[0].astype(_I32)
        own_idx = np.nonzero(colliding & (sw >= 0))[0].astype(_I32)
        own_w = sw[own_idx]
        # segment w = sorted(world w colliding shapes + global colliding shapes); last segment = globals
        order = np.lexsort((own_idx, own_w))
        own_idx, own_w = own_idx[order], own_w[order]
        own_cnt = np.bincount(own_w, minlength=W)[:W] if W else np.zeros(0, np.int64)
        seg_len = np.concatenate([own_cnt + len(glob), [len(glob)]]).astype(np.int64)
        seg_start = np.zeros(W + 2, dtype=np.int64)
        seg_start[1:] = np.cumsum(seg_len)
        seg_shape = np.empty(int(seg_start[-1]), dtype=_I32)
        if W:
            own_start = np.zeros(W + 1, np.int64)
            own_start[1:] = np.cumsum(own_cnt)
            # merged order inside a segment: position of every own shape and every global shape
            rank_own = np.arange(own_idx.size) - own_start[own_w]
            g_before_own = np.searchsorted(glob, own_idx)  # globals with smaller index
            seg_shape[seg_start[own_w] + rank_own + g_before_own] = own_idx
            if len(glob):
                for_w = np.repeat(np.arange(W), len(glob))
                gl = np.tile(glob, W)
                gr = np.tile(np.arange(len(glob)), W)
                # own colliding shapes of world w with a smaller index than global shape g: world shapes
                # occupy one contiguous index block, so a prefix count over colliding world shapes answers it
                pref = np.zeros(len(sw) + 1, np.int64)
                pref[1:] = np.cumsum(colliding & (sw >= 0))
                first = h["world_shape_first"][:W].astype(np.int64)
                own_before = pref[gl] - pref[first[for_w]]
                own_before = np.clip(own_before, 0, own_cnt[for_w])
                seg_shape[seg_start[for_w] + gr + own_before] = gl
        seg_shape[seg_start[W]:] = glob
        h["segment_start"] = seg_start.astype(_I32)
        h["segment_shape"] = seg_shape
        h["global_shape_count"] = np.asarray([len(glob)], dtype=_I32)
        self.global_dynamic_bodies = int(np.count_nonzero((bw == -1) & (h["body_inv_mass"] > 0.0)))

    # ------------------------------------------------------------------ factories
    def state(self, requires_grad: bool | None = None) -> State:
        """model.py:827-878: clone the initial body state, zero ``body_f``."""
        s = State()
        s.body_q = self.body_q.clone()
        s.body_qd = self.body_qd.clone()
        s.body_f = torch.zeros_like(self.body_qd)
        s.joint_q = self.joint_q.clone()
        s.joint_qd = self.joint_qd.clone()
        if "body_parent_f" in self._requested_state_attributes:
            s.body_parent_f = torch.zeros_like(self.body_qd)
        return s

    def request_state_attributes(self, *attributes: str) -> None:
        """model.py:1024-1034: optional State arrays (here: ``body_parent_f``)."""
        unknown = set(attributes) - {"body_parent_f"}
        if unknown:
            raise ValueError(f"unsupported state attributes: {sorted(unknown)}")
        self._requested_state_attributes.update(attributes)

    def get_requested_state_attributes(self) -> list[str]:
        return sorted(self._requested_state_attributes)

    def request_contact_attributes(self, *attributes: str) -> None:
        """model.py:1036-1044: optional Contacts arrays (here: ``force``, filled by update_contacts)."""
        from .contacts import Contacts  # noqa: PLC0415

        unknown = set(attributes) - set(Contacts.EXTENDED_ATTRIBUTES)
        if unknown:
            raise ValueError(f"unsupported contact attributes: {sorted(unknown)}")
        self._requested_contact_attributes.update(attributes)

    def control(self, clone_variables: bool = True) -> Control:
        """model.py:880-921."""
        c = Control()
        if clone_variables:
            c.joint_f = self.joint_f.clone()
            c.joint_target_pos = self.joint_target_pos.clone()
            c.joint_target_vel = self.joint_target_vel.clone()
        else:
            c.joint_f = self.joint_f
            c.joint_target_pos = self.joint_target_pos
            c.joint_target_vel = self.joint_target_vel
        return c

    def _default_pipeline(self):
        if self._collision_pipeline is None:
            from .collide import CollisionPipeline  # noqa: PLC0415

            self._collision_pipeline = CollisionPipeline(self, broad_phase="explicit")
        return self._collision_pipeline

    def contacts(self, collision_pipeline=None):
        """model.py:970-994."""
        p = collision_pipeline or self._default_pipeline()
        return p.contacts()

    def collide(self, state: State, contacts=None, *, collision_pipeline=None):
        """model.py:996-1022."""
        p = collision_pipeline or self._default_pipeline()
        if contacts is None:
            contacts = p.contacts()
        p.collide(state, contacts)
        return contacts

    def get_requested_contact_attributes(self) -> set[str]:
        return set(self._requested_contact_attributes)


def geo_type_counts(model: Model) -> dict[str, int]:
    t = model.host["shape_type"]
    return {g.name: int(np.count_nonzero(t == int(g))) for g in GeoType}


def has_revolute_joints(model: Model) -> bool:
    return bool(np.any(model.host["joint_type"] == int(JointType.REVOLUTE)))
