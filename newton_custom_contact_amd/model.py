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

        m._build_world_tables()
        for name, a in h.items():
            setattr(m, name, torch.from_numpy(np.ascontiguousarray(a)).to(m.device))
        m.shape_contact_pair_count = 0
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

        def csr(ids, n_world, what):
            start = np.zeros(n_world + 1, dtype=_I32)
            local = ids[ids >= 0]
            counts = np.bincount(local, minlength=n_world)[:n_world] if local.size else np.zeros(n_world, int)
            start[1:] = np.cumsum(counts)
            # entities of world w must be contiguous: [start[w], start[w+1])
            idx = np.nonzero(ids >= 0)[0]
            if idx.size and not np.all(np.diff(ids[idx]) >= 0):
                raise ValueError(f"{what} of each world must be contiguous and in world order")
            return start, idx

        self.world_body_start, body_idx = csr(bw, W, "bodies")
        self.world_shape_start, _ = csr(sw, W, "shapes")
        self.world_joint_start, _ = csr(jw, W, "joints")
        # bodies of world w are contiguous; their first global index
        first_body = np.full(W + 1, 0, dtype=_I32)
        for w in range(W):
            sel = np.nonzero(bw == w)[0]
            first_body[w] = sel[0] if sel.size else 0
        h["world_body_first"] = first_body
        sel = np.nonzero(sw >= 0)[0]
        first_shape = np.zeros(W + 1, dtype=_I32)
        for w in range(W):
            s = np.nonzero(sw == w)[0]
            first_shape[w] = s[0] if s.size else 0
        h["world_shape_first"] = first_shape
        first_joint = np.zeros(W + 1, dtype=_I32)
        for w in range(W):
            s = np.nonzero(jw == w)[0]
            first_joint[w] = s[0] if s.size else 0
        h["world_joint_first"] = first_joint
        h["world_body_count"] = np.diff(self.world_body_start).astype(_I32) if W else np.zeros(0, _I32)
        h["world_shape_count"] = np.diff(self.world_shape_start).astype(_I32) if W else np.zeros(0, _I32)
        h["world_joint_count"] = np.diff(self.world_joint_start).astype(_I32) if W else np.zeros(0, _I32)

        colliding = (h["shape_flags"] & int(ShapeFlags.COLLIDE_SHAPES)) != 0
        glob = np.nonzero(colliding & (sw == -1))[0].astype(_I32)
        seg_start = [0]
        seg_shape = []
        for w in range(W):
            own = np.nonzero(colliding & (sw == w))[0].astype(_I32)
            seg = np.sort(np.concatenate([own, glob]))
            seg_shape.append(seg)
            seg_start.append(seg_start[-1] + len(seg))
        seg_shape.append(glob)
        seg_start.append(seg_start[-1] + len(glob))
        h["segment_start"] = np.asarray(seg_start, dtype=_I32)
        h["segment_shape"] = (np.concatenate(seg_shape) if seg_shape else np.zeros(0)).astype(_I32)
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
        return s

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

            self._collision_pipeline = CollisionPipeline(self, broad_phase="sap")
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
        return set()


def geo_type_counts(model: Model) -> dict[str, int]:
    t = model.host["shape_type"]
    return {g.name: int(np.count_nonzero(t == int(g))) for g in GeoType}


def has_revolute_joints(model: Model) -> bool:
    return bool(np.any(model.host["joint_type"] == int(JointType.REVOLUTE)))
