"""Scene construction: the subset of Newton's ``ModelBuilder`` that the rigid-contact step needs.

Mirrors newton/_src/sim/builder.py (class ``ModelBuilder``, line 88):

* ``ShapeConfig`` defaults                       builder.py:230-301
* ``rigid_gap`` = 0.1 m default contact gap      builder.py:1165
* ``add_link`` / ``add_body`` (+FREE joint)       builder.py:3552, 3649
* ``add_joint`` / ``add_joint_free`` / ``add_joint_revolute``  builder.py:3723, 4201, 3898
* ``add_shape`` (same-body and parent filters, mass accumulation)  builder.py:5449-5609
* ``add_shape_plane`` / ``add_ground_plane``     builder.py:5611, 5676
* ``add_shape_sphere`` / ``_box`` / ``_capsule``  builder.py:5703, 5842, 5892
* ``begin_world`` / ``end_world`` / ``add_world`` / ``add_builder`` / ``replicate``
                                                  builder.py:2815-2954, 2116
* ``_update_body_mass``                          builder.py:8534-8566
* ``finalize``                                   builder.py:9745

Host arithmetic (mass, inertia, inverse inertia) is done in float64 and stored as float32; the
reference does it with Warp host types.  Everything after ``finalize`` reads the stored float32
arrays, so the oracle and the HIP path consume identical inputs.
"""

from __future__ import annotations

import copy
import math
from dataclasses import dataclass, field

import numpy as np

from ._types import (
    Axis,
    BodyFlags,
    GeoType,
    JointType,
    ShapeFlags,
    as_transform,
    quat_between_vectors,
    quat_identity,
    transform,
    transform_inertia,
    transform_multiply,
    transform_point,
)

MAXVAL = 1.0e10  # newton/_src/core/types.py:72


def compute_shape_radius(geo_type: int, scale) -> float:
    """Bounding-sphere radius used by the AABB of infinite planes (geometry/utils.py:73-124)."""
    if geo_type == GeoType.SPHERE:
        return float(scale[0])
    if geo_type == GeoType.BOX:
        return float(np.linalg.norm(scale))
    if geo_type in (GeoType.CAPSULE, GeoType.CYLINDER, GeoType.CONE):
        return float(scale[0] + scale[1])
    if geo_type == GeoType.ELLIPSOID:
        return float(max(scale))
    if geo_type == GeoType.PLANE:
        if scale[0] > 0.0 and scale[1] > 0.0:
            return float(np.linalg.norm(scale) * 0.5)
        return 1.0e6
    return 10.0


def compute_inertia_shape(geo_type: int, scale, density: float):
    """Mass, COM and inertia about the COM of a solid primitive (geometry/inertia.py:42-244)."""
    if geo_type == GeoType.SPHERE:
        r = float(scale[0])
        m = density * (4.0 / 3.0 * math.pi * r * r * r)
        ia = 2.0 / 5.0 * m * r * r
        return m, np.zeros(3), np.diag([ia, ia, ia])
    if geo_type == GeoType.BOX:
        hx, hy, hz = (float(s) for s in scale)
        m = density * 8.0 * hx * hy * hz
        return (
            m,
            np.zeros(3),
            np.diag(
                [
                    1.0 / 3.0 * m * (hy * hy + hz * hz),
                    1.0 / 3.0 * m * (hx * hx + hz * hz),
                    1.0 / 3.0 * m * (hx * hx + hy * hy),
                ]
            ),
        )
    if geo_type == GeoType.CAPSULE:
        r = float(scale[0])
        h = 2.0 * float(scale[1])
        ms = density * (4.0 / 3.0) * math.pi * r * r * r
        mc = density * math.pi * r * r * h
        m = ms + mc
        ia = mc * (0.25 * r * r + (1.0 / 12.0) * h * h) + ms * (0.4 * r * r + 0.375 * r * h + 0.25 * h * h)
        ib = (mc * 0.5 + ms * 0.4) * r * r
        return m, np.zeros(3), np.diag([ia, ia, ib])
    if geo_type == GeoType.CYLINDER:
        # inertia.py:115-138 (axis +Z)
        r = float(scale[0])
        h = 2.0 * float(scale[1])
        m = density * math.pi * r * r * h
        ia = 1 / 12 * m * (3 * r * r + h * h)
        ib = 1 / 2 * m * r * r
        return m, np.zeros(3), np.diag([ia, ia, ib])
    if geo_type == GeoType.CONE:
        # inertia.py:141-169 (apex at +Z, base at -half_height; COM a quarter height above the base)
        r = float(scale[0])
        h = 2.0 * float(scale[1])
        m = density * math.pi * r * r * h / 3.0
        ia = 3 / 20 * m * r * r + 3 / 80 * m * h * h
        ib = 3 / 10 * m * r * r
        return m, np.array([0.0, 0.0, -h / 4.0]), np.diag([ia, ia, ib])
    if geo_type == GeoType.ELLIPSOID:
        # inertia.py:172-201 (semi-axes rx, ry, rz)
        rx, ry, rz = (float(v) for v in scale)
        m = density * ((4.0 / 3.0) * math.pi * rx * ry * rz)
        return m, np.zeros(3), np.diag([(1.0 / 5.0) * m * (ry * ry + rz * rz), (1.0 / 5.0) * m * (rx * rx + rz * rz),
                                        (1.0 / 5.0) * m * (rx * rx + ry * ry)])
    if geo_type == GeoType.PLANE:
        return 0.0, np.zeros(3), np.zeros((3, 3))
    raise NotImplementedError(f"inertia of GeoType {geo_type} is not part of the contact-step subset")


class ModelBuilder:
    """Builds a :class:`~newton_custom_contact_amd.Model` (reference: builder.py:88)."""

    @dataclass
    class ShapeConfig:
        """Shape material and collision properties (builder.py:230-301 defaults)."""

        density: float = 1000.0
        ke: float = 2.5e3
        kd: float = 100.0
        kf: float = 1000.0
        ka: float = 0.0
        mu: float = 1.0
        restitution: float = 0.0
        mu_torsional: float = 0.005
        mu_rolling: float = 0.0001
        margin: float = 0.0
        gap: float | None = None
        is_solid: bool = True
        collision_group: int = 1
        collision_filter_parent: bool = True
        has_shape_collision: bool = True
        has_particle_collision: bool = True
        is_visible: bool = True
        is_site: bool = False

        @property
        def flags(self) -> int:
            f = 0
            if self.is_visible:
                f |= int(ShapeFlags.VISIBLE)
            if self.has_shape_collision:
                f |= int(ShapeFlags.COLLIDE_SHAPES)
            if self.has_particle_collision:
                f |= int(ShapeFlags.COLLIDE_PARTICLES)
            if self.is_site:
                f |= int(ShapeFlags.SITE)
            return f

        def copy(self):
            return copy.copy(self)

    @dataclass
    class JointDofConfig:
        """One joint axis (builder.py:440-493)."""

        axis: object = Axis.X
        limit_lower: float = -MAXVAL
        limit_upper: float = MAXVAL
        limit_ke: float = 1e4
        limit_kd: float = 1e1
        target_pos: float = 0.0
        target_vel: float = 0.0
        target_ke: float = 0.0
        target_kd: float = 0.0
        armature: float = 0.0

        def axis_vec(self) -> np.ndarray:
            a = self.axis
            if isinstance(a, (Axis, str, int)) and not isinstance(a, (list, tuple, np.ndarray)):
                v = np.zeros(3)
                v[int(Axis.from_any(a))] = 1.0
                return v
            v = np.asarray(a, dtype=np.float64)
            return v / np.linalg.norm(v)

    def __init__(self, up_axis="Z", gravity: float = -9.81):
        self.up_axis = Axis.from_any(up_axis)
        self.gravity = float(gravity)
        self.rigid_gap = 0.1
        self.default_shape_cfg = ModelBuilder.ShapeConfig()
        self.world_count = 0
        self._current_world = -1
        self.world_gravity: list[tuple[float, float, float]] = []

        # bodies
        self.body_q: list[np.ndarray] = []
        self.body_qd: list[np.ndarray] = []
        self.body_mass: list[float] = []
        self.body_inv_mass: list[float] = []
        self.body_com: list[np.ndarray] = []
        self.body_inertia: list[np.ndarray] = []
        self.body_inv_inertia: list[np.ndarray] = []
        self.body_flags: list[int] = []
        self.body_world: list[int] = []
        self.body_lock_inertia: list[bool] = []
        self.body_label: list[str] = []
        self.body_shapes: dict[int, list[int]] = {-1: []}

        # shapes
        self.shape_body: list[int] = []
        self.shape_transform: list[np.ndarray] = []
        self.shape_type: list[int] = []
        self.shape_scale: list[tuple[float, float, float]] = []
        self.shape_flags: list[int] = []
        self.shape_margin: list[float] = []
        self.shape_gap: list[float] = []
        self.shape_material_mu: list[float] = []
        self.shape_material_mu_torsional: list[float] = []
        self.shape_material_mu_rolling: list[float] = []
        self.shape_material_restitution: list[float] = []
        self.shape_collision_group: list[int] = []
        self.shape_collision_radius: list[float] = []
        self.shape_world: list[int] = []
        self.shape_source: list = []  # Mesh of a MESH shape, else None
        self.shape_label: list[str] = []
        self.shape_collision_filter_pairs: set[tuple[int, int]] = set()

        # joints
        self.joint_type: list[int] = []
        self.joint_parent: list[int] = []
        self.joint_child: list[int] = []
        self.joint_X_p: list[np.ndarray] = []
        self.joint_X_c: list[np.ndarray] = []
        self.joint_dof_dim: list[tuple[int, int]] = []
        self.joint_enabled: list[bool] = []
        self.joint_world: list[int] = []
        self.joint_collision_filter_parent: list[bool] = []
        self.joint_q_start: list[int] = []
        self.joint_qd_start: list[int] = []
        self.joint_q: list[float] = []
        self.joint_qd: list[float] = []
        self.joint_f: list[float] = []
        self.joint_axis: list[np.ndarray] = []
        self.joint_limit_lower: list[float] = []
        self.joint_limit_upper: list[float] = []
        self.joint_target_pos: list[float] = []
        self.joint_target_vel: list[float] = []
        self.joint_target_ke: list[float] = []
        self.joint_target_kd: list[float] = []
        self.joint_parents: dict[int, list[tuple[int, int]]] = {}
        self.joint_children: dict[int, list[tuple[int, int]]] = {}
        self.articulation_start: list[int] = []
        self.articulation_world: list[int] = []

    # ------------------------------------------------------------------ properties
    @property
    def current_world(self) -> int:
        return self._current_world

    @property
    def up_vector(self) -> tuple[float, float, float]:
        v = [0.0, 0.0, 0.0]
        v[int(self.up_axis)] = 1.0
        return tuple(v)

    @property
    def body_count(self) -> int:
        return len(self.body_q)

    @property
    def shape_count(self) -> int:
        return len(self.shape_type)

    @property
    def joint_count(self) -> int:
        return len(self.joint_type)

    @property
    def joint_dof_count(self) -> int:
        return len(self.joint_qd)

    @property
    def joint_coord_count(self) -> int:
        return len(self.joint_q)

    # ------------------------------------------------------------------ worlds
    def begin_world(self, label: str | None = None, gravity=None):
        """builder.py:2815-2883."""
        if self._current_world != -1:
            raise RuntimeError(
                f"Cannot begin a new world: already in world context (current_world={self._current_world})."
            )
        self._current_world = self.world_count
        self.world_count += 1
        if gravity is not None:
            self.world_gravity.append(tuple(float(g) for g in gravity))
        else:
            u = self.up_vector
            self.world_gravity.append((u[0] * self.gravity, u[1] * self.gravity, u[2] * self.gravity))

    def end_world(self):
        """builder.py:2885-2907."""
        if self._current_world == -1:
            raise RuntimeError("Cannot end world: not currently in a world context.")
        self._current_world = -1

    def add_world(self, builder: ModelBuilder, xform=None, label_prefix: str | None = None):
        """builder.py:2909-2952."""
        self.begin_world()
        self.add_builder(builder, xform=xform, label_prefix=label_prefix)
        self.end_world()

    def replicate(self, builder: ModelBuilder, world_count: int, spacing=(0.0, 0.0, 0.0)):
        """builder.py:2116-2152: ``world_count`` copies, copy ``i`` offset by ``i * spacing``."""
        spacing = np.asarray(spacing, dtype=np.float64)
        for i in range(world_count):
            self.add_world(builder, xform=transform(spacing * i, quat_identity()))

    def add_builder(self, builder: ModelBuilder, xform=None, label_prefix: str | None = None):
        """Append every entity of ``builder`` into the current world (builder.py:2954-3240).

        Root bodies (those whose joint parent is the world, or which have no joint) are transformed by
        ``xform``; child bodies of articulations follow the same transform.
        """
        xform = as_transform(xform)
        body_off = self.body_count
        shape_off = self.shape_count
        joint_off = self.joint_count
        q_off = self.joint_coord_count
        qd_off = self.joint_dof_count
        world = self._current_world

        for i in range(builder.body_count):
            self.body_q.append(transform_multiply(xform, builder.body_q[i]))
            self.body_qd.append(builder.body_qd[i].copy())
            self.body_mass.append(builder.body_mass[i])
            self.body_inv_mass.append(builder.body_inv_mass[i])
            self.body_com.append(builder.body_com[i].copy())
            self.body_inertia.append(builder.body_inertia[i].copy())
            self.body_inv_inertia.append(builder.body_inv_inertia[i].copy())
            self.body_flags.append(builder.body_flags[i])
            self.body_world.append(world)
            self.body_lock_inertia.append(builder.body_lock_inertia[i])
            lbl = builder.body_label[i]
            self.body_label.append(f"{label_prefix}/{lbl}" if label_prefix else lbl)
            self.body_shapes[body_off + i] = [s + shape_off for s in builder.body_shapes.get(i, [])]

        for s in range(builder.shape_count):
            b = builder.shape_body[s]
            nb = b + body_off if b >= 0 else -1
            self.shape_body.append(nb)
            t = builder.shape_transform[s]
            if b < 0:
                t = transform_multiply(xform, t)
            self.shape_transform.append(t)
            self.shape_type.append(builder.shape_type[s])
            self.shape_scale.append(builder.shape_scale[s])
            self.shape_flags.append(builder.shape_flags[s])
            self.shape_margin.append(builder.shape_margin[s])
            self.shape_gap.append(builder.shape_gap[s])
            self.shape_material_mu.append(builder.shape_material_mu[s])
            self.shape_material_mu_torsional.append(builder.shape_material_mu_torsional[s])
            self.shape_material_mu_rolling.append(builder.shape_material_mu_rolling[s])
            self.shape_material_restitution.append(builder.shape_material_restitution[s])
            self.shape_collision_group.append(builder.shape_collision_group[s])
            self.shape_collision_radius.append(builder.shape_collision_radius[s])
            self.shape_source.append(builder.shape_source[s])
            self.shape_world.append(world)
            lbl = builder.shape_label[s]
            self.shape_label.append(f"{label_prefix}/{lbl}" if label_prefix else lbl)
            if b < 0:
                self.body_shapes[-1].append(shape_off + s)
        for a, b in builder.shape_collision_filter_pairs:
            self.shape_collision_filter_pairs.add((a + shape_off, b + shape_off))

        for j in range(builder.joint_count):
            p = builder.joint_parent[j]
            c = builder.joint_child[j]
            self.joint_type.append(builder.joint_type[j])
            self.joint_parent.append(p + body_off if p >= 0 else -1)
            self.joint_child.append(c + body_off)
            xp = builder.joint_X_p[j]
            if p < 0:
                xp = transform_multiply(xform, xp)
            self.joint_X_p.append(xp)
            self.joint_X_c.append(builder.joint_X_c[j].copy())
            self.joint_dof_dim.append(builder.joint_dof_dim[j])
            self.joint_enabled.append(builder.joint_enabled[j])
            self.joint_world.append(world)
            self.joint_collision_filter_parent.append(builder.joint_collision_filter_parent[j])
            self.joint_q_start.append(builder.joint_q_start[j] + q_off)
            self.joint_qd_start.append(builder.joint_qd_start[j] + qd_off)
            np_ = p + body_off if p >= 0 else -1
            self.joint_parents.setdefault(c + body_off, []).append((np_, joint_off + j))
            self.joint_children.setdefault(np_, []).append((c + body_off, joint_off + j))
        self.joint_q.extend(builder.joint_q)
        self.joint_qd.extend(builder.joint_qd)
        self.joint_f.extend(builder.joint_f)
        self.joint_axis.extend(a.copy() for a in builder.joint_axis)
        self.joint_limit_lower.extend(builder.joint_limit_lower)
        self.joint_limit_upper.extend(builder.joint_limit_upper)
        self.joint_target_pos.extend(builder.joint_target_pos)
        self.joint_target_vel.extend(builder.joint_target_vel)
        self.joint_target_ke.extend(builder.joint_target_ke)
        self.joint_target_kd.extend(builder.joint_target_kd)
        for a in builder.articulation_start:
            self.articulation_start.append(a + joint_off)
            self.articulation_world.append(world)

    # ------------------------------------------------------------------ bodies
    def add_link(
        self,
        xform=None,
        armature: float | None = None,
        com=None,
        inertia=None,
        mass: float = 0.0,
        label: str | None = None,
        lock_inertia: bool = False,
        is_kinematic: bool = False,
    ) -> int:
        """builder.py:3552-3647."""
        xform = as_transform(xform)
        com = np.zeros(3) if com is None else np.asarray(com, dtype=np.float64)
        inertia = np.zeros((3, 3)) if inertia is None else np.asarray(inertia, dtype=np.float64).reshape(3, 3)
        if armature:
            inertia = inertia + np.eye(3) * armature
        body_id = self.body_count
        self.body_inertia.append(inertia)
        self.body_mass.append(float(mass))
        self.body_com.append(com)
        self.body_lock_inertia.append(lock_inertia)
        self.body_flags.append(int(BodyFlags.KINEMATIC) if is_kinematic else int(BodyFlags.DYNAMIC))
        self.body_inv_mass.append(1.0 / mass if mass > 0.0 else 0.0)
        self.body_inv_inertia.append(np.linalg.inv(inertia) if np.any(inertia) else inertia.copy())
        self.body_q.append(xform)
        self.body_qd.append(np.zeros(6))
        self.body_label.append(label or f"body_{body_id}")
        self.body_shapes[body_id] = []
        self.body_world.append(self._current_world)
        return body_id

    def add_body(self, xform=None, armature=None, com=None, inertia=None, mass: float = 0.0, label=None,
                 lock_inertia: bool = False, is_kinematic: bool = False) -> int:
        """builder.py:3649-3721: link + FREE joint + articulation."""
        body_id = self.add_link(xform=xform, armature=armature, com=com, inertia=inertia, mass=mass, label=label,
                                lock_inertia=lock_inertia, is_kinematic=is_kinematic)
        j = self.add_joint_free(child=body_id, label=f"{label}_free_joint" if label else None)
        self.add_articulation([j])
        return body_id

    def _update_body_mass(self, i: int, m: float, inertia, p, q):
        """builder.py:8534-8566 (float64 here)."""
        if i == -1:
            return
        new_mass = self.body_mass[i] + m
        if new_mass == 0.0:
            return
        new_com = (self.body_com[i] * self.body_mass[i] + np.asarray(p) * m) / new_mass
        com_offset = new_com - self.body_com[i]
        shape_offset = new_com - np.asarray(p)
        new_inertia = transform_inertia(self.body_mass[i], self.body_inertia[i], com_offset, quat_identity()) + \
            transform_inertia(m, inertia, shape_offset, q)
        self.body_mass[i] = new_mass
        self.body_inertia[i] = new_inertia
        self.body_com[i] = new_com
        self.body_inv_mass[i] = 1.0 / new_mass if new_mass > 0.0 else 0.0
        self.body_inv_inertia[i] = np.linalg.inv(new_inertia) if np.any(new_inertia) else new_inertia.copy()

    # ------------------------------------------------------------------ joints
    def _default_filter_parent(self, joint_type, parent) -> bool:
        return not (parent == -1 and joint_type != JointType.FIXED)

    def add_joint(self, joint_type, parent: int, child: int, linear_axes=None, angular_axes=None, label=None,
                  parent_xform=None, child_xform=None, collision_filter_parent=None, enabled: bool = True) -> int:
        """builder.py:3723-3896."""
        linear_axes = linear_axes or []
        angular_axes = angular_axes or []
        joint_type = JointType(joint_type)
        if collision_filter_parent is None:
            collision_filter_parent = self._default_filter_parent(joint_type, parent)
        if parent != -1:
            if parent < 0 or parent >= self.body_count:
                raise ValueError(f"Parent body index {parent} is out of range")
            if self.body_world[parent] != self._current_world:
                raise ValueError("parent body belongs to another world")
        if child < 0 or child >= self.body_count:
            raise ValueError(f"Child body index {child} is out of range")
        if self.body_world[child] != self._current_world:
            raise ValueError("child body belongs to another world")
        joint_idx = self.joint_count
        self.joint_type.append(int(joint_type))
        self.joint_parent.append(parent)
        self.joint_parents.setdefault(child, []).append((parent, joint_idx))
        self.joint_children.setdefault(parent, []).append((child, joint_idx))
        self.joint_child.append(child)
        self.joint_X_p.append(as_transform(parent_xform))
        self.joint_X_c.append(as_transform(child_xform))
        self.joint_dof_dim.append((len(linear_axes), len(angular_axes)))
        self.joint_enabled.append(bool(enabled))
        self.joint_collision_filter_parent.append(bool(collision_filter_parent))
        self.joint_world.append(self._current_world)
        self.joint_q_start.append(self.joint_coord_count)
        self.joint_qd_start.append(self.joint_dof_count)
        for dim in list(linear_axes) + list(angular_axes):
            self.joint_axis.append(dim.axis_vec())
            self.joint_target_pos.append(dim.target_pos)
            self.joint_target_vel.append(dim.target_vel)
            self.joint_target_ke.append(dim.target_ke)
            self.joint_target_kd.append(dim.target_kd)
            self.joint_limit_lower.append(dim.limit_lower if np.isfinite(dim.limit_lower) else -MAXVAL)
            self.joint_limit_upper.append(dim.limit_upper if np.isfinite(dim.limit_upper) else MAXVAL)
        n_axes = len(linear_axes) + len(angular_axes)
        if joint_type in (JointType.FREE, JointType.DISTANCE):
            dof, coord = 6, 7
        elif joint_type == JointType.BALL:
            dof, coord = 3, 4
        elif joint_type == JointType.FIXED:
            dof, coord = 0, 0
        else:
            dof, coord = n_axes, n_axes
        self.joint_q.extend([0.0] * coord)
        self.joint_qd.extend([0.0] * dof)
        self.joint_f.extend([0.0] * dof)
        if joint_type in (JointType.FREE, JointType.DISTANCE, JointType.BALL):
            self.joint_q[-1] = 1.0
        if joint_type == JointType.FREE:
            bq = self.body_q[child]
            self.joint_q[-7:] = list(bq)
        if collision_filter_parent:
            for cs in self.body_shapes.get(child, []):
                if not self.shape_flags[cs] & ShapeFlags.COLLIDE_SHAPES:
                    continue
                for ps in self.body_shapes.get(parent, []):
                    if not self.shape_flags[ps] & ShapeFlags.COLLIDE_SHAPES:
                        continue
                    self.add_shape_collision_filter_pair(ps, cs)
        return joint_idx

    def add_joint_free(self, child: int, parent_xform=None, child_xform=None, parent: int = -1, label=None,
                       collision_filter_parent=None, enabled: bool = True) -> int:
        """builder.py:4201."""
        lin = [ModelBuilder.JointDofConfig(axis=a) for a in (Axis.X, Axis.Y, Axis.Z)]
        ang = [ModelBuilder.JointDofConfig(axis=a) for a in (Axis.X, Axis.Y, Axis.Z)]
        return self.add_joint(JointType.FREE, parent, child, lin, ang, label=label, parent_xform=parent_xform,
                              child_xform=child_xform, collision_filter_parent=collision_filter_parent,
                              enabled=enabled)

    def add_joint_revolute(self, parent: int, child: int, parent_xform=None, child_xform=None, axis=None,
                           target_pos=None, target_vel=None, target_ke=None, target_kd=None, limit_lower=None,
                           limit_upper=None, label=None, collision_filter_parent=None, enabled: bool = True) -> int:
        """builder.py:3898-3999."""
        dof = ModelBuilder.JointDofConfig(axis=Axis.X if axis is None else axis)
        if limit_lower is not None:
            dof.limit_lower = float(limit_lower)
        if limit_upper is not None:
            dof.limit_upper = float(limit_upper)
        if target_ke is not None:
            dof.target_ke = float(target_ke)
        if target_kd is not None:
            dof.target_kd = float(target_kd)
        if target_vel is not None:
            dof.target_vel = float(target_vel)
        if target_pos is not None:
            dof.target_pos = float(target_pos)
        elif dof.limit_lower > -MAXVAL and dof.limit_upper < MAXVAL:
            dof.target_pos = 0.5 * (dof.limit_lower + dof.limit_upper)
        return self.add_joint(JointType.REVOLUTE, parent, child, angular_axes=[dof], label=label,
                              parent_xform=parent_xform, child_xform=child_xform,
                              collision_filter_parent=collision_filter_parent, enabled=enabled)

    def add_articulation(self, joints: list[int], label: str | None = None):
        if joints:
            self.articulation_start.append(int(joints[0]))
            self.articulation_world.append(self._current_world)

    # ------------------------------------------------------------------ shapes
    def add_shape_collision_filter_pair(self, shape_a: int, shape_b: int):
        a, b = (shape_a, shape_b) if shape_a < shape_b else (shape_b, shape_a)
        self.shape_collision_filter_pairs.add((int(a), int(b)))

    def add_shape(self, *, body: int, type: int, xform=None, cfg: ShapeConfig | None = None, scale=None,
                  src=None, is_static: bool = False, label: str | None = None) -> int:
        """builder.py:5449-5609 (``src``: the Mesh of a MESH shape)."""
        if type == GeoType.MESH and src is None:
            raise ValueError("a MESH shape needs its Mesh (src)")
        xform = as_transform(xform)
        cfg = self.default_shape_cfg if cfg is None else cfg
        scale = (1.0, 1.0, 1.0) if scale is None else scale
        shape = self.shape_count
        self.shape_body.append(int(body))
        if cfg.has_shape_collision:
            for same in self.body_shapes.setdefault(body, []):
                self.add_shape_collision_filter_pair(same, shape)
        self.body_shapes.setdefault(body, []).append(shape)
        self.shape_label.append(label or f"shape_{shape}")
        self.shape_transform.append(xform)
        self.shape_flags.append(int(cfg.flags))
        self.shape_type.append(int(type))
        self.shape_scale.append((float(scale[0]), float(scale[1]), float(scale[2])))
        self.shape_margin.append(float(cfg.margin))
        self.shape_material_mu.append(float(cfg.mu))
        self.shape_material_restitution.append(float(cfg.restitution))
        self.shape_material_mu_torsional.append(float(cfg.mu_torsional))
        self.shape_material_mu_rolling.append(float(cfg.mu_rolling))
        self.shape_gap.append(float(cfg.gap) if cfg.gap is not None else float(self.rigid_gap))
        self.shape_collision_group.append(int(cfg.collision_group))
        if type == GeoType.MESH:
            # bounding sphere of the scaled local AABB (geometry/utils.py:86-97)
            lo, hi = src.aabb(scale)
            self.shape_collision_radius.append(float(0.5 * np.linalg.norm(hi - lo)))
        else:
            self.shape_collision_radius.append(compute_shape_radius(type, scale))
        self.shape_source.append(src)
        self.shape_world.append(self._current_world)
        if cfg.has_shape_collision and cfg.collision_filter_parent:
            for parent_body, joint_idx in self.joint_parents.get(body, ()):
                if not self.joint_collision_filter_parent[joint_idx]:
                    continue
                for ps in self.body_shapes.get(parent_body, []):
                    self.add_shape_collision_filter_pair(ps, shape)
            for child_body, joint_idx in self.joint_children.get(body, ()):
                if not self.joint_collision_filter_parent[joint_idx]:
                    continue
                for cs in self.body_shapes.get(child_body, []):
                    self.add_shape_collision_filter_pair(shape, cs)
        if not is_static and cfg.density > 0.0 and body >= 0 and not self.body_lock_inertia[body]:
            if type == GeoType.MESH:
                m, c, inertia = src.shape_mass_properties(cfg.density, scale)
            else:
                m, c, inertia = compute_inertia_shape(type, scale, cfg.density)
            com_body = transform_point(xform, c)
            self._update_body_mass(body, m, inertia, com_body, xform[3:])
        return shape

    def add_shape_mesh(self, body: int, xform=None, mesh=None, scale=None, cfg=None, label=None) -> int:
        """builder.py:6052-6120: a triangle mesh (mass properties of the solid mesh, inertia.py:361-440)."""
        return self.add_shape(body=body, type=GeoType.MESH, xform=xform, cfg=cfg,
                              scale=(1.0, 1.0, 1.0) if scale is None else scale, src=mesh, label=label)

    def add_shape_plane(self, plane=(0.0, 0.0, 1.0, 0.0), xform=None, width: float = 10.0, length: float = 10.0,
                        body: int = -1, cfg=None, label=None) -> int:
        """builder.py:5611-5674."""
        if xform is None:
            normal = np.asarray(plane[:3], dtype=np.float64)
            norm = np.linalg.norm(normal)
            normal = normal / norm
            d = plane[3] / norm
            pos = -d * normal
            rot = quat_between_vectors((0.0, 0.0, 1.0), normal)
            xform = transform(pos, rot)
        return self.add_shape(body=body, type=GeoType.PLANE, xform=xform, cfg=cfg, scale=(width, length, 0.0),
                              is_static=True, label=label)

    def add_ground_plane(self, height: float = 0.0, cfg=None, label=None) -> int:
        """builder.py:5676-5701: infinite plane (width = length = 0) through ``height`` on the up axis."""
        return self.add_shape_plane(plane=(*self.up_vector, -height), width=0.0, length=0.0, cfg=cfg,
                                    label=label or "ground_plane")

    def add_shape_sphere(self, body: int, xform=None, radius: float = 1.0, cfg=None, label=None) -> int:
        return self.add_shape(body=body, type=GeoType.SPHERE, xform=xform, cfg=cfg, scale=(radius, 0.0, 0.0),
                              label=label)

    def add_shape_box(self, body: int, xform=None, hx: float = 0.5, hy: float = 0.5, hz: float = 0.5, cfg=None,
                      label=None) -> int:
        return self.add_shape(body=body, type=GeoType.BOX, xform=xform, cfg=cfg, scale=(hx, hy, hz), label=label)

    def add_shape_capsule(self, body: int, xform=None, radius: float = 1.0, half_height: float = 0.5, cfg=None,
                          label=None) -> int:
        """Capsule along local +Z; ``half_height`` is half the cylinder length (builder.py:5892)."""
        return self.add_shape(body=body, type=GeoType.CAPSULE, xform=xform, cfg=cfg,
                              scale=(radius, half_height, 0.0), label=label)

    def add_shape_ellipsoid(self, body: int, xform=None, rx: float = 1.0, ry: float = 0.75, rz: float = 0.5, cfg=None,
                            label=None) -> int:
        """Ellipsoid with semi-axes rx, ry, rz along its local axes (builder.py:5747)."""
        return self.add_shape(body=body, type=GeoType.ELLIPSOID, xform=xform, cfg=cfg, scale=(rx, ry, rz),
                              label=label)

    def add_shape_cylinder(self, body: int, xform=None, radius: float = 1.0, half_height: float = 0.5, cfg=None,
                           label=None) -> int:
        """Cylinder along local +Z, centred at its origin (builder.py:5945)."""
        return self.add_shape(body=body, type=GeoType.CYLINDER, xform=xform, cfg=cfg,
                              scale=(radius, half_height, 0.0), label=label)

    def add_shape_cone(self, body: int, xform=None, radius: float = 1.0, half_height: float = 0.5, cfg=None,
                       label=None) -> int:
        """Cone along local +Z: base disc at -half_height, apex at +half_height (builder.py:5998)."""
        return self.add_shape(body=body, type=GeoType.CONE, xform=xform, cfg=cfg,
                              scale=(radius, half_height, 0.0), label=label)

    # ------------------------------------------------------------------ finalize
    def request_state_attributes(self, *attributes: str) -> None:
        """builder.py: optional State arrays the finalized model allocates (``body_parent_f``)."""
        self._requested_state_attributes = getattr(self, "_requested_state_attributes", set()) | set(attributes)

    def request_contact_attributes(self, *attributes: str) -> None:
        """builder.py: optional Contacts arrays (``force``)."""
        self._requested_contact_attributes = getattr(self, "_requested_contact_attributes", set()) | set(attributes)

    def finalize(self, device=None, requires_grad: bool = False):
        """builder.py:9745: upload the SoA arrays.  ``device`` is a torch device string."""
        from .model import Model  # noqa: PLC0415

        m = Model._from_builder(self, device)
        if getattr(self, "_requested_state_attributes", None):
            m.request_state_attributes(*self._requested_state_attributes)
        if getattr(self, "_requested_contact_attributes", None):
            m.request_contact_attributes(*self._requested_contact_attributes)
        return m
