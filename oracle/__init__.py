"""CPU oracle -- TEST INFRASTRUCTURE ONLY.

ctypes front end of ``liboracle.so``, the scalar C++ restatement of Newton's rigid-contact step
(see ``oracle.h`` for the per-function reference citations).  Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` import this module; the product
package never does.
"""

from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_i32p = ctypes.POINTER(ctypes.c_int)
_f32p = ctypes.POINTER(ctypes.c_float)
_f64p = ctypes.POINTER(ctypes.c_double)
_i64p = ctypes.POINTER(ctypes.c_int64)


class OracleModel(ctypes.Structure):
    _fields_ = [
        ("body_count", ctypes.c_int), ("shape_count", ctypes.c_int), ("joint_count", ctypes.c_int),
        ("world_count", ctypes.c_int),
        ("body_com", _f32p), ("body_inertia", _f32p), ("body_inv_mass", _f32p), ("body_inv_inertia", _f32p),
        ("body_flags", _i32p), ("body_world", _i32p),
        ("shape_transform", _f32p), ("shape_body", _i32p), ("shape_type", _i32p), ("shape_scale", _f32p),
        ("shape_margin", _f32p), ("shape_gap", _f32p), ("shape_radius", _f32p), ("shape_flags", _i32p),
        ("shape_group", _i32p), ("shape_world", _i32p), ("shape_mu", _f32p), ("shape_mu_torsional", _f32p),
        ("shape_mu_rolling", _f32p),
        ("segment_count", ctypes.c_int), ("segment_start", _i32p), ("segment_shape", _i32p),
        ("filter_count", ctypes.c_int), ("filter_pairs", _i32p),
        ("gravity_count", ctypes.c_int), ("gravity", _f32p),
        ("joint_type", _i32p), ("joint_enabled", _i32p), ("joint_parent", _i32p), ("joint_child", _i32p),
        ("joint_X_p", _f32p), ("joint_X_c", _f32p), ("joint_qd_start", _i32p), ("joint_dof_dim", _i32p),
        ("joint_axis", _f32p), ("joint_limit_lower", _f32p), ("joint_limit_upper", _f32p),
        ("joint_target_ke", _f32p), ("joint_target_kd", _f32p),
        ("shape_mesh", _i32p), ("shape_aabb_lower", _f32p), ("shape_aabb_upper", _f32p), ("mesh_count", ctypes.c_int),
        ("mesh_vertex", _f32p), ("mesh_vertex_start", _i32p), ("mesh_edge", _i32p), ("mesh_edge_start", _i32p),
        ("mesh_sdf_coarse", _f32p), ("mesh_sdf_slots", ctypes.c_void_p), ("mesh_sdf_texels", ctypes.c_void_p),
        ("mesh_sdf_start", _i64p), ("mesh_sdf_box", _f32p), ("mesh_sdf_quant", _f32p), ("mesh_sdf_dims", _i32p),
        ("mesh_sdf_flags", _i32p), ("mesh_reduce", ctypes.c_int),
        ("shape_voxel_res", _i32p),
        ("broad_phase_mode", ctypes.c_int), ("explicit_pairs", _i32p), ("explicit_pair_count", ctypes.c_int),
        ("shape_restitution", _f32p),
    ]


class OracleXpbdParams(ctypes.Structure):
    _fields_ = [
        ("iterations", ctypes.c_int), ("rigid_contact_relaxation", ctypes.c_float),
        ("rigid_contact_con_weighting", ctypes.c_int), ("angular_damping", ctypes.c_float),
        ("joint_linear_relaxation", ctypes.c_float), ("joint_angular_relaxation", ctypes.c_float),
        ("joint_linear_compliance", ctypes.c_float), ("joint_angular_compliance", ctypes.c_float),
        ("enable_restitution", ctypes.c_int),
    ]


def build(force: bool = False) -> str:
    """Compile liboracle.so with the recipe in oracle/Makefile."""
    so = os.path.join(_HERE, "liboracle.so")
    if force or not os.path.exists(so):
        subprocess.run(["make", "-C", _HERE, "-B" if force else "liboracle.so"], check=True,
                       stdout=subprocess.DEVNULL)
    return so


def lib():
    global _LIB
    if _LIB is None:
        so = build()
        L = ctypes.CDLL(so)
        L.oracle_broadphase_f32.restype = ctypes.c_int
        L.oracle_collide_f32.restype = ctypes.c_int
        L.oracle_collide_f64.restype = ctypes.c_int
        L.oracle_xpbd_step_f32.restype = None
        L.oracle_xpbd_step_f64.restype = None
        L.oracle_asin_f32.restype = None
        L.oracle_acos_f32.restype = None
        _LIB = L
    return _LIB


def asin_f32(x: np.ndarray) -> np.ndarray:
    """The restatement's FP32 asin (revolute twist angle, kernels.py:1441-1943)."""
    x = np.ascontiguousarray(x, np.float32)
    out = np.empty_like(x)
    lib().oracle_asin_f32(_p(x, _f32p), ctypes.c_int(x.size), _p(out, _f32p))
    return out


def acos_f32(x: np.ndarray) -> np.ndarray:
    """The restatement's FP32 acos (revolute swing angle)."""
    x = np.ascontiguousarray(x, np.float32)
    out = np.empty_like(x)
    lib().oracle_acos_f32(_p(x, _f32p), ctypes.c_int(x.size), _p(out, _f32p))
    return out


def _p(a: np.ndarray, t):
    return a.ctypes.data_as(t)


class Oracle:
    """Holds a model's host arrays (kept alive) and runs the restatement on them."""

    def __init__(self, model, reduce_contacts: bool = True, deterministic: bool = True, broad_phase: str = "sap",
                 explicit_pairs=None):
        h = model.host
        self.model = model
        keep = {}

        def f32(name):
            a = np.ascontiguousarray(h[name], dtype=np.float32)
            keep[name] = a
            return _p(a, _f32p)

        def i32(name):
            a = np.ascontiguousarray(h[name], dtype=np.int32)
            keep[name] = a
            return _p(a, _i32p)

        m = OracleModel()
        m.body_count = model.body_count
        m.shape_count = model.shape_count
        m.joint_count = model.joint_count
        m.world_count = model.world_count
        m.body_com, m.body_inertia = f32("body_com"), f32("body_inertia")
        m.body_inv_mass, m.body_inv_inertia = f32("body_inv_mass"), f32("body_inv_inertia")
        m.body_flags, m.body_world = i32("body_flags"), i32("body_world")
        m.shape_transform, m.shape_body, m.shape_type = f32("shape_transform"), i32("shape_body"), i32("shape_type")
        m.shape_scale, m.shape_margin, m.shape_gap = f32("shape_scale"), f32("shape_margin"), f32("shape_gap")
        m.shape_radius, m.shape_flags = f32("shape_collision_radius"), i32("shape_flags")
        m.shape_group, m.shape_world = i32("shape_collision_group"), i32("shape_world")
        m.shape_mu, m.shape_mu_torsional = f32("shape_material_mu"), f32("shape_material_mu_torsional")
        m.shape_mu_rolling = f32("shape_material_mu_rolling")
        m.shape_restitution = f32("shape_material_restitution")
        m.segment_count = len(h["segment_start"]) - 1
        m.segment_start, m.segment_shape = i32("segment_start"), i32("segment_shape")
        m.filter_count = len(h["shape_filter_pairs"])
        m.filter_pairs = i32("shape_filter_pairs")
        m.gravity_count = len(h["gravity"])
        m.gravity = f32("gravity")
        m.joint_type, m.joint_enabled = i32("joint_type"), i32("joint_enabled")
        m.joint_parent, m.joint_child = i32("joint_parent"), i32("joint_child")
        m.joint_X_p, m.joint_X_c = f32("joint_X_p"), f32("joint_X_c")
        m.joint_qd_start, m.joint_dof_dim = i32("joint_qd_start"), i32("joint_dof_dim")
        m.joint_axis = f32("joint_axis")
        m.joint_limit_lower, m.joint_limit_upper = f32("joint_limit_lower"), f32("joint_limit_upper")
        m.joint_target_ke, m.joint_target_kd = f32("joint_target_ke"), f32("joint_target_kd")
        m.shape_mesh, m.shape_aabb_lower, m.shape_aabb_upper = (i32("shape_mesh"), f32("shape_aabb_lower"),
                                                                f32("shape_aabb_upper"))
        m.mesh_count = len(getattr(model, "meshes", []))
        m.mesh_vertex, m.mesh_vertex_start = f32("mesh_vertex"), i32("mesh_vertex_start")
        m.mesh_edge, m.mesh_edge_start = i32("mesh_edge"), i32("mesh_edge_start")
        coarse = np.ascontiguousarray(model.mesh_sdf_coarse.detach().cpu().numpy(), np.float32)
        slots = np.ascontiguousarray(model.mesh_sdf_slots.detach().cpu().numpy()).view(np.uint32)
        texels = np.ascontiguousarray(model.mesh_sdf_texels.detach().cpu().numpy(), np.uint8)
        st = np.ascontiguousarray(h["mesh_sdf_start"], np.int64)
        keep.update(mesh_sdf_coarse=coarse, mesh_sdf_slots=slots, mesh_sdf_texels=texels, mesh_sdf_start=st)
        m.mesh_sdf_coarse = _p(coarse, _f32p)
        m.mesh_sdf_slots = slots.ctypes.data
        m.mesh_sdf_texels = texels.ctypes.data
        m.mesh_sdf_start = _p(st, _i64p)
        m.mesh_sdf_box, m.mesh_sdf_quant = f32("mesh_sdf_box"), f32("mesh_sdf_quant")
        m.mesh_sdf_dims, m.mesh_sdf_flags = i32("mesh_sdf_dims"), i32("mesh_sdf_flags")
        # the reference's mesh contact reduction and its value packing (CollisionPipeline reduce_contacts /
        # deterministic)
        m.mesh_reduce = 0 if not reduce_contacts else (2 if deterministic else 1)
        m.shape_voxel_res = i32("shape_voxel_resolution")
        # CollisionPipeline broad_phase: "explicit" without a list is the model's own pair list, the "nxn" set
        if broad_phase == "explicit" and explicit_pairs is None:
            broad_phase = "nxn"
        m.broad_phase_mode = {"sap": 0, "nxn": 1, "explicit": 2}[broad_phase]
        if explicit_pairs is not None:
            ep = np.unique(np.sort(np.asarray(explicit_pairs, np.int64).reshape(-1, 2), axis=1), axis=0)
            ep = np.ascontiguousarray(ep, np.int32)
            keep["explicit_pairs"] = ep
            m.explicit_pairs = _p(ep, _i32p)
            m.explicit_pair_count = len(ep)
        self._keep = keep
        self.m = m
        self.lib = lib()

    def sdf_sample(self, mesh: int, points: np.ndarray):
        """SDF value and gradient of the model's mesh ``mesh`` at local points (n x 3)."""
        p = np.ascontiguousarray(points, np.float32).reshape(-1, 3)
        v = np.zeros(len(p), np.float32)
        g = np.zeros((len(p), 3), np.float32)
        self.lib.oracle_sdf_sample_f32.restype = None
        self.lib.oracle_sdf_sample_f32(ctypes.byref(self.m), ctypes.c_int(mesh), _p(p, _f32p), ctypes.c_int(len(p)),
                                       _p(v, _f32p), _p(g, _f32p))
        return v, g

    def edge_search(self, mesh: int, v0: np.ndarray, v1: np.ndarray):
        """Deepest SDF point of each segment v0 -> v1 (mesh-local coordinates)."""
        a = np.ascontiguousarray(v0, np.float32).reshape(-1, 3)
        b = np.ascontiguousarray(v1, np.float32).reshape(-1, 3)
        v = np.zeros(len(a), np.float32)
        pt = np.zeros((len(a), 3), np.float32)
        self.lib.oracle_edge_search_f32.restype = None
        self.lib.oracle_edge_search_f32(ctypes.byref(self.m), ctypes.c_int(mesh), _p(a, _f32p), _p(b, _f32p),
                                        ctypes.c_int(len(a)), _p(v, _f32p), _p(pt, _f32p))
        return v, pt

    def broadphase(self, body_q: np.ndarray, capacity: int | None = None) -> np.ndarray:
        q = np.ascontiguousarray(body_q, dtype=np.float32)
        cap = capacity if capacity is not None else max(1, self.model.shape_count * 64)
        out = np.zeros((cap, 2), dtype=np.int32)
        n = self.lib.oracle_broadphase_f32(ctypes.byref(self.m), _p(q, _f32p), cap, _p(out, _i32p))
        return out[: min(n, cap)]

    def collide(self, body_q: np.ndarray, capacity: int | None = None, precision: str = "f32",
                layout: str = "canonical") -> dict:
        """CollisionPipeline.collide.  layout "canonical": segment by segment, key order inside a segment;
        "global": the deterministic pipeline's buffer, every contact found sorted by its 64-bit sort key and the
        first `capacity` kept (ContactSorter, contact_sort.py:194-445, as collide.py:1105-1124 applies it)."""
        if layout == "global":
            full = self._collide(body_q, None, precision)
            n = full["count"]
            full = self._collide(body_q, max(n, 1), precision)
            order = np.argsort(full["sort_key"], kind="stable")
            cap = n if capacity is None else min(n, capacity)
            res = {k: v[order][:cap] for k, v in full.items() if isinstance(v, np.ndarray)}
            res["count"] = n
            return res
        return self._collide(body_q, capacity, precision)

    def _collide(self, body_q: np.ndarray, capacity: int | None = None, precision: str = "f32") -> dict:
        dt = np.float32 if precision == "f32" else np.float64
        tp = _f32p if precision == "f32" else _f64p
        q = np.ascontiguousarray(body_q, dtype=dt)
        cap = capacity if capacity is not None else max(1000, self.model.shape_count * 40)
        out = {
            "shape0": np.zeros(cap, np.int32), "shape1": np.zeros(cap, np.int32),
            "point0": np.zeros((cap, 3), dt), "point1": np.zeros((cap, 3), dt), "offset0": np.zeros((cap, 3), dt),
            "offset1": np.zeros((cap, 3), dt), "normal": np.zeros((cap, 3), dt), "margin0": np.zeros(cap, dt),
            "margin1": np.zeros(cap, dt), "sort_key": np.zeros(cap, np.int64),
        }
        fn = self.lib.oracle_collide_f32 if precision == "f32" else self.lib.oracle_collide_f64
        n = fn(ctypes.byref(self.m), _p(q, tp), cap, _p(out["shape0"], _i32p), _p(out["shape1"], _i32p),
               _p(out["point0"], tp), _p(out["point1"], tp), _p(out["offset0"], tp), _p(out["offset1"], tp),
               _p(out["normal"], tp), _p(out["margin0"], tp), _p(out["margin1"], tp), _p(out["sort_key"], _i64p))
        k = min(n, cap)
        res = {key: v[:k] for key, v in out.items()}
        res["count"] = n
        return res

    def step(self, body_q, body_qd, body_f, contacts: dict | None, dt: float, iterations: int = 2,
             rigid_contact_relaxation: float = 0.8, rigid_contact_con_weighting: bool = True,
             angular_damping: float = 0.0, joint_linear_relaxation: float = 0.7,
             joint_angular_relaxation: float = 0.4, joint_linear_compliance: float = 0.0,
             joint_angular_compliance: float = 0.0, joint_f=None, joint_target_pos=None, joint_target_vel=None,
             precision: str = "f32", report: bool = False, enable_restitution: bool = False):
        """One SolverXPBD.step; returns (body_q_out, body_qd_out), plus (contact_force (n, 6), body_parent_f
        (B, 6)) with ``report`` (FP32 only).  ``enable_restitution``: the rigid restitution pass after the
        iterations (solver_xpbd.py:680-764)."""
        h = self.model.host
        dtp = np.float32 if precision == "f32" else np.float64
        tp = _f32p if precision == "f32" else _f64p
        p = OracleXpbdParams(iterations, rigid_contact_relaxation, int(rigid_contact_con_weighting), angular_damping,
                             joint_linear_relaxation, joint_angular_relaxation, joint_linear_compliance,
                             joint_angular_compliance, int(bool(enable_restitution)))
        q = np.ascontiguousarray(body_q, dtype=dtp)
        qd = np.ascontiguousarray(body_qd, dtype=dtp)
        f = np.ascontiguousarray(body_f, dtype=dtp)
        jf = np.ascontiguousarray(h["joint_f"] if joint_f is None else joint_f, dtype=np.float32)
        tpos = np.ascontiguousarray(h["joint_target_pos"] if joint_target_pos is None else joint_target_pos,
                                    dtype=np.float32)
        tvel = np.ascontiguousarray(h["joint_target_vel"] if joint_target_vel is None else joint_target_vel,
                                    dtype=np.float32)
        if jf.size == 0:
            jf = np.zeros(1, np.float32)
        if tpos.size == 0:
            tpos = np.zeros(1, np.float32)
        if tvel.size == 0:
            tvel = np.zeros(1, np.float32)
        if contacts is None:
            nc = -1
            c = {k: np.zeros((1, 3), dtp) for k in ("point0", "point1", "offset0", "offset1", "normal")}
            c.update({"shape0": np.zeros(1, np.int32), "shape1": np.zeros(1, np.int32),
                      "margin0": np.zeros(1, dtp), "margin1": np.zeros(1, dtp)})
        else:
            nc = len(contacts["shape0"])
            c = {}
            for k in ("point0", "point1", "offset0", "offset1", "normal"):
                c[k] = np.ascontiguousarray(contacts[k], dtype=dtp).reshape(-1, 3)
                if c[k].size == 0:
                    c[k] = np.zeros((1, 3), dtp)
            for k in ("margin0", "margin1"):
                c[k] = np.ascontiguousarray(contacts[k], dtype=dtp)
                if c[k].size == 0:
                    c[k] = np.zeros(1, dtp)
            for k in ("shape0", "shape1"):
                c[k] = np.ascontiguousarray(contacts[k], dtype=np.int32)
                if c[k].size == 0:
                    c[k] = np.zeros(1, np.int32)
        qo = np.zeros_like(q)
        qdo = np.zeros_like(qd)
        args = [ctypes.byref(self.m), ctypes.byref(p), _p(q, tp), _p(qd, tp), _p(f, tp), _p(jf, _f32p),
                _p(tpos, _f32p), _p(tvel, _f32p), nc, _p(c["shape0"], _i32p), _p(c["shape1"], _i32p),
                _p(c["point0"], tp), _p(c["point1"], tp), _p(c["offset0"], tp), _p(c["offset1"], tp),
                _p(c["normal"], tp), _p(c["margin0"], tp), _p(c["margin1"], tp)]
        dtc = ctypes.c_float(dt) if precision == "f32" else ctypes.c_double(dt)
        if report:
            assert precision == "f32"
            cf = np.zeros((max(nc, 1), 6), np.float32)
            pf = np.zeros((max(len(q), 1), 6), np.float32)
            self.lib.oracle_xpbd_step_report_f32.restype = None
            self.lib.oracle_xpbd_step_report_f32(*args, dtc, _p(qo, tp), _p(qdo, tp), _p(cf, _f32p), _p(pf, _f32p))
            return qo, qdo, cf[:max(nc, 0)], pf[:len(q)]
        fn = self.lib.oracle_xpbd_step_f32 if precision == "f32" else self.lib.oracle_xpbd_step_f64
        fn(*args, dtc, _p(qo, tp), _p(qdo, tp))
        return qo, qdo

    def step_learn(self, body_q, body_qd, body_f, contacts: dict, dt: float, mode: int, iterations: int = 10,
                   lambda_prev=None, lambda0=None, warm_iterations: int = 3, residual_tol: float = 1e-4,
                   rigid_contact_relaxation: float = 0.8, rigid_contact_con_weighting: bool = True):
        """FP32 step of the learned-warm-start restatement (oracle_learn.h).

        mode 0 plain, 1 plain + features / multiplier logging, 2 warm start.  Returns
        (q, qd, features [n,16], lambda_sum [n], residual [world_count+1], fallback [world_count+1])."""
        h = self.model.host
        L = self.lib
        p = OracleXpbdParams(iterations, rigid_contact_relaxation, int(rigid_contact_con_weighting), 0.0, 0.7, 0.4,
                             0.0, 0.0)
        q = np.ascontiguousarray(body_q, np.float32)
        qd = np.ascontiguousarray(body_qd, np.float32)
        f = np.ascontiguousarray(body_f, np.float32)
        one = np.zeros(1, np.float32)
        jf = np.ascontiguousarray(h["joint_f"], np.float32) if h["joint_f"].size else one
        tpos = np.ascontiguousarray(h["joint_target_pos"], np.float32) if h["joint_target_pos"].size else one
        tvel = np.ascontiguousarray(h["joint_target_vel"], np.float32) if h["joint_target_vel"].size else one
        n = len(contacts["shape0"])
        nn = max(n, 1)
        c = {}
        for k in ("point0", "point1", "offset0", "offset1", "normal"):
            c[k] = np.zeros((nn, 3), np.float32)
            c[k][:n] = np.asarray(contacts[k], np.float32).reshape(-1, 3)
        for k in ("margin0", "margin1"):
            c[k] = np.zeros(nn, np.float32)
            c[k][:n] = contacts[k]
        for k in ("shape0", "shape1"):
            c[k] = np.zeros(nn, np.int32)
            c[k][:n] = contacts[k]
        lp = np.zeros(nn, np.float32) if lambda_prev is None else np.ascontiguousarray(lambda_prev, np.float32)
        l0 = np.zeros(nn, np.float32) if lambda0 is None else np.ascontiguousarray(lambda0, np.float32)
        feat = np.zeros((nn, 16), np.float32)
        lam = np.zeros(nn, np.float32)
        W = self.model.world_count + 1
        res = np.zeros(W, np.float32)
        fb = np.zeros(W, np.int32)
        qo = np.zeros_like(q)
        qdo = np.zeros_like(qd)
        L.oracle_xpbd_step_learn_f32.restype = None
        L.oracle_xpbd_step_learn_f32(
            ctypes.byref(self.m), ctypes.byref(p), _p(q, _f32p), _p(qd, _f32p), _p(f, _f32p), _p(jf, _f32p),
            _p(tpos, _f32p), _p(tvel, _f32p), n, _p(c["shape0"], _i32p), _p(c["shape1"], _i32p),
            _p(c["point0"], _f32p), _p(c["point1"], _f32p), _p(c["offset0"], _f32p), _p(c["offset1"], _f32p),
            _p(c["normal"], _f32p), _p(c["margin0"], _f32p), _p(c["margin1"], _f32p), ctypes.c_float(dt),
            ctypes.c_int(mode), _p(lp, _f32p), _p(l0, _f32p), ctypes.c_int(warm_iterations),
            ctypes.c_float(residual_tol), _p(feat, _f32p), _p(lam, _f32p), _p(res, _f32p), _p(fb, _i32p),
            _p(qo, _f32p), _p(qdo, _f32p))
        return qo, qdo, feat[:n], lam[:n], res, fb


def sdf_fine_grid(vertices, triangles, lower, cell, n, precision: str = "f32",
                  winding_threshold: float = 0.5) -> np.ndarray:
    """The restatement's mesh query at the vertices of a fine grid (n[0..2] vertices per axis from ``lower``
    with per-axis spacing ``cell``); returns (n[2], n[1], n[0])."""
    L = lib()
    v = np.ascontiguousarray(vertices, np.float32).reshape(-1, 3)
    t = np.ascontiguousarray(triangles, np.int32).reshape(-1)
    lo = np.ascontiguousarray(lower, np.float32).reshape(3)
    ce = np.ascontiguousarray(cell, np.float32).reshape(3)
    nn = np.ascontiguousarray(n, np.int32).reshape(3)
    dt, tp, fn = ((np.float32, _f32p, L.oracle_sdf_build_f32) if precision == "f32"
                  else (np.float64, _f64p, L.oracle_sdf_build_f64))
    out = np.zeros((int(nn[2]), int(nn[1]), int(nn[0])), dt)
    fn.restype = None
    fn(_p(v, _f32p), ctypes.c_int(len(v)), _p(t, _i32p), ctypes.c_int(len(t) // 3), _p(lo, _f32p), _p(ce, _f32p),
       _p(nn, _i32p), ctypes.c_float(float(winding_threshold)), _p(out, tp))
    return out


def texture_geometry(points, margin: float = 0.05, narrow_band_range=(-0.1, 0.1), max_resolution=None,
                     target_voxel_size=None):
    """Grid geometry of the reference's texture cook (create_texture_sdf_from_mesh, sdf_texture.py:1436-1480,
    and build_sparse_sdf_from_mesh, :1017-1036, 1134-1136, 1152-1160, 1290): numpy arithmetic on float32
    vertices, as there.  Returns None for a mesh without extent."""
    pts = np.asarray(points, np.float32).reshape(-1, 3)
    mesh_min, mesh_max = np.min(pts, axis=0), np.max(pts, axis=0)
    min_ext, max_ext = mesh_min - float(margin), mesh_max + float(margin)
    ext = max_ext - min_ext
    mx = np.max(ext)
    if mx < 1e-10:
        return None
    if target_voxel_size is not None:
        r = int(np.ceil(mx / float(target_voxel_size)))
        max_resolution = max(8, ((r + 7) // 8) * 8)
    elif max_resolution is None:
        max_resolution = 64
    cs = mx / int(max_resolution)
    dims = np.ceil(ext / cs).astype(int) + 1
    cell = ext / (dims - 1)
    nb = max(abs(narrow_band_range[0]), abs(narrow_band_range[1]))
    w, h, d = ((int(k) - 1 + 7) // 8 for k in dims)
    radius = float(np.linalg.norm(8 * 0.5 * cell))
    lo, hi = -nb - radius, nb + radius
    rng = hi - lo
    if rng < 1e-10:
        rng = 1.0
    return {"min_extents": min_ext, "cell_size": cell, "dims": dims, "coarse_dims": (w, h, d),
            "band": (np.float32(lo), np.float32(hi)),
            "linear_threshold": float(1e-6 * np.linalg.norm(max_ext - min_ext)), "sdf_min": lo, "sdf_range": rng,
            "padded_max": min_ext + np.array([w, h, d], dtype=float) * 8 * cell}


def sdf_build_sparse(vertices, triangles, geo: dict, texture_format: str = "uint16", winding_threshold: float = 0.5,
                     fine: np.ndarray | None = None) -> dict:
    """The restatement's texture cook: the reference's ``build_sparse_sdf_from_mesh`` dictionary for float32
    vertices and the grid geometry ``geo`` (``texture_geometry``).  ``fine`` replaces the mesh queries by a
    given fine grid ((8d+1, 8h+1, 8w+1) values)."""
    L = lib()
    w, h, d = geo["coarse_dims"]
    n = (8 * w + 1, 8 * h + 1, 8 * d + 1)
    if fine is None:
        fine = sdf_fine_grid(vertices, triangles, geo["min_extents"], np.asarray(geo["cell_size"]).astype(np.float32),
                             n, winding_threshold=winding_threshold)
    fine = np.ascontiguousarray(fine, np.float32)
    assert fine.shape == (n[2], n[1], n[0])
    dims = np.array([w, h, d], np.int32)
    total = w * h * d
    coarse = np.zeros((d + 1, h + 1, w + 1), np.float32)
    occ, req, lin, addr = (np.zeros(total, np.int32) for _ in range(4))
    lo, hi = geo["band"]
    L.oracle_sdf_subgrids.restype = ctypes.c_int
    nreq = L.oracle_sdf_subgrids(_p(fine, _f32p), _p(dims, _i32p), ctypes.c_float(float(lo)), ctypes.c_float(float(hi)),
                                 ctypes.c_float(geo["linear_threshold"]), _p(coarse, _f32p), _p(occ, _i32p),
                                 _p(req, _i32p), _p(lin, _i32p), _p(addr, _i32p))
    qmode = {"float32": 4, "uint16": 2, "uint8": 1}[texture_format]
    if nreq == 0:
        blocks, fmt, ts, qmin, qrange = 1, 4, 9, 0.0, 1.0
    else:
        blocks = max(1, int(np.ceil(nreq ** (1.0 / 3.0))))
        while blocks ** 3 < nreq:
            blocks += 1
        fmt, ts = qmode, 9 * blocks
        qmin, qrange = (geo["sdf_min"], geo["sdf_range"]) if qmode != 4 else (0.0, 1.0)
    dt = {4: np.float32, 2: np.uint16, 1: np.uint8}[fmt]
    tex = np.zeros((ts, ts, ts), dt)
    slots = np.zeros((w, h, d), np.uint32)
    L.oracle_sdf_pack.restype = None
    L.oracle_sdf_pack(_p(fine, _f32p), _p(dims, _i32p), _p(req, _i32p), _p(lin, _i32p), _p(addr, _i32p),
                      ctypes.c_int(blocks), ctypes.c_int(fmt), ctypes.c_float(qmin if fmt != 4 else 0.0),
                      ctypes.c_float(1.0 / geo["sdf_range"]), ctypes.c_void_p(slots.ctypes.data),
                      ctypes.c_void_p(tex.ctypes.data))
    if nreq == 0:
        tex = np.zeros((1, 1, 1), np.float32)
        ts = 1
    return {"coarse_sdf": coarse, "subgrid_data": tex, "subgrid_start_slots": slots, "coarse_dims": (w, h, d),
            "subgrid_tex_size": ts, "num_subgrids": int(nreq), "min_extents": geo["min_extents"],
            "max_extents": geo["padded_max"], "cell_size": np.asarray(geo["cell_size"], np.float64),
            "subgrid_size": 8, "quantization_mode": qmode, "subgrids_min_sdf_value": qmin,
            "subgrids_sdf_value_range": qrange, "subgrid_required": req, "subgrid_occupied": occ,
            "subgrid_linear": lin}


def broad_phase(mode: str, lower, upper, gap, group, world, filter_pairs=None, index_map=None, slice_ends=None,
                shape_pairs=None) -> np.ndarray:
    """The restatement's standalone broad phase ("sap", "nxn" or "explicit"); returns the (n, 2) pairs in the
    order one thread finds them."""
    L = lib()
    lo = np.ascontiguousarray(lower, np.float32).reshape(-1, 3)
    up = np.ascontiguousarray(upper, np.float32).reshape(-1, 3)
    ga = None if gap is None else np.ascontiguousarray(gap, np.float32)
    gr = np.ascontiguousarray(group, np.int32)
    wo = np.ascontiguousarray(world, np.int32)
    fp = None if filter_pairs is None or len(filter_pairs) == 0 else np.ascontiguousarray(filter_pairs, np.int32)
    im = np.ascontiguousarray(index_map if index_map is not None else np.zeros(1), np.int32)
    se = np.ascontiguousarray(slice_ends if slice_ends is not None else np.zeros(1), np.int32)
    sp = np.ascontiguousarray(shape_pairs if shape_pairs is not None else np.zeros((1, 2)), np.int32)
    m = {"sap": 0, "nxn": 1, "explicit": 2}[mode]
    fn = L.oracle_broad_phase
    fn.restype = ctypes.c_int

    def run(cap):
        out = np.zeros((max(cap, 1), 2), np.int32)
        n = fn(ctypes.c_int(m), _p(lo, _f32p), _p(up, _f32p), None if ga is None else _p(ga, _f32p), _p(gr, _i32p),
               _p(wo, _i32p), None if fp is None else _p(fp, _i32p), ctypes.c_int(0 if fp is None else len(fp)),
               _p(im, _i32p), _p(se, _i32p), ctypes.c_int(0 if slice_ends is None else len(slice_ends)),
               _p(sp, _i32p), ctypes.c_int(0 if shape_pairs is None else len(shape_pairs)), _p(out, _i32p),
               ctypes.c_int(cap))
        return n, out

    n, out = run(1024)
    if n > 1024:
        n, out = run(n)
    return out[:n].copy()


def mesh_distance_f64(vertices, triangles, points) -> np.ndarray:
    """FP64 signed distance of points to a triangle mesh (nearest triangle, winding-number sign)."""
    L = lib()
    v = np.ascontiguousarray(vertices, np.float32).reshape(-1, 3)
    t = np.ascontiguousarray(triangles, np.int32).reshape(-1)
    p = np.ascontiguousarray(points, np.float64).reshape(-1, 3)
    out = np.zeros(len(p), np.float64)
    L.oracle_mesh_distance_f64.restype = None
    L.oracle_mesh_distance_f64(_p(v, _f32p), ctypes.c_int(len(v)), _p(t, _i32p), ctypes.c_int(len(t) // 3),
                               _p(p, _f64p), ctypes.c_int(len(p)), _p(out, _f64p))
    return out


def closest_on_triangle_f64(p, a, b, c) -> np.ndarray:
    L = lib()
    arrs = [np.ascontiguousarray(x, np.float64).reshape(-1, 3) for x in (p, a, b, c)]
    out = np.zeros_like(arrs[0])
    L.oracle_closest_on_triangle_f64.restype = None
    L.oracle_closest_on_triangle_f64(*[_p(x, _f64p) for x in arrs], ctypes.c_int(len(out)), _p(out, _f64p))
    return out


def solid_angle_f64(a, b, c) -> np.ndarray:
    L = lib()
    arrs = [np.ascontiguousarray(x, np.float64).reshape(-1, 3) for x in (a, b, c)]
    out = np.zeros(len(arrs[0]), np.float64)
    L.oracle_solid_angle_f64.restype = None
    L.oracle_solid_angle_f64(*[_p(x, _f64p) for x in arrs], ctypes.c_int(len(out)), _p(out, _f64p))
    return out


class OracleMatcher:
    """The restatement's frame-to-frame contact matcher (contact_match.py), fed key-sorted contacts."""

    def __init__(self, oracle: "Oracle", pos_threshold: float = 0.0005, normal_dot_threshold: float = 0.995,
                 sticky: bool = False):
        L = lib()
        L.oracle_matcher_create.restype = ctypes.c_void_p
        L.oracle_matcher_destroy.argtypes = [ctypes.c_void_p]
        L.oracle_matcher_reset.argtypes = [ctypes.c_void_p]
        L.oracle_contact_match.restype = ctypes.c_int
        self.L = L
        self.oracle = oracle
        self.sticky = sticky
        self.h = ctypes.c_void_p(L.oracle_matcher_create(ctypes.c_float(pos_threshold),
                                                         ctypes.c_float(normal_dot_threshold), int(sticky)))

    def reset(self):
        self.L.oracle_matcher_reset(self.h)

    def match(self, body_q: np.ndarray, contacts: dict) -> dict:
        """Returns {match_index, matched_key, broken_keys} and, when sticky, the replayed contact fields."""
        order = np.argsort(contacts["sort_key"], kind="stable")
        n = len(order)
        keys = np.ascontiguousarray(contacts["sort_key"][order], np.int64)
        s0 = np.ascontiguousarray(contacts["shape0"][order], np.int32)
        s1 = np.ascontiguousarray(contacts["shape1"][order], np.int32)
        f = {k: np.ascontiguousarray(np.asarray(contacts[k], np.float32).reshape(-1, 3)[order])
             for k in ("point0", "point1", "offset0", "offset1", "normal")}
        q = np.ascontiguousarray(body_q, np.float32)
        mi = np.zeros(max(n, 1), np.int32)
        mk = np.zeros(max(n, 1), np.int64)
        cap = 1 << 22
        bk = np.zeros(cap, np.int64)
        for k in f:
            if f[k].size == 0:
                f[k] = np.zeros((1, 3), np.float32)
        if keys.size == 0:
            keys = np.zeros(1, np.int64)
            s0 = np.zeros(1, np.int32)
            s1 = np.zeros(1, np.int32)
        nb = self.L.oracle_contact_match(self.h, ctypes.byref(self.oracle.m), _p(q, _f32p), ctypes.c_int(n),
                                         _p(keys, _i64p), _p(s0, _i32p), _p(s1, _i32p), _p(f["point0"], _f32p),
                                         _p(f["point1"], _f32p), _p(f["offset0"], _f32p), _p(f["offset1"], _f32p),
                                         _p(f["normal"], _f32p), _p(mi, _i32p), _p(mk, _i64p), _p(bk, _i64p),
                                         ctypes.c_int(cap))
        out = {"sort_key": keys[:n], "match_index": mi[:n], "matched_key": mk[:n], "broken_keys": bk[:min(nb, cap)]}
        if self.sticky:
            out.update({k: v[:n] for k, v in f.items()})
        return out

    def __del__(self):
        try:
            self.L.oracle_matcher_destroy(self.h)
        except Exception:  # noqa: BLE001
            pass


def mlp_forward(params: dict, x: np.ndarray) -> np.ndarray:
    """The restatement's MLP forward (BF16 inputs/weights of layers 1-2, FP32 accumulate)."""
    L = lib()
    L.oracle_mlp_forward.restype = None
    x = np.ascontiguousarray(x, np.float32)
    y = np.zeros(len(x), np.float32)
    P = {k: np.ascontiguousarray(v, np.float32) for k, v in params.items()}
    L.oracle_mlp_forward(*[_p(P[k], _f32p) for k in ("w1", "b1", "w2", "b2", "w3", "b3")], _p(x, _f32p),
                         ctypes.c_int(len(x)), _p(y, _f32p))
    return y


def mlp_backward(params: dict, x: np.ndarray, t: np.ndarray) -> tuple[dict, float]:
    """Gradients of 0.5 * mean((y - t)^2) (FP32 restatement, FP64 sums)."""
    L = lib()
    L.oracle_mlp_backward.restype = None
    x = np.ascontiguousarray(x, np.float32)
    t = np.ascontiguousarray(t, np.float32)
    P = {k: np.ascontiguousarray(v, np.float32) for k, v in params.items()}
    g = {k: np.zeros_like(P[k]) for k in P}
    loss = np.zeros(1, np.float32)
    L.oracle_mlp_backward(*[_p(P[k], _f32p) for k in ("w1", "b1", "w2", "b2", "w3", "b3")], _p(x, _f32p),
                          _p(t, _f32p), ctypes.c_int(len(x)),
                          *[_p(g[k], _f32p) for k in ("w1", "b1", "w2", "b2", "w3", "b3")], _p(loss, _f32p))
    return g, float(loss[0])


# ---------------------------------------------------------------- contact reduction primitives
class reduction:
    """The restatement's contact-reduction primitives (oracle_reduce.h), for tests."""

    @staticmethod
    def get_slot(normals) -> np.ndarray:
        n = np.ascontiguousarray(normals, np.float32).reshape(-1, 3)
        out = np.zeros(len(n), np.int32)
        lib().oracle_red_get_slot(_p(n, _f32p), ctypes.c_int(len(n)), _p(out, _i32p))
        return out

    @staticmethod
    def project(bins, points) -> np.ndarray:
        b = np.ascontiguousarray(bins, np.int32).reshape(-1)
        p = np.ascontiguousarray(points, np.float32).reshape(-1, 3)
        out = np.zeros((len(p), 2), np.float32)
        lib().oracle_red_project(_p(b, _i32p), _p(p, _f32p), ctypes.c_int(len(p)), _p(out, _f32p))
        return out

    @staticmethod
    def directions() -> np.ndarray:
        out = np.zeros((6, 2), np.float32)
        lib().oracle_red_direction(_p(out, _f32p))
        return out

    @staticmethod
    def oct(normals):
        n = np.ascontiguousarray(normals, np.float32).reshape(-1, 3)
        enc = np.zeros((len(n), 2), np.float32)
        dec = np.zeros((len(n), 3), np.float32)
        lib().oracle_red_oct(_p(n, _f32p), ctypes.c_int(len(n)), _p(enc, _f32p), _p(dec, _f32p))
        return enc, dec

    @staticmethod
    def value(score, fingerprint, cid, det: bool) -> np.ndarray:
        s = np.ascontiguousarray(score, np.float32).reshape(-1)
        f = np.ascontiguousarray(fingerprint, np.int32).reshape(-1)
        c = np.ascontiguousarray(cid, np.uint32).reshape(-1)
        out = np.zeros(len(s), np.uint64)
        lib().oracle_red_value(_p(s, _f32p), _p(f, _i32p), c.ctypes.data_as(ctypes.POINTER(ctypes.c_uint)),
                               ctypes.c_int(len(s)), ctypes.c_int(int(det)),
                               out.ctypes.data_as(ctypes.POINTER(ctypes.c_ulonglong)))
        return out

    @staticmethod
    def voxel_index(points, lower, upper, res) -> np.ndarray:
        p = np.ascontiguousarray(points, np.float32).reshape(-1, 3)
        lo = np.ascontiguousarray(lower, np.float32)
        hi = np.ascontiguousarray(upper, np.float32)
        r = np.ascontiguousarray(res, np.int32)
        out = np.zeros(len(p), np.int32)
        lib().oracle_red_voxel_index(_p(p, _f32p), ctypes.c_int(len(p)), _p(lo, _f32p), _p(hi, _f32p),
                                     _p(r, _i32p), _p(out, _i32p))
        return out

    @staticmethod
    def voxel_resolution(lower, upper, budget: int = 100):
        lo = np.ascontiguousarray(lower, np.float64)
        hi = np.ascontiguousarray(upper, np.float64)
        out = np.zeros(3, np.int32)
        lib().oracle_red_voxel_resolution(_p(lo, _f64p), _p(hi, _f64p), ctypes.c_int(budget), _p(out, _i32p))
        return tuple(int(x) for x in out)

    @staticmethod
    def reduce(points, normals, depth, sub, use_beta, voxel, det: bool = True) -> np.ndarray:
        """Candidate indices the reduction keeps (points are both the stored and the projected position)."""
        p = np.ascontiguousarray(points, np.float32).reshape(-1, 3)
        n = np.ascontiguousarray(normals, np.float32).reshape(-1, 3)
        d = np.ascontiguousarray(depth, np.float32).reshape(-1)
        s = np.ascontiguousarray(sub, np.int32).reshape(-1)
        b = np.ascontiguousarray(use_beta, np.int32).reshape(-1)
        v = np.ascontiguousarray(voxel, np.int32).reshape(-1)
        out = np.zeros(max(len(p), 1), np.int32)
        L = lib()
        L.oracle_red_reduce.restype = ctypes.c_int
        k = L.oracle_red_reduce(_p(p, _f32p), _p(n, _f32p), _p(d, _f32p), _p(s, _i32p), _p(b, _i32p), _p(v, _i32p),
                                ctypes.c_int(len(p)), ctypes.c_int(int(det)), _p(out, _i32p))
        return np.sort(out[:k])
