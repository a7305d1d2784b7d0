"""ctypes binding of libnmc.so (declarations in include/nmc.h).

The product path runs only through this library: if it is missing or fails to load, every call raises.
There is no CPU fallback.
"""

from __future__ import annotations

import ctypes
import os

import numpy as np
import torch

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.environ.get("NMC_LIBRARY", os.path.join(_HERE, "libnmc.so"))
_lib = None

_vp = ctypes.c_void_p


class NmcModelDesc(ctypes.Structure):
    _fields_ = [
        ("body_count", ctypes.c_int), ("shape_count", ctypes.c_int), ("joint_count", ctypes.c_int),
        ("world_count", ctypes.c_int), ("joint_dof_count", ctypes.c_int),
        ("body_com", _vp), ("body_inertia", _vp), ("body_inv_mass", _vp), ("body_inv_inertia", _vp),
        ("body_flags", _vp), ("body_world", _vp),
        ("shape_transform", _vp), ("shape_body", _vp), ("shape_type", _vp), ("shape_scale", _vp),
        ("shape_margin", _vp), ("shape_gap", _vp), ("shape_radius", _vp), ("shape_flags", _vp),
        ("shape_group", _vp), ("shape_world", _vp), ("shape_mu", _vp), ("shape_mu_torsional", _vp),
        ("shape_mu_rolling", _vp),
        ("segment_count", ctypes.c_int), ("segment_start", _vp), ("segment_shape", _vp),
        ("max_segment_shapes", ctypes.c_int),
        ("filter_count", ctypes.c_int), ("filter_pairs", _vp),
        ("gravity", _vp),
        ("world_body_first", _vp), ("world_body_count", _vp), ("world_joint_first", _vp),
        ("world_joint_count", _vp), ("max_world_bodies", ctypes.c_int), ("max_world_joints", ctypes.c_int),
        ("body_joint_start", _vp), ("body_joint_entry", _vp),
        ("joint_type", _vp), ("joint_enabled", _vp), ("joint_parent", _vp), ("joint_child", _vp),
        ("joint_X_p", _vp), ("joint_X_c", _vp), ("joint_qd_start", _vp), ("joint_dof_dim", _vp),
        ("joint_axis", _vp), ("joint_limit_lower", _vp), ("joint_limit_upper", _vp),
        ("joint_target_ke", _vp), ("joint_target_kd", _vp),
        ("rigid_contact_max", ctypes.c_int), ("shape_pairs_max", ctypes.c_int),
        ("shape_mesh", _vp), ("shape_aabb_lower", _vp), ("shape_aabb_upper", _vp), ("mesh_count", ctypes.c_int),
        ("mesh_vertex", _vp), ("mesh_vertex_start", _vp), ("mesh_edge", _vp), ("mesh_edge_start", _vp),
        ("mesh_sdf_coarse", _vp), ("mesh_sdf_slots", _vp), ("mesh_sdf_texels", _vp), ("mesh_sdf_start", _vp),
        ("mesh_sdf_box", _vp), ("mesh_sdf_quant", _vp), ("mesh_sdf_dims", _vp), ("mesh_sdf_flags", _vp),
        ("mesh_reduce", ctypes.c_int),
        ("shape_voxel_resolution", _vp),
        ("broad_phase_mode", ctypes.c_int), ("explicit_pairs", _vp), ("explicit_pair_count", ctypes.c_int),
        ("shape_material_restitution", _vp),
    ]


class NmcLaunchConfig(ctypes.Structure):
    _fields_ = [("segment_waves", ctypes.c_int), ("narrow_max_blocks", ctypes.c_int),
                ("mesh_max_blocks", ctypes.c_int)]


class NmcContacts(ctypes.Structure):
    _fields_ = [
        ("capacity", ctypes.c_int), ("count", _vp), ("shape0", _vp), ("shape1", _vp), ("point0", _vp),
        ("point1", _vp), ("offset0", _vp), ("offset1", _vp), ("normal", _vp), ("margin0", _vp),
        ("margin1", _vp), ("sort_key", _vp), ("world_start", _vp), ("world_count", _vp), ("order", _vp),
        ("generation", _vp),
    ]


class NmcXpbdParams(ctypes.Structure):
    _fields_ = [
        ("iterations", ctypes.c_int), ("rigid_contact_relaxation", ctypes.c_float),
        ("rigid_contact_con_weighting", ctypes.c_int), ("angular_damping", ctypes.c_float),
        ("joint_linear_relaxation", ctypes.c_float), ("joint_angular_relaxation", ctypes.c_float),
        ("joint_linear_compliance", ctypes.c_float), ("joint_angular_compliance", ctypes.c_float),
        ("has_contacts", ctypes.c_int), ("block_lanes", ctypes.c_int), ("reference_arithmetic", ctypes.c_int),
        ("enable_restitution", ctypes.c_int),
    ]


class NmcLearnIO(ctypes.Structure):
    _fields_ = [
        ("mode", ctypes.c_int), ("warm_iterations", ctypes.c_int), ("residual_tol", ctypes.c_float),
        ("feat", _vp), ("lambda_sum", _vp), ("lambda0", _vp), ("residual", _vp), ("fallback", _vp),
        ("q_mid", _vp), ("qd_mid", _vp), ("prev_key", _vp), ("prev_start", _vp), ("prev_count", _vp),
        ("prev_lambda", _vp), ("prev_order", _vp),
    ]


class NmcMlp(ctypes.Structure):
    _fields_ = [
        ("w1_bf16", _vp), ("b1", _vp), ("w2_bf16", _vp), ("b2", _vp), ("w3", _vp), ("b3", _vp),
        ("x_mean", _vp), ("x_inv_std", _vp), ("out_scale", ctypes.c_float),
    ]


class NmcMlpTrainBuffers(ctypes.Structure):
    _fields_ = [(n, _vp) for n in ("y", "z1", "z2", "xb", "h1", "dz2b", "dz1b", "dz2f", "dz1f", "g3", "dy")]


class NmcMlpGrads(ctypes.Structure):
    _fields_ = [(n, _vp) for n in ("w1", "b1", "w2", "b2", "w3", "b3")]


NMC_STEP_PLAIN, NMC_STEP_LOG, NMC_STEP_PREP, NMC_STEP_WARM = 0, 1, 2, 3

EXPORTS = ("nmc_model_create", "nmc_model_destroy", "nmc_collide", "nmc_xpbd_step", "nmc_xpbd_step_learn",
           "nmc_mlp_forward", "nmc_mlp_train_grad", "nmc_sdf_build", "nmc_sdf_subgrids", "nmc_sdf_pack",
           "nmc_matcher_create", "nmc_matcher_destroy",
           "nmc_matcher_reset", "nmc_contact_match", "nmc_xpbd_step_report", "nmc_contact_force",
           "nmc_broad_phase_create", "nmc_broad_phase_destroy", "nmc_broad_phase_launch", "nmc_broad_phase_explicit",
           "nmc_broad_phase_explicit_noalloc", "nmc_model_set_launch",
           "nmc_get_counters", "nmc_last_error")


class NmcXpbdReport(ctypes.Structure):
    _fields_ = [("contact_impulse", _vp), ("contact_scratch", _vp), ("joint_impulse", _vp), ("body_parent_f", _vp)]


class NmcContactReport(ctypes.Structure):
    _fields_ = [("new_indices", _vp), ("new_count", _vp), ("broken_indices", _vp), ("broken_count", _vp)]


def lib_path() -> str:
    return _LIB_PATH


def load() -> ctypes.CDLL:
    """Load libnmc.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError(
                f"{_LIB_PATH} is missing: build it with `python -m newton_custom_contact_amd._build` "
                "(or __graft_entry__.build())"
            )
        L = ctypes.CDLL(_LIB_PATH)
        L.nmc_model_create.argtypes = [ctypes.POINTER(NmcModelDesc), ctypes.POINTER(_vp)]
        L.nmc_model_create.restype = ctypes.c_int
        L.nmc_model_destroy.argtypes = [_vp]
        L.nmc_model_destroy.restype = None
        L.nmc_collide.argtypes = [_vp, _vp, ctypes.POINTER(NmcContacts), _vp]
        L.nmc_collide.restype = ctypes.c_int
        L.nmc_xpbd_step.argtypes = [_vp, ctypes.POINTER(NmcXpbdParams), _vp, _vp, _vp, _vp, _vp, _vp,
                                    ctypes.POINTER(NmcContacts), ctypes.c_float, _vp, _vp, _vp]
        L.nmc_xpbd_step.restype = ctypes.c_int
        L.nmc_xpbd_step_learn.argtypes = [_vp, ctypes.POINTER(NmcXpbdParams), _vp, _vp, _vp, _vp, _vp, _vp,
                                          ctypes.POINTER(NmcContacts), ctypes.c_float, _vp, _vp,
                                          ctypes.POINTER(NmcLearnIO), _vp]
        L.nmc_xpbd_step_learn.restype = ctypes.c_int
        L.nmc_mlp_forward.argtypes = [ctypes.POINTER(NmcMlp), _vp, _vp, ctypes.c_int, _vp, _vp]
        L.nmc_mlp_forward.restype = ctypes.c_int
        L.nmc_mlp_train_grad.argtypes = [ctypes.POINTER(NmcMlp), _vp, _vp, ctypes.c_int,
                                         ctypes.POINTER(NmcMlpTrainBuffers), ctypes.POINTER(NmcMlpGrads), _vp]
        L.nmc_mlp_train_grad.restype = ctypes.c_int
        L.nmc_sdf_build.argtypes = [_vp, ctypes.c_int, _vp, ctypes.c_int, ctypes.POINTER(ctypes.c_float),
                                    ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_int), ctypes.c_float, _vp,
                                    _vp]
        L.nmc_sdf_build.restype = ctypes.c_int
        L.nmc_model_set_launch.argtypes = [_vp, ctypes.POINTER(NmcLaunchConfig)]
        L.nmc_model_set_launch.restype = ctypes.c_int
        L.nmc_sdf_subgrids.argtypes = [_vp, ctypes.POINTER(ctypes.c_int), ctypes.c_float, ctypes.c_float,
                                       ctypes.c_float, _vp, _vp, _vp, _vp, _vp]
        L.nmc_sdf_subgrids.restype = ctypes.c_int
        L.nmc_sdf_pack.argtypes = [_vp, ctypes.POINTER(ctypes.c_int), _vp, _vp, ctypes.c_int, ctypes.c_int,
                                   ctypes.c_float, ctypes.c_float, _vp, _vp, _vp]
        L.nmc_sdf_pack.restype = ctypes.c_int
        L.nmc_matcher_create.argtypes = [_vp, ctypes.c_int, ctypes.c_float, ctypes.c_float, ctypes.c_int,
                                         ctypes.c_int, ctypes.POINTER(_vp)]
        L.nmc_matcher_create.restype = ctypes.c_int
        L.nmc_matcher_destroy.argtypes = [_vp]
        L.nmc_matcher_destroy.restype = None
        L.nmc_matcher_reset.argtypes = [_vp, _vp]
        L.nmc_matcher_reset.restype = ctypes.c_int
        L.nmc_contact_match.argtypes = [_vp, _vp, _vp, ctypes.POINTER(NmcContacts), _vp,
                                        ctypes.POINTER(NmcContactReport), _vp]
        L.nmc_contact_match.restype = ctypes.c_int
        L.nmc_xpbd_step_report.argtypes = [_vp, ctypes.POINTER(NmcXpbdParams)] + [_vp] * 6 + \
            [ctypes.POINTER(NmcContacts), ctypes.c_float, _vp, _vp, ctypes.POINTER(NmcXpbdReport), _vp]
        L.nmc_xpbd_step_report.restype = ctypes.c_int
        L.nmc_contact_force.argtypes = [_vp, _vp, ctypes.c_int, ctypes.c_float, _vp, _vp]
        L.nmc_contact_force.restype = ctypes.c_int
        _ip = ctypes.POINTER(ctypes.c_int)
        L.nmc_broad_phase_create.argtypes = [ctypes.c_int, _ip, ctypes.c_int, _ip, ctypes.c_int, ctypes.c_int,
                                             ctypes.POINTER(_vp)]
        L.nmc_broad_phase_create.restype = ctypes.c_int
        L.nmc_broad_phase_destroy.argtypes = [_vp]
        L.nmc_broad_phase_destroy.restype = None
        L.nmc_broad_phase_launch.argtypes = [_vp, ctypes.c_int, _vp, _vp, _vp, _vp, _vp, _vp, ctypes.c_int, _vp,
                                             ctypes.c_int, _vp, ctypes.c_int, _vp]
        L.nmc_broad_phase_launch.restype = ctypes.c_int
        L.nmc_broad_phase_explicit.argtypes = [_vp, ctypes.c_int, _vp, _vp, _vp, _vp, ctypes.c_int, _vp,
                                               ctypes.c_int, _vp, ctypes.c_int, _vp]
        L.nmc_broad_phase_explicit.restype = ctypes.c_int
        L.nmc_broad_phase_explicit_noalloc.argtypes = [ctypes.c_int, _vp, _vp, _vp, _vp, ctypes.c_int, _vp,
                                                       ctypes.c_int, _vp, ctypes.c_int, _vp]
        L.nmc_broad_phase_explicit_noalloc.restype = ctypes.c_int
        L.nmc_get_counters.argtypes = [_vp, ctypes.POINTER(ctypes.c_int), _vp]
        L.nmc_get_counters.restype = ctypes.c_int
        L.nmc_last_error.argtypes = []
        L.nmc_last_error.restype = ctypes.c_char_p
        _lib = L
    return _lib


def check(rc: int, what: str):
    if rc != 0:
        msg = load().nmc_last_error().decode()
        raise RuntimeError(f"{what} failed: {msg}")


def ptr(t: torch.Tensor | None) -> int | None:
    if t is None:
        return None
    if not t.is_contiguous():
        raise ValueError("device arrays passed to libnmc must be contiguous")
    return t.data_ptr() if t.numel() > 0 else None


def stream_handle(device: torch.device) -> int:
    return torch.cuda.current_stream(device).cuda_stream


def check_mesh_pairs(model):
    """Refuse models whose mesh shapes could meet a shape other than a mesh or an infinite plane: the
    reference routes those pairs to its mesh-triangle + GJK/MPR path (narrow_phase.py:315-320), which is
    not part of this library."""
    h = model.host
    if not getattr(model, "meshes", []):
        return
    coll = (h["shape_flags"] & 2) != 0  # ShapeFlags.COLLIDE_SHAPES
    t = h["shape_type"]
    sc = h["shape_scale"]
    infinite_plane = (t == 1) & (sc[:, 0] == 0.0) & (sc[:, 1] == 0.0)
    other = coll & (t != 8) & ~infinite_plane
    if other.any():
        raise NotImplementedError(
            "a model with mesh shapes may only hold other mesh shapes and infinite planes as colliders "
            f"(shape {int(np.nonzero(other)[0][0])} has GeoType {int(t[other][0])}); mesh-vs-primitive "
            "contacts are not supported")


class NativeModel:
    """Owns one NmcModel handle plus the extra device tables it points at."""

    def __init__(self, model, rigid_contact_max: int, shape_pairs_max: int, broad_phase_mode: int = 0,
                 explicit_pairs: torch.Tensor | None = None, reduce_contacts: bool = True,
                 deterministic: bool = False):
        if model.device.type != "cuda":
            raise RuntimeError(
                "newton_custom_contact_amd runs its contact step on the GPU only (HIP kernels in libnmc.so); "
                f"model.device is {model.device}"
            )
        if model.world_count > 0 and int((model.host["body_world"] == -1).sum()) > 0:
            raise NotImplementedError(
                "bodies outside every world (world -1) next to regular worlds are not supported; "
                "place them inside a world (static global geometry uses body -1 and is fine)"
            )
        check_mesh_pairs(model)
        L = load()
        h = model.host
        dev = model.device
        W = model.world_count
        B = model.body_count
        J = model.joint_count
        # per-world body/joint ranges; index W = global entities
        wbf = np.zeros(W + 1, np.int32)
        wbc = np.zeros(W + 1, np.int32)
        wjf = np.zeros(W + 1, np.int32)
        wjc = np.zeros(W + 1, np.int32)
        bw = h["body_world"]
        jw = h["joint_world"]
        for w in list(range(W)) + [-1]:
            idx = W if w == -1 else w
            sel = np.nonzero(bw == w)[0]
            if sel.size:
                if not np.all(np.diff(sel) == 1):
                    raise ValueError(f"bodies of world {w} are not contiguous")
                wbf[idx], wbc[idx] = sel[0], sel.size
            sj = np.nonzero(jw == w)[0]
            if sj.size:
                if not np.all(np.diff(sj) == 1):
                    raise ValueError(f"joints of world {w} are not contiguous")
                wjf[idx], wjc[idx] = sj[0], sj.size
        # body -> (joint, side) in joint order
        lists = [[] for _ in range(B)]
        for j in range(J):
            p = int(h["joint_parent"][j])
            c = int(h["joint_child"][j])
            if p >= 0:
                lists[p].append(2 * j)
            lists[c].append(2 * j + 1)
        bjs = np.zeros(B + 1, np.int32)
        bjs[1:] = np.cumsum([len(x) for x in lists]) if B else []
        bje = np.asarray([e for x in lists for e in x], np.int32)
        seg_start = h["segment_start"]
        max_seg = int(np.max(np.diff(seg_start))) if len(seg_start) > 1 else 0
        self.tables = {
            "world_body_first": torch.from_numpy(wbf).to(dev), "world_body_count": torch.from_numpy(wbc).to(dev),
            "world_joint_first": torch.from_numpy(wjf).to(dev), "world_joint_count": torch.from_numpy(wjc).to(dev),
            "body_joint_start": torch.from_numpy(bjs).to(dev),
            "body_joint_entry": torch.from_numpy(bje if bje.size else np.zeros(1, np.int32)).to(dev),
        }
        d = NmcModelDesc()
        d.body_count, d.shape_count, d.joint_count, d.world_count = B, model.shape_count, J, W
        d.joint_dof_count = model.joint_dof_count
        m = model
        d.body_com, d.body_inertia = ptr(m.body_com), ptr(m.body_inertia)
        d.body_inv_mass, d.body_inv_inertia = ptr(m.body_inv_mass), ptr(m.body_inv_inertia)
        d.body_flags, d.body_world = ptr(m.body_flags), ptr(m.body_world)
        d.shape_transform, d.shape_body, d.shape_type = ptr(m.shape_transform), ptr(m.shape_body), ptr(m.shape_type)
        d.shape_scale, d.shape_margin, d.shape_gap = ptr(m.shape_scale), ptr(m.shape_margin), ptr(m.shape_gap)
        d.shape_radius, d.shape_flags = ptr(m.shape_collision_radius), ptr(m.shape_flags)
        d.shape_group, d.shape_world = ptr(m.shape_collision_group), ptr(m.shape_world)
        d.shape_mu, d.shape_mu_torsional = ptr(m.shape_material_mu), ptr(m.shape_material_mu_torsional)
        d.shape_mu_rolling = ptr(m.shape_material_mu_rolling)
        d.shape_material_restitution = ptr(m.shape_material_restitution)
        d.segment_count = len(seg_start) - 1
        d.segment_start, d.segment_shape = ptr(m.segment_start), ptr(m.segment_shape)
        d.max_segment_shapes = max_seg
        d.filter_count = int(len(h["shape_filter_pairs"]))
        d.filter_pairs = ptr(m.shape_filter_pairs)
        d.gravity = ptr(m.gravity)
        t = self.tables
        d.world_body_first, d.world_body_count = ptr(t["world_body_first"]), ptr(t["world_body_count"])
        d.world_joint_first, d.world_joint_count = ptr(t["world_joint_first"]), ptr(t["world_joint_count"])
        d.max_world_bodies = int(wbc.max()) if wbc.size else 0
        d.max_world_joints = int(wjc.max()) if wjc.size else 0
        d.body_joint_start, d.body_joint_entry = ptr(t["body_joint_start"]), ptr(t["body_joint_entry"])
        d.joint_type, d.joint_enabled = ptr(m.joint_type), ptr(m.joint_enabled)
        d.joint_parent, d.joint_child = ptr(m.joint_parent), ptr(m.joint_child)
        d.joint_X_p, d.joint_X_c = ptr(m.joint_X_p), ptr(m.joint_X_c)
        d.joint_qd_start, d.joint_dof_dim = ptr(m.joint_qd_start), ptr(m.joint_dof_dim)
        d.joint_axis = ptr(m.joint_axis)
        d.joint_limit_lower, d.joint_limit_upper = ptr(m.joint_limit_lower), ptr(m.joint_limit_upper)
        d.joint_target_ke, d.joint_target_kd = ptr(m.joint_target_ke), ptr(m.joint_target_kd)
        d.rigid_contact_max = int(rigid_contact_max)
        d.shape_pairs_max = int(min(shape_pairs_max, 2**31 - 1))
        d.shape_mesh = ptr(m.shape_mesh)
        d.shape_aabb_lower, d.shape_aabb_upper = ptr(m.shape_aabb_lower), ptr(m.shape_aabb_upper)
        d.mesh_count = len(getattr(m, "meshes", []))
        if d.mesh_count and not m.mesh_has_sdf:
            raise RuntimeError("mesh shapes need their SDF (Mesh.build_sdf) before the collision pipeline")
        d.mesh_vertex, d.mesh_vertex_start = ptr(m.mesh_vertex), ptr(m.mesh_vertex_start)
        d.mesh_edge, d.mesh_edge_start = ptr(m.mesh_edge), ptr(m.mesh_edge_start)
        d.mesh_sdf_coarse, d.mesh_sdf_slots = ptr(m.mesh_sdf_coarse), ptr(m.mesh_sdf_slots)
        d.mesh_sdf_texels, d.mesh_sdf_start = ptr(m.mesh_sdf_texels), ptr(m.mesh_sdf_start)
        d.mesh_sdf_box, d.mesh_sdf_quant = ptr(m.mesh_sdf_box), ptr(m.mesh_sdf_quant)
        d.mesh_sdf_dims, d.mesh_sdf_flags = ptr(m.mesh_sdf_dims), ptr(m.mesh_sdf_flags)
        # mesh contact reduction and its value packing (the reducer's deterministic flag is the pipeline's)
        d.mesh_reduce = 0 if not reduce_contacts else (2 if deterministic else 1)
        d.shape_voxel_resolution = ptr(m.shape_voxel_resolution)
        d.broad_phase_mode = int(broad_phase_mode)
        self.explicit_pairs = explicit_pairs
        d.explicit_pairs = ptr(explicit_pairs)
        d.explicit_pair_count = 0 if explicit_pairs is None else int(explicit_pairs.shape[0])
        self.desc = d
        self.model = model
        handle = _vp()
        with torch.cuda.device(dev):
            check(L.nmc_model_create(ctypes.byref(d), ctypes.byref(handle)), "nmc_model_create")
        self.handle = handle
        self._lib = L

    def counters(self) -> np.ndarray:
        out = (ctypes.c_int * 4)()
        check(self._lib.nmc_get_counters(self.handle, out, _vp(stream_handle(self.model.device))), "nmc_get_counters")
        return np.asarray(list(out), np.int32)

    def __del__(self):
        try:
            if getattr(self, "handle", None) and self.handle.value:
                self._lib.nmc_model_destroy(self.handle)
                self.handle = _vp()
        except Exception:
            pass
