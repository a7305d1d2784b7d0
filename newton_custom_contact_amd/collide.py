"""``CollisionPipeline``: AABBs -> SAP broad phase -> narrow phase -> Contacts, on the GPU.

Mirrors newton/_src/sim/collide.py (class ``CollisionPipeline``, line 464): constructor arguments
(``collide.py:483-504``), ``contacts()`` (``:858-885``), ``collide(state, contacts)`` (``:902``) and the
capacity estimators (``_estimate_rigid_contact_max`` ``:278-376``, ``_compute_per_world_shape_pairs_max``
``:379-418``, ``_resolve_shape_pairs_max`` ``:421-439``).

``collide`` launches its kernels from libnmc.so (segment broad phase, route scan, primitive narrow phase,
mesh narrow phase when the model has meshes, contact count / scan / compaction) on the current torch
stream and never synchronises the host, so it can be captured in a graph.  The broad phase runs per
segment out of LDS with the reference's pair sets: ``"sap"`` tests a pair when the sweep on the fixed axis
reaches it (projected intervals overlapping strictly), ``"nxn"`` tests every pair of a segment, and
``"explicit"`` (the default) tests ``model.shape_contact_pairs`` -- the ``"nxn"`` set -- or, given
``shape_pairs_filtered``, exactly the listed pairs (AABB overlap only).  Mesh shapes collide with meshes (SDF edge search) and infinite planes (vertex
test); a model where a mesh could meet any other shape type is refused.
"""

from __future__ import annotations

import ctypes

import numpy as np
import torch

from ._native import NativeModel, NmcContactReport, NmcLaunchConfig, check, load, ptr, stream_handle
from ._types import GeoType, ShapeFlags
from .contacts import Contacts

_SUPPORTED = {int(GeoType.PLANE), int(GeoType.SPHERE), int(GeoType.CAPSULE), int(GeoType.ELLIPSOID),
              int(GeoType.CYLINDER), int(GeoType.BOX), int(GeoType.MESH), int(GeoType.CONE)}


def estimate_rigid_contact_max(model) -> int:
    """collide.py:278-376 (neighbour-budget estimate, per-world plane term)."""
    h = model.host
    st = h["shape_type"]
    PRIMITIVE_CPP, MESH_CPP, MAX_NEIGHBORS = 5, 40, 20
    mesh_mask = (st == int(GeoType.MESH)) | (st == int(GeoType.HFIELD))
    plane_mask = st == int(GeoType.PLANE)
    non_plane = ~plane_mask
    n_mesh = int(np.count_nonzero(mesh_mask))
    n_np = int(np.count_nonzero(non_plane))
    n_prim = n_np - n_mesh
    n_planes = int(np.count_nonzero(plane_mask))
    non_plane_contacts = (n_prim * MAX_NEIGHBORS * PRIMITIVE_CPP + n_mesh * MAX_NEIGHBORS * MESH_CPP) // 2
    avg_cpp = (n_prim * PRIMITIVE_CPP + n_mesh * MESH_CPP) // max(n_np, 1) if n_np > 0 else 0
    plane_contacts = 0
    if n_planes > 0 and n_np > 0:
        sw = h["shape_world"]
        W = model.world_count
        if W > 0:
            gmask = sw == -1
            lmask = ~gmask
            g_planes = int(np.count_nonzero(gmask & plane_mask))
            g_np = int(np.count_nonzero(gmask & non_plane))
            lp = np.bincount(sw[lmask & plane_mask], minlength=W)[:W]
            lnp = np.bincount(sw[lmask & non_plane], minlength=W)[:W]
            ppw = lp + g_planes
            npw = lnp + g_np
            ppc = int(np.sum(ppw.astype(np.int64) * npw.astype(np.int64)))
            if W > 1:
                ppc -= (W - 1) * g_planes * g_np
            plane_contacts = ppc * avg_cpp
        else:
            plane_contacts = n_planes * (n_prim * PRIMITIVE_CPP + n_mesh * MESH_CPP)
    return max(1000, non_plane_contacts + plane_contacts)


def per_world_shape_pairs_max(model) -> int:
    """collide.py:379-418: sum of n(n-1)/2 over the broad-phase segments."""
    seg = np.diff(model.host["segment_start"]).astype(np.int64)
    return int(np.sum(seg * (seg - 1) // 2))


def _explicit_pair_list(model, pairs) -> torch.Tensor:
    """A caller's pair list for broad_phase="explicit": canonical (min, max), sorted, without repeats.  The
    pairs are found inside broad-phase segments, so each must join two colliding shapes of one world (or a
    world and a world -1 shape, or two world -1 shapes); the reference would also test other pairs."""
    p = pairs.detach().cpu().numpy() if isinstance(pairs, torch.Tensor) else np.asarray(pairs)
    p = np.asarray(p, np.int64).reshape(-1, 2)
    S = model.shape_count
    if len(p) and (p.min() < 0 or p.max() >= S):
        raise ValueError("shape_pairs_filtered names shapes outside the model")
    p = p[p[:, 0] != p[:, 1]]
    p = np.unique(np.sort(p, axis=1), axis=0)
    h = model.host
    sw = h["shape_world"]
    coll = (h["shape_flags"] & int(ShapeFlags.COLLIDE_SHAPES)) != 0
    wa, wb = sw[p[:, 0]], sw[p[:, 1]]
    if np.any((wa != -1) & (wb != -1) & (wa != wb)) or not np.all(coll[p[:, 0]] & coll[p[:, 1]]):
        raise NotImplementedError("shape_pairs_filtered may only join colliding shapes of one world "
                                  "(or with world -1 shapes)")
    return torch.from_numpy(np.ascontiguousarray(p, np.int32).reshape(-1, 2))


class CollisionPipeline:
    """GPU collision pipeline with the reference's constructor and ``collide`` API.

    ``collide`` may be captured in a graph; the graph then points into this pipeline's device scratch, so the
    pipeline (and the model, states and contacts it was captured with) must outlive the graph."""

    def __init__(self, model, *, reduce_contacts: bool = True, rigid_contact_max: int | None = None,
                 max_triangle_pairs: int = 1000000, shape_pairs_filtered=None, soft_contact_max: int | None = None,
                 soft_contact_margin: float = 0.01, requires_grad: bool | None = None, broad_phase=None,
                 narrow_phase=None, sdf_hydroelastic_config=None, shape_pairs_max: int | None = None,
                 deterministic: bool = False, contact_matching: str = "disabled",
                 contact_matching_pos_threshold: float = 0.0005, contact_matching_normal_dot_threshold: float = 0.995,
                 contact_report: bool = False, verify_buffers: bool = True):
        if contact_matching not in ("disabled", "latest", "sticky"):
            raise ValueError(f"contact_matching must be one of 'disabled', 'latest', 'sticky', got {contact_matching!r}")
        if contact_matching_pos_threshold < 0.0:
            raise ValueError(
                f"contact_matching_pos_threshold must be non-negative, got {contact_matching_pos_threshold}")
        if not -1.0 <= contact_matching_normal_dot_threshold <= 1.0:
            raise ValueError(
                f"contact_matching_normal_dot_threshold must be in [-1, 1], got {contact_matching_normal_dot_threshold}")
        matching_enabled = contact_matching != "disabled"
        if contact_report and not matching_enabled:
            raise ValueError('contact_report=True requires contact_matching != "disabled"')
        if matching_enabled:
            deterministic = True  # any matching mode implies the deterministic (sort-keyed) contact order
        from .geometry import BroadPhaseAllPairs, BroadPhaseExplicit, BroadPhaseSAP  # noqa: PLC0415

        if broad_phase is None or isinstance(broad_phase, str):
            mode = "explicit" if broad_phase is None else str(broad_phase).lower()
            if mode not in ("nxn", "sap", "explicit"):
                raise ValueError(f"Unsupported broad phase mode: {broad_phase!r}")
            expert = narrow_phase is not None
        else:
            kinds = ((BroadPhaseAllPairs, "nxn"), (BroadPhaseSAP, "sap"), (BroadPhaseExplicit, "explicit"))
            hit = [name for cls, name in kinds if isinstance(broad_phase, cls)]
            if not hit:
                raise TypeError("broad_phase must be a BroadPhaseAllPairs, BroadPhaseSAP, or BroadPhaseExplicit "
                                f"instance (got {type(broad_phase)!r})")
            mode = hit[0]
            expert = True
        if expert:
            if narrow_phase is None or (broad_phase is None or isinstance(broad_phase, str)):
                raise ValueError("Provide both broad_phase and narrow_phase for expert component construction")
            raise NotImplementedError("expert narrow_phase objects are not supported: the narrow phase runs in "
                                      "libnmc.so")
        # "explicit" tests model.shape_contact_pairs, the same set "nxn" visits; a caller's own list is tested
        # as given (AABB overlap only, collide.py:695-709)
        explicit_pairs = None
        bp_mode = {"sap": 0, "nxn": 1, "explicit": 1}[mode]
        if mode == "explicit" and shape_pairs_filtered is not None:
            explicit_pairs = _explicit_pair_list(model, shape_pairs_filtered)
            bp_mode = 2
        types = set(np.unique(model.host["shape_type"]).tolist())
        bad = types - _SUPPORTED
        if bad:
            names = ", ".join(GeoType(t).name for t in sorted(bad))
            raise NotImplementedError(f"shape types not supported by the GPU narrow phase: {names}")
        self.model = model
        self.broad_phase_mode = mode
        self.device = model.device
        self.deterministic = bool(deterministic)
        self.verify_buffers = verify_buffers
        self.reduce_contacts = reduce_contacts
        self.requires_grad = bool(model.requires_grad if requires_grad is None else requires_grad)
        self.soft_contact_margin = soft_contact_margin
        if rigid_contact_max is None:
            rigid_contact_max = int(getattr(model, "rigid_contact_max", 0) or 0) or estimate_rigid_contact_max(model)
        self._rigid_contact_max = int(rigid_contact_max)
        model.rigid_contact_max = self._rigid_contact_max
        if shape_pairs_max is None:
            shape_pairs_max = per_world_shape_pairs_max(model)
        elif shape_pairs_max <= 0:
            raise ValueError(f"shape_pairs_max must be a positive integer or None, got {shape_pairs_max}")
        self.shape_pairs_max = int(shape_pairs_max)
        self._soft_contact_max = int(soft_contact_max) if soft_contact_max is not None else 0
        colliding = (model.host["shape_flags"] & int(ShapeFlags.COLLIDE_SHAPES)) != 0
        self.shape_count = model.shape_count
        self._colliding = int(np.count_nonzero(colliding))
        self.shape_pairs_filtered = explicit_pairs
        self.native = NativeModel(model, self._rigid_contact_max, self.shape_pairs_max, bp_mode,
                                  None if explicit_pairs is None else explicit_pairs.to(model.device),
                                  reduce_contacts=bool(reduce_contacts), deterministic=self.deterministic)
        self._lib = load()
        self._launch = {"segment_waves": 4, "narrow_max_blocks": 2048, "mesh_max_blocks": 2048}
        self.contact_matching = contact_matching
        self._matching_enabled = matching_enabled
        self._matching_sticky = contact_matching == "sticky"
        self.contact_report = bool(contact_report)
        self._contact_matcher = (
            ContactMatcher(self.native, self._rigid_contact_max, contact_matching_pos_threshold,
                           contact_matching_normal_dot_threshold, sticky=self._matching_sticky,
                           contact_report=self.contact_report)
            if matching_enabled else None)

    @property
    def contact_matcher(self):
        """The frame-to-frame matcher (None when contact_matching is "disabled"); ``reset()`` forgets the
        previous frame (after teleporting bodies or an episode reset)."""
        return self._contact_matcher

    @property
    def rigid_contact_max(self) -> int:
        return self._rigid_contact_max

    @property
    def soft_contact_max(self) -> int:
        return self._soft_contact_max

    @property
    def launch_config(self) -> dict:
        """Launch shape of the collide kernels (``autotune.tune_collide`` picks it): ``segment_waves`` (segments
        of at most 64 shapes per broad-phase workgroup: 1, 2 or 4), ``narrow_max_blocks`` and
        ``mesh_max_blocks`` (grid caps of the narrow-phase and mesh kernels).  Contacts are bit-identical for
        every shape."""
        return dict(self._launch)

    def set_launch(self, segment_waves: int | None = None, narrow_max_blocks: int | None = None,
                   mesh_max_blocks: int | None = None) -> None:
        """Set the collide kernels' launch shape (``None`` keeps the current value); raises ``ValueError`` for
        values the kernels do not take."""
        cfg = dict(self._launch)
        for k, v in (("segment_waves", segment_waves), ("narrow_max_blocks", narrow_max_blocks),
                     ("mesh_max_blocks", mesh_max_blocks)):
            if v is not None:
                cfg[k] = int(v)
        if cfg["segment_waves"] not in (1, 2, 4):
            raise ValueError("segment_waves must be 1, 2 or 4")
        if not (1 <= cfg["narrow_max_blocks"] <= 65535 and 1 <= cfg["mesh_max_blocks"] <= 65535):
            raise ValueError("grid caps must lie in [1, 65535]")
        check(self._lib.nmc_model_set_launch(self.native.handle, ctypes.byref(NmcLaunchConfig(**cfg))),
              "nmc_model_set_launch")
        self._launch = cfg

    def contacts(self) -> Contacts:
        """collide.py:858-885."""
        return Contacts(self.rigid_contact_max, self.soft_contact_max, requires_grad=self.requires_grad,
                        device=self.device, segment_count=len(self.model.host["segment_start"]) - 1,
                        deterministic=self.deterministic,
                        requested_attributes=self.model.get_requested_contact_attributes(),
                        contact_matching=self._matching_enabled, contact_report=self.contact_report)

    def collide(self, state, contacts: Contacts, *, soft_contact_margin: float | None = None):
        """collide.py:902-1230 (rigid part): writes ``contacts`` from ``state.body_q``."""
        if contacts.rigid_contact_max != self.rigid_contact_max:
            raise ValueError(
                f"Contacts buffer capacity ({contacts.rigid_contact_max}) does not match the pipeline "
                f"({self.rigid_contact_max}); use CollisionPipeline.contacts()"
            )
        if self.deterministic and contacts.rigid_contact_sort_key is None:
            raise ValueError("deterministic pipeline needs a Contacts buffer with sort keys; use pipeline.contacts()")
        q = state.body_q
        if q is None or q.device != self.device or not q.is_contiguous():
            raise ValueError("state.body_q must be a contiguous tensor on the model device")
        rc = self._lib.nmc_collide(self.native.handle, ctypes.c_void_p(ptr(q)), ctypes.byref(contacts.native_view()),
                                   ctypes.c_void_p(stream_handle(self.device)))
        check(rc, "nmc_collide")
        if self._contact_matcher is not None:
            if contacts.rigid_contact_match_index is None:
                raise ValueError("CollisionPipeline has contact_matching enabled but the Contacts buffer was created "
                                 "without contact_matching. Use pipeline.contacts() to create a compatible buffer.")
            if self.contact_report and contacts.rigid_contact_new_indices is None:
                raise ValueError("CollisionPipeline has contact_report enabled but the Contacts buffer was created "
                                 "without contact_report=True. Use pipeline.contacts() to create a compatible buffer.")
            self._contact_matcher.match(q, contacts)

    def overflow_counters(self) -> dict:
        """Overflow counters of every collide so far (verify_narrow_phase_buffers, narrow_phase.py:1334)."""
        c = self.native.counters()
        return {"pairs_dropped": int(c[0]), "mesh_candidates_dropped": int(c[1]),
                "contacts_over_capacity": int(c[2]), "solver_ordering_overflow": int(c[3])}


class ContactMatcher:
    """Frame-to-frame contact matching (contact_match.py:570-992) over ``nmc_contact_match``.  Owned by the
    pipeline; keeps the previous frame's contacts on the device."""

    def __init__(self, native: NativeModel, capacity: int, pos_threshold: float, normal_dot_threshold: float,
                 sticky: bool, contact_report: bool):
        self._lib = load()
        self._native = native
        self.handle = ctypes.c_void_p()
        self.is_sticky = sticky
        self.has_report = contact_report
        check(self._lib.nmc_matcher_create(native.handle, int(capacity), ctypes.c_float(pos_threshold),
                                           ctypes.c_float(normal_dot_threshold), int(sticky), int(contact_report),
                                           ctypes.byref(self.handle)), "nmc_matcher_create")

    def reset(self) -> None:
        """Forget the previous frame: the next match reports every contact as MATCH_NOT_FOUND."""
        check(self._lib.nmc_matcher_reset(self.handle, ctypes.c_void_p(stream_handle(self._native.model.device))),
              "nmc_matcher_reset")

    def match(self, body_q, contacts) -> None:
        rep = None
        if self.has_report:
            rep = NmcContactReport(ptr(contacts.rigid_contact_new_indices), ptr(contacts.rigid_contact_new_count),
                                   ptr(contacts.rigid_contact_broken_indices), ptr(contacts.rigid_contact_broken_count))
        rc = self._lib.nmc_contact_match(self.handle, self._native.handle, ctypes.c_void_p(ptr(body_q)),
                                         ctypes.byref(contacts.native_view()),
                                         ctypes.c_void_p(ptr(contacts.rigid_contact_match_index)),
                                         ctypes.byref(rep) if rep is not None else None,
                                         ctypes.c_void_p(stream_handle(self._native.model.device)))
        check(rc, "nmc_contact_match")

    def __del__(self):
        try:
            if self.handle:
                self._lib.nmc_matcher_destroy(self.handle)
                self.handle = ctypes.c_void_p()
        except Exception:  # noqa: BLE001 (interpreter shutdown)
            pass


__all__ = ["CollisionPipeline", "ContactMatcher", "estimate_rigid_contact_max", "per_world_shape_pairs_max"]
