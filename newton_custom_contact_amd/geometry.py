"""Standalone broad phases (``newton.geometry``: ``BroadPhaseSAP``, ``BroadPhaseAllPairs``,
``BroadPhaseExplicit``; newton/_src/geometry/broad_phase_sap.py:397-677, broad_phase_nxn.py:217-472).

Each class keeps the reference's constructor and ``launch`` arguments: AABB bounds ``(N, 3)`` float32, optional
per-shape gaps, collision groups and worlds as device tensors; the overlapping pairs go to
``candidate_pair`` ``(M, 2)`` int32 and their number to ``candidate_pair_count[0]`` (it may exceed ``M``; pairs
beyond ``M`` are dropped, as in the reference).  The work runs in libnmc.so (``nmc_broad_phase_launch`` /
``nmc_broad_phase_explicit``): per segment of ``precompute_world_map``, count, scan, write, with no atomics, so
the order of the pairs is the same on every launch (the reference appends through an atomic counter and its
order varies):

* ``BroadPhaseAllPairs``: segment by segment, then (row, column) of the segment's index-map order;
* ``BroadPhaseSAP``: segment by segment, then sweep origin in sorted order, then sweep position.  The pair set
  is the reference's: the earlier shape of a pair in the stable sort of projected lower bounds sweeps the later
  one while ``lower(b) < upper(a)``, so AABBs whose projections only touch are not reported (all pairs
  reports them when the AABBs touch);
* ``BroadPhaseExplicit``: the given pairs whose AABBs overlap, in input order.

``sort_type``, ``tile_block_dim`` and ``sweep_thread_count_multiplier`` select the reference's Warp launch
strategies; they are validated and kept, and do not change the result.  ``skip_count_zero=True`` leaves
``candidate_pair_count`` as the caller set it: the pairs found go after that many and the count grows by their
number (the reference appends through its atomic counter the same way; its callers pass 0 there).
"""

from __future__ import annotations

import ctypes

import numpy as np
import torch

from ._types import ShapeFlags

__all__ = ["BroadPhaseAllPairs", "BroadPhaseExplicit", "BroadPhaseSAP", "precompute_world_map"]


def precompute_world_map(shape_world, shape_flags=None):
    """(index_map, slice_ends) of the segments (broad_phase_common.py:183-300): each world's colliding shapes
    in increasing index followed by every world -1 shape, per world in increasing world id, then one last
    segment with only the world -1 shapes.  ``shape_flags`` keeps only shapes with ``COLLIDE_SHAPES``."""
    world = np.asarray(shape_world.cpu() if isinstance(shape_world, torch.Tensor) else shape_world).astype(np.int64)
    world = world.reshape(-1)
    if shape_flags is not None:
        flags = np.asarray(shape_flags.cpu() if isinstance(shape_flags, torch.Tensor) else shape_flags).reshape(-1)
        if flags.shape[0] != world.shape[0]:
            raise ValueError("shape_flags and shape_world must have the same length")
        keep = (flags.astype(np.int64) & int(ShapeFlags.COLLIDE_SHAPES)) != 0
    else:
        keep = np.ones(world.shape[0], bool)
    bad = np.unique(world[world < -1])
    if bad.size:
        raise ValueError(f"Invalid world IDs detected: {bad.tolist()}. "
                         "Only world ID -1 (global/shared) and non-negative IDs (0, 1, 2, ...) are supported.")
    valid = np.nonzero(keep)[0]
    vw = world[valid]
    shared = valid[vw == -1]
    worlds = np.unique(vw[vw >= 0])
    parts, ends, pos = [], [], 0
    for w in worlds:
        own = valid[vw == w]
        parts += [own, shared]
        pos += len(own) + len(shared)
        ends.append(pos)
    parts.append(shared)
    ends.append(pos + len(shared))
    index_map = np.concatenate(parts).astype(np.int32) if parts else np.zeros(0, np.int32)
    return index_map, np.asarray(ends, np.int32)


def _check_tensor(name, t, dtype, shape_tail=None, device=None):
    if not isinstance(t, torch.Tensor):
        raise TypeError(f"{name} must be a torch tensor on the GPU")
    if t.dtype != dtype:
        raise TypeError(f"{name} must have dtype {dtype}, got {t.dtype}")
    if shape_tail is not None and tuple(t.shape[1:]) != shape_tail:
        raise ValueError(f"{name} must have shape (n, {', '.join(map(str, shape_tail))}), got {tuple(t.shape)}")
    if device is not None and t.device != device:
        raise ValueError(f"{name} is on {t.device}, expected {device}")
    if not t.is_contiguous():
        raise ValueError(f"{name} must be contiguous")


class _Native:
    """One libnmc broad-phase context per device."""

    def __init__(self, mode: int, index_map, slice_ends, pair_capacity: int, device: torch.device):
        from ._native import check, load

        self.L = load()
        self.device = device
        self.h = ctypes.c_void_p()
        im = np.ascontiguousarray(index_map if index_map is not None and len(index_map) else np.zeros(1), np.int32)
        se = np.ascontiguousarray(slice_ends if slice_ends is not None else np.zeros(1), np.int32)
        n_im = 0 if index_map is None else int(len(index_map))
        n_se = 0 if slice_ends is None else int(len(slice_ends))
        ip = ctypes.POINTER(ctypes.c_int)
        with torch.cuda.device(device):
            check(self.L.nmc_broad_phase_create(int(mode), im.ctypes.data_as(ip), n_im, se.ctypes.data_as(ip), n_se,
                                                int(pair_capacity), ctypes.byref(self.h)), "nmc_broad_phase_create")
        self.pair_capacity = pair_capacity

    def __del__(self):
        h = getattr(self, "h", None)
        if h is not None and h.value:
            self.L.nmc_broad_phase_destroy(h)
            self.h = None


class _SegmentBroadPhase:
    _mode = 1

    def __init__(self, shape_world, shape_flags=None, device=None):
        self.world_index_map, self.world_slice_ends = precompute_world_map(shape_world, shape_flags)
        self.num_regular_worlds = max(0, len(self.world_slice_ends) - 1)
        starts = np.concatenate([[0], self.world_slice_ends[:-1]])
        sizes = self.world_slice_ends - starts
        self.max_shapes_per_world = int(sizes.max()) if sizes.size else 0
        self.num_kernel_threads = int((sizes.astype(np.int64) * (sizes - 1) // 2).sum())
        self.device = None if device is None else torch.device(device)
        self._native: _Native | None = None

    def _context(self, device: torch.device) -> _Native:
        if self._native is None or self._native.device != device:
            self._native = _Native(self._mode, self.world_index_map, self.world_slice_ends, 0, device)
        return self._native

    def launch(self, shape_lower, shape_upper, shape_gap, shape_collision_group, shape_world, shape_count,
               candidate_pair, candidate_pair_count, device=None, filter_pairs=None, num_filter_pairs=None,
               skip_count_zero: bool = False) -> None:
        """Find the candidate pairs (broad_phase_sap.py:510-677, broad_phase_nxn.py:305-391)."""
        from ._native import check, ptr, stream_handle

        _check_tensor("shape_lower", shape_lower, torch.float32, (3,))
        dev = shape_lower.device if device is None else torch.device(device)
        if dev.type != "cuda":
            raise RuntimeError("the broad phase runs on the GPU only")
        n = shape_lower.shape[0]
        _check_tensor("shape_upper", shape_upper, torch.float32, (3,), dev)
        if shape_upper.shape[0] != n:
            raise ValueError("shape_lower and shape_upper must have the same length")
        if shape_gap is not None and shape_gap.numel() == 0:
            shape_gap = None
        if shape_gap is not None:
            _check_tensor("shape_gap", shape_gap, torch.float32, (), dev)
            if shape_gap.shape[0] < n:
                raise ValueError("shape_gap is shorter than the AABB arrays")
        for name, t in (("shape_collision_group", shape_collision_group), ("shape_world", shape_world)):
            _check_tensor(name, t, torch.int32, (), dev)
            if t.shape[0] < n:
                raise ValueError(f"{name} is shorter than the AABB arrays")
        _check_tensor("candidate_pair", candidate_pair, torch.int32, (2,), dev)
        _check_tensor("candidate_pair_count", candidate_pair_count, torch.int32, None, dev)
        if candidate_pair_count.numel() < 1:
            raise ValueError("candidate_pair_count needs one element")
        nf = 0
        if filter_pairs is not None and filter_pairs.numel() > 0:
            _check_tensor("filter_pairs", filter_pairs, torch.int32, (2,), dev)
            nf = int(filter_pairs.shape[0]) if num_filter_pairs is None else int(num_filter_pairs)
            if not 0 <= nf <= filter_pairs.shape[0]:
                raise ValueError("num_filter_pairs exceeds filter_pairs")
        else:
            filter_pairs = None
        ctx = self._context(dev)
        check(ctx.L.nmc_broad_phase_launch(ctx.h, int(n), ptr(shape_lower), ptr(shape_upper), ptr(shape_gap),
                                           ptr(shape_collision_group), ptr(shape_world), ptr(filter_pairs), nf,
                                           ptr(candidate_pair), int(candidate_pair.shape[0]),
                                           ptr(candidate_pair_count), int(bool(skip_count_zero)), stream_handle(dev)),
              "nmc_broad_phase_launch")


class BroadPhaseAllPairs(_SegmentBroadPhase):
    """Every pair of every segment (broad_phase_nxn.py:217-391)."""

    _mode = 1


class BroadPhaseSAP(_SegmentBroadPhase):
    """Sweep and prune on the reference's fixed axis (broad_phase_sap.py:397-677)."""

    _mode = 0

    def __init__(self, shape_world, shape_flags=None, sweep_thread_count_multiplier: int = 5,
                 sort_type: str = "segmented", tile_block_dim: int | None = None, device=None):
        mode = str(sort_type).lower()
        if mode not in ("segmented", "tile"):
            raise ValueError(f"Unsupported SAP sort type: {sort_type!r}")
        self.sort_type = mode
        self.sweep_thread_count_multiplier = sweep_thread_count_multiplier
        self.tile_block_dim_override = tile_block_dim
        super().__init__(shape_world, shape_flags, device)


class BroadPhaseExplicit:
    """The AABB-overlapping subset of a given pair list (broad_phase_nxn.py:394-472)."""

    def __init__(self) -> None:
        self._native: _Native | None = None

    def launch(self, shape_lower, shape_upper, shape_gap, shape_pairs, shape_pair_count, candidate_pair,
               candidate_pair_count, device=None, skip_count_zero: bool = False) -> None:
        from ._native import check, ptr, stream_handle

        _check_tensor("shape_lower", shape_lower, torch.float32, (3,))
        dev = shape_lower.device if device is None else torch.device(device)
        if dev.type != "cuda":
            raise RuntimeError("the broad phase runs on the GPU only")
        n = shape_lower.shape[0]
        _check_tensor("shape_upper", shape_upper, torch.float32, (3,), dev)
        if shape_gap is not None and shape_gap.numel() == 0:
            shape_gap = None
        if shape_gap is not None:
            _check_tensor("shape_gap", shape_gap, torch.float32, (), dev)
            if shape_gap.shape[0] < n:
                raise ValueError("shape_gap is shorter than the AABB arrays")
        _check_tensor("shape_pairs", shape_pairs, torch.int32, (2,), dev)
        count = int(shape_pair_count)
        if not 0 <= count <= shape_pairs.shape[0]:
            raise ValueError("shape_pair_count exceeds shape_pairs")
        _check_tensor("candidate_pair", candidate_pair, torch.int32, (2,), dev)
        _check_tensor("candidate_pair_count", candidate_pair_count, torch.int32, None, dev)
        if candidate_pair_count.numel() < 1:
            raise ValueError("candidate_pair_count needs one element")
        if self._native is None or self._native.device != dev or self._native.pair_capacity < count:
            self._native = _Native(2, None, None, max(count, 1), dev)
        ctx = self._native
        check(ctx.L.nmc_broad_phase_explicit(ctx.h, int(n), ptr(shape_lower), ptr(shape_upper), ptr(shape_gap),
                                             ptr(shape_pairs), count, ptr(candidate_pair),
                                             int(candidate_pair.shape[0]), ptr(candidate_pair_count),
                                             int(bool(skip_count_zero)), stream_handle(dev)),
              "nmc_broad_phase_explicit")
