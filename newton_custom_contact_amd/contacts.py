"""``Contacts``: rigid contact buffers (newton/_src/sim/contacts.py:43-358, rigid part).

Field names, dtypes and shapes follow the reference.  With sort keys (a deterministic pipeline) the buffer is
in the reference's deterministic layout: ascending by the 64-bit sort key ``shape0 << 43 | shape1 << 23 |
sub_key`` over the whole buffer (collide.py:1105-1124).  Without sort keys it is in the canonical order:
segment-major (world 0, world 1, ..., then the global-only segment), key order inside each segment.  Two extra
per-segment arrays, ``world_contact_start`` / ``world_contact_count``, index the canonical order for the
per-world solver, and ``rigid_contact_order`` maps canonical positions to buffer indices (with sort keys).
``rigid_contact_tids`` is zero, as the reference's writer always stores (collide.py:134).
"""

from __future__ import annotations

import torch

from ._native import NmcContacts, ptr


MATCH_NOT_FOUND = -1  # contact_match.py:116-120
MATCH_BROKEN = -2


class Contacts:
    EXTENDED_ATTRIBUTES: frozenset[str] = frozenset(("force",))

    def __init__(self, rigid_contact_max: int, soft_contact_max: int = 0, requires_grad: bool = False,
                 device=None, segment_count: int = 1, deterministic: bool = False,
                 requested_attributes: set[str] | None = None, contact_matching: bool = False,
                 contact_report: bool = False, per_contact_shape_properties: bool = False):
        dev = torch.device(device) if device is not None else torch.device("cuda:0")
        C = int(rigid_contact_max)
        self.device = dev
        self.rigid_contact_max = C
        self.soft_contact_max = int(soft_contact_max)
        self.requires_grad = requires_grad
        self.clear_buffers = False
        f = dict(device=dev, dtype=torch.float32)
        i = dict(device=dev, dtype=torch.int32)
        self.contact_counters = torch.zeros(2, **i)
        self.rigid_contact_count = self.contact_counters[0:1]
        self.soft_contact_count = self.contact_counters[1:2]
        self.contact_generation = torch.zeros(1, **i)
        self.rigid_contact_shape0 = torch.full((C,), -1, **i)
        self.rigid_contact_shape1 = torch.full((C,), -1, **i)
        self.rigid_contact_point0 = torch.zeros((C, 3), **f)
        self.rigid_contact_point1 = torch.zeros((C, 3), **f)
        self.rigid_contact_offset0 = torch.zeros((C, 3), **f)
        self.rigid_contact_offset1 = torch.zeros((C, 3), **f)
        self.rigid_contact_normal = torch.zeros((C, 3), **f)
        self.rigid_contact_margin0 = torch.zeros((C,), **f)
        self.rigid_contact_margin1 = torch.zeros((C,), **f)
        self.rigid_contact_tids = torch.zeros((C,), **i)
        # per-contact stiffness / damping / friction (contacts.py:210-224): only the hydroelastic narrow phase
        # writes non-zero values; every collider of this package leaves them at 0, as the reference's do
        self.per_contact_shape_properties = bool(per_contact_shape_properties)
        if self.per_contact_shape_properties:
            self.rigid_contact_stiffness = torch.zeros((C,), **f)
            self.rigid_contact_damping = torch.zeros((C,), **f)
            self.rigid_contact_friction = torch.zeros((C,), **f)
        else:
            self.rigid_contact_stiffness = self.rigid_contact_damping = self.rigid_contact_friction = None
        self.rigid_contact_sort_key = torch.zeros((C,), device=dev, dtype=torch.int64) if deterministic else None
        # contact matching (contacts.py:226-270): match index per contact (>= 0 matched previous-frame
        # index, -1 MATCH_NOT_FOUND, -2 MATCH_BROKEN) and the optional new / broken report
        self.contact_matching = bool(contact_matching)
        self.contact_report = bool(contact_report)
        self.rigid_contact_match_index = torch.full((C,), -1, **i) if contact_matching else None
        if contact_report:
            self.rigid_contact_new_indices = torch.zeros((C,), **i)
            self.rigid_contact_new_count = torch.zeros(1, **i)
            self.rigid_contact_broken_indices = torch.zeros((C,), **i)
            self.rigid_contact_broken_count = torch.zeros(1, **i)
        else:
            self.rigid_contact_new_indices = None
            self.rigid_contact_new_count = None
            self.rigid_contact_broken_indices = None
            self.rigid_contact_broken_count = None
        self.world_contact_start = torch.zeros((segment_count,), **i)
        self.world_contact_count = torch.zeros((segment_count,), **i)
        # with sort keys the buffer is in the reference's deterministic layout, ascending by sort key over the
        # whole buffer (collide.py:1105-1124); rigid_contact_order[j] is the buffer index of the j-th contact in
        # the canonical order (world by world, key order inside a world; -1 when past the capacity), by which
        # the solver walks each world's contacts
        self.rigid_contact_order = torch.full((C,), -1, **i) if deterministic else None
        attrs = requested_attributes or set()
        self.force = torch.zeros((C, 6), **f) if "force" in attrs else None
        self._view = None

    def native_view(self) -> NmcContacts:
        if self._view is None:
            v = NmcContacts()
            v.capacity = self.rigid_contact_max
            v.count = ptr(self.rigid_contact_count)
            v.shape0, v.shape1 = ptr(self.rigid_contact_shape0), ptr(self.rigid_contact_shape1)
            v.point0, v.point1 = ptr(self.rigid_contact_point0), ptr(self.rigid_contact_point1)
            v.offset0, v.offset1 = ptr(self.rigid_contact_offset0), ptr(self.rigid_contact_offset1)
            v.normal = ptr(self.rigid_contact_normal)
            v.margin0, v.margin1 = ptr(self.rigid_contact_margin0), ptr(self.rigid_contact_margin1)
            v.sort_key = ptr(self.rigid_contact_sort_key)
            v.world_start, v.world_count = ptr(self.world_contact_start), ptr(self.world_contact_count)
            v.order = ptr(self.rigid_contact_order)
            v.generation = ptr(self.contact_generation)
            self._view = v
        return self._view

    def clear(self, bump_generation: bool = True):
        """contacts.py: reset the counters (buffers are overwritten by the next collide)."""
        self.contact_counters.zero_()
        self.world_contact_count.zero_()
        if self.per_contact_shape_properties:
            self.rigid_contact_stiffness.zero_()
            self.rigid_contact_damping.zero_()
            self.rigid_contact_friction.zero_()
        if bump_generation:
            self.contact_generation += 1

    def numpy(self) -> dict:
        """Host copy of the stored rigid contacts (``min(count, capacity)`` rows)."""
        n = int(self.rigid_contact_count.item())
        n = min(n, self.rigid_contact_max)
        out = {
            "count": int(self.rigid_contact_count.item()),
            "shape0": self.rigid_contact_shape0[:n].cpu().numpy(),
            "shape1": self.rigid_contact_shape1[:n].cpu().numpy(),
            "point0": self.rigid_contact_point0[:n].cpu().numpy(),
            "point1": self.rigid_contact_point1[:n].cpu().numpy(),
            "offset0": self.rigid_contact_offset0[:n].cpu().numpy(),
            "offset1": self.rigid_contact_offset1[:n].cpu().numpy(),
            "normal": self.rigid_contact_normal[:n].cpu().numpy(),
            "margin0": self.rigid_contact_margin0[:n].cpu().numpy(),
            "margin1": self.rigid_contact_margin1[:n].cpu().numpy(),
        }
        if self.rigid_contact_sort_key is not None:
            out["sort_key"] = self.rigid_contact_sort_key[:n].cpu().numpy()
            out["order"] = self.rigid_contact_order[:n].cpu().numpy()
        if self.rigid_contact_match_index is not None:
            out["match_index"] = self.rigid_contact_match_index[:n].cpu().numpy()
        if self.rigid_contact_new_indices is not None:
            nn = int(self.rigid_contact_new_count.item())
            nbk = int(self.rigid_contact_broken_count.item())
            out["new_indices"] = self.rigid_contact_new_indices[:nn].cpu().numpy()
            out["broken_indices"] = self.rigid_contact_broken_indices[:nbk].cpu().numpy()
        return out


__all__ = ["Contacts", "MATCH_NOT_FOUND", "MATCH_BROKEN"]
