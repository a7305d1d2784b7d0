"""``solvers.SolverXPBD``: the XPBD rigid-body step on the GPU.

Mirrors newton/_src/solvers/xpbd/solver_xpbd.py (``SolverXPBD``, line 33): constructor defaults
(``:95-109``), ``step(state_in, state_out, control, contacts, dt)`` (``:259-764``) and
``notify_model_changed`` (``:141``).  One ``step`` is one launch of the per-world kernel in libnmc.so:
joint forces, integration, ``iterations`` x (contact solve + apply, joint solve + apply) and the store.

Differences in observable behaviour, both deliberate:
* ``state_in`` is read-only here; the reference uses ``state_in.body_q/qd`` as ping-pong scratch
  (solver_xpbd.py:218-230) so its contents after ``step`` are unspecified there.
* per-body sums over contacts and joints run in canonical (contact / joint index) order instead of
  float atomics, so results do not depend on thread scheduling.

``enable_restitution=True`` adds the rigid restitution pass after the iterations (solver_xpbd.py:719-764,
apply_rigid_restitution kernels.py:2472-2615): a second launch that changes only ``state_out.body_qd``.
"""

from __future__ import annotations

import ctypes

import torch

from ._native import NativeModel, NmcXpbdParams, NmcXpbdReport, check, load, ptr, stream_handle


class SolverNotifyFlags:
    """newton/_src/solvers/flags.py:10-56."""

    JOINT_PROPERTIES = 1 << 0
    JOINT_DOF_PROPERTIES = 1 << 1
    BODY_PROPERTIES = 1 << 2
    BODY_INERTIAL_PROPERTIES = 1 << 3
    SHAPE_PROPERTIES = 1 << 4
    MODEL_PROPERTIES = 1 << 5


class SolverBase:
    def __init__(self, model):
        self.model = model

    @property
    def device(self):
        return self.model.device


class SolverXPBD(SolverBase):
    def __init__(self, model, iterations: int = 2, soft_body_relaxation: float = 0.9,
                 soft_contact_relaxation: float = 0.9, joint_linear_relaxation: float = 0.7,
                 joint_angular_relaxation: float = 0.4, joint_linear_compliance: float = 0.0,
                 joint_angular_compliance: float = 0.0, rigid_contact_relaxation: float = 0.8,
                 rigid_contact_con_weighting: bool = True, angular_damping: float = 0.0,
                 enable_restitution: bool = False, block_lanes: int = 0, reference_arithmetic: bool = False):
        super().__init__(model)
        if block_lanes not in (0, 64, 256, 512):
            raise ValueError("block_lanes must be 0 (automatic), 64, 256 or 512 (CSR bins, worlds above 64 bodies)")
        self.block_lanes = int(block_lanes)  # lanes per world of the step kernel (see autotune.tune_block_lanes)
        # False: world-frame contact arithmetic (within rounding of the reference, the fast path); True: the
        # reference's operation order, bit-identical to the CPU restatement
        self.reference_arithmetic = bool(reference_arithmetic)
        self.iterations = int(iterations)
        self.soft_body_relaxation = soft_body_relaxation
        self.soft_contact_relaxation = soft_contact_relaxation
        self.joint_linear_relaxation = joint_linear_relaxation
        self.joint_angular_relaxation = joint_angular_relaxation
        self.joint_linear_compliance = joint_linear_compliance
        self.joint_angular_compliance = joint_angular_compliance
        self.rigid_contact_relaxation = rigid_contact_relaxation
        self.rigid_contact_con_weighting = rigid_contact_con_weighting
        self.angular_damping = angular_damping
        self.enable_restitution = enable_restitution
        # the solver orders each dense world's (contact, body) entries in a buffer of 2 per contact slot,
        # sized for the largest Contacts buffer seen (the pipeline's default capacity to begin with)
        from .collide import estimate_rigid_contact_max  # noqa: PLC0415

        self._contact_capacity = max(1, estimate_rigid_contact_max(model))
        self.native = NativeModel(model, rigid_contact_max=self._contact_capacity, shape_pairs_max=0)
        self._lib = load()
        self._control = None
        self._impulse = None
        self._impulse_scratch = None
        self._joint_impulse = None
        self._impulse_valid = False
        self._impulse_capacity = 0
        self._last_dt = 0.0

    def notify_model_changed(self, flags: int) -> None:
        """Kinematic flags and inverse masses are read on every step; nothing is cached."""

    def params(self, has_contacts: bool) -> NmcXpbdParams:
        return NmcXpbdParams(self.iterations, self.rigid_contact_relaxation, int(self.rigid_contact_con_weighting),
                             self.angular_damping, self.joint_linear_relaxation, self.joint_angular_relaxation,
                             self.joint_linear_compliance, self.joint_angular_compliance, int(has_contacts),
                             int(getattr(self, "block_lanes", 0)), int(getattr(self, "reference_arithmetic", False)),
                             int(bool(getattr(self, "enable_restitution", False))))

    def step(self, state_in, state_out, control, contacts, dt: float) -> None:
        """solver_xpbd.py:259-764."""
        model = self.model
        if control is None:
            if self._control is None:
                self._control = model.control(clone_variables=False)
            control = self._control
        if state_in.body_q is state_out.body_q:
            raise ValueError("state_in and state_out must be different State objects")
        p = self.params(contacts is not None)
        if contacts is not None and contacts.rigid_contact_max > self._contact_capacity:
            # a larger Contacts buffer than any before: regrow the ordering buffer (allocates; do it
            # before capturing the step in a graph)
            self._contact_capacity = contacts.rigid_contact_max
            self.native = NativeModel(model, rigid_contact_max=self._contact_capacity, shape_pairs_max=0)
        cv = ctypes.byref(contacts.native_view()) if contacts is not None else None
        for t in (state_in.body_q, state_in.body_qd, state_in.body_f, state_out.body_q, state_out.body_qd):
            if t.device != model.device or not t.is_contiguous():
                raise ValueError("state arrays must be contiguous tensors on the model device")
        args = (self.native.handle, ctypes.byref(p), ctypes.c_void_p(ptr(state_in.body_q)),
                ctypes.c_void_p(ptr(state_in.body_qd)), ctypes.c_void_p(ptr(state_in.body_f)),
                ctypes.c_void_p(ptr(control.joint_f)), ctypes.c_void_p(ptr(control.joint_target_pos)),
                ctypes.c_void_p(ptr(control.joint_target_vel)), cv, ctypes.c_float(dt),
                ctypes.c_void_p(ptr(state_out.body_q)), ctypes.c_void_p(ptr(state_out.body_qd)))
        want_force = contacts is not None and getattr(contacts, "force", None) is not None
        want_parent = getattr(state_out, "body_parent_f", None) is not None
        if want_force or want_parent:
            # reported forces (solver_xpbd.py:275-296, 632-655): the step also accumulates the per-contact
            # and per-joint impulses
            rep = NmcXpbdReport()
            if want_force:
                C = contacts.rigid_contact_max
                if self._impulse is None or self._impulse.shape[0] != C:
                    self._impulse = torch.zeros((C, 6), device=model.device, dtype=torch.float32)
                    self._impulse_scratch = torch.zeros((C, 8), device=model.device, dtype=torch.float32)
                rep.contact_impulse, rep.contact_scratch = ptr(self._impulse), ptr(self._impulse_scratch)
            if want_parent:
                if self._joint_impulse is None:
                    self._joint_impulse = torch.zeros((max(model.joint_count, 1), 6), device=model.device,
                                                      dtype=torch.float32)
                rep.joint_impulse, rep.body_parent_f = ptr(self._joint_impulse), ptr(state_out.body_parent_f)
            check(self._lib.nmc_xpbd_step_report(*args, ctypes.byref(rep), ctypes.c_void_p(stream_handle(model.device))),
                  "nmc_xpbd_step_report")
            self._impulse_capacity = contacts.rigid_contact_max if want_force else 0
            self._last_dt = float(dt)
            self._impulse_valid = want_force
        else:
            check(self._lib.nmc_xpbd_step(*args, ctypes.c_void_p(stream_handle(model.device))), "nmc_xpbd_step")
            self._impulse_valid = False

    def update_contacts(self, contacts, state=None) -> None:
        """solver_xpbd.py:767-821: ``contacts.force`` [N, N·m] from the contact impulses of the last step
        (linear: normal + friction; angular: torsional + rolling), scaled for the 1/N contact weighting."""
        if contacts.force is None:
            raise ValueError("contacts.force is not allocated. Call model.request_contact_attributes('force') "
                             "before creating the Contacts object.")
        if not self._impulse_valid:
            raise ValueError("No contact impulse data available. Call step() before update_contacts().")
        if contacts.rigid_contact_max != self._impulse_capacity:
            raise ValueError(f"Contacts capacity mismatch: update_contacts() received rigid_contact_max="
                             f"{contacts.rigid_contact_max}, but step() used {self._impulse_capacity}. "
                             "Pass the same Contacts instance to both step() and update_contacts().")
        check(self._lib.nmc_contact_force(ctypes.c_void_p(ptr(self._impulse)),
                                          ctypes.c_void_p(ptr(contacts.rigid_contact_count)),
                                          int(contacts.rigid_contact_max), ctypes.c_float(self._last_dt),
                                          ctypes.c_void_p(ptr(contacts.force)),
                                          ctypes.c_void_p(stream_handle(self.model.device))), "nmc_contact_force")


__all__ = ["SolverBase", "SolverNotifyFlags", "SolverXPBD"]
_ = torch
