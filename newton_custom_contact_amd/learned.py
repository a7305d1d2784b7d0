"""Learned warm start of the XPBD contact solve.

The reference XPBD contact solve keeps no per-contact multiplier: every iteration recomputes the normal
multiplier from the current penetration (newton/_src/solvers/xpbd/kernels.py:1945-1980, 2192-2194).
The warm start adds one.  A per-contact network predicts the total normal multiplier that the plain
10-iteration solve would apply (``lambda_sum`` of a logged plain step), the solver applies it in one
Jacobi pass with the reference's 1/count weighting, then runs a few plain iterations.  A world whose
penetration residual (largest ``max(0, -d)`` over its contacts) is still above ``residual_tol``
restarts from the integrated state with the full plain iteration count.

Network: 16 -> 128 -> 128 -> 1 with SiLU, layers 1-2 in BF16 on the matrix cores with FP32
accumulation (``nmc_mlp_forward``), trained on the GPU by the library's own forward/backward kernels
(``nmc_mlp_train_grad``) with Adam.  Inputs (``FEATURES``) are computed by the solver kernel at the
integrated state; the "previous multiplier" input comes from the previous substep's contact with the
same sort key in the same world (frame-to-frame matching on the canonical contact order).
"""

from __future__ import annotations

import ctypes
import math
import os

import numpy as np
import torch

from ._native import (NMC_STEP_LOG, NMC_STEP_PREP, NMC_STEP_WARM, NmcLearnIO, NmcMlp, NmcMlpGrads,
                      NmcMlpTrainBuffers, check, load, ptr, stream_handle)
from .solvers import SolverXPBD

FEATURES = ("d", "w_n", "lambda_first", "lambda_prev", "v_n", "v_t", "inv_m_a", "inv_m_b", "count_a", "count_b",
            "normal_z", "mu", "static_a", "static_b", "lever_a", "lever_b")
N_FEATURES = 16
HIDDEN = 128
WEIGHTS_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "warmstart_mlp.npz")


def _vp(t) -> ctypes.c_void_p:
    return ctypes.c_void_p(ptr(t))


class WarmStartMLP:
    """FP32 master parameters (torch, on the device) + BF16 copies for the kernels."""

    def __init__(self, device="cuda:0", seed: int = 0):
        self.device = torch.device(device)
        g = torch.Generator().manual_seed(seed)

        def uni(shape, fan_in):
            b = 1.0 / math.sqrt(fan_in)
            return (torch.rand(shape, generator=g) * 2 - 1) * b

        f = dict(device=self.device, dtype=torch.float32)
        self.params = {
            "w1": uni((HIDDEN, N_FEATURES), N_FEATURES).to(**f),
            "b1": uni((HIDDEN,), N_FEATURES).to(**f),
            "w2": uni((HIDDEN, HIDDEN), HIDDEN).to(**f),
            "b2": uni((HIDDEN,), HIDDEN).to(**f),
            "w3": uni((HIDDEN,), HIDDEN).to(**f),
            "b3": torch.zeros(1, **f),
        }
        self.x_mean = torch.zeros(N_FEATURES, **f)
        self.x_inv_std = torch.ones(N_FEATURES, **f)
        self.out_scale = 1.0
        self.w1b = torch.empty((HIDDEN, N_FEATURES), device=self.device, dtype=torch.bfloat16)
        self.w2b = torch.empty((HIDDEN, HIDDEN), device=self.device, dtype=torch.bfloat16)
        self._train = None
        self._grads = None
        self.sync()

    # ---------------------------------------------------------------- parameters
    def sync(self):
        """Refresh the BF16 copies (round to nearest even) after the FP32 parameters changed."""
        self.w1b.copy_(self.params["w1"].to(torch.bfloat16))
        self.w2b.copy_(self.params["w2"].to(torch.bfloat16))

    def native(self) -> NmcMlp:
        p = self.params
        return NmcMlp(ptr(self.w1b), ptr(p["b1"]), ptr(self.w2b), ptr(p["b2"]), ptr(p["w3"]), ptr(p["b3"]),
                      ptr(self.x_mean), ptr(self.x_inv_std), float(self.out_scale))

    def set_normalization(self, feat: torch.Tensor, target: torch.Tensor):
        """Per-feature mean / inverse std and the output scale from a sample of logged data."""
        f = feat.to(torch.float64)
        mean = f.mean(0)
        std = f.std(0)
        self.x_mean.copy_(mean.to(torch.float32))
        self.x_inv_std.copy_(torch.where(std > 1e-8, 1.0 / std, torch.ones_like(std)).to(torch.float32))
        s = float(target.to(torch.float64).std())
        self.out_scale = s if s > 0 else 1.0

    def state_dict(self) -> dict:
        d = {k: v.detach().cpu().numpy() for k, v in self.params.items()}
        d["x_mean"] = self.x_mean.cpu().numpy()
        d["x_inv_std"] = self.x_inv_std.cpu().numpy()
        d["out_scale"] = np.float32(self.out_scale)
        return d

    def save(self, path: str):
        os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
        np.savez(path, **self.state_dict())

    @classmethod
    def load(cls, path: str = WEIGHTS_PATH, device="cuda:0") -> "WarmStartMLP":
        m = cls(device=device)
        with np.load(path, allow_pickle=False) as z:
            for k in m.params:
                m.params[k].copy_(torch.from_numpy(z[k]))
            m.x_mean.copy_(torch.from_numpy(z["x_mean"]))
            m.x_inv_std.copy_(torch.from_numpy(z["x_inv_std"]))
            m.out_scale = float(z["out_scale"])
        m.sync()
        return m

    # ---------------------------------------------------------------- kernels
    def predict(self, feat: torch.Tensor, count: torch.Tensor, capacity: int, out: torch.Tensor):
        """out[r] = network(feat[r]) * out_scale for r < count (count is a device int32 scalar)."""
        L = load()
        w = self.native()
        check(L.nmc_mlp_forward(ctypes.byref(w), _vp(feat), _vp(count), int(capacity), _vp(out),
                                ctypes.c_void_p(stream_handle(self.device))), "nmc_mlp_forward")

    def gradients(self, feat: torch.Tensor, target: torch.Tensor) -> dict:
        """Gradients of 0.5 * mean((y - target)^2) for a batch of raw features and normalized targets."""
        n = int(feat.shape[0])
        if self._train is None or self._train["n"] < n:
            f = dict(device=self.device, dtype=torch.float32)
            b = dict(device=self.device, dtype=torch.bfloat16)
            self._train = {"n": n, "y": torch.empty(n, **f), "z1": torch.empty((n, HIDDEN), **f),
                           "z2": torch.empty((n, HIDDEN), **f), "xb": torch.empty((n, N_FEATURES), **b),
                           "h1": torch.empty((n, HIDDEN), **b), "dz2b": torch.empty((n, HIDDEN), **b),
                           "dz1b": torch.empty((n, HIDDEN), **b), "dz2f": torch.empty((n, HIDDEN), **f),
                           "dz1f": torch.empty((n, HIDDEN), **f), "g3": torch.empty((n, HIDDEN), **f),
                           "dy": torch.empty(n, **f)}
            self._grads = {k: torch.empty_like(v) for k, v in self.params.items()}
        t = self._train
        tb = NmcMlpTrainBuffers(*[ptr(t[k]) for k in ("y", "z1", "z2", "xb", "h1", "dz2b", "dz1b", "dz2f", "dz1f",
                                                        "g3", "dy")])
        g = NmcMlpGrads(*[ptr(self._grads[k]) for k in ("w1", "b1", "w2", "b2", "w3", "b3")])
        w = self.native()
        feat = feat.contiguous()
        target = target.contiguous()
        check(load().nmc_mlp_train_grad(ctypes.byref(w), _vp(feat), _vp(target), n, ctypes.byref(tb),
                                        ctypes.byref(g), ctypes.c_void_p(stream_handle(self.device))),
              "nmc_mlp_train_grad")
        return self._grads

    def loss(self, feat: torch.Tensor, target: torch.Tensor) -> float:
        """0.5 * mean((y - target)^2) of the last gradients() batch's forward (normalized units)."""
        n = int(feat.shape[0])
        return float(0.5 * ((self._train["y"][:n] - target) ** 2).mean())

    def fit(self, feat: torch.Tensor, lam: torch.Tensor, epochs: int = 4, batch: int = 65536, lr: float = 2e-3,
            seed: int = 0, log=None) -> list[float]:
        """Adam on the logged (features, total normal multiplier) pairs; returns the per-epoch loss."""
        target = lam / self.out_scale
        opt = {k: (torch.zeros_like(v), torch.zeros_like(v)) for k, v in self.params.items()}
        b1, b2, eps = 0.9, 0.999, 1e-8
        step = 0
        g = torch.Generator(device="cpu").manual_seed(seed)
        n = int(feat.shape[0])
        hist = []
        for ep in range(epochs):
            perm = torch.randperm(n, generator=g).to(self.device)
            tot, cnt = 0.0, 0
            for s in range(0, n - batch + 1, batch):
                idx = perm[s:s + batch]
                x, t = feat[idx], target[idx]
                grads = self.gradients(x, t)
                step += 1
                lr_t = lr * 0.5 * (1 + math.cos(math.pi * min(1.0, (ep * n + s) / (epochs * n))))
                for k, p in self.params.items():
                    m, v = opt[k]
                    gk = grads[k]
                    m.mul_(b1).add_(gk, alpha=1 - b1)
                    v.mul_(b2).addcmul_(gk, gk, value=1 - b2)
                    mh = m / (1 - b1 ** step)
                    vh = v / (1 - b2 ** step)
                    p.sub_(lr_t * mh / (vh.sqrt() + eps))
                self.sync()
                tot += self.loss(x, t)
                cnt += 1
            hist.append(tot / max(cnt, 1))
            if log:
                log(f"epoch {ep}: loss {hist[-1]:.5f}")
        return hist


class SolverXPBDWarmStart(SolverXPBD):
    """SolverXPBD with the learned warm start: per step a prep launch (joint forces, integration, network
    inputs), the network, and the warm launch (warm pass, ``warm_iterations`` iterations, residual check,
    cold fallback with ``iterations``).  Requires a deterministic CollisionPipeline (sort keys)."""

    def __init__(self, model, mlp: WarmStartMLP, iterations: int = 10, warm_iterations: int = 3,
                 residual_tol: float = 5e-3, **kw):
        super().__init__(model, iterations=iterations, **kw)
        if not 0 <= warm_iterations <= iterations:
            raise ValueError("warm_iterations must be in [0, iterations]")
        self.mlp = mlp
        self.warm_iterations = int(warm_iterations)
        self.residual_tol = float(residual_tol)
        self._buf = None

    def _buffers(self, contacts):
        C = contacts.rigid_contact_max
        if self._buf is not None and self._buf["C"] == C:
            return self._buf
        m = self.model
        dev = m.device
        f = dict(device=dev, dtype=torch.float32)
        i = dict(device=dev, dtype=torch.int32)
        S = m.world_count + 1
        self._buf = {
            "C": C, "feat": torch.zeros((C, N_FEATURES), **f), "lambda0": torch.zeros(C, **f),
            "lam": [torch.zeros(C, **f), torch.zeros(C, **f)], "q_mid": torch.zeros((m.body_count, 7), **f),
            "qd_mid": torch.zeros((m.body_count, 6), **f), "residual": torch.zeros(S, **f),
            "fallback": torch.zeros(S, **i), "snap_key": None, "snap_start": torch.zeros(S, **i),
            "snap_count": torch.zeros(S, **i),
        }
        self._cur = 0
        self._prev = None
        self._snapshot = False
        self._last_remembered = True
        return self._buf

    # History for the "previous multiplier" input: the previous step's contacts (keys, per-world ranges)
    # and its multipliers.  Callers that alternate two Contacts buffers between substeps cost nothing; a
    # caller that reuses one buffer switches the solver to snapshot copies of the keys after each step
    # (the first reuse is detected too late and that step sees no history).
    def _history(self, contacts, b):
        lam_prev = b["lam"][1 - self._cur]
        if self._prev is None:
            return None, b["snap_start"], b["snap_count"], lam_prev, None
        if self._prev is contacts:
            if not self._snapshot:
                self._snapshot = True
                b["snap_count"].zero_()
                b["snap_key"] = torch.zeros_like(contacts.rigid_contact_sort_key)
                b["snap_order"] = torch.zeros_like(contacts.rigid_contact_order)
                return None, b["snap_start"], b["snap_count"], lam_prev, None
            return b["snap_key"], b["snap_start"], b["snap_count"], lam_prev, b["snap_order"]
        pc = self._prev
        return (pc.rigid_contact_sort_key, pc.world_contact_start, pc.world_contact_count, lam_prev,
                pc.rigid_contact_order)

    def _io(self, mode: int, b: dict, contacts, warm_iterations: int = 0,
            residual_tol: float | None = None) -> NmcLearnIO:
        tol = self.residual_tol if residual_tol is None else float(residual_tol)
        key, start, count, lam_prev, order = self._history(contacts, b)
        return NmcLearnIO(mode, warm_iterations, tol, ptr(b["feat"]), ptr(b["lam"][self._cur]),
                          ptr(b["lambda0"]), ptr(b["residual"]), ptr(b["fallback"]), ptr(b["q_mid"]),
                          ptr(b["qd_mid"]), ptr(key), ptr(start), ptr(count), ptr(lam_prev), ptr(order))

    def _launch(self, state_in, state_out, control, contacts, dt, io, iterations: int | None = None):
        model = self.model
        if control is None:
            if self._control is None:
                self._control = model.control(clone_variables=False)
            control = self._control
        if contacts.rigid_contact_max > self._contact_capacity:
            raise ValueError("contacts buffer larger than the solver's capacity")
        p = self.params(True)
        if iterations is not None:
            p.iterations = int(iterations)
        out_q = state_out.body_q if state_out is not None else None
        out_qd = state_out.body_qd if state_out is not None else None
        check(self._lib.nmc_xpbd_step_learn(
            self.native.handle, ctypes.byref(p), _vp(state_in.body_q), _vp(state_in.body_qd), _vp(state_in.body_f),
            _vp(control.joint_f), _vp(control.joint_target_pos), _vp(control.joint_target_vel),
            ctypes.byref(contacts.native_view()), ctypes.c_float(dt), _vp(out_q), _vp(out_qd), ctypes.byref(io),
            ctypes.c_void_p(stream_handle(model.device))), "nmc_xpbd_step_learn")

    def _remember(self, contacts, b):
        if self._snapshot:
            b["snap_key"].copy_(contacts.rigid_contact_sort_key)
            b["snap_order"].copy_(contacts.rigid_contact_order)
            b["snap_start"].copy_(contacts.world_contact_start)
            b["snap_count"].copy_(contacts.world_contact_count)
        self._prev = contacts
        self._cur = 1 - self._cur

    def _check_contacts(self, contacts):
        if contacts is None or contacts.rigid_contact_sort_key is None:
            raise ValueError("the warm-started solver needs contacts from CollisionPipeline(deterministic=True)")

    def step(self, state_in, state_out, control, contacts, dt: float, warm_iterations: int | None = None,
             residual_tol: float | None = None, remember: bool = True, multipliers: torch.Tensor | None = None) -> None:
        """Prep launch, network, warm launch.  ``multipliers`` (one per contact slot, in the buffer's order)
        replaces the network's prediction: with the total normal multipliers a plain step logged on the same
        state and contacts (``log_step``, ``buffers['lambda_sum']``) it measures what the warm start can reach
        with a perfect prediction."""
        if state_in.body_q is state_out.body_q:
            raise ValueError("state_in and state_out must be different State objects")
        self._check_contacts(contacts)
        b = self._buffers(contacts)
        k = self.warm_iterations if warm_iterations is None else int(warm_iterations)
        if not 0 <= k <= self.iterations:
            raise ValueError("warm_iterations must be in [0, iterations]")
        self._launch(state_in, None, control, contacts, dt, self._io(NMC_STEP_PREP, b, contacts))
        if multipliers is None:
            self.mlp.predict(b["feat"], contacts.rigid_contact_count, contacts.rigid_contact_max, b["lambda0"])
        else:
            b["lambda0"].copy_(multipliers[: b["lambda0"].numel()])
        self._launch(state_in, state_out, control, contacts, dt,
                     self._io(NMC_STEP_WARM, b, contacts, k, residual_tol))
        self._last_remembered = remember
        if remember:
            self._remember(contacts, b)

    def log_step(self, state_in, state_out, control, contacts, dt: float, iterations: int | None = None,
                 remember: bool = True) -> None:
        """The plain step that also writes the network inputs (``buffers['feat']``), the total normal
        multiplier of every contact (``buffers['lambda_sum']``) and the per-world penetration residual
        (``buffers['residual']``): training data and the plain reference of the comparison."""
        self._check_contacts(contacts)
        b = self._buffers(contacts)
        self._launch(state_in, state_out, control, contacts, dt, self._io(NMC_STEP_LOG, b, contacts), iterations)
        self._last_remembered = remember
        if remember:
            self._remember(contacts, b)

    @property
    def buffers(self) -> dict:
        """feat, lambda0, residual, fallback, q_mid, qd_mid, and lambda_sum: the total normal multipliers of
        the last step."""
        b = dict(self._buf)
        b["lambda_sum"] = self._buf["lam"][1 - self._cur] if self._last_remembered else self._buf["lam"][self._cur]
        return b

    def reset_history(self):
        """Forget the previous substep's contacts (no multiplier is matched on the next step)."""
        self._prev = None


__all__ = ["FEATURES", "SolverXPBDWarmStart", "WarmStartMLP", "WEIGHTS_PATH"]
