"""GPU (libnmc.so through the C-ABI) vs the CPU oracle on seeded scenes.

Contact sets must match bit-exactly after the canonical key sort (count, shape pairs, sort keys); the
float fields must agree within 1e-5 relative (they are expected to be bit-identical, and the report
says whether they were).  One XPBD step from identical inputs must agree within 1e-5 relative.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

from newton_custom_contact_amd import CollisionPipeline, scenes
from newton_custom_contact_amd.solvers import SolverXPBD
from oracle import Oracle
from parity_helpers import compare_contacts, oracle_states

pytestmark = pytest.mark.gpu

DT = 1.0 / 240.0


def _gpu_collide(pipe, model, q):
    state = model.state()
    state.body_q.copy_(torch.from_numpy(np.ascontiguousarray(q, np.float32)))
    c = pipe.contacts()
    pipe.collide(state, c)
    torch.cuda.synchronize()
    return c, c.numpy(), state


SCENES = [
    ("box_pile_64", lambda dev: scenes.box_pile(64, seed=0, device=dev)),
    ("box_pile_64_s1", lambda dev: scenes.box_pile(64, seed=1, device=dev)),
    ("mixed_8w", lambda dev: scenes.mixed_worlds(8, seed=3, device=dev)),
    # uneven worlds (8-256 bodies, log-uniform) with a 6-link revolute capsule chain each: 256-lane solver
    ("uneven_6w_revolute", lambda dev: scenes.uneven_worlds(6, seed=1, device=dev)),
    # worlds above 256 bodies: two bodies per solver lane
    ("uneven_2w_large", lambda dev: scenes.uneven_worlds(2, seed=4, lo=280, hi=400, device=dev)),
]


@pytest.mark.parametrize("name,make", SCENES, ids=[s[0] for s in SCENES])
def test_collide_parity(cuda_device, name, make):
    print(f"[collide parity] {name}", flush=True)
    model = make(cuda_device)
    oracle = Oracle(model)
    pipe = CollisionPipeline(model, broad_phase="sap", deterministic=True)
    reports = []
    for q, _qd in oracle_states(oracle, model, n_substeps=240, dt=DT, every=40):
        ref = oracle.collide(q)
        _c, got, _s = _gpu_collide(pipe, model, q)
        reports.append(compare_contacts(got, ref))
    print(name, reports, flush=True)
    assert sum(r["count"] for r in reports) > 0
    oc = pipe.overflow_counters()
    assert oc["pairs_dropped"] == 0 and oc["contacts_over_capacity"] == 0


@pytest.mark.parametrize("name,make", SCENES, ids=[s[0] for s in SCENES])
def test_step_parity(cuda_device, name, make):
    print(f"[step parity] {name}", flush=True)
    model = make(cuda_device)
    oracle = Oracle(model)
    pipe = CollisionPipeline(model, broad_phase="sap", deterministic=True)
    solver = SolverXPBD(model, iterations=10)
    worst = 0.0
    exact = True
    for q, qd in oracle_states(oracle, model, n_substeps=240, dt=DT, every=40):
        ref_c = oracle.collide(q)
        c, _got, state_in = _gpu_collide(pipe, model, q)
        state_in.body_qd.copy_(torch.from_numpy(qd))
        state_out = model.state()
        solver.step(state_in, state_out, None, c, DT)
        torch.cuda.synchronize()
        q_ref, qd_ref = oracle.step(q, qd, np.zeros_like(qd), ref_c, DT, iterations=10)
        q_gpu = state_out.body_q.cpu().numpy()
        qd_gpu = state_out.body_qd.cpu().numpy()
        exact = exact and np.array_equal(q_gpu, q_ref) and np.array_equal(qd_gpu, qd_ref)
        worst = max(worst, float(np.max(np.abs(q_gpu - q_ref))), float(np.max(np.abs(qd_gpu - qd_ref))))
        np.testing.assert_allclose(q_gpu, q_ref, rtol=1e-5, atol=1e-6)
        np.testing.assert_allclose(qd_gpu, qd_ref, rtol=1e-5, atol=1e-5)
    print(name, "max abs diff", worst, "bit-exact", exact, flush=True)


def test_trajectory_drift_60_steps(cuda_device):
    """60 frames x 4 substeps of the 64-box pile, GPU and oracle run independently: stacks within 1e-3 m."""
    model = scenes.box_pile(64, seed=0, device=cuda_device)
    oracle = Oracle(model)
    pipe = CollisionPipeline(model, broad_phase="sap")
    solver = SolverXPBD(model, iterations=10)
    s0, s1 = model.state(), model.state()
    c = pipe.contacts()
    h = model.host
    q, qd = h["body_q"].copy(), h["body_qd"].copy()
    for _ in range(60 * 4):
        pipe.collide(s0, c)
        solver.step(s0, s1, None, c, DT)
        s0, s1 = s1, s0
        rc = oracle.collide(q)
        q, qd = oracle.step(q, qd, np.zeros_like(qd), rc, DT, iterations=10)
    torch.cuda.synchronize()
    dq = np.abs(s0.body_q.cpu().numpy()[:, :3] - q[:, :3]).max()
    print("max position drift after 60 steps", dq, flush=True)
    assert dq <= 1e-3


def test_step_parity_joint_forces_and_targets(cuda_device):
    """Revolute chains driven by random joint_f and joint targets: control arrays reach the kernels."""
    model = scenes.uneven_worlds(4, seed=9, device=cuda_device, chain_links=6)
    oracle = Oracle(model)
    pipe = CollisionPipeline(model, broad_phase="sap", deterministic=True)
    solver = SolverXPBD(model, iterations=10)
    rng = np.random.default_rng(0)
    control = model.control()
    jf = rng.normal(0.0, 2.0, control.joint_f.shape).astype(np.float32)
    tp = rng.uniform(-0.5, 0.5, control.joint_target_pos.shape).astype(np.float32)
    tv = rng.uniform(-1.0, 1.0, control.joint_target_vel.shape).astype(np.float32)
    control.joint_f.copy_(torch.from_numpy(jf))
    control.joint_target_pos.copy_(torch.from_numpy(tp))
    control.joint_target_vel.copy_(torch.from_numpy(tv))
    exact = True
    for q, qd in oracle_states(oracle, model, n_substeps=120, dt=DT, every=40):
        ref_c = oracle.collide(q)
        c, _got, state_in = _gpu_collide(pipe, model, q)
        state_in.body_qd.copy_(torch.from_numpy(qd))
        state_out = model.state()
        solver.step(state_in, state_out, control, c, DT)
        torch.cuda.synchronize()
        q_ref, qd_ref = oracle.step(q, qd, np.zeros_like(qd), ref_c, DT, iterations=10, joint_f=jf,
                                    joint_target_pos=tp, joint_target_vel=tv)
        q_gpu = state_out.body_q.cpu().numpy()
        qd_gpu = state_out.body_qd.cpu().numpy()
        exact = exact and np.array_equal(q_gpu, q_ref) and np.array_equal(qd_gpu, qd_ref)
        np.testing.assert_allclose(q_gpu, q_ref, rtol=1e-5, atol=1e-6)
        np.testing.assert_allclose(qd_gpu, qd_ref, rtol=1e-5, atol=1e-5)
    print("joint forces/targets bit-exact", exact, flush=True)
