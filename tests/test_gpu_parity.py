"""GPU (libnmc.so through the C-ABI) vs the CPU oracle on seeded scenes.

Contact sets must match bit-exactly after the canonical key sort (count, shape pairs, sort keys); the
float fields must agree within 1e-5 relative (they are expected to be bit-identical, and the report
says whether they were).  One XPBD step from identical inputs: the reference-order kernel
(reference_arithmetic=True) equals the oracle bit for bit; the world-frame kernel (the default) agrees within
1e-5 relative (floors: 1e-5 m for poses, 1e-5 m per substep for velocities), and against an FP64 rerun of the
oracle from the same state and contacts it is at most 4x as far off as the FP32 oracle (both stated).  Over 60 frames of resting box stacks the default solver stays within
1e-3 m of the oracle and of an FP64 oracle run.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

from newton_custom_contact_amd import CollisionPipeline, scenes
from newton_custom_contact_amd.solvers import SolverXPBD
from oracle import Oracle
from parity_helpers import assert_step_close, bit_equal, compare_contacts, max_abs_diff, oracle_states

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


def _both_steps(model, state_in, c, control=None):
    """One step by each solver kernel: {"reference": (q, qd), "fast": (q, qd)}."""
    out = {}
    for tag, ref_arith in (("reference", True), ("fast", False)):
        solver = SolverXPBD(model, iterations=10, reference_arithmetic=ref_arith)
        s = model.state()
        solver.step(state_in, s, control, c, DT)
        torch.cuda.synchronize()
        out[tag] = (s.body_q.cpu().numpy(), s.body_qd.cpu().numpy())
    return out


@pytest.mark.parametrize("name,make", SCENES, ids=[s[0] for s in SCENES])
def test_step_parity(cuda_device, name, make):
    print(f"[step parity] {name}", flush=True)
    model = make(cuda_device)
    oracle = Oracle(model)
    pipe = CollisionPipeline(model, broad_phase="sap", deterministic=True)
    err = {"fast_vs_oracle": 0.0, "fast_vs_f64": 0.0, "oracle_vs_f64": 0.0}
    for q, qd in oracle_states(oracle, model, n_substeps=240, dt=DT, every=40):
        ref_c = oracle.collide(q)
        c, _got, state_in = _gpu_collide(pipe, model, q)
        state_in.body_qd.copy_(torch.from_numpy(qd))
        steps = _both_steps(model, state_in, c)
        q_ref, qd_ref = oracle.step(q, qd, np.zeros_like(qd), ref_c, DT, iterations=10)
        q64, qd64 = oracle.step(q, qd, np.zeros_like(qd), ref_c, DT, iterations=10, precision="f64")
        qr, qdr = steps["reference"]
        assert bit_equal(qr, q_ref) and bit_equal(qdr, qd_ref), "reference-order kernel differs from the oracle"
        qf, qdf = steps["fast"]
        assert_step_close(qf, qdf, q_ref, qd_ref, DT, name)
        err["fast_vs_oracle"] = max(err["fast_vs_oracle"], max_abs_diff(qf, q_ref), max_abs_diff(qdf, qd_ref))
        err["fast_vs_f64"] = max(err["fast_vs_f64"], max_abs_diff(qf, q64), max_abs_diff(qdf, qd64))
        err["oracle_vs_f64"] = max(err["oracle_vs_f64"], max_abs_diff(q_ref, q64), max_abs_diff(qd_ref, qd64))
    print(name, "reference-order kernel bit-exact; largest |difference| (m, m/s, rad/s, quaternion units):", err,
          flush=True)
    assert err["fast_vs_f64"] <= 4.0 * err["oracle_vs_f64"] + 1e-6


def _trajectories(model, iterations=10, frames=60):
    """frames x 4 substeps from the model's initial state: GPU default solver, GPU reference-order solver,
    FP32 oracle and FP64 oracle (collide and step in FP64), each run on its own.  Returns final positions."""
    oracle = Oracle(model)
    pipe = CollisionPipeline(model, broad_phase="sap")
    out = {}
    for tag, ref_arith in (("fast", False), ("reference", True)):
        solver = SolverXPBD(model, iterations=iterations, reference_arithmetic=ref_arith)
        s0, s1 = model.state(), model.state()
        c = pipe.contacts()
        for _ in range(frames * 4):
            pipe.collide(s0, c)
            solver.step(s0, s1, None, c, DT)
            s0, s1 = s1, s0
        torch.cuda.synchronize()
        out[tag] = s0.body_q.cpu().numpy()[:, :3]
    h = model.host
    for tag, prec in (("oracle", "f32"), ("f64", "f64")):
        dtp = np.float32 if prec == "f32" else np.float64
        q, qd = h["body_q"].astype(dtp), h["body_qd"].astype(dtp)
        for _ in range(frames * 4):
            rc = oracle.collide(q, precision=prec)
            q, qd = oracle.step(q, qd, np.zeros_like(qd), rc, DT, iterations=iterations, precision=prec)
        out[tag] = q[:, :3]
    return out


def _drift(a, b) -> float:
    return float(np.abs(np.asarray(a, np.float64) - np.asarray(b, np.float64)).max())


def test_stack_drift_60_steps(cuda_device):
    """Four towers of four boxes resting on the ground, 60 frames x 4 substeps: the default solver's positions
    stay within 1e-3 m of the FP32 oracle's and of an FP64 oracle run (collide and step in FP64); the
    reference-order kernel's trajectory equals the oracle's bit for bit.  (The towers settle by a few mm; a
    sliding top box would make the comparison chaotic, as in the pile below.)"""
    model = scenes.box_stacks(4, 4, seed=3, device=cuda_device)
    t = _trajectories(model)
    drift = {"fast_vs_oracle": _drift(t["fast"], t["oracle"]), "fast_vs_f64": _drift(t["fast"], t["f64"]),
             "oracle_vs_f64": _drift(t["oracle"], t["f64"])}
    print("stacks, largest position drift after 60 frames (m):", drift, flush=True)
    assert np.array_equal(t["reference"], t["oracle"])
    assert drift["fast_vs_oracle"] <= 1e-3 and drift["fast_vs_f64"] <= 1e-3


def test_pile_drift_60_steps(cuda_device):
    """The 64-box pile dropped from 0.5-2.5 m, 60 frames: a chaotic trajectory (the FP32 oracle ends about 0.1 m
    from its FP64 rerun), so it is held to bit-exactness with the reference-order kernel; the default solver's
    drift against FP64 is stated beside the FP32 oracle's."""
    model = scenes.box_pile(64, seed=0, device=cuda_device)
    t = _trajectories(model)
    drift = {"fast_vs_oracle": _drift(t["fast"], t["oracle"]), "fast_vs_f64": _drift(t["fast"], t["f64"]),
             "oracle_vs_f64": _drift(t["oracle"], t["f64"])}
    print("pile, largest position drift after 60 frames (m):", drift, flush=True)
    assert np.array_equal(t["reference"], t["oracle"])
    assert np.isfinite(t["fast"]).all()


def test_step_parity_joint_forces_and_targets(cuda_device):
    """Revolute chains driven by random joint_f and joint targets: control arrays reach the kernels."""
    model = scenes.uneven_worlds(4, seed=9, device=cuda_device, chain_links=6)
    oracle = Oracle(model)
    pipe = CollisionPipeline(model, broad_phase="sap", deterministic=True)
    rng = np.random.default_rng(0)
    control = model.control()
    jf = rng.normal(0.0, 2.0, control.joint_f.shape).astype(np.float32)
    tp = rng.uniform(-0.5, 0.5, control.joint_target_pos.shape).astype(np.float32)
    tv = rng.uniform(-1.0, 1.0, control.joint_target_vel.shape).astype(np.float32)
    control.joint_f.copy_(torch.from_numpy(jf))
    control.joint_target_pos.copy_(torch.from_numpy(tp))
    control.joint_target_vel.copy_(torch.from_numpy(tv))
    for q, qd in oracle_states(oracle, model, n_substeps=120, dt=DT, every=40):
        ref_c = oracle.collide(q)
        c, _got, state_in = _gpu_collide(pipe, model, q)
        state_in.body_qd.copy_(torch.from_numpy(qd))
        steps = _both_steps(model, state_in, c, control)
        q_ref, qd_ref = oracle.step(q, qd, np.zeros_like(qd), ref_c, DT, iterations=10, joint_f=jf,
                                    joint_target_pos=tp, joint_target_vel=tv)
        assert bit_equal(steps["reference"][0], q_ref) and bit_equal(steps["reference"][1], qd_ref)
        assert_step_close(*steps["fast"], q_ref, qd_ref, DT, "joint forces/targets")
    print("joint forces/targets: reference-order kernel bit-exact, world-frame kernel within tolerance", flush=True)


def _layout_scene(dev):
    """World -1 shapes between and after the worlds: a static sphere after world 0 (resting 1 cm into the ground,
    so the last, world -1 only, segment has a sphere-plane pair) and the ground plane as the last shape; three
    worlds of boxes and spheres piled onto both."""
    from newton_custom_contact_amd.builder import ModelBuilder

    rng = np.random.Generator(np.random.PCG64(17))
    b = ModelBuilder()
    cfg = ModelBuilder.ShapeConfig()
    cfg.mu = 0.5

    def world():
        b.begin_world()
        for k in range(10):
            p = (rng.uniform(-0.5, 0.5), rng.uniform(-0.5, 0.5), rng.uniform(0.05, 0.9))
            body = b.add_body(xform=(*p, 0.0, 0.0, 0.0, 1.0))
            if k % 2:
                b.add_shape_box(body, hx=0.08, hy=0.1, hz=0.06, cfg=cfg)
            else:
                b.add_shape_sphere(body, radius=0.09, cfg=cfg)
        b.end_world()

    world()
    b.add_shape_sphere(-1, xform=(0.1, 0.0, 0.29, 0.0, 0.0, 0.0, 1.0), radius=0.3, cfg=cfg)
    world()
    world()
    b.add_ground_plane(cfg=cfg)
    return b.finalize(dev)


LAYOUT_SCENES = [("mixed_8w_plane_first", lambda dev: scenes.mixed_worlds(8, seed=3, device=dev)),
                 ("global_shapes_between_and_last", _layout_scene)]


@pytest.mark.parametrize("name,make", LAYOUT_SCENES, ids=[s[0] for s in LAYOUT_SCENES])
def test_deterministic_layout_element_by_element(cuda_device, name, make):
    """deterministic=True: the Contacts arrays equal, element by element with no sorting of the GPU side, the
    oracle's contacts in the reference's global sort-key order (collide.py:1105-1124); the order array maps the
    canonical per-world ranges onto that buffer; a step on it (reference-order kernel) equals the oracle's step
    bit for bit."""
    print(f"[deterministic layout] {name}", flush=True)
    model = make(cuda_device)
    oracle = Oracle(model)
    pipe = CollisionPipeline(model, broad_phase="sap", deterministic=True)
    solver = SolverXPBD(model, iterations=10, reference_arithmetic=True)
    sw = model.host["shape_world"]
    exercised_global_shape0 = False
    for q, qd in oracle_states(oracle, model, n_substeps=160, dt=DT, every=40):
        ref = oracle.collide(q, layout="global")
        c, got, state = _gpu_collide(pipe, model, q)
        n = ref["count"]
        assert got["count"] == n and n > 0
        for k in ("sort_key", "shape0", "shape1"):
            np.testing.assert_array_equal(got[k], ref[k], err_msg=k)
        for k in ("point0", "point1", "offset0", "offset1", "normal", "margin0", "margin1"):
            np.testing.assert_array_equal(got[k], ref[k], err_msg=k)
        exercised_global_shape0 |= bool(np.any(sw[got["shape0"]] == -1))
        # the canonical order: world by world (the world -1 only segment last), key order inside a world
        order = got["order"]
        canon = oracle.collide(q)
        cw = np.maximum(sw[canon["shape0"]], sw[canon["shape1"]])
        cw = np.where(cw < 0, model.world_count, cw)
        np.testing.assert_array_equal(got["sort_key"][order], canon["sort_key"][np.lexsort((canon["sort_key"], cw))])
        state.body_qd.copy_(torch.from_numpy(np.ascontiguousarray(qd, np.float32)))
        out = model.state()
        solver.step(state, out, None, c, DT)
        torch.cuda.synchronize()
        q1, qd1 = oracle.step(q, qd, np.zeros_like(qd), canon, DT, iterations=10)
        np.testing.assert_array_equal(out.body_q.cpu().numpy(), q1)
        np.testing.assert_array_equal(out.body_qd.cpu().numpy(), qd1)
    assert exercised_global_shape0
