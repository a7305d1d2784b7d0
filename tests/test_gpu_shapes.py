"""Ellipsoids, cylinders and cones on the GPU against the oracle: worlds of every primitive type falling onto the
ground and onto each other (analytic pairs: plane-ellipsoid, plane-cylinder, sphere-cylinder; the rest through
MPR/GJK, the manifold and the rolling post-process).  Contacts equal the oracle's bit for bit after the
canonical key sort (float fields within 1e-5 relative and in fact bit-identical); one step from identical inputs
equals the oracle bit for bit with the reference-order kernel and within the step bar with the world-frame one.
"""

from __future__ import annotations

import math

import numpy as np
import pytest
import torch

from newton_custom_contact_amd import CollisionPipeline
from newton_custom_contact_amd._types import transform
from newton_custom_contact_amd.builder import ModelBuilder
from newton_custom_contact_amd.solvers import SolverXPBD
from oracle import Oracle
from parity_helpers import assert_step_close, bit_equal, compare_contacts, oracle_states

pytestmark = pytest.mark.gpu

DT = 1.0 / 240.0


def _shape_worlds(dev, worlds=6, seed=21):
    rng = np.random.default_rng(seed)
    b = ModelBuilder()
    cfg = ModelBuilder.ShapeConfig(mu=0.6)
    b.add_ground_plane(cfg=cfg)
    for _w in range(worlds):
        b.begin_world()
        # a static box as a step to land on
        b.add_shape_box(-1, xform=transform((0.3, 0.2, 0.1), (0.0, 0.0, 0.0, 1.0)), hx=0.3, hy=0.25, hz=0.1,
                        cfg=cfg)
        for k in range(14):
            p = (rng.uniform(-0.6, 0.6), rng.uniform(-0.6, 0.6), 0.25 + 0.22 * k)
            ax = rng.normal(size=3)
            ax /= np.linalg.norm(ax)
            ang = rng.uniform(0.0, math.pi)
            q = (*(ax * math.sin(0.5 * ang)), math.cos(0.5 * ang))
            body = b.add_body(xform=transform(p, q))
            kind = k % 6
            if kind == 0:
                b.add_shape_cylinder(body, radius=rng.uniform(0.06, 0.12), half_height=rng.uniform(0.05, 0.15), cfg=cfg)
            elif kind == 1:
                b.add_shape_cone(body, radius=rng.uniform(0.06, 0.12), half_height=rng.uniform(0.06, 0.14), cfg=cfg)
            elif kind == 2:
                b.add_shape_ellipsoid(body, rx=rng.uniform(0.05, 0.12), ry=rng.uniform(0.05, 0.12),
                                      rz=rng.uniform(0.05, 0.12), cfg=cfg)
            elif kind == 3:
                b.add_shape_sphere(body, radius=rng.uniform(0.05, 0.1), cfg=cfg)
            elif kind == 4:
                b.add_shape_box(body, hx=rng.uniform(0.05, 0.1), hy=rng.uniform(0.05, 0.1), hz=rng.uniform(0.05, 0.1),
                                cfg=cfg)
            else:
                b.add_shape_capsule(body, radius=rng.uniform(0.04, 0.08), half_height=rng.uniform(0.05, 0.12), cfg=cfg)
        b.end_world()
    return b.finalize(dev)


def _gpu_collide(pipe, model, q):
    state = model.state()
    state.body_q.copy_(torch.from_numpy(np.ascontiguousarray(q, np.float32)))
    c = pipe.contacts()
    pipe.collide(state, c)
    torch.cuda.synchronize()
    return c, c.numpy(), state


def test_shapes_collide_and_step_parity(cuda_device):
    print("[shapes] 6 worlds of cylinders, cones, ellipsoids, spheres, boxes, capsules", flush=True)
    model = _shape_worlds(cuda_device)
    oracle = Oracle(model)
    pipe = CollisionPipeline(model, broad_phase="sap", deterministic=True)
    types = model.host["shape_type"]
    seen = set()
    reports = []
    for q, qd in oracle_states(oracle, model, n_substeps=360, dt=DT, every=60):
        ref = oracle.collide(q)
        c, got, state = _gpu_collide(pipe, model, q)
        reports.append(compare_contacts(got, ref))
        for a, bb in zip(ref["shape0"], ref["shape1"]):
            seen.add((int(types[a]), int(types[bb])))
        state.body_qd.copy_(torch.from_numpy(np.ascontiguousarray(qd, np.float32)))
        q_ref, qd_ref = oracle.step(q, qd, np.zeros_like(qd), ref, DT, iterations=10)
        for ref_arith in (True, False):
            solver = SolverXPBD(model, iterations=10, reference_arithmetic=ref_arith)
            out = model.state()
            solver.step(state, out, None, c, DT)
            torch.cuda.synchronize()
            qg, qdg = out.body_q.cpu().numpy(), out.body_qd.cpu().numpy()
            if ref_arith:
                assert bit_equal(qg, q_ref) and bit_equal(qdg, qd_ref), "reference-order step differs"
            else:
                assert_step_close(qg, qdg, q_ref, qd_ref, DT, "shapes")
    print(reports, sorted(seen), flush=True)
    assert sum(r["count"] for r in reports) > 0
    # the pairs that exercise the new colliders appear: plane (1) with ellipsoid (5), cylinder (6), cone (9)
    assert {(1, 5), (1, 6), (1, 9)} <= seen
