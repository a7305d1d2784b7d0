"""GPU tests of the mesh path: SDF cook, mesh-mesh / mesh-plane contacts and the step over them.

* The GPU cook of the reference's texture SDF (nmc_sdf_build / nmc_sdf_subgrids / nmc_sdf_pack) against the
  oracle's cook, array for array, in the three texel formats.
* The cooked texture's accuracy against the exact FP64 distance to the mesh, with the reference's own
  tolerances for a 64-cell grid of a radius-0.5 sphere (tests/test_sdf_texture.py:1123-1228: distance
  mean 2e-4 / median 1.5e-4 / p95 7e-4 / max 2e-3 m, gradient angle mean 2 / median 1.5 / p95 5 deg
  within 0.05 m of the surface).
* Contacts of seeded mesh worlds (mesh-mesh and mesh-plane) against the oracle on the same grids, bit-exact
  after the key sort, for reduce_contacts=False (every contact) and for the reference's normal-bin / voxel
  reduction with the deterministic and the fast value packing; one XPBD step over them within 1e-5 relative.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

import oracle
from newton_custom_contact_amd import CollisionPipeline, scenes
from newton_custom_contact_amd.builder import ModelBuilder
from newton_custom_contact_amd.mesh import random_convex_hull
from newton_custom_contact_amd.solvers import SolverXPBD
from parity_helpers import assert_step_close, compare_contacts
from tests.mesh_helpers import box_mesh, icosphere

pytestmark = pytest.mark.gpu

DT = 1.0 / 240.0


def assert_texture_matches(got: dict, ref: dict, fmt: str):
    """A GPU-cooked texture against the oracle's: geometry, flags, slots and texture sizes equal; values equal
    in magnitude, with the winding-number sign free to differ only on the surface itself (|v| < 1e-6), which
    moves a unorm texel by at most one step."""
    for k in ("coarse_dims", "subgrid_tex_size", "num_subgrids", "subgrid_size", "quantization_mode"):
        assert tuple(np.atleast_1d(got[k])) == tuple(np.atleast_1d(ref[k])), k
    for k in ("min_extents", "max_extents", "cell_size"):
        np.testing.assert_array_equal(np.asarray(got[k], np.float64), np.asarray(ref[k], np.float64), err_msg=k)
    for k in ("subgrids_min_sdf_value", "subgrids_sdf_value_range"):
        assert np.float32(got[k]) == np.float32(ref[k]), k
    for k in ("subgrid_required", "subgrid_occupied", "subgrid_start_slots"):
        np.testing.assert_array_equal(got[k], ref[k], err_msg=k)
    gc, rc = got["coarse_sdf"], ref["coarse_sdf"]
    np.testing.assert_array_equal(np.abs(gc), np.abs(rc))
    assert np.all(np.abs(rc[np.sign(gc) != np.sign(rc)]) < 1e-6)
    gt, rt = got["subgrid_data"], ref["subgrid_data"]
    assert gt.dtype == rt.dtype and gt.shape == rt.shape
    if fmt == "float32":
        np.testing.assert_array_equal(np.abs(gt), np.abs(rt))
        assert np.all(np.abs(rt[np.sign(gt) != np.sign(rt)]) < 1e-6)
    else:
        assert np.abs(gt.astype(np.int64) - rt.astype(np.int64)).max() <= 1
        assert np.mean(gt != rt) < 1e-3


def test_sdf_cook_matches_oracle(cuda_device):
    rng = np.random.Generator(np.random.PCG64(11))
    meshes = [box_mesh(0.3, 0.2, 0.1), random_convex_hull(rng), random_convex_hull(rng, 64, 96)]
    for k, m in enumerate(meshes):
        for fmt, res in (("uint16", 24), ("float32", 32), ("uint8", 16)):
            print(f"[sdf cook parity] mesh {k}: {len(m.vertices)} vertices, {fmt}, {res} cells", flush=True)
            m.clear_sdf()
            sdf = m.build_sdf(max_resolution=res, device=cuda_device, texture_format=fmt)
            torch.cuda.synchronize()
            got = sdf.sparse_data()
            geo = oracle.texture_geometry(m.vertices, 0.05, (-0.1, 0.1), res)
            ref = oracle.sdf_build_sparse(m.vertices, m.indices, geo, fmt)
            assert ref["num_subgrids"] > 0
            assert_texture_matches(got, ref, fmt)


def test_sdf_accuracy_vs_exact_distance(cuda_device):
    m = icosphere(0.5, 3)
    m.build_sdf(max_resolution=64, margin=0.05, device=cuda_device)
    torch.cuda.synchronize()
    b = ModelBuilder()
    b.add_shape_mesh(b.add_body(), mesh=m)
    o = oracle.Oracle(b.finalize("cpu"))
    rng = np.random.Generator(np.random.PCG64(12))
    d = rng.normal(size=(3000, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    p = d * rng.uniform(0.42, 0.58, size=(3000, 1))
    val, grad = o.sdf_sample(0, p)
    exact = oracle.mesh_distance_f64(m.vertices, m.indices, p)
    ok = np.abs(exact) < 0.05
    assert ok.sum() > 500
    err = np.abs(val[ok] - exact[ok])
    print("distance error mean/median/p95/max", err.mean(), np.median(err), np.percentile(err, 95), err.max(),
          flush=True)
    assert err.mean() < 2e-4 and np.median(err) < 1.5e-4 and np.percentile(err, 95) < 7e-4 and err.max() < 2e-3
    h = 1e-4
    fd = np.stack([(oracle.mesh_distance_f64(m.vertices, m.indices, p[ok] + h * e) -
                    oracle.mesh_distance_f64(m.vertices, m.indices, p[ok] - h * e)) / (2 * h) for e in np.eye(3)], 1)
    g = grad[ok]
    cos = np.sum(g * fd, 1) / (np.linalg.norm(g, axis=1) * np.linalg.norm(fd, axis=1))
    ang = np.degrees(np.arccos(np.clip(cos, -1, 1)))
    print("gradient angle mean/median/p95", ang.mean(), np.median(ang), np.percentile(ang, 95), flush=True)
    assert ang.mean() < 2.0 and np.median(ang) < 1.5 and np.percentile(ang, 95) < 5.0


def _mesh_scene(dev, worlds=6, seed=21):
    lib = scenes.mesh_library(6, seed=seed, n_min=48, n_max=160)
    for mm in lib:
        mm.build_sdf(max_resolution=32, device=dev)
    return scenes.mesh_worlds(worlds, seed=seed, device=dev, library=lib)


def _poses(model, rng, k):
    """k pose sets: the initial one, then the bodies lowered towards the ground and each other."""
    q0 = model.host["body_q"].copy()
    out = [q0]
    for i in range(1, k):
        q = q0.copy()
        q[:, 2] = 0.35 * q0[:, 2] - 0.05 * i
        q[:, :2] *= 0.35
        out.append(q)
    return out


def _gpu_collide(pipe, model, q):
    state = model.state()
    state.body_q.copy_(torch.from_numpy(np.ascontiguousarray(q, np.float32)))
    c = pipe.contacts()
    pipe.collide(state, c)
    torch.cuda.synchronize()
    return c, c.numpy(), state


@pytest.mark.parametrize("reduce,det", [(True, True), (True, False), (False, True)],
                         ids=["reduced-deterministic", "reduced-fast", "unreduced"])
def test_mesh_collide_parity(cuda_device, reduce, det):
    from newton_custom_contact_amd.contacts import Contacts

    model = _mesh_scene(cuda_device)
    o = oracle.Oracle(model, reduce_contacts=reduce, deterministic=det)
    # unreduced, pairs keep every contact: a capacity beyond the 40-per-mesh-pair estimate
    pipe = CollisionPipeline(model, broad_phase="sap", deterministic=det, reduce_contacts=reduce,
                             rigid_contact_max=None if reduce else 200_000)
    rng = np.random.Generator(np.random.PCG64(3))
    reports = []
    poses = _poses(model, rng, 4)
    for i, q in enumerate(poses):
        print(f"[mesh collide parity] reduce={reduce} deterministic={det} pose {i}", flush=True)
        ref = o.collide(q, capacity=pipe.rigid_contact_max)
        state = model.state()
        state.body_q.copy_(torch.from_numpy(np.ascontiguousarray(q, np.float32)))
        # sort keys requested also from the fast-packing pipeline, to compare after the key sort
        c = Contacts(pipe.rigid_contact_max, 0, device=cuda_device, segment_count=len(model.host["segment_start"]) - 1,
                     deterministic=True)
        pipe.collide(state, c)
        torch.cuda.synchronize()
        reports.append(compare_contacts(c.numpy(), ref))
    print(reports, flush=True)
    assert sum(r["count"] for r in reports) > 0
    oc = pipe.overflow_counters()
    assert oc["pairs_dropped"] == 0 and oc["contacts_over_capacity"] == 0
    # both mesh-mesh and mesh-plane contacts were exercised; a reduced pair keeps at most 240
    last = o.collide(poses[-1], capacity=pipe.rigid_contact_max)
    st = model.host["shape_type"]
    assert np.any(st[last["shape1"]] == 8) and np.any(st[last["shape1"]] == 1)
    _, counts = np.unique(last["shape0"].astype(np.int64) * 100000 + last["shape1"], return_counts=True)
    print("contacts per pair: max", counts.max(), "mean", counts.mean(), flush=True)
    if reduce:
        assert counts.max() <= 240
        full = oracle.Oracle(model, reduce_contacts=False).collide(poses[-1], capacity=pipe.rigid_contact_max)
        assert last["count"] < full["count"]


def test_mesh_step_parity(cuda_device):
    model = _mesh_scene(cuda_device, worlds=4, seed=22)
    o = oracle.Oracle(model)
    pipe = CollisionPipeline(model, broad_phase="sap", deterministic=True)
    solver = SolverXPBD(model, iterations=10)
    rng = np.random.Generator(np.random.PCG64(4))
    for i, q in enumerate(_poses(model, rng, 3)):
        print(f"[mesh step parity] pose {i}", flush=True)
        qd = rng.normal(0, 0.2, size=(model.body_count, 6)).astype(np.float32)
        ref_c = o.collide(q, capacity=pipe.contacts().rigid_contact_max)
        c, _got, state_in = _gpu_collide(pipe, model, q)
        state_in.body_qd.copy_(torch.from_numpy(qd))
        state_out = model.state()
        solver.step(state_in, state_out, None, c, DT)
        torch.cuda.synchronize()
        q_ref, qd_ref = o.step(q, qd, np.zeros_like(qd), ref_c, DT, iterations=10)
        assert_step_close(state_out.body_q.cpu().numpy(), state_out.body_qd.cpu().numpy(), q_ref, qd_ref, DT, "mesh")


def test_mesh_worlds_settle(cuda_device):
    """Dropped hulls come to rest on the ground: none falls through, none keeps moving fast."""
    model = _mesh_scene(cuda_device, worlds=16, seed=23)
    pipe = CollisionPipeline(model, broad_phase="sap")
    solver = SolverXPBD(model, iterations=10)
    s0, s1 = model.state(), model.state()
    c = pipe.contacts()
    for _ in range(240 * 4):
        pipe.collide(s0, c)
        solver.step(s0, s1, None, c, DT)
        s0, s1 = s1, s0
    torch.cuda.synchronize()
    q = s0.body_q.cpu().numpy()
    qd = s0.body_qd.cpu().numpy()
    print("lowest body z", q[:, 2].min(), "max speed", np.abs(qd).max(), flush=True)
    assert q[:, 2].min() > 0.0
    assert np.all(np.isfinite(q)) and np.abs(qd[:, :3]).max() < 2.0
