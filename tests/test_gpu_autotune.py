"""The step and collide kernels' launch shapes give the same results (bit for bit, or within the step bar where noted),
and the autotuner picks one of them."""

from __future__ import annotations

import numpy as np
import pytest
import torch

from newton_custom_contact_amd import CollisionPipeline, scenes
from newton_custom_contact_amd.autotune import candidate_block_lanes, tune_block_lanes
from newton_custom_contact_amd.solvers import SolverXPBD

pytestmark = pytest.mark.gpu

DT = 1.0 / 240.0


def _mesh_model(dev):
    lib = scenes.mesh_library(4, seed=31, n_min=48, n_max=96)
    for mm in lib:
        mm.build_sdf(max_resolution=24, device=dev)
    return scenes.mesh_worlds(4, seed=31, device=dev, library=lib)


MODELS = [
    ("mixed_8w", lambda dev: scenes.mixed_worlds(8, seed=5, device=dev)),
    ("mesh_4w", _mesh_model),
]


@pytest.mark.parametrize("name,make", MODELS, ids=[m[0] for m in MODELS])
def test_block_lanes_bit_identical(cuda_device, name, make):
    from parity_helpers import assert_step_close

    print(f"[block lanes] {name}", flush=True)
    model = make(cuda_device)
    pipe = CollisionPipeline(model, broad_phase="sap")
    solver = SolverXPBD(model, iterations=10)
    assert candidate_block_lanes(solver) == [64, 256, 512]
    c = pipe.contacts()
    s0, s1 = model.state(), model.state()
    # settle a little so that worlds carry contacts
    for _ in range(120):
        pipe.collide(s0, c)
        solver.step(s0, s1, None, c, DT)
        s0, s1 = s1, s0
    pipe.collide(s0, c)
    assert int(c.rigid_contact_count.item()) > 0
    for ref_order in (False, True):
        sv = SolverXPBD(model, iterations=10, reference_arithmetic=ref_order)
        outs = {}
        for lanes in (64, 256, 512):
            sv.block_lanes = lanes
            a, b = model.state(), model.state()
            a.body_q.copy_(s0.body_q)
            a.body_qd.copy_(s0.body_qd)
            sv.step(a, b, None, c, DT)
            torch.cuda.synchronize()
            outs[lanes] = (b.body_q.cpu().numpy(), b.body_qd.cpu().numpy())
        assert np.array_equal(outs[64][0], outs[256][0]) and np.array_equal(outs[64][1], outs[256][1])
        # CSR bins (worlds packed into one 64-body wave): the same with the reference order, within the step bar
        # with the world-frame arithmetic (one lane a body instead of a tree of lanes)
        if ref_order:
            assert np.array_equal(outs[64][0], outs[512][0]) and np.array_equal(outs[64][1], outs[512][1])
        else:
            assert_step_close(outs[512][0], outs[512][1], outs[64][0], outs[64][1], DT, "bins vs worlds")
    res = tune_block_lanes(solver, s0, s1, c, DT, reps=2)
    print(name, res, flush=True)
    assert res["block_lanes"] in (64, 256, 512) and solver.block_lanes == res["block_lanes"]


@pytest.mark.parametrize("big", [False, True], ids=["bins_256", "bins_and_large_worlds"])
def test_tiered_launch_bit_identical_uneven(cuda_device, big):
    """Worlds of 8-262 bodies (or, `big`, 60-306 with some above 256): launching by size tier (automatic) equals
    running every world on 256 lanes, bit for bit; the binned launch (CSR bins of consecutive worlds on 256 lanes,
    beside a 512-lane launch of the worlds above 256 bodies) equals both bit for bit with the reference-order
    arithmetic, and within the step bar with the world-frame arithmetic (its packed worlds sum per body on one
    lane instead of a tree of lanes)."""
    from parity_helpers import assert_step_close

    print(f"[block lanes] uneven tiers and bins, big={big}", flush=True)
    if big:
        model = scenes.uneven_worlds(10, seed=5, lo=60, hi=300, device=cuda_device)
    else:
        model = scenes.uneven_worlds(12, seed=8, device=cuda_device)
    pipe = CollisionPipeline(model, broad_phase="sap")
    solver = SolverXPBD(model, iterations=10)
    assert candidate_block_lanes(solver) == [0, 256, 512]
    counts = model.host["world_body_count"]
    if big:
        assert (counts[:-1] > 256).any() and ((counts[:-1] > 64) & (counts[:-1] <= 256)).any()
    else:
        assert counts.max() > 64 and (counts[:-1] <= 64).any()
    c = pipe.contacts()
    s0, s1 = model.state(), model.state()
    for _ in range(60):
        pipe.collide(s0, c)
        solver.step(s0, s1, None, c, DT)
        s0, s1 = s1, s0
    pipe.collide(s0, c)
    assert int(c.rigid_contact_count.item()) > 0
    for ref_order in (False, True):
        sv = SolverXPBD(model, iterations=10, reference_arithmetic=ref_order)
        outs = {}
        for lanes in (0, 256, 512):
            sv.block_lanes = lanes
            a, b = model.state(), model.state()
            a.body_q.copy_(s0.body_q)
            a.body_qd.copy_(s0.body_qd)
            sv.step(a, b, None, c, DT)
            torch.cuda.synchronize()
            outs[lanes] = (b.body_q.cpu().numpy(), b.body_qd.cpu().numpy())
        assert np.array_equal(outs[0][0], outs[256][0]) and np.array_equal(outs[0][1], outs[256][1])
        if ref_order:
            assert np.array_equal(outs[0][0], outs[512][0]) and np.array_equal(outs[0][1], outs[512][1])
        else:
            assert_step_close(outs[512][0], outs[512][1], outs[0][0], outs[0][1], DT, "bins vs tiers")
    res = tune_block_lanes(solver, s0, s1, c, DT, reps=2)
    print(res, flush=True)
    assert res["block_lanes"] in (0, 256, 512)


@pytest.mark.parametrize("name,make", MODELS, ids=[m[0] for m in MODELS])
def test_collide_launch_shapes_bit_identical(cuda_device, name, make):
    """Every collide launch shape the tuner tries writes the same contacts, bit for bit; the tuner keeps one
    of its candidates and leaves it set on the pipeline."""
    from newton_custom_contact_amd.autotune import collide_candidates, tune_collide

    print(f"[collide launch] {name}", flush=True)
    model = make(cuda_device)
    pipe = CollisionPipeline(model, broad_phase="sap", deterministic=True)
    solver = SolverXPBD(model, iterations=10)
    c = pipe.contacts()
    s0, s1 = model.state(), model.state()
    for _ in range(60):
        pipe.collide(s0, c)
        solver.step(s0, s1, None, c, DT)
        s0, s1 = s1, s0
    pipe.collide(s0, c)
    ref = c.numpy()
    assert ref["count"] > 0
    cand = collide_candidates(pipe)
    assert ("mesh_max_blocks" in cand) == (name == "mesh_4w")
    for key, values in cand.items():
        for v in values:
            pipe.set_launch(**{key: v})
            pipe.collide(s0, c)
            got = c.numpy()
            assert got["count"] == ref["count"], (key, v)
            for k in ("shape0", "shape1", "point0", "point1", "normal", "sort_key", "margin0", "margin1"):
                assert np.array_equal(got[k], ref[k]), (key, v, k)
        pipe.set_launch(**{key: values[-1]})
    with pytest.raises(ValueError):
        pipe.set_launch(segment_waves=3)
    with pytest.raises(ValueError):
        pipe.set_launch(narrow_max_blocks=0)
    res = tune_collide(pipe, s0, c, reps=2)
    print(name, res, flush=True)
    for key, values in cand.items():
        assert res["launch"][key] in values and pipe.launch_config[key] == res["launch"][key]
