"""Standalone broad phases on the GPU (nmc_broad_phase_launch / nmc_broad_phase_explicit) against the oracle.

Every mode reports the oracle's count and pair set (each pair once) on seeded multi-world scenes with shared
shapes, collision groups, per-shape gaps and a filter list; a world above the LDS limit (2 100 shapes) runs
the all-pairs global path and the SAP predicate path; two launches write identical arrays (fixed order); with
too small an output the count is still the number found and the written prefix equals the full run's;
explicit keeps the oracle's (input) order exactly; the reference's per-shape gap known answer holds.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

import oracle
from newton_custom_contact_amd.geometry import (BroadPhaseAllPairs, BroadPhaseExplicit, BroadPhaseSAP,
                                                precompute_world_map)
from tests.test_oracle_broad import _gap_scene, random_scene

pytestmark = pytest.mark.gpu


def _dev(a, dev, dt):
    return torch.from_numpy(np.ascontiguousarray(a, dt)).to(dev)


def _run(bp, dev, lower, upper, gap, group, world, filt, cap):
    out = torch.full((max(cap, 1), 2), -7, dtype=torch.int32, device=dev)
    cnt = torch.zeros(1, dtype=torch.int32, device=dev)
    bp.launch(_dev(lower, dev, np.float32), _dev(upper, dev, np.float32),
              None if gap is None else _dev(gap, dev, np.float32), _dev(group, dev, np.int32),
              _dev(world, dev, np.int32), len(world), out[:cap] if cap else out[:0], cnt,
              filter_pairs=None if filt is None else _dev(filt, dev, np.int32))
    torch.cuda.synchronize()
    return int(cnt.item()), out.cpu().numpy()


SCENES = [("3w_12", dict(worlds=3, per_world=12, shared=3)),
          ("64w_33", dict(worlds=64, per_world=32, shared=1, extent=2.0)),
          ("2w_300", dict(worlds=2, per_world=300, shared=4, extent=4.0)),
          ("1w_2100", dict(worlds=1, per_world=2100, shared=2, extent=12.0))]


@pytest.mark.parametrize("name,kw", SCENES, ids=[s[0] for s in SCENES])
@pytest.mark.parametrize("mode", ["sap", "nxn"])
def test_broad_phase_matches_oracle(cuda_device, name, kw, mode):
    print(f"[broad phase] {mode} {name}", flush=True)
    rng = np.random.Generator(np.random.PCG64(sum(map(ord, name))))
    lower, upper, gap, group, world, filt = random_scene(rng, **kw)
    im, se = precompute_world_map(world)
    ref = oracle.broad_phase(mode, lower, upper, gap, group, world, filt, im, se)
    bp = (BroadPhaseSAP if mode == "sap" else BroadPhaseAllPairs)(world)
    cap = len(ref) + 16
    n, got = _run(bp, cuda_device, lower, upper, gap, group, world, filt, cap)
    assert n == len(ref) and n > 0
    pairs = [tuple(p) for p in got[:n].tolist()]
    assert len(set(pairs)) == n
    assert set(pairs) == {tuple(p) for p in ref.tolist()}
    assert np.all(got[n:cap] == -7)
    n2, got2 = _run(bp, cuda_device, lower, upper, gap, group, world, filt, cap)
    np.testing.assert_array_equal(got2, got)
    # too small an output: full count, the same prefix
    k = n // 3
    n3, got3 = _run(bp, cuda_device, lower, upper, gap, group, world, filt, k)
    assert n3 == n
    np.testing.assert_array_equal(got3[:k], got[:k])
    # no gaps, no filter
    ref0 = oracle.broad_phase(mode, lower, upper, None, group, world, None, im, se)
    n4, got4 = _run(bp, cuda_device, lower, upper, None, group, world, None, len(ref0) + 1)
    assert n4 == len(ref0)
    assert {tuple(p) for p in got4[:n4].tolist()} == {tuple(p) for p in ref0.tolist()}


def test_explicit_matches_oracle(cuda_device):
    print("[broad phase] explicit", flush=True)
    rng = np.random.Generator(np.random.PCG64(5))
    lower, upper, gap, group, world, _ = random_scene(rng, worlds=8, per_world=24, shared=0)
    n = len(world)
    pairs = np.array([(a, b) for a in range(n) for b in range(a + 1, n) if world[a] == world[b]], np.int32)
    rng.shuffle(pairs)
    ref = oracle.broad_phase("explicit", lower, upper, gap, group, world, shape_pairs=pairs)
    bp = BroadPhaseExplicit()
    dev = cuda_device
    out = torch.zeros((len(pairs), 2), dtype=torch.int32, device=dev)
    cnt = torch.zeros(1, dtype=torch.int32, device=dev)
    bp.launch(_dev(lower, dev, np.float32), _dev(upper, dev, np.float32), _dev(gap, dev, np.float32),
              _dev(pairs, dev, np.int32), len(pairs), out, cnt)
    torch.cuda.synchronize()
    k = int(cnt.item())
    assert k == len(ref) and k > 0
    np.testing.assert_array_equal(out[:k].cpu().numpy(), ref)
    # a pair naming a shape past the AABB arrays is skipped, never read
    bad = np.concatenate([pairs[:5], [[0, n + 5]]]).astype(np.int32)
    ref_b = oracle.broad_phase("explicit", lower, upper, gap, group, world, shape_pairs=bad[:5])
    bp.launch(_dev(lower, dev, np.float32), _dev(upper, dev, np.float32), _dev(gap, dev, np.float32),
              _dev(bad, dev, np.int32), len(bad), out, cnt)
    torch.cuda.synchronize()
    k = int(cnt.item())
    np.testing.assert_array_equal(out[:k].cpu().numpy(), ref_b)


def test_per_shape_gap_known_answer_gpu(cuda_device):
    lower, upper, gap, group, world = _gap_scene()
    for cls in (BroadPhaseAllPairs, BroadPhaseSAP):
        n, got = _run(cls(world), cuda_device, lower, upper, gap, group, world, None, 8)
        assert n == 1 and tuple(got[0]) == (0, 2)


# ---------------------------------------------------------------- the pipeline's broad-phase modes
def _pipe_vs_oracle(model, mode, q=None, explicit=None):
    from newton_custom_contact_amd import CollisionPipeline
    from parity_helpers import compare_contacts

    q = model.host["body_q"] if q is None else q
    pipe = CollisionPipeline(model, broad_phase=mode, deterministic=True, shape_pairs_filtered=explicit)
    o = oracle.Oracle(model, broad_phase=mode, explicit_pairs=explicit)
    state = model.state()
    state.body_q.copy_(torch.from_numpy(np.ascontiguousarray(q, np.float32)))
    c = pipe.contacts()
    pipe.collide(state, c)
    torch.cuda.synchronize()
    ref = o.collide(q, capacity=c.rigid_contact_max)
    return compare_contacts(c.numpy(), ref)


def test_pipeline_modes_corner_touching(cuda_device):
    from tests.test_oracle_broad import corner_boxes_model

    print("[pipeline broad phase] corner-touching boxes", flush=True)
    counts = {"sap": 0, "nxn": 0, "explicit": 0}
    for off in np.arange(2, 12) / 8.0:
        model = corner_boxes_model(float(off), cuda_device)
        for mode in counts:
            counts[mode] += _pipe_vs_oracle(model, mode)["count"]
    print(counts, flush=True)
    assert counts["nxn"] == counts["explicit"] == 10 and counts["sap"] < 10


def test_pipeline_explicit_list_matches_oracle(cuda_device):
    from newton_custom_contact_amd import scenes

    print("[pipeline broad phase] explicit list", flush=True)
    model = scenes.mixed_worlds(8, seed=3, device=cuda_device)
    allp = model.shape_contact_pairs.cpu().numpy()
    rng = np.random.Generator(np.random.PCG64(3))
    sub = allp[rng.random(len(allp)) < 0.4]
    q = model.host["body_q"].copy()
    q[:, 2] *= 0.3  # bodies lowered into one another and the ground
    q[:, :2] *= 0.4
    rep_all = _pipe_vs_oracle(model, "explicit", q)
    rep_nxn = _pipe_vs_oracle(model, "nxn", q)
    rep_sub = _pipe_vs_oracle(model, "explicit", q, explicit=sub)
    print(rep_all, rep_nxn, rep_sub, flush=True)
    assert rep_all["count"] == rep_nxn["count"] > rep_sub["count"] > 0


def test_skip_count_zero_appends(cuda_device):
    """skip_count_zero=True: the pairs found go after the count already in the buffer and the count grows by
    their number (two launches chained into one buffer hold both full results, in launch order)."""
    print("[broad phase] skip_count_zero chaining", flush=True)
    dev = cuda_device
    rng = np.random.Generator(np.random.PCG64(11))
    lower, upper, gap, group, world, filt = random_scene(rng, worlds=6, per_world=20, shared=2)
    im, se = precompute_world_map(world)
    bp = BroadPhaseAllPairs(world)
    n1, got1 = _run(bp, dev, lower, upper, gap, group, world, filt, 4096)
    assert n1 > 0
    args = (_dev(lower, dev, np.float32), _dev(upper, dev, np.float32), _dev(gap, dev, np.float32),
            _dev(group, dev, np.int32), _dev(world, dev, np.int32), len(world))
    out = torch.full((2 * n1 + 8, 2), -7, dtype=torch.int32, device=dev)
    cnt = torch.zeros(1, dtype=torch.int32, device=dev)
    fp = _dev(filt, dev, np.int32)
    bp.launch(*args, out, cnt, filter_pairs=fp)
    bp.launch(*args, out, cnt, filter_pairs=fp, skip_count_zero=True)
    torch.cuda.synchronize()
    assert int(cnt.item()) == 2 * n1
    o = out.cpu().numpy()
    np.testing.assert_array_equal(o[:n1], got1[:n1])
    np.testing.assert_array_equal(o[n1:2 * n1], got1[:n1])
    assert np.all(o[2 * n1:] == -7)
    # without skip_count_zero the count restarts from zero
    bp.launch(*args, out, cnt, filter_pairs=fp)
    torch.cuda.synchronize()
    assert int(cnt.item()) == n1
    # explicit: a count of 5 on entry puts the pairs at [5, 5 + k); an output too small keeps the full count
    n = len(world)
    pairs = np.array([(a, b) for a in range(n) for b in range(a + 1, n) if world[a] == world[b]], np.int32)
    ref = oracle.broad_phase("explicit", lower, upper, gap, group, world, shape_pairs=pairs)
    ex = BroadPhaseExplicit()
    out = torch.full((len(ref) + 5, 2), -7, dtype=torch.int32, device=dev)
    cnt = torch.full((1,), 5, dtype=torch.int32, device=dev)
    ex.launch(args[0], args[1], args[2], _dev(pairs, dev, np.int32), len(pairs), out, cnt, skip_count_zero=True)
    torch.cuda.synchronize()
    assert int(cnt.item()) == len(ref) + 5
    o = out.cpu().numpy()
    assert np.all(o[:5] == -7)
    np.testing.assert_array_equal(o[5:], ref)
    small = torch.full((len(ref), 2), -7, dtype=torch.int32, device=dev)
    cnt.fill_(5)
    ex.launch(args[0], args[1], args[2], _dev(pairs, dev, np.int32), len(pairs), small, cnt, skip_count_zero=True)
    torch.cuda.synchronize()
    assert int(cnt.item()) == len(ref) + 5
    np.testing.assert_array_equal(small.cpu().numpy()[5:], ref[:len(ref) - 5])
