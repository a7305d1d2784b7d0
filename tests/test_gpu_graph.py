"""Graph capture of the substep: collide + XPBD step (plain; deterministic with contact matching and
reported forces; worlds of uneven size, whose step launches its size tiers concurrently on side streams forked
from and joined into the capturing stream; the same with per-shape restitution, whose pass follows the joined
tiers) captured once with torch.cuda.CUDAGraph and replayed; every replay,
with unrelated eager work and a synchronisation between replays, equals the eager substep bit for bit.  The
library records no memset nodes (scratch is cleared by its own kernels), so replays after the first are
checked too.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

from newton_custom_contact_amd import CollisionPipeline, scenes
from newton_custom_contact_amd.solvers import SolverXPBD

pytestmark = pytest.mark.gpu

DT = 1.0 / 240.0


def _setup(dev, report, uneven=False, restitution=False):
    model = scenes.uneven_worlds(12, seed=8, device=dev) if uneven else scenes.mixed_worlds(16, seed=5, device=dev)
    if restitution:
        e = np.random.default_rng(4).uniform(0.0, 0.9, model.shape_count).astype(np.float32)
        model.shape_material_restitution.copy_(torch.from_numpy(e))
    if report:
        model.request_contact_attributes("force")
        model.request_state_attributes("body_parent_f")
        pipe = CollisionPipeline(model, broad_phase="sap", contact_matching="latest", contact_report=True)
    else:
        pipe = CollisionPipeline(model, broad_phase="sap")
    solver = SolverXPBD(model, iterations=10, enable_restitution=restitution)
    return model, pipe, solver


def _outputs(c, s, report):
    out = [s.body_q.clone(), s.body_qd.clone(), c.rigid_contact_count.clone()]
    if report:
        out += [c.force.clone(), s.body_parent_f.clone(), c.rigid_contact_match_index.clone()]
    return out


@pytest.mark.parametrize("report,uneven,restitution",
                         [(False, False, False), (True, False, False), (False, True, False), (False, True, True)],
                         ids=["plain", "matching_and_forces", "uneven_tiers", "uneven_tiers_restitution"])
def test_graph_replays_equal_eager(cuda_device, report, uneven, restitution):
    print(f"[graph] report={report} uneven={uneven} restitution={restitution}", flush=True)
    steps = 12
    # eager reference
    model, pipe, solver = _setup(cuda_device, report, uneven, restitution)
    if uneven:
        counts = model.host["world_body_count"][:-1]
        assert counts.max() > 64 and (counts <= 64).any(), "the scene should launch its step in several size tiers"
    c = pipe.contacts()
    a, b = model.state(), model.state()
    q0, qd0 = a.body_q.clone(), a.body_qd.clone()
    ref = []
    for _ in range(steps):
        pipe.collide(a, c)
        solver.step(a, b, None, c, DT)
        if report:
            solver.update_contacts(c)
        a.body_q.copy_(b.body_q)
        a.body_qd.copy_(b.body_qd)
        torch.cuda.synchronize()
        ref.append([t.cpu().numpy() for t in _outputs(c, b, report)])
    # the same substep captured and replayed (fresh objects, same initial state)
    model, pipe, solver = _setup(cuda_device, report, uneven, restitution)
    c = pipe.contacts()
    a, b = model.state(), model.state()

    def substep():
        pipe.collide(a, c)
        solver.step(a, b, None, c, DT)
        if report:
            solver.update_contacts(c)
        a.body_q.copy_(b.body_q)
        a.body_qd.copy_(b.body_qd)

    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        substep()  # warm-up outside the graph (lazy allocations)
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        substep()
    torch.cuda.synchronize()
    if report:
        pipe.contact_matcher.reset()
    a.body_q.copy_(q0)
    a.body_qd.copy_(qd0)
    torch.cuda.synchronize()
    noise = torch.zeros(1 << 16, device=cuda_device)
    for k in range(steps):
        graph.replay()
        torch.cuda.synchronize()
        got = [t.cpu().numpy() for t in _outputs(c, b, report)]
        for g, r in zip(got, ref[k]):
            np.testing.assert_array_equal(g, r, err_msg=f"replay {k}")
        noise.add_(1.0).mul_(0.5)  # unrelated eager work between replays
        torch.cuda.synchronize()
    assert int(ref[-1][2][0]) > 0
