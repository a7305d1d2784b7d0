"""Launch-shape autotuning for the XPBD step kernel.

The step kernel runs one workgroup per world.  Two shapes exist for worlds of up to 64 bodies: one
64-lane wave holding 128 contacts in registers (best when a world has a few hundred contacts at most),
and 256 lanes holding 256 (more lanes per world for contact-heavy worlds, e.g. triangle meshes with up
to 64 contacts per pair).  Models with larger worlds can launch by size tier (automatic, 0) or run every
world in one launch of the large-world shape (256: 512 lanes a world).  All give bit-identical results, so the choice is purely a timing question, answered
here by timing the solver on the caller's own state and contacts.
"""

from __future__ import annotations

import torch


def candidate_block_lanes(solver) -> list[int]:
    """Launch shapes the solver's model admits: 64 lanes needs worlds of at most 64 bodies; 0 (automatic)
    launches worlds by size tier (<= 64 bodies on one wave, larger worlds on 512 lanes), 256 runs every
    world in one launch of the large-world shape."""
    return [64, 256] if solver.native.desc.max_world_bodies <= 64 else [0, 256]


def tune_block_lanes(solver, state_in, state_out, contacts, dt: float, reps: int = 5) -> dict:
    """Time ``reps`` back-to-back steps with every candidate shape and keep the fastest in
    ``solver.block_lanes``.  ``state_out`` is overwritten.  Returns {lanes: ms per step} and the choice."""
    times = {}
    stream = torch.cuda.current_stream(solver.model.device)
    keep = solver.block_lanes
    for lanes in candidate_block_lanes(solver):
        solver.block_lanes = lanes
        solver.step(state_in, state_out, None, contacts, dt)  # warm (first launch of this shape)
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(reps):
            solver.step(state_in, state_out, None, contacts, dt)
        e1.record(stream)
        torch.cuda.synchronize(solver.model.device)
        times[lanes] = e0.elapsed_time(e1) / reps
    solver.block_lanes = min(times, key=times.get) if times else keep
    return {"ms_per_step": times, "block_lanes": solver.block_lanes}


__all__ = ["candidate_block_lanes", "tune_block_lanes"]
