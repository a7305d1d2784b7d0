"""Launch-shape autotuning of the collide and step kernels.

The launch shapes below give the same results (bit for bit, except where the step notes otherwise), so the
choice is a timing question, answered here by timing the kernels on the caller's own state and contacts (HIP
events on the current stream).

Step kernel (``SolverXPBD.block_lanes``): one workgroup per world.  Two shapes exist for worlds of up to 64
bodies: one 64-lane wave holding 128 contacts in registers (best when a world has a few hundred contacts at
most), and 256 lanes holding 256 (more lanes per world for contact-heavy worlds, e.g. triangle meshes with up
to 64 contacts per pair).  Models with larger worlds can launch by size tier (automatic, 0), run every world
in one launch of the large-world shape (256: 512 lanes a world); any multi-world model can run one launch over
CSR bins (512: runs of consecutive worlds packed up to max(largest world, 64) bodies, largest bins first; 512
lanes a bin, or one wave when every world fits one).  The reference-order
arithmetic is bit-identical across all of them; the world-frame arithmetic rounds its per-body sums by the
lanes a body gets, which the bins change for the small worlds they pack.

Collide kernels (``CollisionPipeline.set_launch``): ``segment_waves``, the segments of at most 64 shapes per
workgroup of the broad phase (one wave each: more waves per workgroup mean fewer workgroups, fewer waves
spread the segments over more CUs); ``narrow_max_blocks`` and ``mesh_max_blocks``, the grid caps of the
grid-stride narrow-phase and mesh kernels (more workgroups hide more latency until the tail of uneven pairs
dominates).  They are tuned one after another (coordinate descent from the defaults), each over its candidate
list, keeping the fastest collide.
"""

from __future__ import annotations

import torch

SEGMENT_WAVES = (1, 2, 4)
NARROW_MAX_BLOCKS = (512, 1024, 2048, 4096, 8192)
MESH_MAX_BLOCKS = (256, 512, 1024, 2048, 4096)


def _time(fn, device, reps: int) -> float:
    fn()  # warm (first launch of this shape)
    stream = torch.cuda.current_stream(device)
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(reps):
        fn()
    e1.record(stream)
    torch.cuda.synchronize(device)
    return e0.elapsed_time(e1) / reps


def candidate_block_lanes(solver) -> list[int]:
    """Launch shapes the solver's model admits: 64 lanes needs worlds of at most 64 bodies; 0 (automatic)
    launches worlds by size tier (<= 64 bodies on one wave, larger worlds on 512 lanes), 256 runs every
    world in one launch of the large-world shape, 512 runs one launch over CSR bins of consecutive worlds
    (packed up to max(largest world, 64) bodies: one wave per bin when every world fits one)."""
    return [64, 256, 512] if solver.native.desc.max_world_bodies <= 64 else [0, 256, 512]


def tune_block_lanes(solver, state_in, state_out, contacts, dt: float, reps: int = 5) -> dict:
    """Time ``reps`` back-to-back steps with every candidate shape and keep the fastest in
    ``solver.block_lanes``.  ``state_out`` is overwritten.  Returns {lanes: ms per step} and the choice."""
    times = {}
    keep = solver.block_lanes
    for lanes in candidate_block_lanes(solver):
        solver.block_lanes = lanes
        times[lanes] = _time(lambda: solver.step(state_in, state_out, None, contacts, dt), solver.model.device, reps)
    solver.block_lanes = min(times, key=times.get) if times else keep
    return {"ms_per_step": times, "block_lanes": solver.block_lanes}


def collide_candidates(pipe) -> dict:
    """Candidate values per collide launch parameter for the pipeline's model (the mesh grid cap only when the
    model has meshes)."""
    cand = {"segment_waves": list(SEGMENT_WAVES), "narrow_max_blocks": list(NARROW_MAX_BLOCKS)}
    if getattr(pipe.model, "meshes", []):
        cand["mesh_max_blocks"] = list(MESH_MAX_BLOCKS)
    return cand


def tune_collide(pipe, state, contacts, reps: int = 5) -> dict:
    """Coordinate descent over the collide launch parameters: for each parameter in turn, time ``reps``
    collides at every candidate value (the others held at their current best) and keep the fastest.
    ``contacts`` is overwritten.  Returns {parameter: {value: ms per collide}} and the chosen configuration,
    which stays set on ``pipe``."""
    times = {}
    for name, values in collide_candidates(pipe).items():
        t = {}
        for v in values:
            pipe.set_launch(**{name: v})
            t[v] = _time(lambda: pipe.collide(state, contacts), pipe.model.device, reps)
        pipe.set_launch(**{name: min(t, key=t.get)})
        times[name] = t
    return {"ms_per_collide": times, "launch": pipe.launch_config}


def tune(pipe, solver, state_in, state_out, contacts, dt: float, reps: int = 5) -> dict:
    """Tune the collide launch shape on ``state_in``, collide once with it, then tune the step kernel's shape
    on the resulting contacts.  ``contacts`` and ``state_out`` are overwritten."""
    col = tune_collide(pipe, state_in, contacts, reps)
    pipe.collide(state_in, contacts)
    step = tune_block_lanes(solver, state_in, state_out, contacts, dt, reps)
    return {"collide": col, "step": step}


__all__ = ["candidate_block_lanes", "collide_candidates", "tune", "tune_block_lanes", "tune_collide"]
