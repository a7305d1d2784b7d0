"""World-steps/s of the rigid-contact step on one MI355X.

Workload: 4096 identical-structure worlds x 32 mixed primitives (11 spheres, 11 boxes, 10 capsules) dropped
onto one global ground plane, per-world seeded poses, fp32, dt = 1/60 s as 4 substeps, 10 XPBD
iterations; each substep is CollisionPipeline.collide + SolverXPBD.step, and one step (4 substeps,
16 kernel launches) is captured in a hipGraph and replayed.  Synthetic scene, no external data.

One JSON line is printed.  ``roofline`` prices the dominant kernel against HBM bandwidth with the
algorithmic bytes of the API boundary (each substep's collide/step read their inputs and write their
outputs once):
    B_collide = 28 N_b + 64 N_s + 76 N_c + 4          bytes per world-substep
    B_step    = 76 N_b + 96 N_b + 16 N_s + 76 N_c + 52 N_b + 120 N_j
N_c is the measured mean contact count per world per substep in the timed region.  Every form first runs an
untimed settle phase (``--settle`` steps, 2 s of simulated time by default, whatever ``--warmup`` is) so that the
timed region measures the piled-up scene, not bodies still falling.
``iterations`` sweeps the solver's iteration count at the settled state: the mean over worlds of the largest
remaining penetration after one substep, plain (1..``--iters-max`` iterations) and learned warm start (1..10 warm
iterations, no fallback), and reports the iterations each needs to reach 1e-4 m and the plain 10-iteration
residual ("not reached" when it does not within the sweep).
``cpu_baseline`` times the C++ oracle (scalar restatement of the reference algorithm) on a bounded sample: a
snapshot of the first worlds of the timed state, stepped on the host, one thread and then one thread per core of
the box's CPU share.
``learned`` times the same workload with the learned warm start (SolverXPBDWarmStart: prep launch,
network, warm launch with ``--warm-iterations`` iterations and the per-world residual check) and
reports its penetration residual beside the plain solver's at the same state.
``sdf`` times the triangle-mesh workload (1024 worlds x 8 convex-hull meshes with 64^3 SDFs, mesh-mesh and
mesh-plane contacts, the reference's contact reduction) with the same step, ``sdf_unique_hulls`` the same with a
seeded hull of its own for every body (8192 hulls and SDFs); ``uneven`` the uneven-worlds workload
(2048 worlds of 8-256 primitives plus a 6-link revolute chain each; its scene is built on the host in a separate
process while the other forms run); ``single_world`` the one-world 64-box case.
"""

from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

HBM_PEAK_GBS = 8000.0


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=240)
    p.add_argument("--warmup", type=int, default=20)
    p.add_argument("--settle", type=int, default=120,
                   help="untimed steps before warmup and timing (the pile settles; independent of --warmup)")
    p.add_argument("--iters-max", type=int, default=30, help="largest plain iteration count of the sweep")
    p.add_argument("--worlds", type=int, default=4096)
    p.add_argument("--iterations", type=int, default=10)
    p.add_argument("--substeps", type=int, default=4)
    p.add_argument("--seed", type=int, default=0)
    p.add_argument("--no-graph", action="store_true")
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--cpu-seconds", type=float, default=12.0)
    p.add_argument("--kernel-reps", type=int, default=20)
    p.add_argument("--no-learned", action="store_true", help="skip the learned warm-start form")
    p.add_argument("--warm-iterations", type=int, default=3)
    p.add_argument("--residual-tol", type=float, default=None,
                   help="per-world penetration residual above which the learned form falls back to a cold solve")
    p.add_argument("--no-sdf", action="store_true", help="skip the SDF (triangle-mesh) workload")
    p.add_argument("--no-sdf-unique", action="store_true",
                   help="skip the SDF workload's form with a unique hull per body")
    p.add_argument("--no-uneven", action="store_true",
                   help="skip the uneven-worlds workload (log-uniform 8-256 bodies per world + revolute chains)")
    p.add_argument("--no-single", action="store_true", help="skip the one-world 64-box case")
    p.add_argument("--uneven-worlds", type=int, default=2048)
    p.add_argument("--uneven-steps", type=int, default=60)
    p.add_argument("--uneven-settle", type=int, default=60)
    p.add_argument("--no-components", action="store_true",
                   help="skip the per-launch timings of contact matching, force reporting, standalone broad "
                        "phases and the SDF cache")
    p.add_argument("--sdf-worlds", type=int, default=1024)
    p.add_argument("--sdf-steps", type=int, default=60)
    p.add_argument("--sdf-settle", type=int, default=60)
    p.add_argument("--copy-probe-mib", type=int, default=0,
                   help="only run a device-to-device copy of this many MiB, 5 times (HBM counter calibration)")
    return p.parse_args()


def copy_probe(mib: int) -> None:
    """A plain copy of known size: FETCH_SIZE / WRITE_SIZE read on it give the counters' scale."""
    import torch

    n = mib * (1 << 20) // 4
    src = torch.arange(n, dtype=torch.float32, device="cuda:0")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(5):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes_read": n * 4, "copy_probe_bytes_written": n * 4, "launches": 5}), flush=True)


UNEVEN_BUILD = """
import sys, time
sys.path.insert(0, sys.argv[1])
from newton_custom_contact_amd import scenes
t0 = time.time()
m = scenes.uneven_worlds(int(sys.argv[3]), seed=int(sys.argv[4]), device="cpu",
                         progress=lambda k, n: print(f"[bench] uneven scene: {k}/{n} worlds built", file=sys.stderr,
                                                     flush=True))
m.save(sys.argv[2])
print(f"[bench] uneven scene built on the host in {time.time() - t0:.1f}s", file=sys.stderr, flush=True)
"""


def start_uneven_build(args):
    """The uneven scene takes minutes of host Python to build: start it in a child process before this process
    touches the GPU, so it overlaps the other forms."""
    if args.no_uneven:
        return None
    d = tempfile.mkdtemp(prefix="nmc_bench_")
    path = os.path.join(d, "uneven.npz")
    root = os.path.dirname(os.path.abspath(__file__))
    proc = subprocess.Popen([sys.executable, "-c", UNEVEN_BUILD, root, path, str(args.uneven_worlds), str(args.seed)])
    return proc, path, time.time()


def phase_times(pipe, solver, a, b, contacts, sdt, reps):
    """Average durations on the launch stream (HIP events around graph replays, as in the timed loop, so the
    GPU never waits on the host's launch rate): `reps` back-to-back collides, then `reps` back-to-back steps,
    at the given state."""
    import torch

    pipe.collide(a, contacts)
    torch.cuda.synchronize()
    graphs = []
    for work in (lambda: pipe.collide(a, contacts), lambda: solver.step(a, b, None, contacts, sdt)):
        side = torch.cuda.Stream()
        side.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(side):
            work()
        torch.cuda.current_stream().wait_stream(side)
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            for _ in range(reps):
                work()
        graphs.append(g)
    stream = torch.cuda.current_stream()
    out = []
    for g in graphs:
        g.replay()  # warm
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        g.replay()
        e1.record(stream)
        torch.cuda.synchronize()
        out.append(e0.elapsed_time(e1) / reps)
    del graphs
    return out[0], out[1]


def timed_form(args, model, pipe, solver, worlds, settle, warmup, steps, dev, tune=None):
    """Settle (untimed), warm up, then time `steps` graph replays of one step (4 substeps of collide + step).
    `tune(state_a, state_b, contacts, sdt)` (optional) runs on the settled state and may change the collide and
    step kernels' launch shapes; the graph is captured again after it.  N_c is counted over as many steps after
    the timed ones (a second graph with the counter), so the timed graph holds the step alone.  Returns (states,
    contacts, ms, N_c, t_collide_ms, t_xpbd_ms, tune result)."""
    import torch

    contacts = pipe.contacts()
    states = [model.state(), model.state()]
    sdt = (1.0 / 60.0) / args.substeps
    nc_acc = torch.zeros(1, dtype=torch.int64, device=dev)

    def step(count=False):
        for s in range(args.substeps):
            a, b = states[s % 2], states[(s + 1) % 2]
            pipe.collide(a, contacts)
            solver.step(a, b, None, contacts, sdt)
            if count:
                nc_acc.add_(contacts.rigid_contact_count)
        if args.substeps % 2:
            states[0].body_q.copy_(states[1].body_q)
            states[0].body_qd.copy_(states[1].body_qd)

    def capture(count=False):
        torch.cuda.synchronize()
        if args.no_graph:
            return None
        side = torch.cuda.Stream()
        side.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(side):
            step(count)
        torch.cuda.current_stream().wait_stream(side)
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            step(count)
        torch.cuda.synchronize()
        return g

    graph = capture()

    def run(n, g, count=False):
        for _ in range(n):
            if g is not None:
                g.replay()
            else:
                step(count)

    run(settle, graph)
    tuned = None
    if tune is not None:
        torch.cuda.synchronize()
        pipe.collide(states[0], contacts)
        tuned = tune(states[0], states[1], contacts, sdt)
        del graph
        graph = capture()
    run(warmup, graph)
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record()
    run(steps, graph)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1)
    # contacts per world and substep over as many further steps (the bench's own accounting stays out of the
    # timed graph)
    counting = capture(count=True)
    nc_acc.zero_()
    run(steps, counting, count=True)
    torch.cuda.synchronize()
    n_c = float(nc_acc.item()) / (steps * args.substeps * worlds)
    del counting
    t_col, t_xp = phase_times(pipe, solver, states[0], states[1], contacts, sdt, args.kernel_reps)
    del graph
    return states, contacts, ms, n_c, t_col, t_xp, tuned


COLLIDE_KERNELS = ("collide_segment", "route_scan_kernel", "narrow_kernel", "contact_count_kernel",
                   "contact_scan_kernel", "contact_compact_kernel")


def roofline_block(t_col, t_xp, worlds, b_collide, b_step, traffic=None, traffic_at=None, xpbd_kernel=True):
    """The dominant kernel priced with its share of the algorithmic bytes, the collide phase beside it.  On the
    primitive workloads that is the XPBD step (one launch per substep, the largest single kernel; the collide
    phase is six smaller ones); xpbd_kernel=False prices whichever phase takes longer (the mesh workload's
    collide phase, whose mesh narrow phase is its largest kernel)."""
    col = worlds * b_collide / (t_col / 1e3) / 1e9
    collide = {"kernels": list(COLLIDE_KERNELS), "ms": round(t_col, 4), "achieved": round(col, 2),
               "frac": round(col / HBM_PEAK_GBS, 5), "traffic": traffic(COLLIDE_KERNELS) if traffic else None}
    if xpbd_kernel or t_xp >= t_col:
        achieved = worlds * b_step / (t_xp / 1e3) / 1e9
        out = {"kernel": "xpbd_world_kernel", "launch_ms": round(t_xp, 4), "bound": "hbm",
               "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
               "frac": round(achieved / HBM_PEAK_GBS, 5),
               "traffic": traffic(("xpbd_world_kernel",)) if traffic else None}
    else:
        out = {"kernel": "collide phase: " + " + ".join(COLLIDE_KERNELS), "launch_ms": round(t_col, 4),
               "bound": "hbm", "achieved": collide["achieved"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
               "frac": collide["frac"], "traffic": collide["traffic"]}
    if traffic_at:
        out["traffic_at"] = traffic_at
    out["collide_phase"] = collide
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported (worlds shard across "
              "GPUs with no per-step exchange, so the multi-GPU figure is N x the one-GPU figure by construction)",
              file=sys.stderr)
        sys.exit(2)
    if args.copy_probe_mib > 0:
        copy_probe(args.copy_probe_mib)
        return
    uneven_job = start_uneven_build(args)
    import torch

    from newton_custom_contact_amd import CollisionPipeline, scenes
    from newton_custom_contact_amd.solvers import SolverXPBD

    dev = "cuda:0"
    t0 = time.time()
    model = scenes.mixed_worlds(args.worlds, seed=args.seed, device=dev)
    print(f"[bench] scene: {args.worlds} worlds, {model.body_count} bodies, {model.shape_count} shapes "
          f"({time.time() - t0:.1f}s)", file=sys.stderr, flush=True)
    pipe = CollisionPipeline(model, broad_phase="sap")
    solver = SolverXPBD(model, iterations=args.iterations)
    sdt = (1.0 / 60.0) / args.substeps
    from newton_custom_contact_amd import autotune

    tune = lambda a, b, c, sdt: autotune.tune(pipe, solver, a, b, c, sdt)  # noqa: E731
    states, contacts, ms, n_c, t_col_ms, t_xp_ms, tuned = timed_form(args, model, pipe, solver, args.worlds,
                                                                      args.settle, args.warmup, args.steps, dev,
                                                                      tune=tune)
    ms_per_step = ms / args.steps
    ws_per_s = args.worlds * args.steps / (ms / 1e3)
    nonfinite = int((~torch.isfinite(states[0].body_qd).all(dim=1)).sum())
    print(f"[bench] timed {args.steps} steps after {args.settle} settle + {args.warmup} warmup: {ms:.2f} ms, "
          f"N_c={n_c:.1f}, non-finite bodies {nonfinite}", file=sys.stderr, flush=True)

    N_b, N_s, N_j = 32, 33, 0
    b_collide = 28 * N_b + 64 * N_s + 76 * n_c + 4
    b_step = 76 * N_b + 96 * N_b + 16 * N_s + 76 * n_c + 52 * N_b + 120 * N_j
    b_ws = args.substeps * (b_collide + b_step)
    roof_ws = HBM_PEAK_GBS * 1e9 / b_ws
    roof = roofline_block(t_col_ms, t_xp_ms, args.worlds, b_collide, b_step,
                          traffic=lambda k: measured_traffic(args, k),
                          traffic_at="rocprofv3 FETCH_SIZE + WRITE_SIZE passes at the settled state (N_c 120.8), "
                                     "profiles/mixed4096_hbm_traffic.md")

    iterations = None
    if not args.no_learned:
        iterations = iterations_form(args, model, states[0], dev)

    learned = None
    if not args.no_learned:
        learned = learned_form(args, model, dev, ws_per_s)

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(model, states[0], args, sdt)

    single = None
    if not args.no_single:
        single = single_world_form(args, dev)

    sdf = None
    sdf_unique = None
    if not args.no_sdf:
        sdf = sdf_form(args, dev)
        if not args.no_sdf_unique:
            sdf_unique = sdf_form(args, dev, unique=True)

    uneven = None
    if uneven_job is not None:
        uneven = uneven_form(args, dev, uneven_job)

    components = None
    if not args.no_components:
        components = components_form(args, model, states[0], sdt, dev)

    line = {
        "metric": "world-steps/s @4096 worlds; iters to 1e-4 m pen residual, learned vs plain",
        "value": round(ws_per_s, 1),
        "unit": "world-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"mixed primitives: {args.worlds} worlds x 32 (11 spheres, 11 boxes, 10 capsules) on a "
                        f"ground plane, dt=1/60 s, {args.substeps} substeps, {args.iterations} XPBD iterations, "
                        f"{'hipGraph' if not args.no_graph else 'eager'} step, timed after {args.settle} untimed "
                        f"settle steps + {args.warmup} warmup steps",
            "worlds": args.worlds,
            "bodies_per_world": N_b,
            "settle_steps": args.settle,
            "mean_contacts_per_world_substep": round(n_c, 2),
            "non_finite_bodies_at_end": nonfinite,
            "bytes_per_world_step": round(b_ws, 1),
            "bytes_per_world_substep": {"collide": round(b_collide, 1), "step": round(b_step, 1)},
            "hbm_roofline_world_steps_per_s": round(roof_ws, 1),
            "fraction_of_step_roofline": round(ws_per_s / roof_ws, 5),
            "collide_ms_per_substep": round(t_col_ms, 4),
            "xpbd_ms_per_substep": round(t_xp_ms, 4),
            "launch_shapes": tuned,
        },
        "roofline": roof,
        "cpu_baseline": cpu,
        "iterations": iterations,
        "learned": learned,
        "single_world": single,
        "sdf": sdf,
        "sdf_unique_hulls": sdf_unique,
        "uneven": uneven,
        "components": components,
    }
    print(json.dumps(line), flush=True)


def iterations_form(args, model, state, dev):
    """Iterations to a penetration residual at the settled state.  Residual of one substep = mean over worlds
    of the world's largest remaining penetration max(0, -d) after the solve, on the same state and contacts for
    every iteration count: plain 1..iters_max iterations, learned warm start 1..10 warm iterations (no fallback,
    the trained network), and the warm start fed the exact total multipliers of the plain 10-iteration step with
    0..10 iterations after it (what a perfect network would reach)."""
    import torch

    from newton_custom_contact_amd import CollisionPipeline
    from newton_custom_contact_amd.learned import WEIGHTS_PATH, SolverXPBDWarmStart, WarmStartMLP

    if not os.path.exists(WEIGHTS_PATH):
        return None
    mlp = WarmStartMLP.load(WEIGHTS_PATH, device=dev)
    pipe = CollisionPipeline(model, broad_phase="sap", deterministic=True)
    kmax = max(args.iters_max, args.iterations)
    solver = SolverXPBDWarmStart(model, mlp, iterations=kmax, warm_iterations=3, residual_tol=float("inf"))
    sdt = (1.0 / 60.0) / args.substeps
    W = model.world_count
    c_prev, c = pipe.contacts(), pipe.contacts()
    tmp = model.state()
    # the previous substep's contacts and multipliers for the "previous multiplier" feature
    a = model.state()
    a.body_q.copy_(state.body_q)
    a.body_qd.copy_(state.body_qd)
    pipe.collide(a, c_prev)
    solver.log_step(a, tmp, None, c_prev, sdt, iterations=args.iterations)
    a.body_q.copy_(tmp.body_q)
    a.body_qd.copy_(tmp.body_qd)
    pipe.collide(a, c)
    plain = []
    for k in range(1, kmax + 1):
        solver.log_step(a, tmp, None, c, sdt, iterations=k, remember=False)
        plain.append(float(solver.buffers["residual"][:W].mean()))
    learned = []
    for k in range(1, min(10, kmax) + 1):
        solver.step(a, tmp, None, c, sdt, warm_iterations=k, residual_tol=float("inf"), remember=False)
        learned.append(float(solver.buffers["residual"][:W].mean()))
    # the method's ceiling: the warm start fed the exact total multipliers of the plain 10-iteration step
    solver.log_step(a, tmp, None, c, sdt, iterations=args.iterations, remember=False)
    exact = solver.buffers["lambda_sum"].clone()
    ceiling = []
    for k in range(0, min(10, kmax) + 1):
        solver.step(a, tmp, None, c, sdt, warm_iterations=k, residual_tol=float("inf"), remember=False,
                    multipliers=exact)
        ceiling.append(float(solver.buffers["residual"][:W].mean()))
    torch.cuda.synchronize()
    p10 = plain[args.iterations - 1]

    def first_below(seq, thr):
        for k, r in enumerate(seq, 1):
            if r <= thr:
                return k
        return f"not reached within {len(seq)} iterations (smallest {min(seq):.3e} m)"

    out = {
        "state": f"settled plain workload, {W} worlds, one substep of dt {sdt:.6f} s from the same state and contacts",
        "residual_m_by_iterations": {"plain": plain, "learned_warm_start": learned,
                                     "exact_multiplier_warm_start_from_0": ceiling},
        "plain_10_iteration_residual_m": p10,
        "iters_to_1e-4_m": {"plain": first_below(plain, 1e-4), "learned": first_below(learned, 1e-4),
                            "exact_multipliers": first_below(ceiling[1:], 1e-4)},
        "iters_to_plain10_residual": {"plain": first_below(plain, p10), "learned": first_below(learned, p10),
                                      "exact_multipliers": first_below(ceiling[1:], p10)},
    }
    print(f"[bench] iterations: {json.dumps(out)}", file=sys.stderr, flush=True)
    return out


def single_world_form(args, dev):
    """The one-world case: 64 boxes (edge 0.1-0.5 m) dropped on the ground, 120 steps timed after settling (a
    latency-bound launch sequence; the throughput metric is for the batched forms)."""
    from newton_custom_contact_amd import CollisionPipeline, scenes
    from newton_custom_contact_amd.solvers import SolverXPBD

    model = scenes.box_pile(64, seed=args.seed, device=dev)
    pipe = CollisionPipeline(model, broad_phase="sap")
    solver = SolverXPBD(model, iterations=args.iterations)
    steps = 120
    _st, _c, ms, n_c, t_col, t_xp, _ = timed_form(args, model, pipe, solver, 1, args.settle, 5, steps, dev)
    return {"value": round(steps / (ms / 1e3), 1), "unit": "world-steps/s", "ms_per_step": round(ms / steps, 4),
            "workload": "1 world x 64 boxes (edge 0.1-0.5 m) on a ground plane, dt=1/60 s, 4 substeps, "
                        f"{args.iterations} iterations, hipGraph step",
            "mean_contacts_per_world_substep": round(n_c, 2),
            "collide_ms_per_substep": round(t_col, 4), "xpbd_ms_per_substep": round(t_xp, 4)}


def learned_form(args, model, dev, plain_ws_per_s):
    """The same workload with the learned warm start (prep launch, network, warm launch with
    ``warm_iterations`` iterations and the per-world residual check), hipGraph-captured like the plain step;
    plus the residual comparison: mean over worlds of the largest penetration after each step."""
    import torch

    from newton_custom_contact_amd import CollisionPipeline
    from newton_custom_contact_amd.learned import WEIGHTS_PATH, SolverXPBDWarmStart, WarmStartMLP

    if not os.path.exists(WEIGHTS_PATH):
        return None
    mlp = WarmStartMLP.load(WEIGHTS_PATH, device=dev)
    pipe = CollisionPipeline(model, broad_phase="sap", deterministic=True)
    tol = args.residual_tol
    solver = SolverXPBDWarmStart(model, mlp, iterations=args.iterations, warm_iterations=args.warm_iterations,
                                 residual_tol=float("inf") if tol is None else tol)
    # two Contacts buffers alternate, so the previous substep's contacts stay readable for the matching
    cbufs = [pipe.contacts(), pipe.contacts()]
    states = [model.state(), model.state()]
    sdt = (1.0 / 60.0) / args.substeps
    W = model.world_count
    res_acc = torch.zeros(1, dtype=torch.float64, device=dev)
    fb_acc = torch.zeros(1, dtype=torch.float64, device=dev)
    stats = [False]

    def step():
        for s in range(args.substeps):
            a, b = states[s % 2], states[(s + 1) % 2]
            contacts = cbufs[s % 2]
            pipe.collide(a, contacts)
            solver.step(a, b, None, contacts, sdt)
            if stats[0]:
                buf = solver.buffers
                res_acc.add_(buf["residual"][:W].mean())
                fb_acc.add_(buf["fallback"][:W].sum())
        if args.substeps % 2:
            states[0].body_q.copy_(states[1].body_q)
            states[0].body_qd.copy_(states[1].body_qd)

    def capture():
        side = torch.cuda.Stream()
        side.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(side):
            step()
        torch.cuda.current_stream().wait_stream(side)
        torch.cuda.synchronize()
        if args.no_graph:
            return None
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            step()
        torch.cuda.synchronize()
        return g

    graph = capture()
    for _ in range(args.settle):
        graph.replay() if graph is not None else step()
    torch.cuda.synchronize()
    tol_rule = "given (--residual-tol)"
    if tol is None:
        # the fallback tolerance tied to the plain solver: twice the mean over worlds of the plain 10-iteration
        # residual at the settled state (the learned run's own state and contacts)
        a, tmp = states[0], model.state()
        pipe.collide(a, cbufs[0])
        solver.log_step(a, tmp, None, cbufs[0], sdt, remember=False)
        tol = 2.0 * float(solver.buffers["residual"][:W].mean())
        solver.residual_tol = tol
        tol_rule = "2 x the plain 10-iteration residual at the settled state"
        del graph
        graph = capture()
    for _ in range(args.warmup):
        graph.replay() if graph is not None else step()
    torch.cuda.synchronize()
    res_acc.zero_()
    fb_acc.zero_()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(args.steps):
        graph.replay() if graph is not None else step()
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1)
    ws = args.worlds * args.steps / (ms / 1e3)
    # residual and fallback statistics over further (untimed, eager) steps of the same run
    stats[0] = True
    n_stat = 30
    for _ in range(n_stat):
        step()
    torch.cuda.synchronize()
    n_sub = n_stat * args.substeps
    warm_res = float(res_acc.item()) / n_sub
    fb_frac = float(fb_acc.item()) / (n_sub * W)
    # plain residual at the learned run's end state: the plain step with 10 and with warm_iterations iterations
    a, tmp = states[0], states[1]
    contacts = cbufs[0]
    pipe.collide(a, contacts)
    solver.log_step(a, tmp, None, contacts, sdt, remember=False)
    plain10 = float(solver.buffers["residual"][:W].mean())
    solver.log_step(a, tmp, None, contacts, sdt, iterations=args.warm_iterations, remember=False)
    plain_k = float(solver.buffers["residual"][:W].mean())
    solver.step(a, tmp, None, contacts, sdt, remember=False)
    warm_here = float(solver.buffers["residual"][:W].mean())
    print(f"[bench] learned: {ws:.1f} world-steps/s, residual {warm_res:.3e}, fallback {fb_frac:.3f}",
          file=sys.stderr, flush=True)
    return {
        "value": round(ws, 1),
        "unit": "world-steps/s",
        "ms_per_step": round(ms / args.steps, 4),
        "speedup_vs_plain": round(ws / plain_ws_per_s, 3),
        "warm_iterations": args.warm_iterations,
        "fallback_iterations": args.iterations,
        "residual_tol_m": tol,
        "residual_tol_rule": tol_rule,
        "fallback_world_fraction": round(fb_frac, 4),
        "mean_world_residual_m": warm_res,
        "end_state_residual_m": {"learned": warm_here, f"plain_{args.warm_iterations}": plain_k,
                                 f"plain_{args.iterations}": plain10},
        "network": "16-128-128-1 SiLU, BF16 MFMA / FP32 accumulate, newton_custom_contact_amd/data/warmstart_mlp.npz",
    }


def components_form(args, model, state, sdt, dev):
    """Per-launch cost of the path's other entry points on the plain workload's end state (HIP events on the
    launch stream, ``--kernel-reps`` back-to-back launches): contact matching (collide with matching minus the
    deterministic collide), the reporting XPBD step and ``update_contacts``, the standalone broad phases over
    the end state's bounding boxes; and the SDF cache (host wall time of a 64^3 cook + write and of a load).
    Runs last: it requests ``Contacts.force`` and ``State.body_parent_f`` on the model."""
    import tempfile

    import torch

    from newton_custom_contact_amd import CollisionPipeline, scenes
    from newton_custom_contact_amd.geometry import BroadPhaseAllPairs, BroadPhaseExplicit, BroadPhaseSAP
    from newton_custom_contact_amd.solvers import SolverXPBD

    stream = torch.cuda.current_stream()
    reps = args.kernel_reps

    def timed_us(fn):
        fn()
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(reps):
            fn()
        e1.record(stream)
        torch.cuda.synchronize()
        return round(e0.elapsed_time(e1) / reps * 1e3, 1)

    out = {"workload": f"end state of the plain workload ({args.worlds} worlds x 32 bodies)"}
    # contact matching
    det = CollisionPipeline(model, broad_phase="sap", deterministic=True)
    cd = det.contacts()
    t_det = timed_us(lambda: det.collide(state, cd))
    match = {"deterministic_collide_us": t_det}
    for mode in ("latest", "sticky"):
        mp = CollisionPipeline(model, broad_phase="sap", contact_matching=mode, contact_report=True)
        cm = mp.contacts()
        match[f"{mode}_with_report_us"] = round(timed_us(lambda: mp.collide(state, cm)) - t_det, 1)
    out["contact_matching_us_per_substep"] = match
    # reported forces
    solver = SolverXPBD(model, iterations=args.iterations)
    c_plain = det.contacts()
    det.collide(state, c_plain)
    s_out = model.state()
    t_plain = timed_us(lambda: solver.step(state, s_out, None, c_plain, sdt))
    model.request_contact_attributes("force")
    model.request_state_attributes("body_parent_f")
    c_rep = det.contacts()
    det.collide(state, c_rep)
    s_rep = model.state()
    t_rep = timed_us(lambda: solver.step(state, s_rep, None, c_rep, sdt))
    t_upd = timed_us(lambda: solver.update_contacts(c_rep))
    out["force_reporting_us_per_substep"] = {"plain_step": t_plain, "reporting_step": t_rep,
                                             "update_contacts": t_upd}
    # standalone broad phases over the end state's bounding spheres' boxes
    h = model.host
    sb = torch.from_numpy(h["shape_body"].astype(np.int64)).to(dev)
    xf = torch.from_numpy(np.ascontiguousarray(h["shape_transform"][:, :3], np.float32)).to(dev)
    pos = torch.where((sb >= 0)[:, None], state.body_q[sb.clamp(min=0), :3], xf)
    r = torch.from_numpy(np.ascontiguousarray(h["shape_collision_radius"], np.float32)).to(dev)[:, None]
    lower, upper = (pos - r).contiguous(), (pos + r).contiguous()
    gap = model.shape_gap
    group, world = model.shape_collision_group, model.shape_world
    fpairs = model.shape_filter_pairs if len(h["shape_filter_pairs"]) else None
    cap = 4 << 20
    pairs = torch.zeros((cap, 2), dtype=torch.int32, device=dev)
    cnt = torch.zeros(1, dtype=torch.int32, device=dev)
    bp = {}
    for name, cls in (("sap", BroadPhaseSAP), ("nxn", BroadPhaseAllPairs)):
        inst = cls(h["shape_world"], h["shape_flags"])
        t = timed_us(lambda: inst.launch(lower, upper, gap, group, world, model.shape_count, pairs, cnt,
                                         filter_pairs=fpairs))
        bp[name] = {"us": t, "pairs": int(cnt.item())}
    ex = BroadPhaseExplicit()
    sp = model.shape_contact_pairs
    t = timed_us(lambda: ex.launch(lower, upper, gap, sp, sp.shape[0], pairs, cnt))
    bp["explicit"] = {"us": t, "pairs": int(cnt.item()), "input_pairs": int(sp.shape[0])}
    out["standalone_broad_phase_us"] = bp
    # SDF cache: cook + publish, then load, of one 64^3 grid
    with tempfile.TemporaryDirectory() as d:
        m1 = scenes.mesh_library(1, seed=args.seed)[0]
        m2 = scenes.mesh_library(1, seed=args.seed)[0]
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        m1.build_sdf(max_resolution=64, device=dev, cache_dir=d)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        m2.build_sdf(max_resolution=64, device=dev, cache_dir=d)
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        same = all(bool(torch.equal(a, b)) for a, b in ((m1.sdf.coarse, m2.sdf.coarse), (m1.sdf.slots, m2.sdf.slots),
                                                         (m1.sdf.texels, m2.sdf.texels)))
    out["sdf_cache_ms"] = {"cook_and_write": round((t1 - t0) * 1e3, 2), "load": round((t2 - t1) * 1e3, 2),
                           "grid": "texture SDF, 64 cells on the longest axis, 8^3 subgrids, uint16 texels",
                           "load_equals_cook": same}
    print(f"[bench] components: {json.dumps(out)}", file=sys.stderr, flush=True)
    return out


def uneven_form(args, dev, job):
    """Uneven worlds: ``--uneven-worlds`` worlds of log-uniform 8-256 primitives (spheres, boxes, capsules) plus
    a 6-link revolute capsule chain with +-90 degree limits each, on a ground plane; 4 substeps, 10 iterations,
    the collide and step kernels' launch shapes autotuned on the settled state, hipGraph step, timed after
    ``--uneven-settle`` steps.  The scene was built on the host by a child process started at the beginning."""
    from newton_custom_contact_amd import CollisionPipeline
    from newton_custom_contact_amd import autotune
    from newton_custom_contact_amd.model import Model
    from newton_custom_contact_amd.solvers import SolverXPBD

    proc, path, t_start = job
    t0 = time.time()
    rc = proc.wait()
    if rc != 0 or not os.path.exists(path):
        print(f"[bench] uneven scene build failed (exit {rc})", file=sys.stderr, flush=True)
        return None
    model = Model.load(path, dev)
    try:
        os.remove(path)
        os.rmdir(os.path.dirname(path))
    except OSError:
        pass
    W = args.uneven_worlds
    print(f"[bench] uneven scene: {W} worlds, {model.body_count} bodies, {model.joint_count} joints (host build "
          f"started {t0 - t_start:.1f}s before this form; waited {time.time() - t0:.1f}s for it and the upload)",
          file=sys.stderr, flush=True)
    pipe = CollisionPipeline(model, broad_phase="sap")
    solver = SolverXPBD(model, iterations=args.iterations)
    tune = lambda a, b, c, sdt: autotune.tune(pipe, solver, a, b, c, sdt)  # noqa: E731
    states, contacts, ms, n_c, t_col, t_xp, tuned = timed_form(args, model, pipe, solver, W, args.uneven_settle,
                                                                3, args.uneven_steps, dev, tune=tune)
    ws = W * args.uneven_steps / (ms / 1e3)
    h = model.host
    n_b = model.body_count / W
    n_s = (model.shape_count - int((h["shape_world"] == -1).sum())) / W + int((h["shape_world"] == -1).sum())
    n_j = model.joint_count / W
    revolute = int((h["joint_type"] == 1).sum()) / W
    b_collide = 28 * n_b + 64 * n_s + 76 * n_c + 4
    b_step = 76 * n_b + 96 * n_b + 16 * n_s + 76 * n_c + 52 * n_b + 120 * revolute
    b_ws = args.substeps * (b_collide + b_step)
    roof_ws = HBM_PEAK_GBS * 1e9 / b_ws
    oc = pipe.overflow_counters()
    print(f"[bench] uneven: {ws:.1f} world-steps/s, N_c={n_c:.1f}", file=sys.stderr, flush=True)
    return {
        "value": round(ws, 1),
        "unit": "world-steps/s",
        "ms_per_step": round(ms / args.uneven_steps, 4),
        "workload": f"{W} worlds of log-uniform 8-256 primitives + a 6-link revolute chain each "
                    f"({model.body_count} bodies, {model.joint_count} joints) on a ground plane, dt=1/60 s, "
                    f"{args.substeps} substeps, {args.iterations} iterations, timed after {args.uneven_settle} "
                    "untimed settle steps",
        "mean_bodies_per_world": round(n_b, 2),
        "mean_revolute_joints_per_world": round(revolute, 2),
        "mean_contacts_per_world_substep": round(n_c, 2),
        "bytes_per_world_step": round(b_ws, 1),
        "hbm_roofline_world_steps_per_s": round(roof_ws, 1),
        "fraction_of_step_roofline": round(ws / roof_ws, 5),
        "collide_ms_per_substep": round(t_col, 4),
        "xpbd_ms_per_substep": round(t_xp, 4),
        "roofline": roofline_block(t_col, t_xp, W, b_collide, b_step),
        "overflow": oc,
        "launch_shapes": tuned,
    }


def sdf_form(args, dev, unique: bool = False):
    """The triangle-mesh workload: ``--sdf-worlds`` worlds x 8 convex-hull meshes (64-256 points on a sphere
    of radius 0.1-0.3 m, 64^3 FP32 texture SDF each) dropped on a ground plane, mesh-mesh and mesh-plane contacts
    with the reference's contact reduction (at most 240 per pair), 4 substeps, 10 iterations, hipGraph step; the
    collide and step kernels' launch shapes autotuned on the settled state.  The hulls come from a seeded
    library of 64, or with ``unique`` every body has its own seeded hull (8192 hulls and SDFs for 1024 worlds).
    Timed after ``--sdf-settle`` untimed steps, when the hulls have landed and pile up."""
    import torch

    from newton_custom_contact_amd import CollisionPipeline, scenes
    from newton_custom_contact_amd import autotune
    from newton_custom_contact_amd.solvers import SolverXPBD

    t0 = time.time()
    if unique:
        lib = scenes.mesh_library(args.sdf_worlds * 8, seed=args.seed)
    else:
        lib = scenes.mesh_library(64, seed=args.seed)
    t_hull = time.time() - t0
    from newton_custom_contact_amd.mesh import cook_texture_sdfs

    cook_texture_sdfs(lib, dev, max_resolution=64, texture_format="float32")
    torch.cuda.synchronize()
    t_cook = time.time() - t0 - t_hull
    model = scenes.mesh_worlds(args.sdf_worlds, seed=args.seed, device=dev, library=lib, unique=unique)
    sdf_bytes = sum(m.sdf.coarse.numel() * 4 + m.sdf.slots.numel() * 4 + m.sdf.texels.numel() for m in lib)
    print(f"[bench] sdf scene: {args.sdf_worlds} worlds, {model.body_count} mesh bodies, {len(lib)} hulls "
          f"(hulls {t_hull:.1f}s, SDF cook {t_cook:.1f}s, {sdf_bytes / 2**30:.2f} GiB of SDF textures, total "
          f"{time.time() - t0:.1f}s)", file=sys.stderr, flush=True)
    pipe = CollisionPipeline(model, broad_phase="sap")
    solver = SolverXPBD(model, iterations=args.iterations)
    tune = lambda a, b, c, sdt: autotune.tune(pipe, solver, a, b, c, sdt)  # noqa: E731
    states, contacts, ms, n_c, t_col, t_xp, tuned = timed_form(args, model, pipe, solver, args.sdf_worlds,
                                                                args.sdf_settle, 5, args.sdf_steps, dev, tune=tune)
    ws = args.sdf_worlds * args.sdf_steps / (ms / 1e3)
    n_b, n_s = 8, 9
    b_collide = 28 * n_b + 64 * n_s + 76 * n_c + 4
    b_step = 76 * n_b + 96 * n_b + 16 * n_s + 76 * n_c + 52 * n_b
    b_ws = args.substeps * (b_collide + b_step)
    roof_ws = HBM_PEAK_GBS * 1e9 / b_ws
    oc = pipe.overflow_counters()
    print(f"[bench] sdf{' unique hulls' if unique else ''}: {ws:.1f} world-steps/s, N_c={n_c:.1f}", file=sys.stderr,
          flush=True)
    hulls = f"{len(lib)} unique hulls, one per body" if unique else "64-hull library"
    return {
        "value": round(ws, 1),
        "unit": "world-steps/s",
        "ms_per_step": round(ms / args.sdf_steps, 4),
        "workload": f"{args.sdf_worlds} worlds x 8 convex-hull meshes ({hulls}, 64^3 fp32 texture SDF, "
                    f"{sdf_bytes / 2**30:.2f} GiB of SDFs) on a ground plane, reference contact reduction, dt=1/60 s, "
                    f"{args.substeps} substeps, {args.iterations} iterations, timed after {args.sdf_settle} untimed "
                    f"settle steps",
        "hulls": len(lib),
        "sdf_texture_bytes": int(sdf_bytes),
        "mean_contacts_per_world_substep": round(n_c, 2),
        "bytes_per_world_step": round(b_ws, 1),
        "hbm_roofline_world_steps_per_s": round(roof_ws, 1),
        "fraction_of_step_roofline": round(ws / roof_ws, 5),
        "collide_ms_per_substep": round(t_col, 4),
        "xpbd_ms_per_substep": round(t_xp, 4),
        "roofline": roofline_block(t_col, t_xp, args.sdf_worlds, b_collide, b_step, xpbd_kernel=False),
        "overflow": oc,
        "launch_shapes": tuned,
        "lowest_body_z_m": round(float(states[0].body_q[:, 2].min()), 4),
    }


def measured_traffic(args, kernels):
    """HBM bytes per launch of the dominant kernel(s) from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE
    passes (copy-calibrated), valid for the default workload only; None otherwise."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "mixed4096_hbm_traffic.json")
    if not (args.worlds == 4096 and args.iterations == 10 and args.substeps == 4 and os.path.exists(path)):
        return None
    with open(path) as f:
        prof = json.load(f)["kernels"]
    total = 0
    for k in kernels:
        hit = [v for name, v in prof.items() if name.startswith(k)]
        if not hit:
            return None
        total += hit[0]["traffic_bytes"]
    return total


def cpu_baseline(model, state, args, sdt):
    """Oracle (C++ scalar restatement) on snapshots of the timed state: one thread, then one thread per host
    core over disjoint world subsets (ctypes releases the GIL, the oracle keeps no global state)."""
    import threading

    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    from newton_custom_contact_amd.model import Model
    from oracle import Oracle

    n_w, nb = 32, 32
    q_all = state.body_q.cpu().numpy()
    qd_all = state.body_qd.cpu().numpy()
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    # one GPU's share of the host is 16 cores, whatever the machine has
    threads = max(1, min(16, cores, args.worlds // n_w))
    model_name = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    model_name = ln.split(":", 1)[1].strip()
                    break
    except OSError:
        pass

    def make(k):
        sl = slice(k * n_w * nb, (k + 1) * n_w * nb)
        sub = Model._replicate(model.template, n_w, q_all[sl].reshape(n_w, nb, 7), "cpu")
        return Oracle(sub), sub.host["body_q"].copy(), qd_all[sl].copy()

    def run(n_thr, seconds):
        jobs = [make(k) for k in range(n_thr)]
        steps = [0] * n_thr
        go = threading.Barrier(n_thr + 1)
        stop = [False]

        def work(k):
            o, qq, qd = jobs[k]
            f = np.zeros_like(qd)
            go.wait()
            while not stop[0]:
                for _ in range(args.substeps):
                    c = o.collide(qq)
                    qq, qd = o.step(qq, qd, f, c, sdt, iterations=args.iterations)
                steps[k] += 1

        ts = [threading.Thread(target=work, args=(k,)) for k in range(n_thr)]
        for t in ts:
            t.start()
        go.wait()
        t0 = time.perf_counter()
        time.sleep(seconds)
        stop[0] = True
        for t in ts:
            t.join()
        el = time.perf_counter() - t0
        return n_w * sum(steps) / el, sum(steps), el

    v1, s1, e1 = run(1, args.cpu_seconds / 2)
    vn, sn, en = run(threads, args.cpu_seconds / 2)
    return {
        "value": round(vn, 2),
        "unit": "world-steps/s",
        "cores": threads,
        "host": {"cpu_model": model_name, "logical_cpus": os.cpu_count(), "usable_cpus": cores,
                 "threads_used": threads, "note": "one GPU's share of the host is 16 cores"},
        "kind": "port",
        "sample": f"{threads} threads x {n_w} worlds each from the GPU run's timed state, C++ oracle, "
                  f"{sn} world-batch steps in {en:.1f} s; one thread: {v1:.2f} world-steps/s "
                  f"({s1} steps of {n_w} worlds in {e1:.1f} s)",
        "single_thread_value": round(v1, 2),
    }


if __name__ == "__main__":
    main()
