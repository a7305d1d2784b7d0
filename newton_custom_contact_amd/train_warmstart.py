"""Train and evaluate the warm-start network on the GPU.

    python -m newton_custom_contact_amd.train_warmstart --out newton_custom_contact_amd/data/warmstart_mlp.npz

1. Log: plain XPBD steps (10 iterations) of seeded mixed-primitive scenes (the 4096-world workload's
   generator, training seeds), recording every contact's 16 inputs and total normal multiplier (contacts
   with a non-finite input or multiplier, from a world whose body diverged, are left out).
2. Fit: Adam on the library's forward/backward kernels.
3. Evaluate on held-out seeds, from the states of a plain trajectory: the per-world penetration
   residual (mean over worlds of the largest max(0, -d)) after k plain iterations (k = 1..10) and after
   the warm start + k iterations (k = 0..4, no fallback), and the iterations each needs to reach the
   plain 10-iteration residual.  Prints one JSON line with the curves.
"""

from __future__ import annotations

import argparse
import json
import sys
import time

import torch


def log_contacts(scene_seed: int, worlds: int, steps: int, skip: int, every: int, dev: str, mlp, out: list):
    from . import CollisionPipeline, scenes
    from .learned import SolverXPBDWarmStart

    model = scenes.mixed_worlds(worlds, seed=scene_seed, device=dev)
    pipe = CollisionPipeline(model, broad_phase="sap", deterministic=True)
    solver = SolverXPBDWarmStart(model, mlp, iterations=10)
    c = pipe.contacts()
    s0, s1 = model.state(), model.state()
    sdt = 1.0 / 240.0
    n_rows = 0
    for i in range(steps * 4):
        pipe.collide(s0, c)
        solver.log_step(s0, s1, None, c, sdt)
        if i >= skip * 4 and (i % every) == 0:
            n = min(int(c.rigid_contact_count.item()), c.rigid_contact_max)
            b = solver.buffers
            f, lam = b["feat"][:n], b["lambda_sum"][:n]
            # contacts of a world whose thin box diverged (non-finite state) carry no training signal
            ok = torch.isfinite(f).all(dim=1) & torch.isfinite(lam)
            out.append((f[ok].clone(), lam[ok].clone()))
            n_rows += int(ok.sum())
        s0, s1 = s1, s0
    return n_rows


def _finite_mean(r: torch.Tensor) -> float:
    """Mean of the per-world residuals over the worlds whose state is finite."""
    return float(r[torch.isfinite(r)].mean())


def evaluate(mlp, seeds, worlds: int, steps: int, every: int, dev: str) -> dict:
    from . import CollisionPipeline, scenes
    from .learned import SolverXPBDWarmStart

    plain_k = list(range(1, 11))
    warm_k = [0, 1, 2, 3, 4]
    acc_p = {k: [] for k in plain_k}
    acc_w = {k: [] for k in warm_k}
    acc_wd = []
    fallback = []
    for seed in seeds:
        model = scenes.mixed_worlds(worlds, seed=seed, device=dev)
        pipe = CollisionPipeline(model, broad_phase="sap", deterministic=True)
        solver = SolverXPBDWarmStart(model, mlp, iterations=10, warm_iterations=3)
        c = pipe.contacts()
        s0, s1, tmp = model.state(), model.state(), model.state()
        sdt = 1.0 / 240.0
        W = model.world_count
        for i in range(steps * 4):
            pipe.collide(s0, c)
            if i % every == every - 1:
                for k in plain_k:
                    solver.log_step(s0, tmp, None, c, sdt, iterations=k, remember=False)
                    acc_p[k].append(_finite_mean(solver.buffers["residual"][:W]))
                for k in warm_k:
                    solver.step(s0, tmp, None, c, sdt, warm_iterations=k, residual_tol=1e30, remember=False)
                    acc_w[k].append(_finite_mean(solver.buffers["residual"][:W]))
                solver.step(s0, tmp, None, c, sdt, remember=False)
                acc_wd.append(_finite_mean(solver.buffers["residual"][:W]))
                fallback.append(float(solver.buffers["fallback"][:W].float().mean()))
            # the trajectory itself: the plain 10-iteration step (which also keeps the matching history)
            solver.log_step(s0, s1, None, c, sdt)
            s0, s1 = s1, s0
    plain = {k: sum(v) / len(v) for k, v in acc_p.items()}
    warm = {k: sum(v) / len(v) for k, v in acc_w.items()}
    ref = plain[10]
    it_plain = next((k for k in plain_k if plain[k] <= ref), 10)
    it_warm = next((k for k in warm_k if warm[k] <= ref), None)
    tgt = 1e-4
    return {"plain_residual": plain, "warm_residual": warm, "warm3_with_fallback_residual": sum(acc_wd) / len(acc_wd),
            "fallback_fraction": sum(fallback) / len(fallback), "plain10_residual": ref,
            "iters_to_plain10_residual": {"plain": it_plain, "learned": it_warm},
            "iters_to_1e-4_m": {"plain": next((k for k in plain_k if plain[k] <= tgt), None),
                                "learned": next((k for k in warm_k if warm[k] <= tgt), None)}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=None)
    ap.add_argument("--worlds", type=int, default=1024)
    ap.add_argument("--train-seeds", type=int, nargs="+", default=[101, 102, 103, 104, 105, 106, 107, 108])
    ap.add_argument("--eval-seeds", type=int, nargs="+", default=[201, 202])
    ap.add_argument("--steps", type=int, default=120)
    ap.add_argument("--skip", type=int, default=4)
    ap.add_argument("--every", type=int, default=8)
    ap.add_argument("--epochs", type=int, default=8)
    ap.add_argument("--batch", type=int, default=65536)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--eval-worlds", type=int, default=512)
    ap.add_argument("--eval-steps", type=int, default=120)
    ap.add_argument("--eval-every", type=int, default=40)
    ap.add_argument("--eval-only", default=None, help="weights file to evaluate instead of training")
    a = ap.parse_args()
    dev = "cuda:0"
    from .learned import WarmStartMLP

    t0 = time.time()
    if a.eval_only:
        mlp = WarmStartMLP.load(a.eval_only, device=dev)
        hist = []
        n_rows = 0
    else:
        mlp = WarmStartMLP(device=dev, seed=0)
        rows = []
        n_rows = 0
        for s in a.train_seeds:
            n_rows += log_contacts(s, a.worlds, a.steps, a.skip, a.every, dev, mlp, rows)
            print(f"[train] logged seed {s}: {n_rows} contacts ({time.time() - t0:.1f}s)", file=sys.stderr, flush=True)
        feat = torch.cat([r[0] for r in rows])
        lam = torch.cat([r[1] for r in rows])
        del rows
        mlp.set_normalization(feat[: min(len(feat), 1 << 20)], lam[: min(len(lam), 1 << 20)])
        hist = mlp.fit(feat, lam, epochs=a.epochs, batch=a.batch, lr=a.lr,
                       log=lambda m: print(f"[train] {m} ({time.time() - t0:.1f}s)", file=sys.stderr, flush=True))
        if a.out:
            mlp.save(a.out)
        del feat, lam
    ev = evaluate(mlp, a.eval_seeds, a.eval_worlds, a.eval_steps, a.eval_every, dev)
    print(json.dumps({"contacts_logged": n_rows, "loss_per_epoch": hist, "evaluation": ev,
                      "seconds": round(time.time() - t0, 1)}), flush=True)


if __name__ == "__main__":
    main()
