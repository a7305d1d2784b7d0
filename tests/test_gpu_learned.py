"""GPU learned warm start (nmc_xpbd_step_learn, nmc_mlp_forward, nmc_mlp_train_grad) vs the oracle.

* logging step: network inputs (with the frame-to-frame matched previous multiplier), total normal
  multipliers, per-world residual and the state are bit-identical to the restatement; the state equals
  the plain step;
* warm step from a given lambda0: state, multipliers, residual and fallback flags bit-identical;
  lambda0 = 0 with warm_iterations = iterations is the plain step, also with the restitution pass on;
* network: the MFMA forward within 1e-2 of the output scale of the restatement (BF16 operands, FP32
  accumulation in a different order), gradients within 2e-2 relative (Frobenius) of the restatement;
* the committed trained network: warm start + k iterations leaves less penetration than k plain
  iterations, and at most as much as k + 1 plain iterations (k = 1..4).
"""

from __future__ import annotations

import os

import numpy as np
import pytest
import torch

import oracle
from newton_custom_contact_amd import CollisionPipeline, scenes
from newton_custom_contact_amd.learned import WEIGHTS_PATH, SolverXPBDWarmStart, WarmStartMLP
from newton_custom_contact_amd.solvers import SolverXPBD
from newton_custom_contact_amd._native import NMC_STEP_PREP, NMC_STEP_WARM
from oracle import Oracle

pytestmark = pytest.mark.gpu

DT = 1.0 / 240.0


def _settled(dev, worlds=8, seed=3, substeps=240):
    model = scenes.mixed_worlds(worlds, seed=seed, device=dev)
    o = Oracle(model)
    h = model.host
    q, qd = h["body_q"].copy(), h["body_qd"].copy()
    f = np.zeros_like(qd)
    for _ in range(substeps):
        c = o.collide(q)
        q, qd = o.step(q, qd, f, c, DT, iterations=10)
    return model, o, q, qd


def _state(model, q, qd):
    s = model.state()
    s.body_q.copy_(torch.from_numpy(np.ascontiguousarray(q, np.float32)))
    s.body_qd.copy_(torch.from_numpy(np.ascontiguousarray(qd, np.float32)))
    return s


def test_log_step_matches_oracle(cuda_device):
    print("[learned] log step", flush=True)
    model, o, q, qd = _settled(cuda_device)
    f = np.zeros_like(qd)
    pipe = CollisionPipeline(model, broad_phase="sap", deterministic=True)
    solver = SolverXPBDWarmStart(model, WarmStartMLP(device=cuda_device), iterations=10,
                                 reference_arithmetic=True)
    c, c_next = pipe.contacts(), pipe.contacts()  # alternating buffers keep step A's contacts for matching
    # step A (history), then step B with the matched previous multipliers
    sa, sb, sc = _state(model, q, qd), model.state(), model.state()
    pipe.collide(sa, c)
    solver.log_step(sa, sb, None, c, DT)
    torch.cuda.synchronize()
    ca = c.numpy()
    lam_a = solver.buffers["lambda_sum"][:ca["count"]].cpu().numpy()
    ref_a = o.collide(q)
    qa, qda, feat_a, lam_ref_a, res_a, _ = o.step_learn(q, qd, f, ref_a, DT, mode=1)
    assert np.array_equal(sb.body_q.cpu().numpy(), qa) and np.array_equal(sb.body_qd.cpu().numpy(), qda)
    og = np.argsort(ca["sort_key"], kind="stable")
    orf = np.argsort(ref_a["sort_key"], kind="stable")
    assert np.array_equal(ca["sort_key"][og], ref_a["sort_key"][orf])
    assert np.array_equal(lam_a[og], lam_ref_a[orf])
    fa = solver.buffers["feat"][:ca["count"]].cpu().numpy()
    assert np.array_equal(fa[og], feat_a[orf])
    assert np.array_equal(solver.buffers["residual"].cpu().numpy(), res_a)
    # step B
    c = c_next
    pipe.collide(sb, c)
    solver.log_step(sb, sc, None, c, DT)
    torch.cuda.synchronize()
    cb = c.numpy()
    ref_b = o.collide(qa)
    prev = dict(zip(ref_a["sort_key"].tolist(), lam_ref_a.tolist()))
    lam_prev = np.asarray([prev.get(k, 0.0) for k in ref_b["sort_key"].tolist()], np.float32)
    qb, qdb, feat_b, lam_ref_b, _, _ = o.step_learn(qa, qda, f, ref_b, DT, mode=1, lambda_prev=lam_prev)
    og = np.argsort(cb["sort_key"], kind="stable")
    orf = np.argsort(ref_b["sort_key"], kind="stable")
    fb = solver.buffers["feat"][:cb["count"]].cpu().numpy()
    assert (feat_b[:, 3] > 0).sum() > 0
    assert np.array_equal(fb[og], feat_b[orf])
    assert np.array_equal(sc.body_q.cpu().numpy(), qb)


def _warm(solver, s_in, s_out, c, lam0, k, tol):
    """The warm launch with a given lambda0 in place of the network's prediction."""
    b = solver._buffers(c)
    solver._launch(s_in, None, None, c, DT, solver._io(NMC_STEP_PREP, b, c))
    b["lambda0"][:len(lam0)].copy_(torch.from_numpy(lam0))
    solver._launch(s_in, s_out, None, c, DT, solver._io(NMC_STEP_WARM, b, c, k, tol))
    solver._last_remembered = False
    torch.cuda.synchronize()
    return solver.buffers


def test_warm_step_matches_oracle(cuda_device):
    print("[learned] warm step", flush=True)
    model, o, q, qd = _settled(cuda_device, seed=5)
    f = np.zeros_like(qd)
    pipe = CollisionPipeline(model, broad_phase="sap", deterministic=True)
    solver = SolverXPBDWarmStart(model, WarmStartMLP(device=cuda_device), iterations=10,
                                 reference_arithmetic=True)
    c = pipe.contacts()
    s_in, s_out = _state(model, q, qd), model.state()
    pipe.collide(s_in, c)
    torch.cuda.synchronize()
    cg = c.numpy()
    ref = o.collide(q)
    og = np.argsort(cg["sort_key"], kind="stable")
    orf = np.argsort(ref["sort_key"], kind="stable")
    assert np.array_equal(cg["sort_key"][og], ref["sort_key"][orf])
    _, _, _, lam_true, _, _ = o.step_learn(q, qd, f, ref, DT, mode=1)
    lam0_ref = (lam_true * 0.7).astype(np.float32)
    lam0_gpu = np.empty_like(lam0_ref)
    lam0_gpu[og] = lam0_ref[orf]
    W = model.world_count
    _, _, _, _, res_probe, _ = o.step_learn(q, qd, f, ref, DT, mode=2, lambda0=lam0_ref, warm_iterations=2,
                                            residual_tol=1e30)
    tol = float(np.median(res_probe[:W]))  # about half of the worlds fall back
    for k, t in ((2, tol), (0, 1e30), (3, 0.0)):
        qw, qdw, _, lam_w, res_w, fb_w = o.step_learn(q, qd, f, ref, DT, mode=2, lambda0=lam0_ref, warm_iterations=k,
                                                      residual_tol=t)
        b = _warm(solver, s_in, s_out, c, lam0_gpu, k, t)
        assert np.array_equal(s_out.body_q.cpu().numpy(), qw), k
        assert np.array_equal(s_out.body_qd.cpu().numpy(), qdw), k
        lg = b["lambda_sum"][:cg["count"]].cpu().numpy()
        assert np.array_equal(lg[og], lam_w[orf]), k
        assert np.array_equal(b["residual"].cpu().numpy(), res_w), k
        assert np.array_equal(b["fallback"].cpu().numpy(), fb_w), k
        print("warm", k, "tol", t, "fallback worlds", int(fb_w.sum()), flush=True)
    # lambda0 = 0 and warm_iterations = iterations: the plain step
    plain = SolverXPBD(model, iterations=10, reference_arithmetic=True)
    s_plain = model.state()
    plain.step(s_in, s_plain, None, c, DT)
    _warm(solver, s_in, s_out, c, np.zeros_like(lam0_gpu), 10, 1e30)
    assert np.array_equal(s_out.body_q.cpu().numpy(), s_plain.body_q.cpu().numpy())
    assert np.array_equal(s_out.body_qd.cpu().numpy(), s_plain.body_qd.cpu().numpy())
    # the same identity for the world-frame kernels (the default)
    fast_warm = SolverXPBDWarmStart(model, WarmStartMLP(device=cuda_device), iterations=10)
    SolverXPBD(model, iterations=10).step(s_in, s_plain, None, c, DT)
    _warm(fast_warm, s_in, s_out, c, np.zeros_like(lam0_gpu), 10, 1e30)
    assert np.array_equal(s_out.body_q.cpu().numpy(), s_plain.body_q.cpu().numpy())
    assert np.array_equal(s_out.body_qd.cpu().numpy(), s_plain.body_qd.cpu().numpy())


def _random_mlp(dev, seed=1):
    mlp = WarmStartMLP(device=dev, seed=seed)
    g = torch.Generator().manual_seed(seed)
    mlp.x_mean.copy_(torch.randn(16, generator=g) * 0.1)
    mlp.x_inv_std.copy_(torch.rand(16, generator=g) + 0.5)
    mlp.out_scale = 0.7
    mlp.params["b3"].fill_(0.05)
    mlp.sync()
    return mlp


def _oracle_params(mlp):
    return {k: v.detach().cpu().numpy() for k, v in mlp.params.items()}


def test_mlp_forward_matches_oracle(cuda_device):
    print("[learned] mlp forward", flush=True)
    mlp = _random_mlp(cuda_device)
    rng = np.random.default_rng(2)
    for n in (1, 37, 4096 + 5):
        feat = rng.normal(0, 1, (n, 16)).astype(np.float32)
        ft = torch.from_numpy(feat).to(cuda_device)
        count = torch.tensor([n], dtype=torch.int32, device=cuda_device)
        out = torch.full((n + 3,), -7.0, device=cuda_device)
        mlp.predict(ft, count, n + 3, out)
        torch.cuda.synchronize()
        got = out.cpu().numpy()
        assert (got[n:] == -7.0).all()  # rows at and beyond the count are not written
        x = (feat - mlp.x_mean.cpu().numpy()) * mlp.x_inv_std.cpu().numpy()
        ref = oracle.mlp_forward(_oracle_params(mlp), x) * mlp.out_scale
        scale = max(1e-3, float(np.abs(ref).max()))
        assert float(np.abs(got[:n] - ref).max()) <= 1e-2 * scale, n


def test_mlp_gradients_match_oracle(cuda_device):
    print("[learned] mlp gradients", flush=True)
    mlp = _random_mlp(cuda_device, seed=3)
    rng = np.random.default_rng(4)
    n = 1000
    feat = rng.normal(0, 1, (n, 16)).astype(np.float32)
    t = rng.normal(0, 1, n).astype(np.float32)
    g = mlp.gradients(torch.from_numpy(feat).to(cuda_device), torch.from_numpy(t).to(cuda_device))
    torch.cuda.synchronize()
    x = (feat - mlp.x_mean.cpu().numpy()) * mlp.x_inv_std.cpu().numpy()
    ref, loss = oracle.mlp_backward(_oracle_params(mlp), x, t)
    assert abs(mlp.loss(torch.from_numpy(feat), torch.from_numpy(t).to(cuda_device)) - loss) <= 1e-2 * abs(loss)
    for k, r in ref.items():
        gk = g[k].cpu().numpy()
        err = np.linalg.norm(gk - r) / max(np.linalg.norm(r), 1e-12)
        print(k, "relative gradient error", err, flush=True)
        assert err < 2e-2, (k, err)


@pytest.mark.skipif(not os.path.exists(WEIGHTS_PATH), reason="no trained weights committed")
def test_trained_network_gains_over_plain(cuda_device):
    """Held-out scene: warm start + k iterations vs k and k + 1 plain iterations."""
    print("[learned] trained network", flush=True)
    from newton_custom_contact_amd.train_warmstart import evaluate

    mlp = WarmStartMLP.load(WEIGHTS_PATH, device=cuda_device)
    ev = evaluate(mlp, seeds=[301], worlds=256, steps=24, every=8, dev=cuda_device)
    print(ev, flush=True)
    for k in (1, 2, 3, 4):
        assert ev["warm_residual"][k] < ev["plain_residual"][k], k
    # the warm start is worth at least one more plain iteration at every k
    for k in (1, 2, 3, 4):
        assert ev["warm_residual"][k] <= ev["plain_residual"][k + 1], k


@pytest.mark.parametrize("reference_arithmetic", [True, False], ids=["reference_order", "world_frame"])
def test_warm_step_applies_restitution(cuda_device, reference_arithmetic):
    """The warm launch runs the restitution pass on the start-of-step state: with lambda0 = 0 and
    warm_iterations = iterations it equals the plain step with the pass, bit for bit, and the pass changes
    the velocities of this state (so the identity does not hold by the pass doing nothing)."""
    print(f"[learned] warm step with restitution, reference_arithmetic={reference_arithmetic}", flush=True)
    model, o, q, qd = _settled(cuda_device, seed=5)
    e = np.random.default_rng(2).uniform(0.2, 0.9, model.shape_count).astype(np.float32)
    model.host["shape_material_restitution"] = e
    model.shape_material_restitution.copy_(torch.from_numpy(e))
    # a downward kick, so contacts approach at the start of the step and the pass has work to do
    qd = qd.copy()
    qd[:, 2] -= 0.5
    pipe = CollisionPipeline(model, broad_phase="sap", deterministic=True)
    c = pipe.contacts()
    s_in, s_out, s_plain, s_norest = _state(model, q, qd), model.state(), model.state(), model.state()
    pipe.collide(s_in, c)
    kw = dict(iterations=10, reference_arithmetic=reference_arithmetic)
    SolverXPBD(model, enable_restitution=True, **kw).step(s_in, s_plain, None, c, DT)
    SolverXPBD(model, **kw).step(s_in, s_norest, None, c, DT)
    warm = SolverXPBDWarmStart(model, WarmStartMLP(device=cuda_device), enable_restitution=True, **kw)
    zeros = np.zeros(c.rigid_contact_max, np.float32)
    _warm(warm, s_in, s_out, c, zeros, 10, 1e30)
    assert not np.array_equal(s_plain.body_qd.cpu().numpy(), s_norest.body_qd.cpu().numpy())
    assert np.array_equal(s_out.body_q.cpu().numpy(), s_plain.body_q.cpu().numpy())
    assert np.array_equal(s_out.body_qd.cpu().numpy(), s_plain.body_qd.cpu().numpy())
