"""CPU restatement of the learned warm start (oracle_learn.h), pinned by its defining identities:

* the logging step and the warm step with lambda0 = 0, warm_iterations = iterations are the plain step,
  bit for bit; a zero residual tolerance sends every world with contacts to the cold fallback, which
  is again the plain step;
* the logged total multipliers, applied as the warm start, leave less penetration than the same
  number of plain iterations (what the network is trained to reproduce);
* the MLP restatement's gradients equal torch autograd (float64) of the same forward with BF16
  rounding as a straight-through operation.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

import oracle
from newton_custom_contact_amd import scenes
from oracle import Oracle

DT = 1.0 / 240.0


@pytest.fixture(scope="module")
def settled():
    model = scenes.mixed_worlds(6, seed=4, device="cpu")
    o = Oracle(model)
    h = model.host
    q, qd = h["body_q"].copy(), h["body_qd"].copy()
    f = np.zeros_like(qd)
    for _ in range(300):
        c = o.collide(q)
        q, qd = o.step(q, qd, f, c, DT, iterations=10)
    return model, o, q, qd, o.collide(q)


def test_log_and_trivial_warm_are_the_plain_step(settled):
    model, o, q, qd, c = settled
    f = np.zeros_like(qd)
    qp, qdp = o.step(q, qd, f, c, DT, iterations=10)
    q1, qd1, feat, lam, res, fb = o.step_learn(q, qd, f, c, DT, mode=1)
    assert np.array_equal(q1, qp) and np.array_equal(qd1, qdp)
    assert feat.shape == (c["count"], 16) and (lam >= 0).all() and lam.sum() > 0
    q2, qd2, _, lam2, _, fb2 = o.step_learn(q, qd, f, c, DT, mode=2, lambda0=np.zeros(c["count"], np.float32),
                                            warm_iterations=10, residual_tol=1e30)
    assert np.array_equal(q2, qp) and np.array_equal(qd2, qdp) and not fb2.any()
    assert np.array_equal(lam2, lam)
    q3, qd3, _, lam3, res3, fb3 = o.step_learn(q, qd, f, c, DT, mode=2, lambda0=lam, warm_iterations=2,
                                               residual_tol=0.0)
    worlds_with_pen = res > 0
    assert np.array_equal(fb3.astype(bool)[:-1], worlds_with_pen[:-1] | (res3[:-1] > 0))
    assert np.array_equal(q3, qp) and np.array_equal(qd3, qdp) and np.array_equal(lam3, lam)


def test_logged_multipliers_warm_start_beats_plain(settled):
    model, o, q, qd, c = settled
    f = np.zeros_like(qd)
    _, _, _, lam, res10, _ = o.step_learn(q, qd, f, c, DT, mode=1)
    W = model.world_count
    for k in (0, 1, 2):
        _, _, _, _, res_w, _ = o.step_learn(q, qd, f, c, DT, mode=2, lambda0=lam, warm_iterations=k,
                                            residual_tol=1e30)
        _, _, _, _, res_p, _ = o.step_learn(q, qd, f, c, DT, mode=1, iterations=max(k, 1))
        assert res_w[:W].mean() < res_p[:W].mean()
    assert res_w[:W].mean() <= res10[:W].mean() * 1.05


def test_features_first_multiplier_matches_contact_solve(settled):
    """Feature 2 is the first iteration's normal multiplier before relaxation: -d / (dt w) for d < 0."""
    model, o, q, qd, c = settled
    _, _, feat, _, _, _ = o.step_learn(q, qd, np.zeros_like(qd), c, DT, mode=1)
    d, w, l1 = feat[:, 0], feat[:, 1], feat[:, 2]
    pen = d < 0
    assert pen.any()
    np.testing.assert_allclose(l1[pen], -d[pen] / (DT * w[pen]), rtol=1e-5)
    assert (l1[~pen] == 0).all()
    assert set(np.unique(feat[:, 12])) <= {0.0, 1.0}



def test_mlp_backward_equals_autograd():
    rng = np.random.default_rng(0)
    P = {"w1": rng.normal(0, 0.3, (128, 16)), "b1": rng.normal(0, 0.1, 128), "w2": rng.normal(0, 0.1, (128, 128)),
         "b2": rng.normal(0, 0.1, 128), "w3": rng.normal(0, 0.1, 128), "b3": rng.normal(0, 0.1, 1)}
    P = {k: v.astype(np.float32) for k, v in P.items()}
    x = rng.normal(0, 1, (24, 16)).astype(np.float32)
    t = rng.normal(0, 1, 24).astype(np.float32)
    g, loss = oracle.mlp_backward(P, x, t)

    def bf(v):  # round to nearest even BF16 (value), straight-through gradient
        u = v.detach().to(torch.float32).numpy().view(np.uint32)
        r = ((u >> 16) & 1) + 0x7FFF
        rv = ((u + r) & 0xFFFF0000).view(np.float32)
        return v + (torch.from_numpy(rv.astype(np.float64)) - v).detach()

    T = {k: torch.tensor(v.astype(np.float64), requires_grad=True) for k, v in P.items()}
    X = torch.tensor(x.astype(np.float64))
    z1 = bf(X) @ bf(T["w1"]).T + T["b1"]
    h1 = bf(torch.nn.functional.silu(z1))
    z2 = h1 @ bf(T["w2"]).T + T["b2"]
    y = torch.nn.functional.silu(z2) @ T["w3"] + T["b3"]
    L = 0.5 * ((y - torch.tensor(t.astype(np.float64))) ** 2).mean()
    L.backward()
    assert abs(L.item() - loss) < 1e-5 * max(1.0, abs(L.item()))
    for k in P:
        ref = T[k].grad.numpy()
        err = np.linalg.norm(g[k] - ref) / max(np.linalg.norm(ref), 1e-12)
        assert err < 5e-3, (k, err)
