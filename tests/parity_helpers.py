"""Shared helpers for the GPU-vs-oracle parity tests."""

from __future__ import annotations

import numpy as np

FIELDS_F = ("point0", "point1", "offset0", "offset1", "normal", "margin0", "margin1")


def canonical(c: dict) -> dict:
    """Contacts sorted by their 64-bit key (keys are unique per contact)."""
    order = np.argsort(c["sort_key"], kind="stable")
    return {k: (np.asarray(v)[order] if isinstance(v, np.ndarray) else v) for k, v in c.items()}


def compare_contacts(gpu: dict, ref: dict, rtol: float = 1e-5, atol: float = 1e-6) -> dict:
    """Returns a report; raises AssertionError on a contact-set mismatch."""
    g = canonical(gpu)
    r = canonical(ref)
    assert gpu["count"] == ref["count"], f"contact count {gpu['count']} != oracle {ref['count']}"
    assert np.array_equal(g["sort_key"], r["sort_key"]), "contact sort keys differ"
    assert np.array_equal(g["shape0"], r["shape0"]) and np.array_equal(g["shape1"], r["shape1"])
    report = {"count": int(gpu["count"]), "bit_exact": True, "max_abs": 0.0}
    for f in FIELDS_F:
        a = np.asarray(g[f], np.float32)
        b = np.asarray(r[f], np.float32)
        if not np.array_equal(a, b):
            report["bit_exact"] = False
        if a.size:
            report["max_abs"] = max(report["max_abs"], float(np.max(np.abs(a - b))))
        np.testing.assert_allclose(a, b, rtol=rtol, atol=atol, err_msg=f"contact field {f}")
    return report


def oracle_states(oracle, model, n_substeps: int, dt: float, iterations: int = 10, every: int = 1):
    """Advance the oracle from the model's initial state and yield (q, qd) every `every` substeps."""
    h = model.host
    q = h["body_q"].copy()
    qd = h["body_qd"].copy()
    f = np.zeros_like(qd)
    out = [(q.copy(), qd.copy())]
    for i in range(n_substeps):
        c = oracle.collide(q)
        q, qd = oracle.step(q, qd, f, c, dt, iterations=iterations)
        if (i + 1) % every == 0:
            out.append((q.copy(), qd.copy()))
    return out
