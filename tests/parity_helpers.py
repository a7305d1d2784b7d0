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


# The solver's parity bar per step from identical inputs: 1e-5 relative, with the scene's length scale (1 m)
# as the floor for components near zero -- poses (positions, quaternion components) within 1e-5 relative or
# 1e-5 (m, quaternion units); velocities within 1e-5 relative or 1e-5 m per substep (XPBD velocities are
# position corrections over dt; the FP32 oracle itself differs from its FP64 rerun by up to 1.5e-6 m and 5e-4 m/s
# on these scenes).  Non-finite entries must be non-finite on both sides.
STEP_RTOL = 1e-5
Q_ATOL = 1e-5


def assert_step_close(q, qd, q_ref, qd_ref, dt: float, what: str = "") -> None:
    np.testing.assert_allclose(q, q_ref, rtol=STEP_RTOL, atol=Q_ATOL, equal_nan=True, err_msg=f"body_q {what}")
    np.testing.assert_allclose(qd, qd_ref, rtol=STEP_RTOL, atol=Q_ATOL / dt, equal_nan=True,
                               err_msg=f"body_qd {what}")


def max_abs_diff(a, b) -> float:
    """Largest |a - b| over the entries finite on both sides (0 when there are none)."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    fin = np.isfinite(a) & np.isfinite(b)
    return float(np.abs(a[fin] - b[fin]).max()) if fin.any() else 0.0


def bit_equal(a, b) -> bool:
    return bool(np.array_equal(a, b, equal_nan=True))


def step_report_full_size(q, qd, q32, qd32, q64, qd64, dt: float, what: str = "") -> dict:
    """The step bar at full size (10^4 bodies and more), where a few bodies sit on a contact whose activation
    or friction clamp flips with the rounding of the operation order: every body within the per-step
    tolerance except at most 0.1% of them, and against the FP64 rerun the step is at most 4x as far off as the
    FP32 oracle (largest difference over the finite entries).  Returns the counts and largest differences."""
    rep = {"what": what}
    outside = None
    for name, a, r, r64, atol in (("body_q", q, q32, q64, Q_ATOL), ("body_qd", qd, qd32, qd64, Q_ATOL / dt)):
        a = np.asarray(a, np.float64)
        r = np.asarray(r, np.float64)
        r64 = np.asarray(r64, np.float64)
        assert np.array_equal(np.isfinite(a), np.isfinite(r)), f"{name} {what}: non-finite entries differ"
        fin = np.isfinite(a) & np.isfinite(r)
        bad = fin & ~(np.abs(a - r) <= atol + STEP_RTOL * np.abs(r))
        rows = bad.any(axis=1)
        outside = rows if outside is None else (outside | rows)
        fin64 = fin & np.isfinite(r64)
        e_gpu = float(np.abs(a - r64)[fin64].max()) if fin64.any() else 0.0
        e_cpu = float(np.abs(r - r64)[fin64].max()) if fin64.any() else 0.0
        rep[name] = {"entries_outside_tolerance": int(bad.sum()),
                     "max_vs_fp32_oracle": float(np.abs(a - r)[fin].max()) if fin.any() else 0.0,
                     "max_vs_fp64": e_gpu, "fp32_oracle_max_vs_fp64": e_cpu}
        assert e_gpu <= 4.0 * e_cpu + atol, f"{name} {what}: {rep[name]}"
    bodies = len(outside)
    rep["bodies"] = bodies
    rep["bodies_outside_tolerance"] = int(outside.sum())
    assert outside.sum() <= max(2, int(1e-3 * bodies)), rep
    return rep
