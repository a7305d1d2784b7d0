"""CPU tests of the contact-matching restatement (contact_match.py) on constructed contact sets.

Pins: a repeated frame matches every contact to itself; the first frame and unseen shape pairs give
MATCH_NOT_FOUND; a midpoint moved beyond the threshold or a normal turned beyond the dot threshold gives
MATCH_BROKEN and reports the old contact broken; two new contacts closest to one old contact resolve to the
nearer one (ties: the smaller low 32 key bits) and the loser is MATCH_BROKEN; sticky mode replays the old
body-frame points, offsets and normal of matched rows only; brute force over random contact sets (pair
ranges found by scanning every old contact, claims by explicit minimum) gives the same answers.
"""

from __future__ import annotations

import numpy as np

import oracle
from newton_custom_contact_amd.builder import ModelBuilder


def _model():
    b = ModelBuilder()
    b.add_ground_plane()
    for x in (0.0, 1.0, 2.0):
        body = b.add_body(xform=(x, 0.0, 0.5, 0.0, 0.0, 0.0, 1.0))
        b.add_shape_box(body, hx=0.1, hy=0.1, hz=0.1)
    return b.finalize("cpu")


def _key(a, b, sub):
    return (a << 43) | (b << 23) | sub


def _contacts(rows):
    """rows: (shape0, shape1, sub, point0 (body frame), point1, normal)."""
    n = len(rows)
    c = {"sort_key": np.array([_key(r[0], r[1], r[2]) for r in rows], np.int64),
         "shape0": np.array([r[0] for r in rows], np.int32), "shape1": np.array([r[1] for r in rows], np.int32)}
    for k, idx in (("point0", 3), ("point1", 4), ("normal", 5)):
        c[k] = np.array([r[idx] for r in rows], np.float32).reshape(n, 3)
    c["offset0"] = np.full((n, 3), 0.25, np.float32)
    c["offset1"] = np.full((n, 3), -0.25, np.float32)
    return c


UP = (0.0, 0.0, 1.0)


def test_repeat_frame_matches_itself_and_first_frame_not_found():
    model = _model()
    o = oracle.Oracle(model)
    m = oracle.OracleMatcher(o)
    q = model.host["body_q"]
    # shape 0 = ground, shapes 1..3 = boxes on bodies 0..2
    c = _contacts([(0, 1, k, (0.0, 0.0, 0.0), (0.1 * k, 0.0, 0.0), UP) for k in range(4)] +
                  [(1, 2, 0, (0.1, 0.0, 0.0), (-0.1, 0.0, 0.0), (1.0, 0.0, 0.0))])
    r1 = m.match(q, c)
    assert np.all(r1["match_index"] == -1) and r1["broken_keys"].size == 0
    r2 = m.match(q, c)
    np.testing.assert_array_equal(r2["matched_key"], r2["sort_key"])
    np.testing.assert_array_equal(r2["match_index"], np.arange(5))
    assert r2["broken_keys"].size == 0


def test_moved_turned_and_new_pairs():
    model = _model()
    o = oracle.Oracle(model)
    m = oracle.OracleMatcher(o, pos_threshold=0.01, normal_dot_threshold=0.99)
    q = model.host["body_q"]
    base = [(0, 1, 0, (0.0, 0.0, 0.0), (0.0, 0.0, 0.0), UP),        # stays
            (0, 2, 0, (0.0, 0.0, 0.0), (1.0, 0.0, 0.0), UP),        # moves 2 cm
            (0, 3, 0, (0.0, 0.0, 0.0), (2.0, 0.0, 0.0), UP)]        # normal turns 20 degrees
    m.match(q, _contacts(base))
    t = np.radians(20.0)
    nxt = [(0, 1, 1, (0.0, 0.0, 0.0), (0.0, 0.0, 0.0), UP),         # same pair, new sub key: still matched
           (0, 2, 0, (0.02, 0.0, 0.0), (1.02, 0.0, 0.0), UP),
           (0, 3, 0, (0.0, 0.0, 0.0), (2.0, 0.0, 0.0), (np.sin(t), 0.0, np.cos(t))),
           (1, 2, 0, (0.1, 0.0, 0.0), (-0.1, 0.0, 0.0), (1.0, 0.0, 0.0))]  # pair not seen before
    r = m.match(q, _contacts(nxt))
    np.testing.assert_array_equal(r["match_index"], [0, -2, -2, -1])
    assert r["matched_key"][0] == _key(0, 1, 0)
    np.testing.assert_array_equal(np.sort(r["broken_keys"]), [_key(0, 2, 0), _key(0, 3, 0)])


def test_claim_resolution_nearest_then_smaller_key():
    model = _model()
    o = oracle.Oracle(model)
    m = oracle.OracleMatcher(o, pos_threshold=0.05)
    q = model.host["body_q"]
    m.match(q, _contacts([(0, 1, 0, (0.0, 0.0, 0.0), (0.0, 0.0, 0.0), UP)]))
    # two new contacts of the pair pick the one old contact: the nearer (sub 5, 1 cm) wins
    r = m.match(q, _contacts([(0, 1, 3, (0.0, 0.0, 0.0), (0.02, 0.0, 0.0), UP),
                              (0, 1, 5, (0.0, 0.0, 0.0), (0.01, 0.0, 0.0), UP)]))
    np.testing.assert_array_equal(r["match_index"], [-2, 0])
    # equal distances: the smaller sort key wins
    m.reset()
    m.match(q, _contacts([(0, 1, 0, (0.0, 0.0, 0.0), (0.0, 0.0, 0.0), UP)]))
    r = m.match(q, _contacts([(0, 1, 3, (0.0, 0.0, 0.0), (0.01, 0.0, 0.0), UP),
                              (0, 1, 5, (0.0, 0.0, 0.0), (-0.01, 0.0, 0.0), UP)]))
    np.testing.assert_array_equal(r["match_index"], [0, -2])


def test_sticky_replays_matched_rows_only():
    model = _model()
    o = oracle.Oracle(model)
    m = oracle.OracleMatcher(o, pos_threshold=0.01, sticky=True)
    q = model.host["body_q"]
    first = _contacts([(0, 1, 0, (0.0, 0.0, 0.0), (0.0, 0.0, 0.0), UP),
                       (0, 2, 0, (0.0, 0.0, 0.0), (1.0, 0.0, 0.0), UP)])
    m.match(q, first)
    second = _contacts([(0, 1, 0, (0.001, 0.0, 0.0), (0.001, 0.0, 0.0), UP),   # matched: replayed
                        (0, 2, 0, (0.5, 0.0, 0.0), (1.5, 0.0, 0.0), UP)])      # broken: kept
    second["offset0"][:] = 9.0
    r = m.match(q, second)
    np.testing.assert_array_equal(r["match_index"], [0, -2])
    np.testing.assert_array_equal(r["point0"][0], first["point0"][0])
    np.testing.assert_array_equal(r["offset0"][0], first["offset0"][0])
    np.testing.assert_array_equal(r["point0"][1], second["point0"][1])
    np.testing.assert_array_equal(r["offset0"][1], [9.0, 9.0, 9.0])


def _brute(prev, cur, q, sb, thr_sq, ndot):
    """Independent restatement: every old contact is tested (no binary search), claims by explicit min."""
    def world(shape, p):
        b = sb[shape]
        if b < 0:
            return p.astype(np.float64)
        qq = q[b]
        u, w = qq[3:6].astype(np.float64), float(qq[6])
        v = p.astype(np.float64)
        return v * (2 * w * w - 1) + u * 2 * np.dot(u, v) + np.cross(u, v) * 2 * w + qq[:3]

    def mid(c, i):
        return 0.5 * (world(c["shape0"][i], c["point0"][i]) + world(c["shape1"][i], c["point1"][i]))

    pm = [mid(prev, j) for j in range(len(prev["sort_key"]))]
    res, cand = [], []
    for i in range(len(cur["sort_key"])):
        pair = cur["sort_key"][i] >> 23
        same = [j for j in range(len(prev["sort_key"])) if prev["sort_key"][j] >> 23 == pair]
        if not same:
            res.append(-1)
            cand.append(None)
            continue
        ci = mid(cur, i)
        best, bd = None, thr_sq
        for j in same:
            d2 = float(np.dot(ci - pm[j], ci - pm[j]))
            if d2 <= bd and float(np.dot(cur["normal"][i], prev["normal"][j])) >= ndot:
                best, bd = j, d2
        res.append(-2 if best is None else best)
        cand.append(None if best is None else (bd, int(cur["sort_key"][i]) & 0xFFFFFFFF))
    for i, c in enumerate(res):
        if c >= 0:
            rivals = [cand[k] for k in range(len(res)) if res[k] == c]
            if min(rivals) != cand[i]:
                res[i] = -2
    return np.array(res)


def test_random_sets_match_brute_force():
    model = _model()
    o = oracle.Oracle(model)
    sb = model.host["shape_body"]
    q = model.host["body_q"]
    rng = np.random.Generator(np.random.PCG64(9))
    pairs = [(0, 1), (0, 2), (0, 3), (1, 2), (2, 3)]
    for trial in range(20):
        m = oracle.OracleMatcher(o, pos_threshold=0.05, normal_dot_threshold=0.9)

        def frame(n):
            rows, used = [], set()
            for _ in range(n):
                a, b = pairs[rng.integers(len(pairs))]
                sub = int(rng.integers(0, 6))
                if (a, b, sub) in used:
                    continue
                used.add((a, b, sub))
                nrm = rng.normal(size=3) * 0.15 + np.array([0.0, 0.0, 1.0])
                rows.append((a, b, sub, tuple(rng.uniform(-0.06, 0.06, 3)), tuple(rng.uniform(-0.06, 0.06, 3)),
                             tuple(nrm / np.linalg.norm(nrm))))
            rows.sort(key=lambda r: _key(r[0], r[1], r[2]))
            return _contacts(rows)

        prev = frame(14)
        m.match(q, prev)
        cur = frame(14)
        r = m.match(q, cur)
        exp = _brute(prev, cur, q, sb, 0.05 ** 2, 0.9)
        np.testing.assert_array_equal(r["match_index"], exp, err_msg=f"trial {trial}")
