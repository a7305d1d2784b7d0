"""nmc_broad_phase_explicit_noalloc (the explicit broad phase with its scratch in a static device array, for
callers that cannot allocate) against nmc_broad_phase_explicit, through the C-ABI: two calls in a row, the
second with a smaller input, each bit-identical to the allocating entry point (count and pairs), plus the
oracle's pair list, append mode, and the size limit refused with an error before any launch."""

from __future__ import annotations

import ctypes

import numpy as np
import pytest
import torch

import oracle
from tests.test_oracle_broad import random_scene

pytestmark = pytest.mark.gpu

_vp = ctypes.c_void_p
LIMIT = 16384 * 256 * 4  # input pairs one call accepts


def _lib():
    from newton_custom_contact_amd._native import load

    L = load()
    L.nmc_broad_phase_explicit_noalloc.argtypes = [ctypes.c_int, _vp, _vp, _vp, _vp, ctypes.c_int, _vp,
                                                   ctypes.c_int, _vp, ctypes.c_int, _vp]
    L.nmc_broad_phase_explicit_noalloc.restype = ctypes.c_int
    return L


def _p(t):
    return ctypes.c_void_p(t.data_ptr())


def _scene(seed, worlds, per_world):
    rng = np.random.Generator(np.random.PCG64(seed))
    lower, upper, gap, group, world, _ = random_scene(rng, worlds=worlds, per_world=per_world, shared=0)
    n = len(world)
    pairs = np.array([(a, b) for a in range(n) for b in range(a + 1, n) if world[a] == world[b]], np.int32)
    rng.shuffle(pairs)
    return lower, upper, gap, group, world, pairs


def _both(L, dev, lower, upper, gap, pairs, cap, append_base=None):
    """(count, pairs) from the allocating entry point and from the no-allocation one, same inputs."""
    from newton_custom_contact_amd._native import check, stream_handle

    lo = torch.from_numpy(np.ascontiguousarray(lower, np.float32)).to(dev)
    hi = torch.from_numpy(np.ascontiguousarray(upper, np.float32)).to(dev)
    gp = torch.from_numpy(np.ascontiguousarray(gap, np.float32)).to(dev)
    pr = torch.from_numpy(np.ascontiguousarray(pairs, np.int32)).to(dev)
    n = len(lower)
    append = 0 if append_base is None else 1
    out = []
    for noalloc in (False, True):
        cand = torch.full((cap, 2), -7, dtype=torch.int32, device=dev)
        cnt = torch.full((1,), 0 if append_base is None else append_base, dtype=torch.int32, device=dev)
        s = stream_handle(dev)
        if noalloc:
            rc = L.nmc_broad_phase_explicit_noalloc(n, _p(lo), _p(hi), _p(gp), _p(pr), len(pairs), _p(cand), cap,
                                                    _p(cnt), append, s)
            check(rc, "nmc_broad_phase_explicit_noalloc")
        else:
            h = ctypes.c_void_p()
            check(L.nmc_broad_phase_create(2, None, 0, None, 0, max(len(pairs), 1), ctypes.byref(h)),
                  "nmc_broad_phase_create")
            try:
                check(L.nmc_broad_phase_explicit(h, n, _p(lo), _p(hi), _p(gp), _p(pr), len(pairs), _p(cand), cap,
                                                 _p(cnt), append, s), "nmc_broad_phase_explicit")
                torch.cuda.synchronize()
            finally:
                L.nmc_broad_phase_destroy(h)
        torch.cuda.synchronize()
        out.append((int(cnt.item()), cand.cpu().numpy()))
    return out


def test_noalloc_matches_allocating_entry_point(cuda_device):
    L = _lib()
    dev = cuda_device
    # first call: 48 worlds x 40 shapes (37 440 listed pairs, 37 blocks of scratch)
    lower, upper, gap, group, world, pairs = _scene(21, 48, 40)
    print(f"[noalloc] call 1: {len(pairs)} pairs", flush=True)
    (n_a, a), (n_b, b) = _both(L, dev, lower, upper, gap, pairs, len(pairs))
    ref = oracle.broad_phase("explicit", lower, upper, gap, group, world, shape_pairs=pairs)
    assert n_a == n_b == len(ref) > 0
    np.testing.assert_array_equal(b, a)
    np.testing.assert_array_equal(b[:n_b], ref)
    # second call, smaller input: the scratch keeps the first call's entries past this call's blocks
    lower2, upper2, gap2, group2, world2, pairs2 = _scene(22, 5, 12)
    print(f"[noalloc] call 2: {len(pairs2)} pairs", flush=True)
    (n_a, a), (n_b, b) = _both(L, dev, lower2, upper2, gap2, pairs2, len(pairs2))
    ref2 = oracle.broad_phase("explicit", lower2, upper2, gap2, group2, world2, shape_pairs=pairs2)
    assert n_a == n_b == len(ref2) > 0
    np.testing.assert_array_equal(b, a)
    np.testing.assert_array_equal(b[:n_b], ref2)
    # append onto a count of 3, into an output too small for all of them
    print("[noalloc] append", flush=True)
    (n_a, a), (n_b, b) = _both(L, dev, lower2, upper2, gap2, pairs2, len(ref2), append_base=3)
    assert n_a == n_b == len(ref2) + 3
    np.testing.assert_array_equal(b, a)
    np.testing.assert_array_equal(b[3:], ref2[:len(ref2) - 3])
    # no pairs at all
    (n_a, a), (n_b, b) = _both(L, dev, lower2, upper2, gap2, pairs2[:0], 4)
    assert n_a == n_b == 0


def test_noalloc_refuses_beyond_limit(cuda_device):
    """A size beyond the scratch's limit is refused by the host check, before any launch or access."""
    from newton_custom_contact_amd._native import stream_handle

    L = _lib()
    dev = cuda_device
    lo = torch.zeros((2, 3), dtype=torch.float32, device=dev)
    pr = torch.zeros((1, 2), dtype=torch.int32, device=dev)
    cand = torch.zeros((1, 2), dtype=torch.int32, device=dev)
    cnt = torch.zeros(1, dtype=torch.int32, device=dev)
    rc = L.nmc_broad_phase_explicit_noalloc(2, _p(lo), _p(lo), None, _p(pr), LIMIT + 1, _p(cand), 1, _p(cnt), 0,
                                            stream_handle(dev))
    assert rc != 0
    msg = L.nmc_last_error().decode()
    assert "exceed the limit" in msg, msg
    rc = L.nmc_broad_phase_explicit_noalloc(2, _p(lo), _p(lo), None, _p(pr), -1, _p(cand), 1, _p(cnt), 0,
                                            stream_handle(dev))
    assert rc != 0
