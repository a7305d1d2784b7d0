"""CPU tests of the on-disk SDF cache (_sdf_cache) and of Mesh.build_sdf's argument handling.

Mirrors the reference's tests/test_sdf_disk_cache.py:153-392 for this cache's dense schema: the key is
stable, 32 hex digits, and changes with every input that changes a cook (resolution, margin, scale,
is_solid, a vertex, the triangles); a saved grid loads back bit for bit; the file carries the version, kind,
package version and UTC timestamp; a missing, corrupt or truncated file, another format version and another
kind (a reference texture cache file) are misses; concurrent writers publish one loadable file and leave no
temporary files; every save uses its own temporary name.
"""

from __future__ import annotations

import os
import threading
import zipfile

import numpy as np
import pytest

from newton_custom_contact_amd import _sdf_cache
from tests.mesh_helpers import box_mesh, cook_cpu


def _kwargs(mesh, **over):
    kw = {"vertices": mesh.vertices, "indices": mesh.indices, "is_solid": True, "max_resolution": 64,
          "margin": 0.05, "scale": None}
    kw.update(over)
    return kw


def _data(mesh):
    cook_cpu(mesh, 8)
    s = mesh.sdf
    return {"grid": s.grid.numpy(), "lower": s.lower, "voxel_size": np.float32(s.voxel_size), "cells": s.cells,
            "scale_baked": False}


def test_hash_is_stable():
    m = box_mesh()
    h1 = _sdf_cache.hash_inputs(**_kwargs(m))
    assert h1 == _sdf_cache.hash_inputs(**_kwargs(box_mesh()))
    assert len(h1) == 32 and int(h1, 16) >= 0


@pytest.mark.parametrize("name,value", [("max_resolution", 32), ("margin", 0.1), ("scale", (2.0, 1.0, 1.0)),
                                        ("is_solid", False)])
def test_hash_changes_with_params(name, value):
    m = box_mesh()
    assert _sdf_cache.hash_inputs(**_kwargs(m, **{name: value})) != _sdf_cache.hash_inputs(**_kwargs(m))


def test_hash_changes_with_mesh():
    m = box_mesh()
    base = _sdf_cache.hash_inputs(**_kwargs(m))
    moved = m.vertices.copy()
    moved[0, 0] += 0.1
    assert _sdf_cache.hash_inputs(**_kwargs(m, vertices=moved)) != base
    flipped = m.indices.reshape(-1, 3)[:, [0, 2, 1]].reshape(-1)
    assert _sdf_cache.hash_inputs(**_kwargs(m, indices=flipped)) != base


def test_round_trip_and_provenance(tmp_path):
    m = box_mesh(0.3, 0.2, 0.1)
    data = _data(m)
    h = _sdf_cache.hash_inputs(**_kwargs(m))
    path = _sdf_cache.save_sdf_data(tmp_path, h, data, newton_version="test")
    assert path == _sdf_cache.cache_path(tmp_path, h) and path.name == f"{h}.sdf.npz"
    got = _sdf_cache.try_load_sdf_data(tmp_path, h)
    assert got is not None
    np.testing.assert_array_equal(got["grid"], data["grid"])
    assert got["grid"].dtype == np.float32
    np.testing.assert_array_equal(got["lower"], data["lower"])
    assert got["voxel_size"] == data["voxel_size"]
    np.testing.assert_array_equal(got["cells"], data["cells"])
    assert got["scale_baked"] is False
    with np.load(path, allow_pickle=False) as npz:
        for k in ("__cache_format_version__", "__kind__", "__newton_version__", "__created_utc__", "grid",
                  "lower", "voxel_size", "cells", "scale_baked"):
            assert k in npz.files
        assert int(npz["__cache_format_version__"].item()) == _sdf_cache.CACHE_FORMAT_VERSION
        assert str(npz["__kind__"].item()) == "newton_custom_contact_amd.dense_sdf"
        assert str(npz["__newton_version__"].item()) == "test"
        assert str(npz["__created_utc__"].item()).endswith("+00:00")
    # the package version when none is given
    import newton_custom_contact_amd

    _sdf_cache.save_sdf_data(tmp_path, h, data)
    with np.load(path, allow_pickle=False) as npz:
        assert str(npz["__newton_version__"].item()) == newton_custom_contact_amd.__version__


def test_misses(tmp_path):
    m = box_mesh()
    data = _data(m)
    h = _sdf_cache.hash_inputs(**_kwargs(m))
    assert _sdf_cache.try_load_sdf_data(tmp_path, "deadbeef") is None
    path = _sdf_cache.save_sdf_data(tmp_path, h, data, newton_version="test")
    raw = path.read_bytes()
    path.write_bytes(b"not an npz")
    assert _sdf_cache.try_load_sdf_data(tmp_path, h) is None
    path.write_bytes(raw[: len(raw) // 2])  # truncated zip
    assert _sdf_cache.try_load_sdf_data(tmp_path, h) is None
    path.write_bytes(raw)
    assert _sdf_cache.try_load_sdf_data(tmp_path, h) is not None
    with np.load(path) as npz:
        contents = {k: npz[k] for k in npz.files}
    for key, value in (("__cache_format_version__", np.asarray(_sdf_cache.CACHE_FORMAT_VERSION + 999, np.int32)),
                       ("__kind__", np.asarray("newton.texture_sdf", dtype=np.str_))):
        bad = dict(contents)
        bad[key] = value
        np.savez(path, **bad)
        assert _sdf_cache.try_load_sdf_data(tmp_path, h) is None, key
    bad = {k: v for k, v in contents.items() if k != "__cache_format_version__"}
    np.savez(path, **bad)
    assert _sdf_cache.try_load_sdf_data(tmp_path, h) is None
    bad = dict(contents)
    bad["cells"] = np.asarray(contents["cells"]) + 1  # grid no longer matches
    np.savez(path, **bad)
    assert _sdf_cache.try_load_sdf_data(tmp_path, h) is None
    with pytest.raises(ValueError):
        _sdf_cache.save_sdf_data(tmp_path, h, dict(data, cells=np.asarray(data["cells"]) + 1))


def test_concurrent_writers(tmp_path):
    m = box_mesh()
    data = _data(m)
    h = _sdf_cache.hash_inputs(**_kwargs(m))
    errors = []

    def worker():
        try:
            _sdf_cache.save_sdf_data(tmp_path, h, data, newton_version="test")
        except BaseException as exc:  # noqa: BLE001
            errors.append(exc)

    threads = [threading.Thread(target=worker) for _ in range(8)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    assert errors == []
    got = _sdf_cache.try_load_sdf_data(tmp_path, h)
    assert got is not None
    np.testing.assert_array_equal(got["grid"], data["grid"])
    assert list(tmp_path.glob("*.tmp.npz")) == []
    assert zipfile.is_zipfile(_sdf_cache.cache_path(tmp_path, h))


def test_unique_tmp_names(tmp_path, monkeypatch):
    m = box_mesh()
    data = _data(m)
    h = _sdf_cache.hash_inputs(**_kwargs(m))
    seen = set()
    real = os.replace

    def spy(src, dst):
        seen.add(os.fspath(src))
        return real(src, dst)

    monkeypatch.setattr(os, "replace", spy)
    for _ in range(4):
        _sdf_cache.save_sdf_data(tmp_path, h, data, newton_version="test")
    assert len(seen) == 4
    assert all(s.endswith(".tmp.npz") and f"{h}.sdf.npz.{os.getpid()}." in s for s in seen)


def test_write_is_best_effort(tmp_path):
    m = box_mesh()
    blocker = tmp_path / "file"
    blocker.write_text("x")
    # cache_dir below a regular file: mkdir fails, write() logs and returns
    _sdf_cache.write(blocker / "sub", "00", _data(m))
    with pytest.raises(OSError):
        _sdf_cache.save_sdf_data(blocker / "sub", "00", _data(box_mesh()))


def test_build_sdf_argument_checks():
    m = box_mesh()
    with pytest.raises(RuntimeError, match="GPU"):
        m.build_sdf(device="cpu")
    with pytest.raises(ValueError, match="texture_format"):
        m.build_sdf(device="cpu", texture_format="half")
    with pytest.raises(ValueError, match="narrow_band_range"):
        m.build_sdf(device="cpu", narrow_band_range=(0.1, -0.1))
    with pytest.raises(ValueError, match="divisible by 8"):
        m.sdf_resolution(max_resolution=20)
    with pytest.raises(ValueError, match="target_voxel_size"):
        m.sdf_resolution(target_voxel_size=0.0)
    cook_cpu(m, 8)
    with pytest.raises(RuntimeError, match="clear_sdf"):
        m.build_sdf(device="cpu")
    m.clear_sdf()
    assert m.sdf is None


def test_resolution_from_target_voxel_size():
    m = box_mesh()  # 1 m cube; padded extent 1.1 m
    assert m.sdf_resolution() == 64
    assert m.sdf_resolution(max_resolution=32) == 32
    assert m.sdf_resolution(target_voxel_size=0.05) == 24      # ceil(22) -> multiple of 8
    assert m.sdf_resolution(target_voxel_size=0.3) == 8        # ceil(3.67) = 4 -> at least 8
    assert m.sdf_resolution(target_voxel_size=0.01, max_resolution=8) == 112  # the voxel size wins
    assert m.sdf_resolution(target_voxel_size=0.05, scale=(2.0, 1.0, 1.0)) == 48  # ceil(2.1 / 0.05) = 42
