"""On-disk cache of cooked texture SDFs in the reference's format (newton/_src/geometry/_sdf_cache.py).

One ``{hash}.sdf.npz`` per cook under ``cache_dir`` (an uncompressed ``np.savez`` bundle), written to a
per-writer ``{hash}.sdf.npz.{pid}.{token}.tmp.npz`` and published with ``os.replace`` so that a reader never
sees a partial file and concurrent writers of one key never share a temporary file (_sdf_cache.py:260-330); a
replace that loses to a peer publishing the same key is benign.  Load treats a missing file, a file that does
not parse and a file whose embedded int32 ``__cache_format_version__`` differs from
:data:`CACHE_FORMAT_VERSION` as a miss (:334-385); a failed write is logged and ignored (:388-402).

The arrays are the cooked dictionary of ``build_sparse_sdf_from_mesh`` (sdf_texture.py:1292-1308) as the
reference stores it (:21-60): ``coarse_sdf`` float32 (d+1, h+1, w+1), ``subgrid_data`` float32 / uint16 /
uint8 (ts, ts, ts), ``subgrid_start_slots`` uint32 (w, h, d), ``subgrid_required`` and ``subgrid_occupied``
int32 (w*h*d,), ``coarse_dims`` int32 (3,), the int32 scalars ``subgrid_tex_size``, ``num_subgrids``,
``subgrid_size``, ``quantization_mode``, the float32 scalars ``subgrids_min_sdf_value`` and
``subgrids_sdf_value_range``, the float64 (3,) vectors ``min_extents``, ``max_extents``, ``cell_size``, and the
diagnostic strings ``__kind__`` (``"newton.texture_sdf"``), ``__newton_version__``, ``__created_utc__``.  The
key is a 128-bit BLAKE2b digest of canonical JSON holding SHA-256 digests of the vertices (float32) and indices
(int32) and the build parameters (:144-215), so a directory of the reference's cache files serves this build
and the reverse.
"""

from __future__ import annotations

import contextlib
import hashlib
import json
import logging
import os
import secrets
import zipfile
from collections.abc import Mapping
from datetime import datetime, timezone
from pathlib import Path
from typing import Any

import numpy as np

logger = logging.getLogger(__name__)

CACHE_FORMAT_VERSION: int = 1
"""Version of the on-disk format; bump it when the arrays' shapes, dtypes or meaning change."""

_VERSION_KEY = "__cache_format_version__"
_KIND_KEY = "__kind__"
_NEWTON_VERSION_KEY = "__newton_version__"
_CREATED_UTC_KEY = "__created_utc__"
_NPZ_SUFFIX = ".sdf.npz"
_KIND = "newton.texture_sdf"
_ARRAYS = ("coarse_sdf", "subgrid_data", "subgrid_start_slots", "subgrid_required", "subgrid_occupied")
_INTS = ("subgrid_tex_size", "num_subgrids", "subgrid_size", "quantization_mode")
_FLOATS = ("subgrids_min_sdf_value", "subgrids_sdf_value_range")
_VEC3 = ("min_extents", "max_extents", "cell_size")


def _digest(arr, dtype) -> str:
    return hashlib.sha256(np.ascontiguousarray(arr, dtype=dtype).tobytes()).hexdigest()


def _package_version() -> str:
    try:
        from . import __version__  # noqa: PLC0415

        return str(__version__)
    except ImportError:
        return "unknown"


def hash_inputs(*, vertices: np.ndarray, indices: np.ndarray, is_solid: bool, narrow_band_range,
                target_voxel_size: float | None, max_resolution: int | None, margin: float, texture_format: str,
                sign_method_resolved: str, winding_threshold: float, scale) -> str:
    """Cache key of a texture SDF cook: 32 hexadecimal digits (_sdf_cache.py:144-215).

    Args:
        vertices: Mesh vertices ``(N, 3)`` with any baked scale applied (hashed as float32).
        indices: Triangle indices (hashed as int32).
        is_solid: ``Mesh.is_solid``.
        narrow_band_range: Signed narrow band (inner, outer) [m].
        target_voxel_size: Target voxel size [m] or ``None``.
        max_resolution: Effective resolution [voxel] or ``None`` (when ``target_voxel_size`` decides).
        margin: AABB padding [m].
        texture_format: ``"float32"``, ``"uint16"`` or ``"uint8"``.
        sign_method_resolved: ``"parity"`` or ``"winding"``.
        winding_threshold: only its sign enters the key.
        scale: Baked scale, or ``None``.
    """
    v = np.asarray(vertices).reshape(-1, 3)
    payload = {
        "kind": _KIND,
        "cache_format_version": CACHE_FORMAT_VERSION,
        "mesh": {
            "vertices_sha256": _digest(v, np.float32),
            "indices_sha256": _digest(indices, np.int32),
            "num_vertices": int(v.shape[0]),
            "num_triangles": int(np.asarray(indices).size // 3),
            "is_solid": bool(is_solid),
        },
        "build_params": {
            "narrow_band_range": [float(narrow_band_range[0]), float(narrow_band_range[1])],
            "target_voxel_size": None if target_voxel_size is None else float(target_voxel_size),
            "max_resolution": None if max_resolution is None else int(max_resolution),
            "margin": float(margin),
            "texture_format": str(texture_format),
            "sign_method_resolved": str(sign_method_resolved),
            "winding_threshold_sign": 1 if winding_threshold >= 0.0 else -1,
            "scale": None if scale is None else [float(scale[0]), float(scale[1]), float(scale[2])],
        },
    }
    encoded = json.dumps(payload, sort_keys=True, separators=(",", ":")).encode("utf-8")
    return hashlib.blake2b(encoded, digest_size=16).hexdigest()


def cache_path(cache_dir: str | os.PathLike[str], hash_hex: str) -> Path:
    """Path of the ``.npz`` file of one key."""
    return Path(cache_dir) / f"{hash_hex}{_NPZ_SUFFIX}"


def save_sparse_data(cache_dir: str | os.PathLike[str], hash_hex: str, sparse: Mapping[str, Any], *,
                     newton_version: str | None = None) -> Path:
    """Write one cooked texture (a ``build_sparse_sdf_from_mesh`` dictionary).

    Creates ``cache_dir`` when missing and raises ``OSError`` on filesystem errors (see :func:`write` for
    the best-effort form).  Returns the published path.
    """
    d = Path(cache_dir)
    d.mkdir(parents=True, exist_ok=True)
    npz_path = cache_path(d, hash_hex)
    arrays = {_VERSION_KEY: np.asarray(CACHE_FORMAT_VERSION, np.int32)}
    for k in _ARRAYS:
        arrays[k] = np.asarray(sparse[k])
    arrays["coarse_dims"] = np.asarray([int(x) for x in sparse["coarse_dims"]], np.int32)
    for k in _INTS:
        arrays[k] = np.asarray(int(sparse[k]), np.int32)
    for k in _FLOATS:
        arrays[k] = np.asarray(float(sparse[k]), np.float32)
    for k in _VEC3:
        arrays[k] = np.asarray(sparse[k], np.float64).reshape(3)
    arrays[_KIND_KEY] = np.asarray(_KIND, dtype=np.str_)
    arrays[_NEWTON_VERSION_KEY] = np.asarray(newton_version if newton_version is not None else _package_version(),
                                             dtype=np.str_)
    arrays[_CREATED_UTC_KEY] = np.asarray(datetime.now(timezone.utc).isoformat(), dtype=np.str_)
    # np.savez appends ".npz" unless the name already ends in it
    tmp = npz_path.parent / f"{npz_path.name}.{os.getpid()}.{secrets.token_hex(8)}.tmp.npz"
    try:
        np.savez(tmp, **arrays)
        try:
            os.replace(tmp, npz_path)
        except OSError as exc:
            if not npz_path.exists():
                raise
            # a peer published the same content-addressed file first
            logger.debug("SDF cache: concurrent publish of %s won by a peer (%s)", npz_path.name, exc)
            with contextlib.suppress(OSError):
                tmp.unlink()
    except BaseException:
        with contextlib.suppress(OSError):
            tmp.unlink()
        raise
    return npz_path


def try_load_sparse_data(cache_dir: str | os.PathLike[str], hash_hex: str) -> dict[str, Any] | None:
    """The cooked texture of one key, or ``None`` when the file is missing, unreadable or of another format
    version.  The arrays are checked against the dimensions they claim, so a truncated or inconsistent file is
    a miss rather than an out-of-range read later."""
    npz_path = cache_path(cache_dir, hash_hex)
    if not npz_path.exists():
        return None
    try:
        with np.load(npz_path, allow_pickle=False) as npz:
            if _VERSION_KEY not in npz.files:
                logger.info("SDF cache: no embedded version, treated as a miss (%s)", npz_path)
                return None
            version = int(npz[_VERSION_KEY].item())
            if version != CACHE_FORMAT_VERSION:
                logger.info("SDF cache: embedded version %d != %d, treated as a miss (%s)", version,
                            CACHE_FORMAT_VERSION, npz_path)
                return None
            data: dict[str, Any] = {k: np.asarray(npz[k]) for k in _ARRAYS}
            data["coarse_dims"] = tuple(int(x) for x in npz["coarse_dims"].reshape(-1))
            for k in _INTS:
                data[k] = int(npz[k].item())
            for k in _FLOATS:
                data[k] = float(npz[k].item())
            for k in _VEC3:
                data[k] = np.asarray(npz[k], np.float64).reshape(3)
    except (OSError, ValueError, KeyError, EOFError, zipfile.BadZipFile) as exc:
        logger.warning("SDF cache: failed to load %s: %s", npz_path, exc)
        return None
    w, h, d = data["coarse_dims"]
    ts = data["subgrid_tex_size"]
    ok = (min(w, h, d) > 0 and data["coarse_sdf"].shape == (d + 1, h + 1, w + 1)
          and data["subgrid_start_slots"].shape == (w, h, d) and data["subgrid_data"].shape == (ts, ts, ts)
          and data["subgrid_data"].dtype in (np.float32, np.uint16, np.uint8)
          and data["subgrid_required"].size == w * h * d and data["subgrid_occupied"].size == w * h * d)
    if ok:
        slots = data["subgrid_start_slots"].astype(np.uint32).reshape(-1)
        used = slots[slots < 0xFFFFFFFE]
        bx, by, bz = used & 0x3FF, (used >> 10) & 0x3FF, (used >> 20) & 0x3FF
        ok = used.size == 0 or int(max(bx.max(), by.max(), bz.max())) * 9 + 9 <= ts
    if not ok:
        logger.warning("SDF cache: arrays inconsistent with their dimensions (%s)", npz_path)
        return None
    return data


def write(cache_dir: str | os.PathLike[str], hash_hex: str, sparse: Mapping[str, Any]) -> None:
    """Best-effort :func:`save_sparse_data`: a filesystem error is logged, never raised."""
    try:
        save_sparse_data(cache_dir, hash_hex, sparse)
    except OSError as exc:
        logger.warning("SDF cache: failed to write %s: %s", cache_dir, exc)


__all__ = ["CACHE_FORMAT_VERSION", "cache_path", "hash_inputs", "save_sparse_data", "try_load_sparse_data", "write"]
