"""On-disk cache of cooked mesh signed distance grids (newton/_src/geometry/_sdf_cache.py).

The protocol is the reference's: one ``{hash}.sdf.npz`` per cooked grid under ``cache_dir`` (an uncompressed
``np.savez`` bundle), written to a per-writer ``{hash}.sdf.npz.{pid}.{token}.tmp.npz`` and published with
``os.replace`` so that a reader never sees a partial file and concurrent writers of one key never share a
temporary file (_sdf_cache.py:121-212); a replace that loses to a peer publishing the same key is benign.
Load treats a missing file, a file that does not parse and a file whose embedded int32
``__cache_format_version__`` differs from :data:`CACHE_FORMAT_VERSION` as a miss (:215-266); a failed
write is logged and ignored (:269-282).  ``__newton_version__`` and ``__created_utc__`` are diagnostic
strings.  The key is a 128-bit BLAKE2b digest of canonical JSON holding SHA-256 digests of the vertices
(float32) and indices (int32) and the build parameters (:48-112).

The arrays differ because the cooked data differs.  The reference snapshots its sparse narrow-band texture
(coarse grid and 8^3 subgrids, optionally quantized); the cook here is a dense float32 grid:

* ``grid`` -- float32 ``(cells_z + 1, cells_y + 1, cells_x + 1)``, x fastest, negative inside;
* ``lower`` -- float32 ``(3,)``: position of grid vertex (0, 0, 0) in the mesh frame [m];
* ``voxel_size`` -- float32 ``()`` [m];
* ``cells`` -- int32 ``(3,)``: cells along x, y, z;
* ``scale_baked`` -- int32 ``()``: 1 when the shape scale was multiplied into the vertices before the cook.

Since the schema differs, the kind marker is ``"newton_custom_contact_amd.dense_sdf"``; it is part of the
key and, unlike in the reference, load checks it too, so a reference texture cache file kept in the same
directory neither collides with nor loads as one of these.
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
_KIND = "newton_custom_contact_amd.dense_sdf"


def _digest(arr, dtype) -> str:
    return hashlib.sha256(np.ascontiguousarray(arr, dtype=dtype).tobytes()).hexdigest()


def _package_version() -> str:
    try:
        from . import __version__  # noqa: PLC0415

        return str(__version__)
    except ImportError:
        return "unknown"


def hash_inputs(*, vertices: np.ndarray, indices: np.ndarray, is_solid: bool, max_resolution: int, margin: float,
                scale: tuple[float, float, float] | None) -> str:
    """Cache key of a dense SDF cook: 32 hexadecimal digits.

    Args:
        vertices: Mesh vertices ``(N, 3)`` with any baked scale already applied (hashed as float32).
        indices: Triangle indices (hashed as int32).
        is_solid: ``Mesh.is_solid``.
        max_resolution: Effective resolution along the longest padded axis [voxel] (``target_voxel_size``
            resolved into it).
        margin: AABB padding [m].
        scale: Baked scale, or ``None``.

    The sign rule (generalized winding number, threshold sign from the mesh's signed volume) follows from
    the vertices and indices, so it adds nothing to the key; ``narrow_band_range``, ``texture_format`` and
    ``shape_margin`` do not change a dense cook and are left out.
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
            "max_resolution": int(max_resolution),
            "margin": float(margin),
            "scale": None if scale is None else [float(scale[0]), float(scale[1]), float(scale[2])],
        },
    }
    encoded = json.dumps(payload, sort_keys=True, separators=(",", ":")).encode("utf-8")
    return hashlib.blake2b(encoded, digest_size=16).hexdigest()


def cache_path(cache_dir: str | os.PathLike[str], hash_hex: str) -> Path:
    """Path of the ``.npz`` file of one key."""
    return Path(cache_dir) / f"{hash_hex}{_NPZ_SUFFIX}"


def save_sdf_data(cache_dir: str | os.PathLike[str], hash_hex: str, data: Mapping[str, Any], *,
                  newton_version: str | None = None) -> Path:
    """Write one cooked grid (keys ``grid``, ``lower``, ``voxel_size``, ``cells``, ``scale_baked``).

    Creates ``cache_dir`` when missing and raises ``OSError`` on filesystem errors (see :func:`write` for
    the best-effort form).  Returns the published path.
    """
    grid = np.ascontiguousarray(data["grid"], np.float32)
    cells = np.asarray(data["cells"], np.int32).reshape(3)
    if grid.shape != tuple(int(c) + 1 for c in cells[::-1]):
        raise ValueError(f"grid shape {grid.shape} does not match cells {tuple(cells)}")
    d = Path(cache_dir)
    d.mkdir(parents=True, exist_ok=True)
    npz_path = cache_path(d, hash_hex)
    arrays = {
        _VERSION_KEY: np.asarray(CACHE_FORMAT_VERSION, np.int32),
        "grid": grid,
        "lower": np.asarray(data["lower"], np.float32).reshape(3),
        "voxel_size": np.asarray(data["voxel_size"], np.float32),
        "cells": cells,
        "scale_baked": np.asarray(1 if data.get("scale_baked", False) else 0, np.int32),
        _KIND_KEY: np.asarray(_KIND, dtype=np.str_),
        _NEWTON_VERSION_KEY: np.asarray(newton_version if newton_version is not None else _package_version(),
                                        dtype=np.str_),
        _CREATED_UTC_KEY: np.asarray(datetime.now(timezone.utc).isoformat(), dtype=np.str_),
    }
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


def try_load_sdf_data(cache_dir: str | os.PathLike[str], hash_hex: str) -> dict[str, Any] | None:
    """The cooked grid of one key, or ``None`` when the file is missing, unreadable, of another kind or of
    another format version."""
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
            kind = str(npz[_KIND_KEY].item()) if _KIND_KEY in npz.files else ""
            if kind != _KIND:
                logger.info("SDF cache: kind %r is not %r, treated as a miss (%s)", kind, _KIND, npz_path)
                return None
            cells = np.asarray(npz["cells"], np.int32).reshape(3)
            grid = np.ascontiguousarray(npz["grid"], np.float32)
            if grid.shape != tuple(int(c) + 1 for c in cells[::-1]) or np.any(cells <= 0):
                logger.warning("SDF cache: grid shape %s does not match cells %s (%s)", grid.shape, tuple(cells),
                               npz_path)
                return None
            return {
                "grid": grid,
                "lower": np.asarray(npz["lower"], np.float32).reshape(3),
                "voxel_size": np.float32(npz["voxel_size"].item()),
                "cells": cells,
                "scale_baked": bool(int(npz["scale_baked"].item())),
            }
    except (OSError, ValueError, KeyError, EOFError, zipfile.BadZipFile) as exc:
        logger.warning("SDF cache: failed to load %s: %s", npz_path, exc)
        return None


def write(cache_dir: str | os.PathLike[str], hash_hex: str, data: Mapping[str, Any]) -> None:
    """Best-effort :func:`save_sdf_data`: a filesystem error is logged, never raised."""
    try:
        save_sdf_data(cache_dir, hash_hex, data)
    except OSError as exc:
        logger.warning("SDF cache: failed to write %s: %s", cache_dir, exc)


__all__ = ["CACHE_FORMAT_VERSION", "cache_path", "hash_inputs", "save_sdf_data", "try_load_sdf_data", "write"]
