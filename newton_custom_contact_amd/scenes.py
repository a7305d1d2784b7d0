"""Seeded synthetic scenes for the rigid-contact workloads.

Every scene is built through :class:`ModelBuilder` from a ``numpy.random.Generator(PCG64(seed))``.

* ``box_pile``      1 world, N boxes (edges U[0.1, 0.5] m per axis, density 1000 kg/m^3) at uniform
                    positions in [-1, 1]^2 x [0.5, 2.5] m with uniform random orientations (Shoemake),
                    friction 0.5 on every shape including the infinite ground plane.

Positions are rejection-sampled so that no two bounding spheres overlap at t = 0 (a dropped pile
starts separated; position-based contact resolves any initial overlap d by injecting a velocity of
about d/dt, so overlapping starts are not a drop test).  When 2000 draws find no free spot the
sampling box grows 0.25 m upward.
* ``mixed_worlds``  W worlds sharing one set of 32 primitive sizes (11 spheres r U[0.05, 0.2], 11 boxes
                    as above, 10 capsules r U[0.05, 0.15], full cylinder length h U[0.1, 0.4] so
                    half_height = h/2), each world with its own seeded poses in the same volume, one
                    global ground plane, friction 0.5.
* ``uneven_worlds`` W worlds with log-uniform body counts in [lo, hi] (same primitive mix and
                    distributions), optionally a revolute chain of capsule links per world.
* ``mesh_worlds``   W worlds of 8 rigid bodies each, every body a convex-hull triangle mesh drawn from a
                    seeded library (hull of 64-256 points on a sphere of radius U[0.1, 0.3] m, 64^3 SDF),
                    placed like the primitives above an infinite ground plane, friction 0.5.
"""

from __future__ import annotations

import math

import numpy as np

from ._types import quat_from_axis_angle, quat_mul, transform
from .builder import ModelBuilder


def shoemake_quat(rng: np.random.Generator) -> np.ndarray:
    """Uniform random unit quaternion (Shoemake 1992), (x, y, z, w)."""
    u1, u2, u3 = rng.random(3)
    a = math.sqrt(1.0 - u1)
    b = math.sqrt(u1)
    return np.array(
        [a * math.sin(2 * math.pi * u2), a * math.cos(2 * math.pi * u2), b * math.sin(2 * math.pi * u3),
         b * math.cos(2 * math.pi * u3)]
    )


class _Placer:
    """Sequential rejection sampling of non-overlapping bounding spheres in [-1,1]^2 x [0.5, z_hi]."""

    def __init__(self, rng: np.random.Generator, z_hi: float = 2.5):
        self.rng = rng
        self.z_hi = z_hi
        self._pts = np.zeros((16, 3))
        self._rad = np.zeros(16)
        self.n = 0

    @property
    def pts(self) -> np.ndarray:
        return self._pts[:self.n]

    @property
    def rad(self) -> np.ndarray:
        return self._rad[:self.n]

    _MOD = 1 << 128  # PCG64 advances modulo its 128-bit state: advancing by _MOD - n steps it back n draws

    def place(self, radius: float) -> np.ndarray:
        """Up to 2000 candidates (x, y, z uniform, drawn in that order) per height band; the first whose sphere
        clears every placed one is taken.  Candidates are tested in batches; the generator is then stepped back to
        just after the taken candidate's three draws, as if they had been drawn one by one."""
        bg = self.rng.bit_generator
        while True:
            tried = 0
            while tried < 2000:
                k = min(64, 2000 - tried)
                u = self.rng.random((k, 3))
                cand = np.empty((k, 3))
                cand[:, 0] = -1.0 + 2.0 * u[:, 0]
                cand[:, 1] = -1.0 + 2.0 * u[:, 1]
                cand[:, 2] = 0.5 + (self.z_hi - 0.5) * u[:, 2]
                if self.n:
                    diff = self.pts[None, :, :] - cand[:, None, :]
                    ok = np.all(np.sqrt(np.einsum("kni,kni->kn", diff, diff)) > radius + self.rad, axis=1)
                else:
                    ok = np.ones(k, bool)
                hit = np.flatnonzero(ok)
                if hit.size:
                    i = int(hit[0])
                    back = 3 * (k - i - 1)
                    if back:
                        bg.advance(self._MOD - back)
                    p = cand[i].copy()
                    if self.n == len(self._rad):
                        self._pts = np.concatenate([self._pts, np.zeros_like(self._pts)])
                        self._rad = np.concatenate([self._rad, np.zeros_like(self._rad)])
                    self._pts[self.n] = p
                    self._rad[self.n] = radius
                    self.n += 1
                    return p
                tried += k
            self.z_hi += 0.25


def _bound_radius(kind, dims) -> float:
    if kind == "sphere":
        return float(dims[0])
    if kind == "box":
        return float(np.linalg.norm(dims))
    return float(dims[0] + dims[1])


def _cfg(mu: float) -> ModelBuilder.ShapeConfig:
    c = ModelBuilder.ShapeConfig()
    c.mu = mu
    return c


def box_pile(n_boxes: int = 64, seed: int = 0, mu: float = 0.5, device=None, world: bool = True):
    """One world of ``n_boxes`` seeded boxes above an infinite ground plane."""
    rng = np.random.Generator(np.random.PCG64(seed))
    b = ModelBuilder()
    b.add_ground_plane(cfg=_cfg(mu))
    if world:
        b.begin_world()
    cfg = _cfg(mu)
    placer = _Placer(rng)
    for _ in range(n_boxes):
        he = 0.5 * rng.uniform(0.1, 0.5, size=3)
        p = placer.place(float(np.linalg.norm(he)))
        q = shoemake_quat(rng)
        body = b.add_body(xform=transform(p, q))
        b.add_shape_box(body, hx=he[0], hy=he[1], hz=he[2], cfg=cfg)
    if world:
        b.end_world()
    return b.finalize(device)


def box_stacks(n_stacks: int = 4, height: int = 6, seed: int = 0, mu: float = 0.5, device=None):
    """One world of ``n_stacks`` towers of ``height`` seeded boxes resting on an infinite ground plane, 1 m apart:
    each box (half extents 0.06-0.15 m across, 0.04-0.10 m tall) sits on the one below at a seeded horizontal
    offset of up to a fifth of the smaller half extent and a seeded turn about the vertical."""
    rng = np.random.Generator(np.random.PCG64(seed))
    b = ModelBuilder()
    b.add_ground_plane(cfg=_cfg(mu))
    b.begin_world()
    cfg = _cfg(mu)
    for k in range(n_stacks):
        x0, y0, z = float(k % 2), float(k // 2), 0.0
        for _ in range(height):
            hx, hy = rng.uniform(0.06, 0.15, size=2)
            hz = rng.uniform(0.04, 0.10)
            dx, dy = rng.uniform(-0.2, 0.2, size=2) * min(hx, hy)
            yaw = rng.uniform(-np.pi, np.pi)
            q = (0.0, 0.0, float(np.sin(0.5 * yaw)), float(np.cos(0.5 * yaw)))
            body = b.add_body(xform=transform((x0 + dx, y0 + dy, z + hz), q))
            b.add_shape_box(body, hx=hx, hy=hy, hz=hz, cfg=cfg)
            z += 2.0 * hz
    b.end_world()
    return b.finalize(device)


def _primitive_sizes(rng, n_sphere, n_box, n_capsule):
    sizes = []
    for _ in range(n_sphere):
        sizes.append(("sphere", (rng.uniform(0.05, 0.2),)))
    for _ in range(n_box):
        sizes.append(("box", tuple(0.5 * rng.uniform(0.1, 0.5, size=3))))
    for _ in range(n_capsule):
        sizes.append(("capsule", (rng.uniform(0.05, 0.15), 0.5 * rng.uniform(0.1, 0.4))))
    return sizes


def _add_primitive(b: ModelBuilder, kind, dims, p, q, cfg):
    body = b.add_body(xform=transform(p, q))
    if kind == "sphere":
        b.add_shape_sphere(body, radius=dims[0], cfg=cfg)
    elif kind == "box":
        b.add_shape_box(body, hx=dims[0], hy=dims[1], hz=dims[2], cfg=cfg)
    else:
        b.add_shape_capsule(body, radius=dims[0], half_height=dims[1], cfg=cfg)
    return body


def mixed_worlds(world_count: int = 4096, seed: int = 0, mu: float = 0.5, device=None,
                 n_sphere: int = 11, n_box: int = 11, n_capsule: int = 10, fast: bool = True):
    """``world_count`` worlds of the same 32 primitives with per-world seeded poses.

    ``fast`` builds one template world through the builder and tiles it (Model._replicate) with the
    poses drawn in exactly the order the per-world builder loop draws them, so both paths give the
    same model.
    """
    rng = np.random.Generator(np.random.PCG64(seed))
    sizes = _primitive_sizes(rng, n_sphere, n_box, n_capsule)
    if fast:
        from .model import Model  # noqa: PLC0415

        poses = np.zeros((world_count, len(sizes), 7))
        for w in range(world_count):
            placer = _Placer(rng)
            for i, (kind, dims) in enumerate(sizes):
                poses[w, i, :3] = placer.place(_bound_radius(kind, dims))
                poses[w, i, 3:] = shoemake_quat(rng)
        t = ModelBuilder()
        cfg = _cfg(mu)
        t.add_ground_plane(cfg=_cfg(mu))
        t.begin_world()
        for kind, dims in sizes:
            _add_primitive(t, kind, dims, np.zeros(3), np.array([0.0, 0.0, 0.0, 1.0]), cfg)
        t.end_world()
        template = t.finalize("cpu")
        model = Model._replicate(template, world_count, poses, device)
        model.template = template
        return model
    b = ModelBuilder()
    cfg = _cfg(mu)
    b.add_ground_plane(cfg=_cfg(mu))
    for _w in range(world_count):
        b.begin_world()
        placer = _Placer(rng)
        for kind, dims in sizes:
            p = placer.place(_bound_radius(kind, dims))
            _add_primitive(b, kind, dims, p, shoemake_quat(rng), cfg)
        b.end_world()
    return b.finalize(device)


def uneven_worlds(world_count: int = 2048, seed: int = 0, mu: float = 0.5, device=None, lo: int = 8,
                  hi: int = 256, chain_links: int = 6, chain_limit: float = 0.5 * math.pi, progress=None):
    """Worlds with log-uniform body counts in [lo, hi] plus an optional revolute capsule chain.
    ``progress(worlds_done, world_count)`` is called every 128 worlds when given."""
    rng = np.random.Generator(np.random.PCG64(seed))
    b = ModelBuilder()
    cfg = _cfg(mu)
    b.add_ground_plane(cfg=_cfg(mu))
    for _w in range(world_count):
        n = int(math.floor(math.exp(rng.uniform(math.log(lo), math.log(hi)))))
        n = max(lo, min(hi, n))
        ns = n // 3 + (1 if n % 3 > 0 else 0)
        nb = n // 3 + (1 if n % 3 > 1 else 0)
        nc = n - ns - nb
        b.begin_world()
        placer = _Placer(rng)
        for kind, dims in _primitive_sizes(rng, ns, nb, nc):
            p = placer.place(_bound_radius(kind, dims))
            _add_primitive(b, kind, dims, p, shoemake_quat(rng), cfg)
        if chain_links > 0:
            # hang the chain clear of the pile: a dense world's placer raises its band above z = 3
            top = float(np.max(placer.pts[:, 2] + placer.rad)) if len(placer.rad) else 0.0
            _add_chain(b, rng, chain_links, chain_limit, cfg, z=max(3.0, top + 0.1))
        b.end_world()
        if progress is not None and (_w + 1) % 128 == 0:
            progress(_w + 1, world_count)
    return b.finalize(device)


def _add_chain(b: ModelBuilder, rng, links: int, limit: float, cfg, z: float = 3.0):
    """A revolute chain of capsule links hanging from a world anchor at height ``z`` above the pile."""
    r, hh = 0.04, 0.1
    base = np.array([rng.uniform(-0.5, 0.5), rng.uniform(-0.5, 0.5), z])
    # links lie along +X; each capsule's local +Z axis is rotated onto X
    q_link = quat_from_axis_angle((0.0, 1.0, 0.0), 0.5 * math.pi)
    prev = -1
    seg = 2.0 * (hh + r)
    for i in range(links):
        c = base + np.array([(i + 0.5) * seg, 0.0, 0.0])
        body = b.add_link(xform=transform(c, (0.0, 0.0, 0.0, 1.0)))
        b.add_shape_capsule(body, xform=transform((0.0, 0.0, 0.0), q_link), radius=r, half_height=hh, cfg=cfg)
        parent_xform = transform(base if prev < 0 else (0.5 * seg, 0.0, 0.0), (0.0, 0.0, 0.0, 1.0))
        j = b.add_joint_revolute(prev, body, parent_xform=parent_xform,
                                 child_xform=transform((-0.5 * seg, 0.0, 0.0), (0.0, 0.0, 0.0, 1.0)),
                                 axis=(0.0, 1.0, 0.0), limit_lower=-limit, limit_upper=limit)
        if i == 0:
            b.add_articulation([j])
        prev = body
    return prev


__all__ = ["box_pile", "mixed_worlds", "uneven_worlds", "shoemake_quat", "quat_mul"]


def mesh_library(count: int = 64, seed: int = 0, n_min: int = 64, n_max: int = 256, r_min: float = 0.1,
                 r_max: float = 0.3):
    """``count`` seeded convex-hull meshes (points on a sphere of radius U[r_min, r_max])."""
    from .mesh import random_convex_hull  # noqa: PLC0415

    rng = np.random.Generator(np.random.PCG64(seed))
    return [random_convex_hull(rng, n_min, n_max, r_min, r_max) for _ in range(count)]


def mesh_worlds(world_count: int = 1024, seed: int = 0, mu: float = 0.5, device=None, bodies: int = 8,
                library=None, library_size: int = 64, sdf_resolution: int = 64, texture_format: str = "float32",
                unique: bool = False):
    """``world_count`` worlds of ``bodies`` convex-hull mesh bodies above an infinite ground plane.

    Each body's mesh is drawn from ``library`` (default: ``mesh_library(library_size, seed)``); worlds
    share the library's meshes and SDFs.  ``unique=True`` gives every body a seeded hull of its own instead
    (``mesh_library(world_count * bodies, seed)``, body i takes hull i).  On a GPU device the SDFs are cooked
    at ``sdf_resolution`` cells along each mesh's longest axis, with ``texture_format`` texels (float32: the
    64^3 FP32 voxel SDF of the 1024-world hull workload), before the model is finalized.
    """
    if unique and library is None:
        library = mesh_library(world_count * bodies, seed)
    lib = mesh_library(library_size, seed) if library is None else library
    dev = device if device is not None else "cpu"
    if str(dev).startswith("cuda"):
        from .mesh import cook_texture_sdfs  # noqa: PLC0415

        cook_texture_sdfs(lib, dev, max_resolution=sdf_resolution, texture_format=texture_format)
    rng = np.random.Generator(np.random.PCG64(seed + 1))
    b = ModelBuilder()
    cfg = _cfg(mu)
    b.add_ground_plane(cfg=_cfg(mu))
    radius = [float(np.linalg.norm(mm.vertices, axis=1).max()) for mm in lib]
    for _w in range(world_count):
        b.begin_world()
        placer = _Placer(rng)
        pick = rng.integers(0, len(lib), size=bodies)
        if unique:
            pick = np.arange(_w * bodies, (_w + 1) * bodies)
        for k in pick:
            p = placer.place(radius[k])
            body = b.add_body(xform=transform(p, shoemake_quat(rng)))
            b.add_shape_mesh(body, mesh=lib[k], cfg=cfg)
        b.end_world()
    return b.finalize(device)
