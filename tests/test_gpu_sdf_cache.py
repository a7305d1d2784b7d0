"""GPU tests of Mesh.build_sdf's cache, winding threshold and baked scale.

* A cook with ``cache_dir`` publishes one ``{hash}.sdf.npz``; a second mesh with the same content loads it and
  its texture is bit-identical to the live cook (the reference's test_sdf_disk_cache.py:394-447 compares
  samples within 1e-5); another resolution or texel format is a separate entry.
* An inward-wound mesh cooks (threshold -0.5) to the oracle's texture and to the outward mesh's values.
* A baked-scale SDF (``scale=``) gives contacts identical to the oracle's on the same textures, and the true
  depth of the known stacked-box case to within one unorm16 quantisation step.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

import oracle
from newton_custom_contact_amd import CollisionPipeline
from newton_custom_contact_amd.builder import ModelBuilder
from newton_custom_contact_amd.mesh import Mesh
from parity_helpers import compare_contacts
from tests.mesh_helpers import box_mesh

pytestmark = pytest.mark.gpu


def _arrays(sdf):
    return (sdf.coarse.cpu().numpy(), sdf.slots.cpu().numpy(), sdf.texels.cpu().numpy())


def test_cache_hit_matches_live_cook(cuda_device, tmp_path):
    print("[sdf cache] hit vs live", flush=True)
    live = box_mesh(0.3, 0.2, 0.1).build_sdf(device=cuda_device, cache_dir=tmp_path, max_resolution=32)
    torch.cuda.synchronize()
    files = sorted(tmp_path.glob("*.sdf.npz"))
    assert len(files) == 1 and not list(tmp_path.glob("*.tmp.npz"))
    cached = box_mesh(0.3, 0.2, 0.1).build_sdf(device=cuda_device, cache_dir=tmp_path, max_resolution=32)
    assert cached.coarse.device.type == "cuda" and cached.texels.device.type == "cuda"
    for a, b in zip(_arrays(cached), _arrays(live)):
        np.testing.assert_array_equal(a, b)
    for k in ("box", "quant", "dims"):
        np.testing.assert_array_equal(cached.params[k], live.params[k])
    # a hit does not cook: a corrupted published coarse grid comes back as stored
    with np.load(files[0]) as npz:
        contents = {k: npz[k] for k in npz.files}
    contents["coarse_sdf"] = contents["coarse_sdf"] + np.float32(1.0)
    np.savez(files[0], **contents)
    poisoned = box_mesh(0.3, 0.2, 0.1).build_sdf(device=cuda_device, cache_dir=tmp_path, max_resolution=32)
    np.testing.assert_array_equal(poisoned.coarse.cpu().numpy(), live.coarse.cpu().numpy() + np.float32(1.0))
    box_mesh(0.3, 0.2, 0.1).build_sdf(device=cuda_device, cache_dir=tmp_path, max_resolution=64)
    box_mesh(0.3, 0.2, 0.1).build_sdf(device=cuda_device, cache_dir=tmp_path, max_resolution=32,
                                      texture_format="float32")
    assert len(sorted(tmp_path.glob("*.sdf.npz"))) == 3


def test_inward_winding_cook(cuda_device):
    print("[sdf cook] inward winding", flush=True)
    m = box_mesh(0.3, 0.2, 0.1)
    inv = Mesh(m.vertices, m.indices.reshape(-1, 3)[:, [0, 2, 1]].reshape(-1))
    g = m.build_sdf(device=cuda_device, max_resolution=24, texture_format="float32").sparse_data()
    gi = inv.build_sdf(device=cuda_device, max_resolution=24, texture_format="float32").sparse_data()
    geo = oracle.texture_geometry(inv.vertices, 0.05, (-0.1, 0.1), 24)
    ref = oracle.sdf_build_sparse(inv.vertices, inv.indices, geo, "float32", winding_threshold=-0.5)
    np.testing.assert_array_equal(gi["subgrid_data"], ref["subgrid_data"])
    np.testing.assert_array_equal(gi["coarse_sdf"], ref["coarse_sdf"])
    np.testing.assert_allclose(gi["subgrid_data"], g["subgrid_data"], rtol=0, atol=1e-7)
    np.testing.assert_array_equal(gi["subgrid_start_slots"], g["subgrid_start_slots"])
    assert (gi["coarse_sdf"] < 0).sum() > 0


def test_baked_scale_contacts(cuda_device):
    print("[sdf cook] baked scale contacts", flush=True)
    for baked, depth in ((True, -0.01), (False, -0.005)):
        a = box_mesh(0.2, 0.2, 0.05)
        a.build_sdf(device=cuda_device, max_resolution=16, scale=(1.0, 1.0, 2.0) if baked else None)
        bm = box_mesh(0.1, 0.1, 0.1)
        bm.build_sdf(device=cuda_device, max_resolution=16)
        assert a.sdf.scale_baked == baked
        b = ModelBuilder()
        cfg = ModelBuilder.ShapeConfig()
        cfg.gap = 0.005
        b.add_shape_mesh(-1, mesh=a, scale=(1.0, 1.0, 2.0), cfg=cfg)
        body = b.add_body(xform=(0.0, 0.0, 0.19, 0.0, 0.0, 0.0, 1.0))
        b.add_shape_mesh(body, mesh=bm, cfg=cfg)
        model = b.finalize(cuda_device)
        o = oracle.Oracle(model)
        pipe = CollisionPipeline(model, broad_phase="sap", deterministic=True)
        state = model.state()
        c = pipe.contacts()
        pipe.collide(state, c)
        torch.cuda.synchronize()
        got = c.numpy()
        q = model.host["body_q"]
        ref = o.collide(q, capacity=c.rigid_contact_max)
        rep = compare_contacts(got, ref)
        print("baked" if baked else "unbaked", rep, flush=True)
        sub = ref["sort_key"] & 0x7FFFFF
        mode1 = (sub >> 1) & 1 == 1
        p0 = ref["point0"][mode1].astype(np.float64)
        p1 = ref["point1"][mode1].astype(np.float64) + q[0, :3]
        d = np.einsum("ij,ij->i", p1 - p0, ref["normal"][mode1]) - ref["margin0"][mode1] - ref["margin1"][mode1]
        # the default unorm16 texels truncate: the depth may read up to one quantisation step deeper
        step = float(a.sdf.subgrids_sdf_value_range) / 65535.0
        assert depth - step - 2e-6 <= d.min() <= depth + 2e-6, (d.min(), depth, step)
