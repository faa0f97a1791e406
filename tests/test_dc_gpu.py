"""GPU parity of dual contouring: the device soup must equal the oracle's
bit for bit, in the same (sequential) order."""
import numpy as np
import pytest

import baby_shark_amd as bs
from baby_shark_amd import _lib
from baby_shark_amd import meshgen as mg

from test_offset_gpu import _import
from test_volume_ops_gpu import assert_same_volume

pytestmark = pytest.mark.gpu


def _same(g, r):
    assert g.shape == r.shape, (g.shape, r.shape)
    bad = np.nonzero(np.any(g.view(np.uint32) != r.view(np.uint32), axis=(1, 2)))[0]
    assert bad.size == 0, f"{bad.size} triangles differ, first {bad[:3]}: {g[bad[:1]]} vs {r[bad[:1]]}"


@pytest.mark.parametrize("r,c,h", [(0.5, (0.01, 0.02, -0.03), 0.02), (0.3, (-1.3, 0.7, 2.1), 0.01),
                                   (0.12, (0.0, 0.0, 0.0), 0.05)])
def test_sphere_dc_matches_oracle(gpu_ctx, oracle_lib, r, c, h):
    g = bs.DualContouringMesher(gpu_ctx).with_voxel_size(h).mesh(bs.VolumeBuilder(gpu_ctx, h).sphere(r, c))
    ref = oracle_lib.sphere(r, c, h).dual_contour(h)
    assert g.shape[0] > 0
    _same(g, ref)


def test_mesh_volume_dc_and_device_output(gpu_ctx, oracle_lib):
    import torch
    soup = mg.torus(1.0, 0.35, 48, 24, 0.01, seed=1)
    ov = oracle_lib.mesh_to_volume(soup, 0.03, 3)
    gv = _import(gpu_ctx, ov)
    m = bs.DualContouringMesher(gpu_ctx).with_voxel_size(0.03)
    g = m.mesh(gv)
    _same(g, ov.dual_contour(0.03))
    d = m.mesh_device(gv)
    t = torch.as_tensor(d, device="cuda").clone().cpu().numpy()
    d.free()
    _same(t, g)
    st = gpu_ctx.last_dc_stats()
    assert st["n_tris"] == g.shape[0]


def test_dc_after_offset_and_hash_map(gpu_ctx, oracle_lib):
    h = 0.02
    B = bs.VolumeBuilder(gpu_ctx, h)
    g = B.sphere(0.5, (0.0, 0.0, 0.0)).subtract(B.sphere(0.3, (0.3, 0.0, 0.0))).offset(0.03)
    r = oracle_lib.sphere(0.5, (0.0, 0.0, 0.0), h).subtract(oracle_lib.sphere(0.3, (0.3, 0.0, 0.0), h)).offset(0.03)
    ref = r.dual_contour(h)
    _same(bs.DualContouringMesher(gpu_ctx).with_voxel_size(h).mesh(g), ref)
    try:
        gpu_ctx.set_paths(_lib.BS_PATH_FORCE_HASH)
        got = bs.DualContouringMesher(gpu_ctx).with_voxel_size(h).mesh(g)
        assert gpu_ctx.last_dc_stats()["used_hash"] == 1
    finally:
        gpu_ctx.set_paths(0)
    _same(got, ref)


def test_dc_error_cases(gpu_ctx, oracle_lib):
    m = bs.DualContouringMesher(gpu_ctx).with_voxel_size(0.05)
    # lone crossing voxels: the reference reaches unreachable!()
    v = bs.volume_from_voxels([[0, 0, 0], [1, 0, 0]], [-1.0, 1.0], 0.05, gpu_ctx)
    with pytest.raises(bs.BSError) as e:
        m.mesh(v)
    assert e.value.status == _lib.BS_ERR_INTERNAL
    # active tiles (a union with a nested sphere): the reference stops at todo!()
    B = bs.VolumeBuilder(gpu_ctx, 0.05)
    u = B.sphere(0.3, (0.1, 0.0, 0.0)).union(B.sphere(1.0, (0.0, 0.0, 0.0)))
    with pytest.raises(bs.BSError) as e:
        m.mesh(u)
    assert e.value.status == _lib.BS_ERR_UNSUPPORTED_TILE
    # nothing to contour
    assert m.mesh(bs.volume_from_voxels([[0, 0, 0], [1, 0, 0]], [1.0, 1.0], 0.05, gpu_ctx)).shape == (0, 3, 3)
    assert m.mesh(B.sphere(-1.0, (0.0, 0.0, 0.0))).shape == (0, 3, 3)


def test_densify_then_dc_matches_oracle(gpu_ctx, oracle_lib):
    """Two nested spheres and the torus-and-sphere workload (torus 128x64,
    h = 0.02, band 3, united with the sphere r = 0.5 at (1, 0, 0)): the union
    leaves active 8^3 tiles; densified, both paths contour the same soup."""
    h = 0.05
    B = bs.VolumeBuilder(gpu_ctx, h)
    g = B.sphere(0.3, (0.1, 0.0, 0.0)).union(B.sphere(1.0, (0.0, 0.0, 0.0)))
    r = oracle_lib.sphere(0.3, (0.1, 0.0, 0.0), h).union(oracle_lib.sphere(1.0, (0.0, 0.0, 0.0), h))
    assert g.export().tile_sizes.size > 0
    g.densify_tiles()
    r.densify_tiles()
    assert_same_volume(g.export(), r.export())
    _same(bs.DualContouringMesher(gpu_ctx).with_voxel_size(h).mesh(g), r.dual_contour(h))

    soup = mg.torus(1.0, 0.35, 128, 64, 0.01, seed=1)
    ov = oracle_lib.mesh_to_volume(soup, 0.02, 3)
    g = _import(gpu_ctx, ov).union(bs.VolumeBuilder(gpu_ctx, 0.02).sphere(0.5, (1.0, 0.0, 0.0))).densify_tiles()
    r = ov.union(oracle_lib.sphere(0.5, (1.0, 0.0, 0.0), 0.02)).densify_tiles()
    assert_same_volume(g.export(), r.export())
    got = bs.DualContouringMesher(gpu_ctx).with_voxel_size(0.02).mesh(g)
    assert got.shape[0] > 100000
    _same(got, r.dual_contour(0.02))


def test_densify_node_level_tiles(gpu_ctx, oracle_lib):
    """A 128^3 tile (the union of a small sphere with one of 225 voxels
    radius) becomes 4096 leaves; a volume without tiles is left alone."""
    h = 0.004
    B = bs.VolumeBuilder(gpu_ctx, h)
    g = B.sphere(0.2, (0.1, 0.0, 0.0)).union(B.sphere(0.9, (0.0, 0.0, 0.0)))
    r = oracle_lib.sphere(0.2, (0.1, 0.0, 0.0), h).union(oracle_lib.sphere(0.9, (0.0, 0.0, 0.0), h))
    assert set(g.export().tile_sizes.tolist()) == {128}
    assert_same_volume(g.densify_tiles().export(), r.densify_tiles().export())
    s = bs.VolumeBuilder(gpu_ctx, 0.05).sphere(0.4, (0.0, 0.0, 0.0))
    before = s.export()
    assert_same_volume(s.densify_tiles().export(), before)
