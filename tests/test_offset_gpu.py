"""GPU parity of Volume::offset (remove_if, fast sweeping, shift): leaf keys,
masks and all 512 values per leaf must match the oracle bit for bit."""
import numpy as np
import pytest

import baby_shark_amd as bs
from baby_shark_amd import _lib
from baby_shark_amd import meshgen as mg

from test_volume_ops_gpu import assert_same_volume

pytestmark = pytest.mark.gpu


def _import(ctx, ov):
    """The oracle volume's active voxels inserted into a device volume."""
    e = ov.export()
    li, off = np.nonzero(e.active())
    idx = e.origins[li] + np.stack([off >> 6, (off >> 3) & 7, off & 7], 1)
    return bs.volume_from_voxels(idx, e.values[li, off], e.voxel_size, ctx)


@pytest.mark.parametrize("r,h,D", [(0.5, 0.02, 0.06), (0.5, 0.02, -0.05), (0.4, 0.01, 0.1), (0.3, 0.02, 0.0),
                                   (0.3, 0.02, -0.0), (0.2, 0.05, 0.5)])
def test_sphere_offset_matches_oracle(gpu_ctx, oracle_lib, r, h, D):
    c = (0.013, -0.02, 0.007)
    g = bs.VolumeBuilder(gpu_ctx, h).sphere(r, c).offset(D)
    st = gpu_ctx.last_offset_stats()
    ref = oracle_lib.sphere(r, c, h).offset(D)
    ge = g.export()
    assert_same_volume(ge, ref.export())
    assert st["n_out"] == ge.origins.shape[0] and st["n_pool"] >= st["n_out"]


def test_torus_offset_both_ways(gpu_ctx, oracle_lib):
    soup = mg.torus(1.0, 0.35, 48, 24, 0.01, seed=1)
    for D in (0.04, -0.03):
        ov = oracle_lib.mesh_to_volume(soup, 0.02, 3)
        g = _import(gpu_ctx, ov).offset(D)
        assert_same_volume(g.export(), ov.offset(D).export())


def test_offset_after_csg_and_flood_fill_signs(gpu_ctx, oracle_lib):
    """Subtract then offset (the CSG's -MAX tiles are dropped by remove_if),
    then sign_at after a flood fill of the result."""
    h = 0.02
    B = bs.VolumeBuilder(gpu_ctx, h)
    g = B.sphere(0.5, (0.0, 0.0, 0.0)).subtract(B.sphere(0.3, (0.3, 0.0, 0.0))).offset(0.05)
    r = oracle_lib.sphere(0.5, (0.0, 0.0, 0.0), h).subtract(oracle_lib.sphere(0.3, (0.3, 0.0, 0.0), h)).offset(0.05)
    assert_same_volume(g.export(), r.export())
    g.flood_fill()
    r.flood_fill()
    ax = np.arange(-40, 40, 3, dtype=np.int32)
    q = np.stack(np.meshgrid(ax, ax, ax, indexing="ij"), -1).reshape(-1, 3)
    assert np.array_equal(g.signs_at(q), r.signs_at(q))
    # union with a nested sphere leaves active -MAX tiles; offset removes them
    g = B.sphere(0.3, (0.1, 0.0, 0.0)).union(B.sphere(1.0, (0.0, 0.0, 0.0)))
    r = oracle_lib.sphere(0.3, (0.1, 0.0, 0.0), h).union(oracle_lib.sphere(1.0, (0.0, 0.0, 0.0), h))
    assert g.export().tile_values.size > 0
    g = g.offset(-0.04)
    r = r.offset(-0.04)
    ge = g.export()
    assert ge.tile_values.size == 0
    assert_same_volume(ge, r.export())


def test_offset_hash_leaf_map_matches_grid(gpu_ctx, oracle_lib):
    h = 0.02
    ref = oracle_lib.sphere(0.45, (0.0, 0.1, 0.0), h).offset(0.07).export()
    try:
        gpu_ctx.set_paths(_lib.BS_PATH_FORCE_HASH)
        g = bs.VolumeBuilder(gpu_ctx, h).sphere(0.45, (0.0, 0.1, 0.0)).offset(0.07)
        assert gpu_ctx.last_offset_stats()["used_hash"] == 1
    finally:
        gpu_ctx.set_paths(0)
    assert_same_volume(g.export(), ref)


def test_offset_of_subtracted_icospheres(gpu_ctx, oracle_lib):
    """The two-icosphere workload at reduced size: S = A - B offset by +D and
    I = A n B offset by -D/2 (displaced icospheres of 5 subdivisions, the
    second moved by 0.3 in x, h = 0.01).  Both sides start from the oracle's
    mesh volumes, so the comparison is exact."""
    h = 0.01
    a = mg.displaced_icosphere(5, 1.0, 0.05, seed=3)
    b = mg.displaced_icosphere(5, 1.0, 0.05, seed=4, center=(0.3, 0.0, 0.0))
    oa, ob = oracle_lib.mesh_to_volume(a, h, 3), oracle_lib.mesh_to_volume(b, h, 3)
    S = _import(gpu_ctx, oa).subtract(_import(gpu_ctx, ob)).offset(0.02)
    rS = oa.clone().subtract(ob.clone()).offset(0.02)
    assert_same_volume(S.export(), rS.export())
    I = _import(gpu_ctx, oa).intersect(_import(gpu_ctx, ob)).offset(-0.01)
    rI = oa.intersect(ob).offset(-0.01)
    assert_same_volume(I.export(), rI.export())


@pytest.mark.parametrize("r,h,c,D,bx", [
    (0.3, 0.02, (2.2, 0.013, -0.011), 0.1, 128),      # grows into a fresh 128^3 internal node
    (3.0, 0.25, (1020.5, 0.1, -0.2), 1.0, 4096),      # grows into a fresh 4096^3 internal node
    (0.44, 0.02, (-0.5, 0.2, 0.1), 0.1, 0),           # grows across the origin plane
])
def test_offset_grows_into_new_blocks(gpu_ctx, oracle_lib, r, h, c, D, bx):
    """A sphere whose surface sits a few voxels below a block boundary on the
    x axis, offset outward across it: the output tree gains nodes the input did
    not have.  The grown volume, its marching-cubes soup, and signs after a
    flood fill over a grid covering the new block all equal the oracle's."""
    from test_dc_gpu import _same
    src = oracle_lib.sphere(r, c, h).export()
    assert src.origins[:, 0].max() < bx
    g = bs.VolumeBuilder(gpu_ctx, h).sphere(r, c).offset(D)
    ref = oracle_lib.sphere(r, c, h).offset(D)
    ge = g.export()
    assert ge.origins[:, 0].max() >= bx
    assert_same_volume(ge, ref.export())
    _same(bs.MarchingCubesMesher(gpu_ctx).with_voxel_size(h).mesh(g), ref.marching_cubes(h))
    g.flood_fill()
    ref.flood_fill()
    rv = int(round((r + abs(D)) / h)) + 6
    cx, cy, cz = (int(round(v / h)) for v in c)
    ax = np.arange(bx - 20, bx + 12, dtype=np.int32)
    ay = np.arange(cy - rv, cy + rv, 2, dtype=np.int32)
    az = np.arange(cz - rv, cz + rv, 2, dtype=np.int32)
    q = np.stack(np.meshgrid(ax, ay, az, indexing="ij"), -1).reshape(-1, 3)
    gs, rs = g.signs_at(q), ref.signs_at(q)
    assert np.array_equal(gs, rs)
    assert gs.min() != gs.max()      # the grid holds inside and outside voxels
