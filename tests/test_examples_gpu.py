"""The reference's examples (examples/{dual_contouring,boolean,offset,voxel_remeshing}.rs)
on the GPU path against the oracle, each run whole: volumes and soups are
compared bit for bit (dual contouring for dual_contouring.rs, marching cubes
for boolean.rs and offset.rs).  Mesh conversions start from the oracle's
bunny volume (the GPU's agrees to rounding, checked elsewhere) so that every
later step is compared exactly."""
import os

import numpy as np
import pytest

import baby_shark_amd as bs

from test_offset_gpu import _import
from test_volume_ops_gpu import assert_same_volume
from test_dc_gpu import _same

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
H = 0.2


@pytest.fixture(scope="module")
def bunny(oracle_lib):
    return oracle_lib.mesh_to_volume(bs.read_stl(os.path.join(GOLDEN, "bunny.stl")), H, 0)


def test_dual_contouring_example(gpu_ctx, oracle_lib):
    """examples/dual_contouring.rs: cuboid (0..10)^3 minus sphere r 3 at (8, 8, 8), DC."""
    B = bs.VolumeBuilder(gpu_ctx, H)
    g = B.cuboid((0.0, 0.0, 0.0), (10.0, 10.0, 10.0)).subtract(B.sphere(3.0, (8.0, 8.0, 8.0)))
    r = oracle_lib.cuboid((0.0, 0.0, 0.0), (10.0, 10.0, 10.0), H).subtract(oracle_lib.sphere(3.0, (8.0, 8.0, 8.0), H))
    assert_same_volume(g.export(), r.export())
    got = bs.DualContouringMesher(gpu_ctx).with_voxel_size(H).mesh(g)
    assert got.shape[0] > 10000
    _same(got, r.dual_contour(H))


def test_boolean_example_volumes(gpu_ctx, oracle_lib, bunny):
    """examples/boolean.rs: an empty volume united with 17 boxes, then
    intersect / union / subtract with the bunny, each meshed by marching cubes."""
    B = bs.VolumeBuilder(gpu_ctx, H)
    boxes = bs.Volume.with_voxel_size(H, gpu_ctx)
    rboxes = oracle_lib.from_voxels(np.zeros((0, 3), np.int32), np.zeros(0, np.float32), H)
    for x in range(-25, 26, 3):
        lo, hi = (float(x), -20.0, 0.0), (float(x) + 1.0, 20.0, 50.0)
        boxes = boxes.union(B.cuboid(lo, hi))
        rboxes = rboxes.union(oracle_lib.cuboid(lo, hi, H))
    assert_same_volume(boxes.export(), rboxes.export())
    gb = _import(gpu_ctx, bunny)
    for op in ("intersect", "union", "subtract"):
        g = getattr(gb.clone(), op)(boxes.clone())
        r = getattr(bunny.clone(), op)(rboxes.clone())
        assert_same_volume(g.export(), r.export())
        got = bs.MarchingCubesMesher(gpu_ctx).with_voxel_size(H).mesh(g)
        assert got.shape[0] > 1000
        _same(got, r.marching_cubes(H))


def test_offset_example_volumes(gpu_ctx, oracle_lib, bunny):
    """examples/offset.rs: the bunny offset by -2.5, -1.5, 1.5 and 3.0, each
    meshed by marching cubes."""
    gb = _import(gpu_ctx, bunny)
    for d in (-2.5, -1.5, 1.5, 3.0):
        g, r = gb.clone().offset(d), bunny.clone().offset(d)
        assert_same_volume(g.export(), r.export())
        got = bs.MarchingCubesMesher(gpu_ctx).with_voxel_size(H).mesh(g)
        assert got.shape[0] > 1000
        _same(got, r.marching_cubes(H))


def test_voxel_remeshing_example(gpu_ctx, oracle_lib, tmp_path):
    """examples/voxel_remeshing.rs: box.stl remeshed at voxel 0.01 with the
    default method (marching cubes), written as STL and read back."""
    soup = bs.read_stl(os.path.join(GOLDEN, "box.stl"))
    got = bs.VoxelRemesher(gpu_ctx).with_voxel_size(0.01).remesh(soup)
    e = bs.MeshToVolume(gpu_ctx).with_voxel_size(0.01).with_narrow_band_width(0).convert(soup).export()
    li, off = np.nonzero(e.active())
    idx = e.origins[li] + np.stack([off >> 6, (off >> 3) & 7, off & 7], 1)
    _same(got, oracle_lib.from_voxels(idx, e.values[li, off], 0.01).marching_cubes(0.01))
    path = str(tmp_path / "remeshed.stl")
    bs.write_stl(path, got)
    back = bs.read_stl(path)
    assert np.array_equal(back.reshape(-1).view(np.uint32), got.reshape(-1).view(np.uint32))
