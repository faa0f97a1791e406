"""CSG and offset at the full size of the two-icosphere workload: displaced
icospheres of 7 subdivisions (327,680 triangles each, seeds 3 and 4, the
second moved by (0.3, 0, 0)) voxelized on the GPU at 0.001 with band 3
(about 570,000 leaves each).  The GPU's operands are handed to the oracle
leaf for leaf (masks and all 512 values), and every later result must be
bit-exact: S = A - B and I = A n B (leaves, masks, all values, tiles), then
S.offset(+0.02) and I.offset(-0.01) on each side's own CSG result."""
import time

import numpy as np
import pytest

import baby_shark_amd as bs
from baby_shark_amd import meshgen as mg

from test_volume_ops_gpu import assert_same_volume

pytestmark = pytest.mark.gpu
H = 0.001


@pytest.fixture(scope="module")
def operands(gpu_ctx, oracle_lib):
    m2v = bs.MeshToVolume(gpu_ctx).with_voxel_size(H).with_narrow_band_width(3)
    A = m2v.convert(mg.displaced_icosphere(7, 1.0, 0.05, seed=3))
    B = m2v.convert(mg.displaced_icosphere(7, 1.0, 0.05, seed=4, center=(0.3, 0.0, 0.0)))
    ea, eb = A.export(), B.export()
    print(f"operands: {ea.origins.shape[0]} and {eb.origins.shape[0]} leaves", flush=True)
    assert ea.origins.shape[0] > 500000 and eb.origins.shape[0] > 500000
    yield A, B, ea, eb
    A.free()
    B.free()


def _csg_both(oracle_lib, operands, op):
    A, B, ea, eb = operands
    t0 = time.perf_counter()
    r = getattr(oracle_lib.from_leaves(ea, H), op)(oracle_lib.from_leaves(eb, H))
    t1 = time.perf_counter()
    g = getattr(A.clone(), op)(B.clone())
    print(f"{op}: oracle {t1 - t0:.1f} s, GPU {time.perf_counter() - t1:.3f} s", flush=True)
    return g, r


@pytest.mark.timeout(600)
@pytest.mark.parametrize("op", ["subtract", "intersect"])
def test_full_size_csg(gpu_ctx, oracle_lib, operands, op):
    g, r = _csg_both(oracle_lib, operands, op)
    ge, re_ = g.export(), r.export()
    print(f"{op}: {ge.origins.shape[0]} leaves, {ge.tile_origins.shape[0]} active tiles", flush=True)
    assert_same_volume(ge, re_)


@pytest.mark.timeout(900)
@pytest.mark.parametrize("op,distance", [("intersect", -0.01), ("subtract", 0.02)])
def test_full_size_offset(gpu_ctx, oracle_lib, operands, op, distance):
    g, r = _csg_both(oracle_lib, operands, op)
    t0 = time.perf_counter()
    r = r.offset(distance)
    t1 = time.perf_counter()
    g = g.offset(distance)
    print(f"{op} then offset {distance:+}: oracle {t1 - t0:.1f} s, GPU {time.perf_counter() - t1:.3f} s", flush=True)
    ge, re_ = g.export(), r.export()
    print(f"  {ge.origins.shape[0]} leaves", flush=True)
    assert_same_volume(ge, re_)
