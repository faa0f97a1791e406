"""GPU parity of marching cubes (MarchingCubesMesher, meshing/marching_cubes.rs):
the device soup must equal the oracle's bit for bit, in the same order --
spheres, a mesh volume, CSG and offset results, volumes with active 8^3 and
128^3 tiles (the reference meshes tiles, unlike dual contouring), random
fields that reach every ambiguous case, the reference's own known answer
(box2.stl, h = 0.2, band 0, offset 0.5: 7944 vertices, volume/mod.rs:134-152)
and the voxel remesher's default (Manifold) method."""
import os

import numpy as np
import pytest

import baby_shark_amd as bs
from baby_shark_amd import _lib
from baby_shark_amd import meshgen as mg

from test_offset_gpu import _import

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _same(g, r):
    assert g.shape == r.shape, (g.shape, r.shape)
    bad = np.nonzero(np.any(g.view(np.uint32) != r.view(np.uint32), axis=(1, 2)))[0]
    assert bad.size == 0, f"{bad.size} triangles differ, first {bad[:3]}: {g[bad[:1]]} vs {r[bad[:1]]}"


def _mc(ctx, vol, h):
    return bs.MarchingCubesMesher(ctx).with_voxel_size(h).mesh(vol)


@pytest.mark.parametrize("r,c,h", [(0.5, (0.01, 0.02, -0.03), 0.02), (0.3, (-1.3, 0.7, 2.1), 0.01),
                                   (0.12, (0.0, 0.0, 0.0), 0.05)])
def test_sphere_mc_matches_oracle(gpu_ctx, oracle_lib, r, c, h):
    print(f"  sphere r={r} c={c} h={h}", flush=True)
    g = _mc(gpu_ctx, bs.VolumeBuilder(gpu_ctx, h).sphere(r, c), h)
    assert g.shape[0] > 0
    _same(g, oracle_lib.sphere(r, c, h).marching_cubes(h))


def test_known_answer_box2_offset(gpu_ctx, oracle_lib):
    """volume/mod.rs:134-152: box2.stl, MeshToVolume::default() (band 0) at
    voxel 0.2, offset 0.5, marching cubes: 7944 vertices."""
    soup = bs.read_stl(os.path.join(GOLDEN, "box2.stl"))
    g = bs.MeshToVolume(gpu_ctx).with_voxel_size(0.2).with_narrow_band_width(0).convert(soup).offset(0.5)
    got = _mc(gpu_ctx, g, 0.2)
    assert got.shape[0] * 3 == 7944
    r = oracle_lib.mesh_to_volume(soup, 0.2, 0)
    r.offset(0.5)
    assert r.marching_cubes(0.2).shape[0] * 3 == 7944
    # the GPU's own volume through the oracle's mesher
    e = g.export()
    act = e.active()
    li, off = np.nonzero(act)
    idx = e.origins[li] + np.stack([off >> 6, (off >> 3) & 7, off & 7], 1)
    if e.tile_sizes.size == 0:
        _same(got, oracle_lib.from_voxels(idx, e.values[li, off], 0.2).marching_cubes(0.2))


def test_mesh_volume_csg_offset_and_device_output(gpu_ctx, oracle_lib):
    import torch
    soup = mg.torus(1.0, 0.35, 48, 24, 0.01, seed=1)
    ov = oracle_lib.mesh_to_volume(soup, 0.03, 3)
    gv = _import(gpu_ctx, ov)
    m = bs.MarchingCubesMesher(gpu_ctx).with_voxel_size(0.03)
    g = m.mesh(gv)
    _same(g, ov.marching_cubes(0.03))
    d = m.mesh_device(gv)
    t = torch.as_tensor(d, device="cuda").clone().cpu().numpy()
    d.free()
    _same(t, g)
    assert gpu_ctx.last_mc_stats()["n_tris"] == g.shape[0]
    h = 0.02
    B = bs.VolumeBuilder(gpu_ctx, h)
    gs = B.sphere(0.5, (0.0, 0.0, 0.0)).subtract(B.sphere(0.3, (0.3, 0.0, 0.0))).offset(0.03)
    rs = oracle_lib.sphere(0.5, (0.0, 0.0, 0.0), h).subtract(oracle_lib.sphere(0.3, (0.3, 0.0, 0.0), h)).offset(0.03)
    ref = rs.marching_cubes(h)
    _same(_mc(gpu_ctx, gs, h), ref)
    try:
        gpu_ctx.set_paths(_lib.BS_PATH_FORCE_HASH)
        got = _mc(gpu_ctx, gs, h)
        assert gpu_ctx.last_mc_stats()["used_hash"] == 1
    finally:
        gpu_ctx.set_paths(0)
    _same(got, ref)


@pytest.mark.parametrize("case", ["nested8", "nested128"])
def test_volumes_with_active_tiles(gpu_ctx, oracle_lib, case):
    """Unions that leave active -MAX tiles (8^3, and 128^3 at h = 0.004): the
    reference's tile visitor meets the cubes of each tile's faces."""
    h, r_small, r_big = (0.05, 0.3, 1.0) if case == "nested8" else (0.004, 0.2, 0.9)
    print(f"  {case}", flush=True)
    B = bs.VolumeBuilder(gpu_ctx, h)
    g = B.sphere(r_small, (0.1, 0.0, 0.0)).union(B.sphere(r_big, (0.0, 0.0, 0.0)))
    r = oracle_lib.sphere(r_small, (0.1, 0.0, 0.0), h).union(oracle_lib.sphere(r_big, (0.0, 0.0, 0.0), h))
    assert g.export().tile_sizes.size > 0
    got = _mc(gpu_ctx, g, h)
    assert got.shape[0] > 0
    _same(got, r.marching_cubes(h))


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_random_fields_reach_every_case(gpu_ctx, oracle_lib, seed):
    """Values of random sign and magnitude on a 24^3 block (and a sparse
    scatter of voxels with gaps): every configuration, ambiguous face and
    interior test appears; ties from repeated magnitudes are included."""
    rng = np.random.default_rng(seed)
    n = 24
    g = np.stack(np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij"), -1).reshape(-1, 3) - 5
    mags = rng.choice([0.25, 0.5, 1.0, 2.0], g.shape[0]) if seed == 3 else np.exp(rng.uniform(-6, 2, g.shape[0]))
    vals = (mags * np.where(rng.random(g.shape[0]) < 0.5, -1.0, 1.0)).astype(np.float32)
    keep = rng.random(g.shape[0]) < (0.97 if seed == 2 else 1.0)
    idx, vals = g[keep].astype(np.int32), vals[keep]
    got = _mc(gpu_ctx, bs.volume_from_voxels(idx, vals, 0.1, gpu_ctx), 0.1)
    ref = oracle_lib.from_voxels(idx, vals, 0.1).marching_cubes(0.1)
    print(f"  seed {seed}: {got.shape[0]} triangles", flush=True)
    assert got.shape[0] > 10000
    _same(got, ref)


def test_manifold_remesh_matches_oracle(gpu_ctx, oracle_lib):
    """VoxelRemesher's default method (remeshing/voxel.rs:75-78,92): the soup
    equals the oracle's marching cubes of the GPU's own band-0 volume."""
    for name, h in (("bunny.stl", 0.2), ("box.stl", 0.01)):
        soup = bs.read_stl(os.path.join(GOLDEN, name))
        got = bs.VoxelRemesher(gpu_ctx).with_voxel_size(h).remesh(soup)
        assert got.shape[0] > 0
        e = bs.MeshToVolume(gpu_ctx).with_voxel_size(h).with_narrow_band_width(0).convert(soup).export()
        li, off = np.nonzero(e.active())
        idx = e.origins[li] + np.stack([off >> 6, (off >> 3) & 7, off & 7], 1)
        _same(got, oracle_lib.from_voxels(idx, e.values[li, off], h).marching_cubes(h))


def test_mc_empty_and_errors(gpu_ctx):
    m = bs.MarchingCubesMesher(gpu_ctx).with_voxel_size(0.05)
    assert m.mesh(bs.volume_from_voxels([[0, 0, 0], [1, 0, 0]], [1.0, 1.0], 0.05, gpu_ctx)).shape == (0, 3, 3)
    assert m.mesh(bs.VolumeBuilder(gpu_ctx, 0.05).sphere(-1.0, (0.0, 0.0, 0.0))).shape == (0, 3, 3)
    # a lone pair of voxels: no cube has all eight corners
    assert m.mesh(bs.volume_from_voxels([[0, 0, 0], [1, 0, 0]], [-1.0, 1.0], 0.05, gpu_ctx)).shape == (0, 3, 3)


def test_hash_maps_tiles_far_clusters_and_empty(gpu_ctx, oracle_lib):
    """The leaf map as a hash (forced, with the tile maps beside it, and
    chosen by itself for two clusters 2^21 voxels apart), and an empty
    volume."""
    h = 0.05
    B = bs.VolumeBuilder(gpu_ctx, h)
    g = B.sphere(0.3, (0.1, 0.0, 0.0)).union(B.sphere(1.0, (0.0, 0.0, 0.0)))
    r = oracle_lib.sphere(0.3, (0.1, 0.0, 0.0), h).union(oracle_lib.sphere(1.0, (0.0, 0.0, 0.0), h))
    ref = r.marching_cubes(h)
    try:
        gpu_ctx.set_paths(_lib.BS_PATH_FORCE_HASH)
        got = _mc(gpu_ctx, g, h)
        assert gpu_ctx.last_mc_stats()["used_hash"] == 1
    finally:
        gpu_ctx.set_paths(0)
    _same(got, ref)
    # two clusters far apart: the dense box would exceed 2^28 cells
    n = 12
    grid = np.stack(np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij"), -1).reshape(-1, 3)
    d = np.sqrt(((grid - 5.5) ** 2).sum(1)) - 4.0
    idx = np.concatenate([grid - 1000000, grid + 1000000]).astype(np.int32)
    vals = np.concatenate([d, d]).astype(np.float32)
    got = _mc(gpu_ctx, bs.volume_from_voxels(idx, vals, h, gpu_ctx), h)
    assert gpu_ctx.last_mc_stats()["used_hash"] == 1
    assert got.shape[0] > 0
    _same(got, oracle_lib.from_voxels(idx, vals, h).marching_cubes(h))
    assert _mc(gpu_ctx, bs.Volume.with_voxel_size(h, gpu_ctx), h).shape == (0, 3, 3)
