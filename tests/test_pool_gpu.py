"""The per-device pool of released leaf arrays: a volume built from arrays
that a dropped volume left behind (fully active leaves of a junk value) must
equal, bit for bit, the same volume built in fresh memory -- every producer
writes all of its origins, masks and values -- and emptying the pool keeps
later results unchanged."""
import numpy as np
import pytest

import baby_shark_amd as bs
from baby_shark_amd import meshgen as mg

pytestmark = pytest.mark.gpu
JUNK = np.float32(12345.0)


def _poison(ctx, caps):
    """Leaves the pool holding one set per capacity in `caps`, every mask bit
    set and every value JUNK."""
    g = np.stack(np.meshgrid(np.arange(8), np.arange(8), np.arange(8), indexing="ij"), -1).reshape(-1, 3)
    vols = []
    for n in caps:
        k = np.arange(n)
        idx = (g[None, :, :] + np.stack([8 * k, 0 * k, 0 * k], 1)[:, None, :]).reshape(-1, 3).astype(np.int32)
        vols.append(bs.volume_from_voxels(idx, np.full(idx.shape[0], JUNK, np.float32), 0.05, ctx))
    for v in vols:
        assert v.info()["n_leaves"] > 0
        v.free()


def _same(a, b):
    assert np.array_equal(a.origins, b.origins)
    assert np.array_equal(a.masks, b.masks)
    assert np.array_equal(a.values.view(np.uint32), b.values.view(np.uint32))
    assert np.array_equal(a.tile_origins, b.tile_origins)
    assert np.array_equal(a.tile_sizes, b.tile_sizes)
    assert np.array_equal(a.tile_values.view(np.uint32), b.tile_values.view(np.uint32))


def _producers(ctx):
    torus = mg.torus(1.0, 0.35, 48, 24, 0.01, seed=1)
    B = bs.VolumeBuilder(ctx, 0.02)

    def clone():
        v = B.sphere(0.45, (0.0, 0.1, 0.0))
        c = v.clone()
        v.free()
        return c

    def voxels():
        rng = np.random.default_rng(5)
        idx = rng.integers(-60, 60, (20000, 3)).astype(np.int32)
        idx = np.unique(idx, axis=0)
        return bs.volume_from_voxels(idx, rng.uniform(-1, 1, idx.shape[0]).astype(np.float32), 0.02, ctx)

    return {
        "mesh_to_volume": lambda: bs.MeshToVolume(ctx).with_voxel_size(0.03).with_narrow_band_width(2).convert(torus),
        "mesh_to_volume_shared_signs": lambda: bs.MeshToVolume(ctx).with_voxel_size(0.03).with_narrow_band_width(2)
        .with_shared_signs().convert(torus),
        "sphere": lambda: B.sphere(0.5, (0.0, 0.0, 0.0)),
        "subtract": lambda: B.sphere(0.5, (0.0, 0.0, 0.0)).subtract(B.sphere(0.3, (0.3, 0.0, 0.0))),
        "union_with_tiles": lambda: bs.VolumeBuilder(ctx, 0.05).sphere(0.3, (0.1, 0.0, 0.0))
        .union(bs.VolumeBuilder(ctx, 0.05).sphere(1.0, (0.0, 0.0, 0.0))),
        "densify_tiles": lambda: bs.VolumeBuilder(ctx, 0.05).sphere(0.3, (0.1, 0.0, 0.0))
        .union(bs.VolumeBuilder(ctx, 0.05).sphere(1.0, (0.0, 0.0, 0.0))).densify_tiles(),
        "offset": lambda: B.sphere(0.4, (0.0, 0.0, 0.0)).offset(0.05),
        "clone": clone,
        "from_voxels": voxels,
    }


@pytest.mark.parametrize("name", ["mesh_to_volume", "mesh_to_volume_shared_signs", "sphere", "subtract",
                                  "union_with_tiles", "densify_tiles", "offset", "clone", "from_voxels"])
def test_pooled_arrays_give_identical_volumes(gpu_ctx, name):
    f = _producers(gpu_ctx)[name]
    print(f"  {name}", flush=True)
    bs.release_cached_memory()
    v = f()
    ref = v.export()
    v.free()
    n = ref.origins.shape[0]
    assert n > 0
    bs.release_cached_memory()
    _poison(gpu_ctx, [n, (5 * n + 3) // 4, (8 * n + 4) // 5, 2 * n])
    v = f()
    got = v.export()
    v.free()
    _same(got, ref)
    assert not np.any(got.values == JUNK)
    # a second run from whatever the pool holds now, then one after emptying it
    v = f()
    _same(v.export(), ref)
    v.free()
    bs.release_cached_memory()
    v = f()
    _same(v.export(), ref)
    v.free()
