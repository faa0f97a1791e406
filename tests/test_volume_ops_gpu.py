"""GPU parity of the volume operations: sphere sampling, flood fill, CSG,
sign_at, import of inserted voxels.  Leaf keys, masks, every one of the 512
values per leaf (active or not) and the active tiles must match the oracle
bit for bit."""
import numpy as np
import pytest

import baby_shark_amd as bs
from baby_shark_amd import meshgen as mg

pytestmark = pytest.mark.gpu

FMAX = np.float32(3.4028235e38)


def _tiles(e):
    if e.tile_origins.shape[0] == 0:
        return np.zeros((0, 5), np.int64)
    t = np.concatenate([e.tile_origins.astype(np.int64), e.tile_sizes[:, None].astype(np.int64),
                        e.tile_values.view(np.uint32)[:, None].astype(np.int64)], axis=1)
    return t[np.lexsort(t.T[::-1])]


def assert_same_volume(g, r):
    """Bit-exact equality of two exports (GPU and oracle)."""
    assert g.origins.shape == r.origins.shape, (g.origins.shape, r.origins.shape)
    assert np.array_equal(g.origins, r.origins), "leaf origins differ"
    assert np.array_equal(g.masks, r.masks), "masks differ"
    bad = np.nonzero(g.values.view(np.uint32) != r.values.view(np.uint32))
    assert bad[0].size == 0, f"{bad[0].size} values differ, first leaf {bad[0][:3]} offset {bad[1][:3]}: " \
                             f"{g.values[bad][:3]} vs {r.values[bad][:3]}"
    assert np.array_equal(_tiles(g), _tiles(r)), "tiles differ"


SPHERES = [
    (0.5, (1.0, 0.0, 0.0), 0.02),        # the sphere of the torus-and-sphere workload
    (0.4, (0.1, -0.2, 0.3), 0.05),
    (0.9, (0.01, -0.02, 0.03), 0.004),
    (0.5, (-1.3, -0.7, -2.1), 0.02),
    (0.015, (0.0, 0.0, 0.0), 0.02),      # radius below one voxel
    (-0.1, (0.0, 0.0, 0.0), 0.05),       # negative radius: no leaf at all
    (3.0, (40.96, 40.96, 40.96), 0.01),  # straddles the 4096-voxel root blocks
]


@pytest.mark.parametrize("r,c,h", SPHERES)
def test_sphere_matches_oracle(gpu_ctx, oracle_lib, r, c, h):
    g = bs.VolumeBuilder(gpu_ctx, h).sphere(r, c).export()
    ref = oracle_lib.sphere(r, c, h).export()
    assert_same_volume(g, ref)
    if r < 0:
        assert g.origins.shape[0] == 0


def test_flood_fill_sphere_and_mesh(gpu_ctx, oracle_lib):
    g = bs.VolumeBuilder(gpu_ctx, 0.02).sphere(0.5, (1.0, 0.0, 0.0)).flood_fill().export()
    ref = oracle_lib.sphere(0.5, (1.0, 0.0, 0.0), 0.02).flood_fill().export()
    assert_same_volume(g, ref)
    # a mesh volume: its values differ from the oracle's by rounding only, so
    # flood fill the oracle's own values imported into the GPU path
    soup = mg.torus(1.0, 0.35, 48, 24, 0.01, seed=1)
    ov = oracle_lib.mesh_to_volume(soup, 0.04, 3)
    e = ov.export()
    act = e.active()
    li, off = np.nonzero(act)
    idx = e.origins[li] + np.stack([off >> 6, (off >> 3) & 7, off & 7], 1)
    gv = bs.volume_from_voxels(idx, e.values[li, off], 0.04, gpu_ctx).flood_fill().export()
    assert_same_volume(gv, ov.flood_fill().export())


def _pairs():
    return [
        ("overlap", 0.05, (0.4, (0.0, 0.0, 0.0)), (0.4, (0.3, 0.0, 0.0))),
        ("nested", 0.05, (1.0, (0.0, 0.0, 0.0)), (0.3, (0.1, 0.0, 0.0))),
        ("nested-rev", 0.05, (0.3, (0.1, 0.0, 0.0)), (1.0, (0.0, 0.0, 0.0))),
        ("disjoint-root", 0.01, (0.3, (0.0, 0.0, 0.0)), (0.3, (50.0, 0.0, 0.0))),
        ("big-small-h", 0.004, (0.9, (0.01, -0.02, 0.03)), (0.5, (0.6, 0.2, -0.1))),
        ("neg-coords", 0.02, (0.5, (-1.3, -0.7, -2.1)), (0.45, (-1.0, -0.5, -2.0))),
    ]


@pytest.mark.parametrize("name,h,a,b", _pairs(), ids=[p[0] for p in _pairs()])
@pytest.mark.parametrize("op", ["union", "subtract", "intersect"])
def test_csg_of_spheres_matches_oracle(gpu_ctx, oracle_lib, name, h, a, b, op):
    B = bs.VolumeBuilder(gpu_ctx, h)
    ga, gb = B.sphere(*a), B.sphere(*b)
    out = getattr(ga, op)(gb)
    assert ga._h is None and gb._h is None   # both operands consumed
    ra = oracle_lib.sphere(a[0], a[1], h)
    rb = oracle_lib.sphere(b[0], b[1], h)
    ref = getattr(ra, op)(rb)
    assert_same_volume(out.export(), ref.export())
    # sign_at over a lattice after one more flood fill (every inactive slot set)
    out.flood_fill()
    ref.flood_fill()
    span = int((max(a[0], b[0]) + 60.0 * h) / h)
    ax = np.arange(-span, span, max(1, span // 25), dtype=np.int32)
    q = np.stack(np.meshgrid(ax, ax, ax, indexing="ij"), -1).reshape(-1, 3)
    assert np.array_equal(out.signs_at(q), ref.signs_at(q))


def test_csg_chains_near_root_blocks(gpu_ctx, oracle_lib):
    """Random sphere chains (h = 1, coordinates in voxels) placed around the
    corners of the 4096-voxel root blocks, mixing the three operations."""
    rng = np.random.default_rng(5)
    for trial in range(6):
        ops = []
        g = r = None
        for k in range(2 + trial % 3):
            rad = float(30 + 90 * rng.random())
            c = (4096.0 * rng.integers(0, 2) + rng.uniform(-200, 200), 4096.0 * rng.integers(0, 2) + rng.uniform(-200, 200),
                 4096.0 * rng.integers(0, 3) + rng.uniform(-200, 200))
            gs = bs.VolumeBuilder(gpu_ctx, 1.0).sphere(rad, c)
            rs = oracle_lib.sphere(rad, c, 1.0)
            if g is None:
                g, r = gs, rs
                continue
            op = ["union", "subtract", "intersect"][int(rng.integers(0, 3))]
            ops.append(op)
            g = getattr(g, op)(gs)
            r = getattr(r, op)(rs)
            assert_same_volume(g.export(), r.export())
        g.flood_fill()
        r.flood_fill()
        assert_same_volume(g.export(), r.export())
        ax = np.arange(-4096, 3 * 4096 + 4096, 512, dtype=np.int32)
        q = np.stack(np.meshgrid(ax, ax, ax, indexing="ij"), -1).reshape(-1, 3)
        assert np.array_equal(g.signs_at(q), r.signs_at(q)), (trial, ops)


def test_root_flood_fill_fills_z_gap(gpu_ctx, oracle_lib):
    """root_node/flood_fill.rs:74-95 on the 5-4-3 grid: two negative voxels in
    the blocks at z = 0 and z = 8192 of one (x, y) column make the missing block
    at z = 4096 an inside node."""
    idx = np.array([[3968, 3968, 3968], [0, 0, 8192]], np.int32)
    vals = np.array([-1.0, -1.0], np.float32)
    g = bs.volume_from_voxels(idx, vals, 1.0, gpu_ctx).flood_fill()
    r = oracle_lib.from_voxels(idx, vals, 1.0).flood_fill()
    ax = np.arange(0, 4096, 97, dtype=np.int32)
    q = np.stack(np.meshgrid(ax, ax, ax + 4096, indexing="ij"), -1).reshape(-1, 3)
    assert g.signs_at(q).all()
    assert np.array_equal(g.signs_at(q), r.signs_at(q))
    assert_same_volume(g.export(), r.export())
    # positive second voxel: no gap fill, the missing block reads positive
    g2 = bs.volume_from_voxels(idx, np.array([-1.0, 1.0], np.float32), 1.0, gpu_ctx).flood_fill()
    assert not g2.signs_at(q).any()


def test_leaf_flood_fill_patterns(gpu_ctx, oracle_lib):
    """leaf_node/flood_fill.rs:82-113 scaled to 8^3 leaves, plus random masks."""
    cases = [
        (np.array([[2, 2, 2]]), [1.0]),
        (np.array([[2, 2, 2]]), [-1.0]),
        (np.array([[x, y, z] for x in (3, 4) for y in range(8) for z in range(8)]),
         [-1.0 if x == 3 else 1.0 for x in (3, 4) for y in range(8) for z in range(8)]),
    ]
    rng = np.random.default_rng(9)
    for _ in range(4):
        n = int(rng.integers(1, 400))
        idx = rng.integers(-20, 20, size=(n, 3))
        cases.append((idx, rng.choice([-1.0, 1.0], n) * rng.uniform(0.1, 2.0, n)))
    for idx, vals in cases:
        g = bs.volume_from_voxels(idx, vals, 0.5, gpu_ctx).flood_fill().export()
        r = oracle_lib.from_voxels(idx, vals, 0.5).flood_fill().export()
        assert_same_volume(g, r)
    g = bs.volume_from_voxels(cases[2][0], cases[2][1], 0.5, gpu_ctx).flood_fill().export()
    neg = np.signbit(g.values[0])
    assert neg[:4 * 64].all() and not neg[4 * 64:].any()


def test_torus_and_sphere_union(gpu_ctx, oracle_lib):
    """The torus-and-sphere workload (torus 128x64, h = 0.02, band 3, united
    with the sphere r = 0.5 at (1, 0, 0)).  The torus values carry the
    mesh-to-volume rounding differences (<= 1e-4 voxel); structure and signs
    must match exactly."""
    soup = mg.torus(1.0, 0.35, 128, 64, 0.01, seed=1)
    g = bs.MeshToVolume(gpu_ctx).with_voxel_size(0.02).with_narrow_band_width(3).convert(soup)
    g = g.union(bs.VolumeBuilder(gpu_ctx, 0.02).sphere(0.5, (1.0, 0.0, 0.0)))
    r = oracle_lib.mesh_to_volume(soup, 0.02, 3).union(oracle_lib.sphere(0.5, (1.0, 0.0, 0.0), 0.02))
    ge, re_ = g.export(), r.export()
    assert np.array_equal(ge.origins, re_.origins) and np.array_equal(ge.masks, re_.masks)
    assert np.array_equal(_tiles(ge)[:, :4], _tiles(re_)[:, :4])
    gv, rv = ge.values.astype(np.float64), re_.values.astype(np.float64)
    big = np.abs(rv) > 1e30
    assert np.array_equal(gv[big], rv[big])
    assert np.max(np.abs(gv[~big] - rv[~big])) / 0.02 < 1e-4
    assert np.array_equal(np.signbit(ge.values[~big]), np.signbit(re_.values[~big]))


def test_csg_with_empty_and_self(gpu_ctx, oracle_lib):
    B = bs.VolumeBuilder(gpu_ctx, 0.05)
    a = B.sphere(0.4, (0.0, 0.0, 0.0))
    base = a.clone().export()
    a = a.union(B.sphere(-1.0, (0.0, 0.0, 0.0)))      # union with an empty volume
    r = oracle_lib.sphere(0.4, (0.0, 0.0, 0.0), 0.05).union(oracle_lib.sphere(-1.0, (0.0, 0.0, 0.0), 0.05))
    e = a.export()
    assert_same_volume(e, r.export())
    assert np.array_equal(e.origins, base.origins) and np.array_equal(e.masks, base.masks)
    empty = B.sphere(-1.0, (0.0, 0.0, 0.0)).intersect(a)
    assert empty.export().origins.shape[0] == 0
    c = B.sphere(0.4, (0.0, 0.0, 0.0))
    with pytest.raises(ValueError):
        c.union(c)
    with pytest.raises(ValueError):
        a.union(c)                                       # a was consumed above
