"""Shared signs (one winding-number evaluation per group of voxels that no
surface point separates) against the per-voxel signs of the reference's
compute_sings (mesh_to_volume.rs:198-215): on closed meshes the two volumes
must be identical bit for bit, and the number of evaluations must equal the
number of leaf-local groups of the unsigned field."""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.csgraph as cg

import baby_shark_amd as bs
from baby_shark_amd import meshgen as mg

pytestmark = pytest.mark.gpu


def _convert(ctx, soup, h, band, shared):
    ctx.set_profiling(True, counters=False)
    try:
        v = bs.MeshToVolume(ctx).with_voxel_size(h).with_narrow_band_width(band).with_shared_signs(shared).convert(soup)
        return v, ctx.last_stats()
    finally:
        ctx.set_profiling(False)


def leaf_groups(e, thr):
    """Groups of each leaf's active voxels joined by face links with
    |d1| + |d2| > thr (scipy connected components, leaf by leaf)."""
    act = e.active()
    li, off = np.nonzero(act)
    d = np.abs(e.values[li, off]).astype(np.float32)
    slot = li.astype(np.int64) * 512 + off
    pos = np.full(e.values.size, -1, np.int64)
    pos[slot] = np.arange(slot.size)
    a_list, b_list = [], []
    for step, coord in ((64, off >> 6), (8, (off >> 3) & 7), (1, off & 7)):
        ok = coord < 7
        nb = pos[slot[ok] + step]
        a = np.nonzero(ok)[0][nb >= 0]
        b = nb[nb >= 0]
        link = (d[a] + d[b]) > np.float32(thr)
        a_list.append(a[link])
        b_list.append(b[link])
    A, B = np.concatenate(a_list), np.concatenate(b_list)
    g = sp.coo_matrix((np.ones(A.size), (A, B)), shape=(slot.size, slot.size))
    return cg.connected_components(g, directed=False)[0]


CASES = [
    ("torus 128x64, h 0.02", lambda: mg.torus(1.0, 0.35, 128, 64, 0.01, seed=1), 0.02, 3),
    ("icosphere L5, h 0.004", lambda: mg.displaced_icosphere(5, 1.0, 0.05, seed=7), 0.004, 3),
    ("icosphere L2 (many subdivisions), h 0.01", lambda: mg.displaced_icosphere(2, 0.7, 0.03, seed=9), 0.01, 2),
    ("two tori, band 5", lambda: np.concatenate([mg.torus(0.6, 0.2, 64, 32, 0.005, seed=2),
                                                mg.torus(0.6, 0.2, 64, 32, 0.005, seed=3) + np.float32(0.75)]),
     0.01, 5),
]


@pytest.mark.parametrize("name,make,h,band", CASES, ids=[c[0] for c in CASES])
def test_shared_signs_equal_per_voxel_signs(gpu_ctx, name, make, h, band):
    soup = make()
    p, sp_ = _convert(gpu_ctx, soup, h, band, False)
    s, ss = _convert(gpu_ctx, soup, h, band, True)
    pe, se = p.export(), s.export()
    assert np.array_equal(pe.origins, se.origins) and np.array_equal(pe.masks, se.masks)
    diff = np.count_nonzero(pe.values.view(np.uint32) != se.values.view(np.uint32))
    n_act = int(pe.active().sum())
    print(f"{name}: {n_act} voxels, {ss['n_sign_evals']} evaluations "
          f"({ss['n_sign_evals'] / n_act:.4f}), winding {sp_['ms_winding']:.3f} -> {ss['ms_winding']:.3f} ms, "
          f"groups {ss['ms_sign_share']:.3f} ms, differing values {diff}", flush=True)
    assert diff == 0
    assert sp_["n_sign_evals"] == n_act
    # the group count lies between the counts at the bare spacing and at a
    # looser threshold than any rounding margin the library adds
    lo, hi = leaf_groups(pe, h), leaf_groups(pe, 1.1 * h)
    assert lo <= ss["n_sign_evals"] <= hi, (lo, ss["n_sign_evals"], hi)
    assert ss["n_sign_evals"] < 0.1 * n_act


def test_shared_signs_batch_and_params(gpu_ctx):
    soups = [mg.torus(1.0, 0.35, 32, 16, 0.01, seed=4), mg.displaced_icosphere(3, 1.0, 0.05, seed=5)]
    a = bs.batch_mesh_to_volume(soups, [0.02, 0.01], 3, ctx=gpu_ctx, shared_signs=True)
    b = bs.batch_mesh_to_volume(soups, [0.02, 0.01], 3, ctx=gpu_ctx)
    for x, y in zip(a, b):
        assert np.array_equal(x.export().values.view(np.uint32), y.export().values.view(np.uint32))
    m = bs.MeshToVolume(gpu_ctx).with_voxel_size(0.02).with_narrow_band_width(3)
    m.params.sign_mode = 7
    with pytest.raises(bs.BSError) as e:
        m.convert(soups[0])
    assert e.value.status == bs._lib.BS_ERR_INVALID_ARG


def test_shared_signs_on_alternate_paths(gpu_ctx):
    """Shared signs through the GPU leaf hash and the LDS traversal stack
    equal the per-voxel signs on the default paths."""
    soup = mg.displaced_icosphere(4, 0.8, 0.05, seed=6)
    p, _ = _convert(gpu_ctx, soup, 0.01, 3, False)
    pe = p.export()
    gpu_ctx.set_paths(bs._lib.BS_PATH_FORCE_HASH | bs._lib.BS_PATH_FORCE_LDS_STACK)
    try:
        s, st = _convert(gpu_ctx, soup, 0.01, 3, True)
        assert st["used_hash"] == 1 and st["used_lds_stack"] == 1
    finally:
        gpu_ctx.set_paths(0)
    assert np.array_equal(pe.values.view(np.uint32), s.export().values.view(np.uint32))
