"""Mesh -> narrow-band SDF on the GPU against the CPU oracle (the reference's
algorithm restated).  Leaf keys and active masks bit-exact; SDF within 1e-4
voxel; signs bit-exact wherever |d| > 1e-6 voxel."""
import numpy as np
import pytest

import baby_shark_amd as bs
from baby_shark_amd import meshgen as mg
from tests.helpers import compare_volumes, on_surface, voxel_points

pytestmark = pytest.mark.gpu
TOL_VOX = 1e-4   # SDF tolerance, voxels
RNG = np.random.default_rng(3)


def run_gpu(ctx, soup, h, band):
    return bs.MeshToVolume(ctx).with_voxel_size(h).with_narrow_band_width(band).convert(soup)


# An open surface's sign may differ from the oracle's only where the exact
# winding number is within the far-field (order-2) approximation error of the
# threshold: the GPU's tree and the reference's SAH tree group far triangles
# differently.  0.1 bounds that error (the open-cap test measures the same).
WN_AMBIGUITY = 0.1


def sign_mismatch_wn(oracle_lib, soup, ge, re_, h):
    """Exact (brute-force) winding numbers at the voxels whose signs differ."""
    act = re_.active()
    li, off = np.nonzero(act)
    r = re_.values[li, off]
    g = ge.values[li, off]
    bad = ~on_surface(r, voxel_points(re_.origins, li, off, h), h) & (np.signbit(g) != np.signbit(r))
    vox = re_.origins[li[bad]] + np.stack([off[bad] >> 6, (off[bad] >> 3) & 7, off[bad] & 7], 1)
    pts = vox.astype(np.float32) * np.float32(h)
    print("  mismatch |d|/h:", np.abs(r[bad]) / h, " |d|/ulp(max|p|):",
          np.abs(r[bad]) / np.spacing(np.max(np.abs(pts), axis=1)) if len(pts) else [], flush=True)
    return oracle_lib.winding_numbers(soup, pts, 2.0, exact=True)[2] if len(pts) else np.zeros(0, np.float32)


def check_case(ctx, oracle_lib, soup, h, band, fp64=False, closed=True):
    g = run_gpu(ctx, soup, h, band)
    r = oracle_lib.mesh_to_volume(soup, h, band)
    if r is None:
        assert g is None
        return None
    ge, re_ = g.export(), r.export()
    res = compare_volumes(ge, re_, h)
    if not closed:
        # an open surface's sign can flip where w is within rounding of 0.2
        # (checked below): the distance tolerance applies to |d|
        act = re_.active()
        res["max_err_vox"] = float(np.max(np.abs(np.abs(ge.values[act]).astype(np.float64) -
                                                 np.abs(re_.values[act]))) / h)
    if fp64:
        d64 = oracle_lib.unsigned_distance_f64(r, soup, band)
        act = re_.active()
        res["max_err_vs_fp64_vox"] = float(np.max(np.abs(np.abs(ge.values[act]) - d64[act])) / h)
        res["rms_err_vs_fp64_vox"] = float(np.sqrt(np.mean((np.abs(ge.values[act]) - d64[act]) ** 2)) / h)
        res["ref_max_err_vs_fp64_vox"] = float(np.max(np.abs(np.abs(re_.values[act]) - d64[act])) / h)
    print(f"case h={h} band={band} n={len(soup)}: {res}", flush=True)
    assert res["max_err_vox"] <= TOL_VOX, res
    if closed or res["sign_mismatches"] == 0:
        assert res["sign_mismatches"] == 0, res
    else:
        w = sign_mismatch_wn(oracle_lib, soup, ge, re_, h)
        print(f"  open surface: {w.size} sign mismatches, exact |w-0.2| max {np.max(np.abs(w - 0.2)):.2e}",
              flush=True)
        assert np.all(np.abs(w - 0.2) < WN_AMBIGUITY), w
    return res


def test_torus_128x64_parity(gpu_ctx, oracle_lib):
    """Torus R=1, r=0.35, 128x64 segments, radial noise +-0.01 (seed 1), h=0.02, band 3."""
    soup = mg.torus(1.0, 0.35, 128, 64, 0.01, seed=1)
    res = check_case(gpu_ctx, oracle_lib, soup, 0.02, 3, fp64=True)
    assert res["n_active"] > 300000
    assert res["max_err_vs_fp64_vox"] <= TOL_VOX


@pytest.mark.parametrize("band", [0, 1, 2, 5])
def test_small_sphere_bands(gpu_ctx, oracle_lib, band):
    check_case(gpu_ctx, oracle_lib, mg.displaced_icosphere(3, radius=0.7, amp=0.05, seed=9), 0.03, band)


def test_wide_band_candidate_lists(gpu_ctx, oracle_lib):
    """Band 12: each sub-triangle's box covers 64 or more leaves, so a block's
    candidate pairs overflow the LDS staging of the pair kernel and are written
    directly; the volume must still equal the oracle's."""
    check_case(gpu_ctx, oracle_lib, mg.displaced_icosphere(2, radius=0.7, amp=0.05, seed=9), 0.03, 12)


def test_single_triangle_and_axis_aligned(gpu_ctx, oracle_lib):
    # one triangle; a triangle flat in z (zero box extent -> +-1 expansion,
    # mesh_to_volume.rs:136-141); negative coordinates across leaf boundaries
    tri = np.array([[[-0.31, 0.2, 0.05], [0.4, -0.27, 0.11], [0.05, 0.33, -0.2]]], np.float32)
    check_case(gpu_ctx, oracle_lib, tri, 0.05, 2, closed=False)
    flat = np.array([[[-0.5, -0.5, 0.0], [0.5, -0.5, 0.0], [0.0, 0.5, 0.0]]], np.float32)
    check_case(gpu_ctx, oracle_lib, flat, 0.1, 0, closed=False)
    check_case(gpu_ctx, oracle_lib, flat, 0.1, 2, closed=False)


def test_large_triangle_many_subdivisions(gpu_ctx, oracle_lib):
    tri = np.array([[[-1.0, -0.8, 0.1], [1.1, -0.7, -0.05], [0.0, 1.2, 0.2]]], np.float32)
    check_case(gpu_ctx, oracle_lib, tri, 0.02, 1, closed=False)   # ~110 rows, ~1.2e4 sub-triangles


def test_degenerate_triangles(gpu_ctx, oracle_lib):
    soup = mg.sphere_soup(2, radius=0.6)
    deg = np.array([[[0.1, 0.1, 0.1], [0.1, 0.1, 0.1], [0.1, 0.1, 0.1]],          # a point
                    [[0.0, 0.0, 0.0], [0.3, 0.3, 0.3], [0.6, 0.6, 0.6]]], np.float32)  # a segment
    check_case(gpu_ctx, oracle_lib, np.concatenate([soup, deg]), 0.05, 2, closed=False)


def test_empty_mesh_is_none(gpu_ctx):
    assert run_gpu(gpu_ctx, np.zeros((0, 3, 3), np.float32), 0.1, 3) is None


def test_device_input_matches_host_input(gpu_ctx):
    torch = pytest.importorskip("torch")
    soup = mg.torus(1.0, 0.35, 64, 32, 0.01, seed=4)
    a = run_gpu(gpu_ctx, soup, 0.03, 2).export()
    b = run_gpu(gpu_ctx, torch.from_numpy(soup).cuda(), 0.03, 2).export()
    assert np.array_equal(a.origins, b.origins) and np.array_equal(a.masks, b.masks)
    assert np.array_equal(a.values, b.values)


def test_repeatable(gpu_ctx):
    soup = mg.displaced_icosphere(4, seed=5)
    a = run_gpu(gpu_ctx, soup, 0.01, 3).export()
    b = run_gpu(gpu_ctx, soup, 0.01, 3).export()
    assert np.array_equal(a.values, b.values) and np.array_equal(a.masks, b.masks)


@pytest.mark.timeout(900)
def test_icosphere_L8_full_size(gpu_ctx, oracle_lib):
    """Displaced icosphere L8 (1,310,720 tris), h=0.002, band 3, the headline
    workload, against the oracle's whole volume: leaf keys and active masks
    bit-exact, every value within 1e-4 voxel, every sign off the surface
    exact, for per-voxel signs and for shared signs; values at 20000 sampled
    voxels also against FP64 distances."""
    import os
    import time
    h, band = 0.002, 3
    soup = mg.displaced_icosphere(8, 1.0, 0.05, seed=2)
    ctx = gpu_ctx
    ctx.set_profiling(True)
    g = run_gpu(ctx, soup, h, band)
    stats = ctx.last_stats()
    ctx.set_profiling(False)
    print("icosphere L8 stats:", stats, flush=True)
    ge = g.export()
    g.free()
    threads = min(16, os.cpu_count() or 1)
    t0 = time.perf_counter()
    r = oracle_lib.mesh_to_volume(soup, h, band, threads=threads)
    re_ = r.export()
    r = None
    print(f"icosphere L8 oracle convert: {time.perf_counter() - t0:.1f} s on {threads} threads", flush=True)
    res = compare_volumes(ge, re_, h)
    print("icosphere L8 whole volume, per-voxel signs:", res, flush=True)
    assert res["n_active"] == 40131672
    assert res["max_err_vox"] <= TOL_VOX and res["sign_mismatches"] == 0, res
    shared = bs.MeshToVolume(ctx).with_voxel_size(h).with_narrow_band_width(band).with_shared_signs().convert(soup)
    se = shared.export()
    shared.free()
    res_s = compare_volumes(se, re_, h)
    print("icosphere L8 whole volume, shared signs:", res_s, flush=True)
    assert res_s["max_err_vox"] <= TOL_VOX and res_s["sign_mismatches"] == 0, res_s
    act = ge.active()
    li, off = np.nonzero(act)
    pick = RNG.choice(li.size, 20000, replace=False)
    vox = ge.origins[li[pick]] + np.stack([off[pick] >> 6, (off[pick] >> 3) & 7, off[pick] & 7], 1)
    v = ge.values[li[pick], off[pick]]
    r32, r64 = oracle_lib.sample_unsigned(soup, h, band, vox)
    err64 = np.abs(np.abs(v).astype(np.float64) - r64) / h
    print(f"icosphere L8 sampled vs FP64: max {err64.max():.3g} rms {np.sqrt(np.mean(err64**2)):.3g} vox", flush=True)
    assert err64.max() <= TOL_VOX


def test_hash_leaf_allocation_matches_dense(gpu_ctx, oracle_lib):
    """The GPU hash (used for extents too large for a dense leaf grid) gives the
    same volume as the dense leaf grid, and both match the oracle."""
    soup = mg.displaced_icosphere(4, radius=0.8, seed=6)
    gpu_ctx.set_paths(bs._lib.BS_PATH_FORCE_HASH)
    try:
        a = run_gpu(gpu_ctx, soup, 0.01, 3).export()
        assert gpu_ctx.last_stats()["used_hash"] == 1
    finally:
        gpu_ctx.set_paths(0)
    b = run_gpu(gpu_ctx, soup, 0.01, 3).export()
    assert gpu_ctx.last_stats()["used_hash"] == 0
    assert np.array_equal(a.origins, b.origins) and np.array_equal(a.masks, b.masks)
    assert np.array_equal(a.values, b.values)
    check_case(gpu_ctx, oracle_lib, soup, 0.01, 3)


def test_lds_stack_traversal_matches_register_stack(gpu_ctx, oracle_lib):
    """The LDS-stack traversal (trees of height >= 62) signs exactly like the
    register-stack one used for lower trees."""
    soup = mg.displaced_icosphere(5, radius=0.8, seed=8)
    b = run_gpu(gpu_ctx, soup, 0.006, 3).export()
    assert gpu_ctx.last_stats()["used_lds_stack"] == 0
    gpu_ctx.set_paths(bs._lib.BS_PATH_FORCE_LDS_STACK)
    try:
        a = run_gpu(gpu_ctx, soup, 0.006, 3).export()
        assert gpu_ctx.last_stats()["used_lds_stack"] == 1
    finally:
        gpu_ctx.set_paths(0)
    assert np.array_equal(a.masks, b.masks)
    assert np.array_equal(np.signbit(a.values), np.signbit(b.values))
    assert np.max(np.abs(np.abs(a.values) - np.abs(b.values))) == 0.0


def test_open_cap_takes_the_fp64_recheck_path(gpu_ctx, oracle_lib):
    """An open spherical cap: beside its rim the winding number passes through
    the 0.2 threshold, so voxels whose f32 value falls in the ambiguity band
    are re-evaluated in FP64 (k_winding_fp64).  Wherever the exact
    (brute-force) winding number is at least 0.1 from 0.2 -- more than the
    order-2 approximation error of either tree -- the GPU's sign and the
    oracle's are both that side's sign."""
    soup = mg.sphere_soup(4, radius=0.6)
    cap = soup[soup.mean(axis=1)[:, 2] > 0.15]
    h, band = 0.02, 3
    g = run_gpu(gpu_ctx, cap, h, band)
    n_re = gpu_ctx.last_stats()["n_fp64_rechecks"]
    ge = g.export()
    re_ = oracle_lib.mesh_to_volume(cap, h, band).export()
    assert np.array_equal(ge.origins, re_.origins) and np.array_equal(ge.masks, re_.masks)
    li, off = np.nonzero(ge.active())
    pts = voxel_points(ge.origins, li, off, h)
    _, _, w_ex = oracle_lib.winding_numbers(cap, pts, 2.0, exact=True)
    w_ex = w_ex.astype(np.float64)
    g_neg = np.signbit(ge.values[li, off])
    r_neg = np.signbit(re_.values[li, off])
    sure = (np.abs(w_ex - 0.2) >= 0.1) & ~on_surface(ge.values[li, off], pts, h)
    band_vox = int(np.sum(np.abs(w_ex - 0.2) < 0.1))
    print(f"open cap: {li.size} voxels, {n_re} FP64 re-evaluations, {band_vox} within 0.1 of the threshold",
          flush=True)
    assert n_re > 0, "the FP64 re-evaluation did not run"
    want_neg = ~(w_ex < 0.2)
    assert np.array_equal(g_neg[sure], want_neg[sure])
    assert np.array_equal(r_neg[sure], want_neg[sure])


def _visits(st):
    return st["wn_node_visits"] + st["wn_far_evals"] + st["wn_leaf_visits"]


def _tree_runs(gpu_ctx, soup, h, band, trees):
    runs = {}
    for name, flags in trees:
        gpu_ctx.set_paths(flags)
        gpu_ctx.set_profiling(True, counters=True)
        try:
            v = run_gpu(gpu_ctx, soup, h, band)
            runs[name] = (v.export(), dict(gpu_ctx.last_stats()))
            v.free()
        finally:
            gpu_ctx.set_profiling(False)
            gpu_ctx.set_paths(0)
    return runs


def _same_but_signs_near_surface(a, b, h):
    assert np.array_equal(a.origins, b.origins) and np.array_equal(a.masks, b.masks)
    assert np.array_equal(np.abs(a.values), np.abs(b.values))
    li, off = np.nonzero(a.active())
    off_surface = ~on_surface(a.values[li, off], voxel_points(a.origins, li, off, h), h)
    assert np.array_equal(np.signbit(a.values[li, off])[off_surface], np.signbit(b.values[li, off])[off_surface])


def test_tree_builds_agree(gpu_ctx, oracle_lib):
    """The three winding-number trees of a mesh of >= 2^18 triangles -- the
    top-down binned tree (default), the clustered one (BS_PATH_PLOC) and the
    Morton split (BS_PATH_LBVH) -- give identical distances and masks and the
    same signs off the surface; the top-down tree needs the fewest node
    evaluations, the Morton split the most.  The default matches the oracle
    (the reference's own SAH tree)."""
    soup = mg.displaced_icosphere(7, radius=0.8, seed=12)
    h, band = 0.005, 3
    runs = _tree_runs(gpu_ctx, soup, h, band, (("topdown", 0), ("ploc", bs._lib.BS_PATH_PLOC),
                                               ("lbvh", bs._lib.BS_PATH_LBVH)))
    n_act = runs["topdown"][1]["n_active"]
    for name, (_, st) in runs.items():
        print(f"{name}: {st['n_bvh_iterations']} passes, height {st['bvh_height']}, "
              f"{_visits(st) / n_act:.1f} visits/voxel (near {st['wn_node_visits'] / n_act:.1f}, "
              f"far {st['wn_far_evals'] / n_act:.1f}, leaves {st['wn_leaf_visits'] / n_act:.2f})", flush=True)
    st_td, st_pl, st_lb = runs["topdown"][1], runs["ploc"][1], runs["lbvh"][1]
    assert st_td["n_bvh_iterations"] > 0 and st_pl["n_bvh_iterations"] > 0 and st_lb["n_bvh_iterations"] == 0
    _same_but_signs_near_surface(runs["topdown"][0], runs["lbvh"][0], h)
    _same_but_signs_near_surface(runs["ploc"][0], runs["lbvh"][0], h)
    assert _visits(st_td) < _visits(st_pl) < _visits(st_lb)
    check_case(gpu_ctx, oracle_lib, soup, h, band)


def test_topdown_tree_is_deterministic(gpu_ctx):
    """The top-down build takes node ids with atomics, so ids differ from run
    to run; the tree, the triangle order and the layout must not: two
    conversions of the same mesh give the same volume bit for bit and the
    same visit counts."""
    soup = mg.displaced_icosphere(7, radius=0.8, seed=3)
    h, band = 0.004, 3
    runs = _tree_runs(gpu_ctx, soup, h, band, (("a", 0), ("b", 0)))
    (a, sa), (b, sb) = runs["a"], runs["b"]
    assert sa["n_bvh_iterations"] > 0
    assert np.array_equal(a.origins, b.origins) and np.array_equal(a.masks, b.masks)
    assert np.array_equal(a.values.view(np.uint32), b.values.view(np.uint32))
    for k in ("wn_node_visits", "wn_far_evals", "wn_leaf_visits", "wn_leaf_tris", "bvh_height"):
        assert sa[k] == sb[k], k


def test_topdown_tree_coincident_centroids(gpu_ctx):
    """300 copies of each triangle of a closed 1280-triangle mesh (384,000
    triangles): nodes whose triangles share one centre cannot be cut by bins
    and are cut at the middle of their range, both in the level passes and in
    the per-wave subtrees.  The build must finish with a full tree; distances
    and masks do not depend on the tree and equal the Morton split's (the
    signs are not compared: the far-field error adds up over the copies)."""
    base = mg.displaced_icosphere(3, radius=0.8, seed=4)
    soup = np.ascontiguousarray(np.repeat(base, 300, axis=0))
    assert soup.shape[0] >= 1 << 18
    h, band = 0.02, 2
    runs = _tree_runs(gpu_ctx, soup, h, band, (("topdown", 0), ("lbvh", bs._lib.BS_PATH_LBVH)))
    st = runs["topdown"][1]
    print(f"topdown: height {st['bvh_height']}, {st['n_bvh_iterations']} passes", flush=True)
    assert 0 < st["bvh_height"] < 100
    a, b = runs["topdown"][0], runs["lbvh"][0]
    assert np.array_equal(a.origins, b.origins) and np.array_equal(a.masks, b.masks)
    assert np.array_equal(np.abs(a.values), np.abs(b.values))


@pytest.mark.parametrize("n", [2, 3, 4, 5, 9])
def test_tiny_meshes(gpu_ctx, oracle_lib, n):
    """Trees of a handful of triangles (a root evaluated as a leaf, one or two
    levels): the first n faces of a level-1 icosphere, open, against the oracle."""
    soup = mg.sphere_soup(1, radius=0.5)[:n]
    check_case(gpu_ctx, oracle_lib, soup, 0.05, 2, closed=False)
