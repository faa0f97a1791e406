"""The batched mesh-to-volume call equals per-mesh calls bit for bit."""
import numpy as np
import pytest

import baby_shark_amd as bs
from baby_shark_amd import _lib
from baby_shark_amd import meshgen as mg

from test_volume_ops_gpu import assert_same_volume

pytestmark = pytest.mark.gpu


def _meshes():
    ms = mg.batch_meshes(count=10, seed=7, lo=500, hi=20000, span_voxels=96)
    ms.insert(3, (np.zeros((0, 3, 3), np.float32), 0.05))   # an empty mesh: None, like convert
    return ms


@pytest.mark.parametrize("device_input", [True, False])
def test_batch_matches_single_calls(gpu_ctx, device_input):
    import torch
    ms = _meshes()
    soups = [torch.from_numpy(s).cuda() if device_input else s for s, _ in ms]
    vols = bs.batch_mesh_to_volume(soups, [h for _, h in ms], 3, ctx=gpu_ctx, workers=4)
    assert vols[3] is None
    for (soup, h), v in zip(ms, vols):
        single = bs.MeshToVolume(gpu_ctx).with_voxel_size(h).with_narrow_band_width(3).convert(soup)
        if single is None:
            assert v is None
            continue
        assert_same_volume(v.export(), single.export())


def test_autotuned_workers(gpu_ctx, tmp_path, monkeypatch):
    """workers="auto": the tuner times the batch at each candidate count,
    caches the fastest for this device and batch shape, and reads it back the
    next time; the volumes equal the one-stream call's."""
    import json
    import torch
    from baby_shark_amd import tune
    monkeypatch.setenv("BS_TUNE_CACHE", str(tmp_path / "tune.json"))
    ms = [m for m in _meshes() if m[0].shape[0] > 0]
    soups = [torch.from_numpy(s).cuda() for s, _ in ms]
    hs = [h for _, h in ms]
    w = tune.batch_workers(soups, hs, 3, ctx=gpu_ctx, candidates=(1, 2, 4))
    cache = json.loads((tmp_path / "tune.json").read_text())
    key = tune.batch_key(torch.cuda.get_device_name(0), [s.shape[0] for s in soups], 3, False)
    print(f"autotuned workers {w}: {cache[key]}", flush=True)
    assert cache[key]["workers"] == w and w in (1, 2, 4)
    assert set(cache[key]["ms"]) == {"1", "2", "4"}
    assert tune.batch_workers(soups, hs, 3, ctx=gpu_ctx, candidates=(1, 2, 4)) == w  # from the cache
    a = bs.batch_mesh_to_volume(soups, hs, 3, ctx=gpu_ctx, workers="auto")
    b = bs.batch_mesh_to_volume(soups, hs, 3, ctx=gpu_ctx, workers=1)
    for va, vb in zip(a, b):
        assert_same_volume(va.export(), vb.export())


def test_batch_reports_per_mesh_errors(gpu_ctx):
    L = _lib.load()
    import ctypes as C
    soup = np.ascontiguousarray(mg.torus(1.0, 0.35, 16, 8, 0.0, seed=1).reshape(-1, 9))
    ptrs = (C.c_void_p * 2)(soup.ctypes.data, soup.ctypes.data)
    counts = (C.c_size_t * 2)(soup.shape[0], soup.shape[0])
    params = (_lib.M2VParams * 2)()
    for i in range(2):
        L.bs_m2v_params_default(C.byref(params[i]))
        params[i].voxel_size = 0.1
        params[i].band = 1
    params[1].voxel_size = -1.0          # invalid
    out = (C.c_void_p * 2)()
    status = (C.c_int * 2)()
    rc = L.bs_batch_mesh_to_volume(gpu_ctx.handle, 2, ptrs, counts, params, 1, 2, out, status)
    assert rc == _lib.BS_ERR_INVALID_ARG and status[0] == _lib.BS_OK and status[1] == _lib.BS_ERR_INVALID_ARG
    assert out[0] and not out[1]
    L.bs_volume_free(out[0])


@pytest.mark.timeout(600)
def test_full_uneven_batch_matches_single_calls_and_oracle(gpu_ctx, oracle_lib):
    """The whole uneven-batch workload: 256 seeded tori and displaced
    icospheres, 1k-200k triangles each (8.5M in all), each voxelized so its
    bounding box spans 256 voxels, band 3, soups in device memory.  Every
    volume of the batched call equals the per-mesh call's bit for bit (leaf
    arrays compared on the device), and a sample of five meshes across the size
    range equals the oracle's convert (keys and masks exact, values within
    1e-4 voxel, signs off the surface exact)."""
    import os
    import torch
    from tests.helpers import compare_volumes
    ms = mg.batch_meshes(count=256, seed=5)
    soups = [torch.from_numpy(s).cuda() for s, _ in ms]
    hs = [h for _, h in ms]
    torch.cuda.synchronize()
    vols = bs.batch_mesh_to_volume(soups, hs, 3, ctx=gpu_ctx, workers=8)
    n_act = 0
    for soup, h, v in zip(soups, hs, vols):
        single = bs.MeshToVolume(gpu_ctx).with_voxel_size(h).with_narrow_band_width(3).convert(soup)
        a, b = v.device_arrays(), single.device_arrays()
        assert all(torch.equal(x, y) for x, y in zip(a[:2], b[:2]))
        assert torch.equal(a[2].view(torch.int32), b[2].view(torch.int32))
        n_act += v.info()["n_active"]
        single.free()
    print(f"batch of {len(ms)} meshes, {sum(s.shape[0] for s, _ in ms)} triangles, {n_act} active voxels: "
          "identical to per-mesh calls", flush=True)
    sizes = np.array([s.shape[0] for s, _ in ms])
    order = np.argsort(sizes)
    sample = [order[0], order[1], order[len(order) // 2], order[(3 * len(order)) // 4], order[-1]]
    threads = min(16, os.cpu_count() or 1)
    for i in sample:
        soup, h = ms[i]
        ref = oracle_lib.mesh_to_volume(soup, h, 3, threads=threads).export()
        res = compare_volumes(vols[i].export(), ref, h)
        print(f"  mesh {i} ({soup.shape[0]} tris) vs oracle: {res}", flush=True)
        assert res["max_err_vox"] <= 1e-4 and res["sign_mismatches"] == 0, res
    for v in vols:
        v.free()
