"""bs_dual_contour_into: the no-allocation variant of bs_dual_contour_device.

Through the C ABI it must give the same soup, bit for bit, as the allocating
entry point, on two calls in a row (the second on a smaller volume, so stale
workspace contents from the first must not leak), allocate no device memory,
and reject inputs beyond its fixed workspace and outputs beyond the caller's
capacity with BS_ERR_CAPACITY and nothing written."""
import ctypes as C

import numpy as np
import pytest

import baby_shark_amd as bs
from baby_shark_amd import _lib
from baby_shark_amd import meshgen as mg

pytestmark = pytest.mark.gpu


def _into(ctx, vol, h, buf, capacity):
    import torch
    n = C.c_size_t(0)
    torch.cuda.synchronize()
    st = _lib.load().bs_dual_contour_into(ctx.handle, vol._h, C.c_float(h), C.c_void_p(buf.data_ptr()),
                                          C.c_size_t(capacity), C.byref(n))
    return st, n.value


def _reference(ctx, vol, h):
    import torch
    d = bs.DualContouringMesher(ctx).with_voxel_size(h).mesh_device(vol)
    if d.n == 0:
        return np.zeros((0, 9), np.float32)
    t = torch.as_tensor(d, device="cuda").clone().cpu().numpy().reshape(-1, 9)
    d.free()
    return t


def _check_call(ctx, vol, h):
    import torch
    ref = _reference(ctx, vol, h)
    n_ref = ref.shape[0]
    assert n_ref > 0
    spare = 7
    buf = torch.full(((n_ref + spare) * 9,), float("nan"), device="cuda", dtype=torch.float32)
    torch.cuda.synchronize()
    free0 = torch.cuda.mem_get_info()[0]
    st, n = _into(ctx, vol, h, buf, n_ref + spare)
    free1 = torch.cuda.mem_get_info()[0]
    assert st == _lib.BS_OK, _lib.load().bs_last_error()
    assert n == n_ref
    assert free1 == free0, f"device memory changed by {free0 - free1} bytes"
    got = buf.cpu().numpy().reshape(-1, 9)
    assert np.array_equal(got[:n].view(np.uint32), ref.view(np.uint32))
    assert np.all(np.isnan(got[n:])), "wrote past the soup"
    return n_ref


def test_into_matches_allocating_entry_point_on_two_calls(gpu_ctx):
    ctx = bs.Context(0)  # a fresh context: its scratch arena is empty
    h = 0.03
    big = bs.MeshToVolume(ctx).with_voxel_size(h).with_narrow_band_width(3).convert(
        mg.torus(1.0, 0.35, 48, 24, 0.01, seed=1))
    small = bs.VolumeBuilder(ctx, 0.05).sphere(0.3, (0.1, -0.2, 0.05))
    assert small.info()["n_leaves"] < big.info()["n_leaves"] <= 8192
    n_big = _check_call(ctx, big, h)
    n_small = _check_call(ctx, small, 0.05)
    assert n_small < n_big
    # and the large one again after the small one
    _check_call(ctx, big, h)


def test_into_capacity_and_limits(gpu_ctx):
    import torch
    ctx = gpu_ctx
    h = 0.05
    vol = bs.VolumeBuilder(ctx, h).sphere(0.4, (0.0, 0.0, 0.0))
    n_ref = _reference(ctx, vol, h).shape[0]
    buf = torch.full(((n_ref + 1) * 9,), float("nan"), device="cuda", dtype=torch.float32)
    st, n = _into(ctx, vol, h, buf, n_ref - 1)
    assert st == _lib.BS_ERR_CAPACITY and n == n_ref
    assert torch.isnan(buf).all().item(), "a rejected call wrote output"
    # more leaves than the workspace holds: rejected on the host before any launch
    many = bs.VolumeBuilder(ctx, 0.01).sphere(2.0, (0.0, 0.0, 0.0))
    assert many.info()["n_leaves"] > 8192
    st, n = _into(ctx, many, 0.01, buf, n_ref + 1)
    assert st == _lib.BS_ERR_CAPACITY and n == 0
    assert b"8192" in _lib.load().bs_last_error()
    # an empty volume: nothing to write
    empty = bs.Volume.with_voxel_size(h, ctx)
    st, n = _into(ctx, empty, h, buf, 0)
    assert st == _lib.BS_OK and n == 0
