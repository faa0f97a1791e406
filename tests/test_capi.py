"""The C ABI library loads and exports every function include/bs_voxel.h
declares; host-side argument checks and defaults.  No GPU compute."""
import ctypes as C

import pytest

from baby_shark_amd import _lib


def test_library_exports_every_header_symbol():
    L = _lib.load()
    names = _lib.header_functions()
    assert "bs_mesh_to_volume" in names and "bs_ctx_create" in names
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_params_default_match_reference():
    # mesh_to_volume.rs:223-236 (voxel 1.0, band 0), :264 (accuracy 2.0), :266 (0.2)
    p = _lib.M2VParams()
    _lib.load().bs_m2v_params_default(C.byref(p))
    assert p.voxel_size == 1.0 and p.band == 0
    assert p.wn_accuracy == 2.0 and abs(p.sign_threshold - 0.2) < 1e-7


def test_null_arguments_are_rejected():
    L = _lib.load()
    assert L.bs_mesh_to_volume(None, None, 0, None, None) == _lib.BS_ERR_INVALID_ARG
    assert L.bs_volume_export(None, None, None, None) == _lib.BS_ERR_INVALID_ARG
    assert L.bs_ctx_create(0, None) == _lib.BS_ERR_INVALID_ARG


def test_ctx_without_device_reports_no_device():
    torch = pytest.importorskip("torch")
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    L = _lib.load()
    h = C.c_void_p()
    assert L.bs_ctx_create(0, C.byref(h)) == _lib.BS_ERR_NO_DEVICE
    assert b"device" in L.bs_last_error()


def test_new_entry_points_reject_null_arguments():
    L = _lib.load()
    out = C.c_void_p()
    n = C.c_size_t(0)
    fp = C.POINTER(C.c_float)()
    assert L.bs_volume_cuboid(None, 0, 0, 0, 1, 1, 1, 0.1, C.byref(out)) == _lib.BS_ERR_INVALID_ARG
    assert L.bs_volume_iwp(None, 0, 0, 0, 1, 1, 1, 0.5, 0.1, C.byref(out)) == _lib.BS_ERR_INVALID_ARG
    assert L.bs_volume_densify_tiles(None, None) == _lib.BS_ERR_INVALID_ARG
    assert L.bs_voxel_remesh(None, None, 0, 0, 0.1, 0, C.byref(fp), C.byref(n)) == _lib.BS_ERR_INVALID_ARG
    assert L.bs_stl_read(None, C.byref(fp), C.byref(n)) == _lib.BS_ERR_INVALID_ARG
    assert L.bs_stl_write(b"/tmp/x.stl", None, 3) == _lib.BS_ERR_INVALID_ARG
    assert L.bs_ctx_set_paths(None, 0) == _lib.BS_ERR_INVALID_ARG
    assert L.bs_ctx_set_profiling(None, 1) == _lib.BS_ERR_INVALID_ARG


def test_m2v_params_struct_matches_header():
    """The ctypes mirror has the header's fields in order (sign_mode last)."""
    names = [f for f, _ in _lib.M2VParams._fields_]
    assert names == ["voxel_size", "band", "wn_accuracy", "sign_threshold", "fp64_band", "sign_mode"]
    p = _lib.M2VParams()
    _lib.load().bs_m2v_params_default(C.byref(p))
    assert p.sign_mode == _lib.BS_SIGN_PER_VOXEL and abs(p.fp64_band - 0.1) < 1e-7
