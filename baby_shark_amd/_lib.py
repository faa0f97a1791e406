"""ctypes binding of libbs_voxel.so (the C ABI of include/bs_voxel.h).

The library is built in-tree (``python __graft_entry__.py`` or
``make -C baby_shark_amd/csrc``).  Loading fails loudly when it is missing:
there is no fallback implementation.
"""
from __future__ import annotations

import ctypes as C
import os
import threading

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libbs_voxel.so")

BS_OK = 0
BS_ERR_EMPTY = 1
BS_ERR_INVALID_ARG = 2
BS_ERR_HIP = 3
BS_ERR_UNSUPPORTED_TILE = 4
BS_ERR_INTERNAL = 5
BS_ERR_NO_DEVICE = 6
BS_ERR_IO = 7
BS_ERR_UNSUPPORTED = 8
BS_ERR_CAPACITY = 9
BS_MESHING_FEATURE_PRESERVING = 0
BS_MESHING_MANIFOLD = 1
BS_PATH_FORCE_HASH = 1
BS_PATH_FORCE_LDS_STACK = 2
BS_PATH_LBVH = 4
BS_PATH_PLOC = 8
BS_DC_VERTEX_PARTICLE = 0
BS_DC_VERTEX_QEF_F64 = 1
BS_SIGN_PER_VOXEL = 0
BS_SIGN_SHARED = 1
BS_SIGN_SURROGATE = 2
BS_MLP_PARAMS = 8641
BS_PROFILE_EVENTS = 1
BS_PROFILE_COUNTERS = 2

STATUS_NAMES = {0: "BS_OK", 1: "BS_ERR_EMPTY", 2: "BS_ERR_INVALID_ARG", 3: "BS_ERR_HIP",
                4: "BS_ERR_UNSUPPORTED_TILE", 5: "BS_ERR_INTERNAL", 6: "BS_ERR_NO_DEVICE",
                7: "BS_ERR_IO", 8: "BS_ERR_UNSUPPORTED", 9: "BS_ERR_CAPACITY"}


class M2VParams(C.Structure):
    _fields_ = [("voxel_size", C.c_float), ("band", C.c_int64), ("wn_accuracy", C.c_float),
                ("sign_threshold", C.c_float), ("fp64_band", C.c_float), ("sign_mode", C.c_int32)]


class VolumeInfo(C.Structure):
    _fields_ = [("voxel_size", C.c_float), ("n_leaves", C.c_uint64), ("n_active", C.c_uint64),
                ("n_tiles", C.c_uint64)]


class M2VStats(C.Structure):
    _fields_ = [(k, C.c_uint64) for k in ("n_tris", "n_sub", "n_pairs", "n_leaves", "n_active", "n_candidates",
                                          "n_bvh_nodes", "wn_node_visits", "wn_leaf_tris", "wn_far_evals",
                                          "n_fp64_rechecks", "wn_leaf_visits", "used_hash", "used_lds_stack",
                                          "bvh_height")] + \
              [(k, C.c_double) for k in ("ms_total", "ms_subdivide", "ms_leaves", "ms_distance", "ms_bvh",
                                         "ms_winding")] + \
              [("n_sign_evals", C.c_uint64), ("ms_sign_share", C.c_double)] + \
              [(k, C.c_uint64) for k in ("n_sg_tiles", "n_sg_culled", "n_sg_misclassified", "n_sg_fallback",
                                         "n_bvh_iterations")]


class MlpTrainParams(C.Structure):
    _fields_ = [("epochs", C.c_uint32), ("batch", C.c_uint32), ("lr", C.c_float), ("lr_final", C.c_float),
                ("beta1", C.c_float), ("beta2", C.c_float), ("eps", C.c_float), ("seed", C.c_uint64)]


class OffsetStats(C.Structure):
    _fields_ = [(k, C.c_uint64) for k in ("n_kept", "n_pool", "n_out", "n_planes", "n_launches", "n_capacity",
                                          "used_hash", "n_swept")] + \
              [("ms_device", C.c_double), ("ms_wall", C.c_double)]


class CsgStats(C.Structure):
    _fields_ = [(k, C.c_uint64) for k in ("n_a", "n_b", "n_out", "n_paired")] + \
              [(k, C.c_double) for k in ("ms_flood", "ms_assemble", "ms_wall", "ms_topology", "ms_plan",
                                         "ms_memory", "ms_compact", "ms_plan_fill", "ms_plan_tree",
                                         "ms_plan_order")]


class DcStats(C.Structure):
    _fields_ = [("n_leaves", C.c_uint64), ("n_tris", C.c_uint64), ("used_hash", C.c_uint64),
                ("ms_total", C.c_double)]


class BSError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {msg}")
        self.status = status


_lib = None
_lock = threading.Lock()


def load():
    """Load libbs_voxel.so; raises if it was not built."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        # One HIP runtime per process: torch ships its own libamdhip64.so.7, and
        # a second copy (the system one this library links) initialised first
        # leaves torch without devices.  Loading torch first makes the dynamic
        # linker bind this library to torch's runtime (same SONAME).
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"bs_voxel HIP library not built: {LIB_PATH} is missing "
                               "(run `python __graft_entry__.py` or `make -C baby_shark_amd/csrc`)")
        L = C.CDLL(LIB_PATH)
        vp, pp = C.c_void_p, C.POINTER(C.c_void_p)
        f32p, i32p, u64p = C.POINTER(C.c_float), C.POINTER(C.c_int32), C.POINTER(C.c_uint64)
        sig = {
            "bs_last_error": (C.c_char_p, []),
            "bs_ctx_create": (C.c_int, [C.c_int, pp]),
            "bs_ctx_destroy": (None, [vp]),
            "bs_ctx_stream": (vp, [vp]),
            "bs_ctx_set_profiling": (C.c_int, [vp, C.c_int]),
            "bs_ctx_last_m2v_stats": (C.c_int, [vp, C.POINTER(M2VStats)]),
            "bs_ctx_set_paths": (C.c_int, [vp, C.c_int]),
            "bs_ctx_set_dc_vertex": (C.c_int, [vp, C.c_int]),
            "bs_m2v_params_default": (None, [C.POINTER(M2VParams)]),
            "bs_mesh_to_volume": (C.c_int, [vp, f32p, C.c_size_t, C.POINTER(M2VParams), pp]),
            "bs_mesh_to_volume_device": (C.c_int, [vp, vp, C.c_size_t, C.POINTER(M2VParams), pp]),
            "bs_batch_mesh_to_volume": (C.c_int, [vp, C.c_size_t, C.POINTER(vp), C.POINTER(C.c_size_t),
                                                  C.POINTER(M2VParams), C.c_int, C.c_int, pp,
                                                  C.POINTER(C.c_int)]),
            "bs_volume_clone": (C.c_int, [vp, pp]),
            "bs_volume_free": (None, [vp]),
            "bs_release_cached_memory": (None, []),
            "bs_volume_info_get": (C.c_int, [vp, C.POINTER(VolumeInfo)]),
            "bs_volume_export": (C.c_int, [vp, i32p, u64p, f32p]),
            "bs_volume_export_tiles": (C.c_int, [vp, i32p, i32p, f32p]),
            "bs_volume_sphere": (C.c_int, [vp, C.c_float, C.c_float, C.c_float, C.c_float, C.c_float, pp]),
            "bs_volume_cuboid": (C.c_int, [vp] + [C.c_float] * 7 + [pp]),
            "bs_volume_iwp": (C.c_int, [vp] + [C.c_float] * 8 + [pp]),
            "bs_volume_from_voxels": (C.c_int, [vp, i32p, f32p, C.c_size_t, C.c_float, pp]),
            "bs_volume_union": (C.c_int, [vp, vp, vp]),
            "bs_volume_subtract": (C.c_int, [vp, vp, vp]),
            "bs_volume_intersect": (C.c_int, [vp, vp, vp]),
            "bs_ctx_last_csg_stats": (C.c_int, [vp, C.POINTER(CsgStats)]),
            "bs_volume_flood_fill": (C.c_int, [vp, vp]),
            "bs_volume_densify_tiles": (C.c_int, [vp, vp]),
            "bs_volume_offset": (C.c_int, [vp, vp, C.c_float]),
            "bs_dual_contour": (C.c_int, [vp, vp, C.c_float, C.POINTER(f32p), C.POINTER(C.c_size_t)]),
            "bs_dual_contour_device": (C.c_int, [vp, vp, C.c_float, C.POINTER(vp), C.POINTER(C.c_size_t)]),
            "bs_dual_contour_into": (C.c_int, [vp, vp, C.c_float, vp, C.c_size_t, C.POINTER(C.c_size_t)]),
            "bs_ctx_last_dc_stats": (C.c_int, [vp, C.POINTER(DcStats)]),
            "bs_marching_cubes": (C.c_int, [vp, vp, C.c_float, C.POINTER(f32p), C.POINTER(C.c_size_t)]),
            "bs_marching_cubes_device": (C.c_int, [vp, vp, C.c_float, C.POINTER(vp), C.POINTER(C.c_size_t)]),
            "bs_ctx_last_mc_stats": (C.c_int, [vp, C.POINTER(DcStats)]),
            "bs_free": (None, [vp]),
            "bs_voxel_remesh": (C.c_int, [vp, vp, C.c_size_t, C.c_int, C.c_float, C.c_int,
                                          C.POINTER(C.POINTER(C.c_float)), C.POINTER(C.c_size_t)]),
            "bs_stl_read": (C.c_int, [C.c_char_p, C.POINTER(C.POINTER(C.c_float)), C.POINTER(C.c_size_t)]),
            "bs_stl_write": (C.c_int, [C.c_char_p, vp, C.c_size_t]),
            "bs_device_free": (None, [vp]),
            "bs_ctx_last_offset_stats": (C.c_int, [vp, C.POINTER(OffsetStats)]),
            "bs_volume_signs_at": (C.c_int, [vp, vp, i32p, C.c_size_t, C.POINTER(C.c_uint8)]),
            "bs_volume_device_leaves": (C.c_int, [vp, C.POINTER(i32p), C.POINTER(u64p), C.POINTER(f32p),
                                                  C.POINTER(C.c_uint64)]),
            "bs_mlp_create": (C.c_int, [vp, C.c_uint64, pp]),
            "bs_mlp_destroy": (None, [vp]),
            "bs_mlp_set_params": (C.c_int, [vp, vp, f32p]),
            "bs_mlp_get_params": (C.c_int, [vp, vp, f32p]),
            "bs_mlp_set_input_map": (C.c_int, [vp, C.c_float, C.c_float, C.c_float, C.c_float]),
            "bs_mlp_forward_device": (C.c_int, [vp, vp, vp, C.c_size_t, vp]),
            "bs_mlp_gradient_device": (C.c_int, [vp, vp, vp, vp, C.c_size_t, f32p, C.POINTER(C.c_double)]),
            "bs_mlp_train_device": (C.c_int, [vp, vp, vp, vp, C.c_size_t, C.POINTER(MlpTrainParams), f32p]),
            "bs_volume_tile_samples": (C.c_int, [vp, vp, vp, vp, C.c_size_t, C.POINTER(C.c_size_t)]),
            "bs_ctx_set_surrogate": (C.c_int, [vp, vp, C.c_float]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
        return L


def check(status: int) -> None:
    if status != BS_OK:
        msg = load().bs_last_error()
        raise BSError(status, msg.decode() if msg else "")


def header_functions() -> list[str]:
    """Names of the functions declared in include/bs_voxel.h."""
    import re
    hdr = os.path.join(os.path.dirname(HERE), "include", "bs_voxel.h")
    src = open(hdr).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(bs_[a-z0-9_]+)\s*\(", src)))
