"""ctypes front end of the CPU oracle (test infrastructure).

Only tests/, ``__graft_entry__.smoke()`` and bench.py's ``cpu_baseline`` leg
import this module, and only to check or time the reference algorithm; no
product path calls it.  The library is the C++ restatement under this
directory, built by ``make -C oracle``.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# BS_ORACLE_LIB names another build of the same library (the sanitizer build
# of `make -C oracle asan`, for tests/test_sanitizers.py)
LIB_PATH = os.environ.get("BS_ORACLE_LIB") or os.path.join(_HERE, "_build", "liboracle.so")
_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library missing: {LIB_PATH} (run `make -C oracle`)")
        L = C.CDLL(LIB_PATH)
        vp, f32p, f64p, i32p, u64p = C.c_void_p, C.POINTER(C.c_float), C.POINTER(C.c_double), \
            C.POINTER(C.c_int32), C.POINTER(C.c_uint64)
        L.or_mesh_to_volume.restype = vp
        L.or_mesh_to_volume.argtypes = [f32p, C.c_size_t, C.c_float, C.c_int64, C.c_int, f64p]
        L.or_unsigned_distance.restype = vp
        L.or_unsigned_distance.argtypes = [f32p, C.c_size_t, C.c_float, C.c_int64, C.c_int]
        L.or_volume_free.argtypes = [vp]
        L.or_volume_leaf_count.restype = C.c_size_t
        L.or_volume_leaf_count.argtypes = [vp]
        L.or_volume_tile_count.restype = C.c_size_t
        L.or_volume_tile_count.argtypes = [vp]
        L.or_volume_export.argtypes = [vp, i32p, u64p, f32p]
        L.or_volume_export_tiles.argtypes = [vp, i32p, i32p, f32p]
        L.or_subdivide.restype = C.c_size_t
        L.or_subdivide.argtypes = [f32p, C.c_size_t, C.c_float, f32p]
        L.or_pair_count.restype = C.c_uint64
        L.or_pair_count.argtypes = [f32p, C.c_size_t, C.c_float, C.c_int64]
        L.or_winding_numbers.argtypes = [f32p, C.c_size_t, f32p, C.c_size_t, C.c_float, f32p, f64p, f32p, C.c_int]
        L.or_solid_angle.restype = C.c_float
        L.or_bvh_export.restype = C.c_size_t
        L.or_bvh_export.argtypes = [f32p, C.c_size_t, C.POINTER(C.c_uint8), u64p, u64p, f32p, f32p, f32p, f32p]
        L.or_dc_feature_point.argtypes = [f32p, f32p, C.c_size_t, f32p]
        L.or_solid_angle.argtypes = [f32p, f32p]
        L.or_closest_point.argtypes = [f32p, f32p, f32p]
        L.or_unsigned_distance_f64.argtypes = [vp, f32p, C.c_size_t, C.c_int64, f64p, C.c_int]
        L.or_active_set.restype = C.c_size_t
        L.or_active_set.argtypes = [f32p, C.c_size_t, C.c_float, C.c_int64, C.c_int]
        L.or_active_set_export.argtypes = [i32p, u64p]
        L.or_sample_unsigned.argtypes = [f32p, C.c_size_t, C.c_float, C.c_int64, i32p, C.c_size_t, f32p, f64p,
                                         C.c_int]
        u8p = C.POINTER(C.c_uint8)
        L.or_volume_sphere.restype = vp
        L.or_volume_sphere.argtypes = [C.c_float] * 5
        L.or_volume_cuboid.restype = vp
        L.or_volume_cuboid.argtypes = [f32p, f32p, C.c_float]
        L.or_volume_iwp.restype = vp
        L.or_volume_iwp.argtypes = [f32p, f32p, C.c_float, C.c_float]
        L.or_volume_densify_tiles.argtypes = [vp]
        L.or_dual_contour.restype = C.c_int
        L.or_dual_contour.argtypes = [vp, C.c_float, C.POINTER(C.POINTER(C.c_float)), C.POINTER(C.c_size_t)]
        L.or_marching_cubes.restype = C.c_int
        L.or_marching_cubes.argtypes = [vp, C.c_float, C.POINTER(C.POINTER(C.c_float)), C.POINTER(C.c_size_t)]
        L.or_mc_cube_tiling.restype = C.c_int
        L.or_mc_cube_tiling.argtypes = [C.POINTER(C.c_float), C.POINTER(C.c_int)]
        L.or_free.argtypes = [vp]
        L.or_fs_compute_distance.restype = C.c_float
        L.or_fs_compute_distance.argtypes = [C.c_float] * 4
        L.or_volume_offset.restype = C.c_int
        L.or_volume_offset.argtypes = [vp, C.c_float]
        L.or_volume_from_voxels.restype = vp
        L.or_volume_from_voxels.argtypes = [C.POINTER(C.c_int32), C.POINTER(C.c_float), C.c_size_t, C.c_float]
        L.or_volume_from_leaves.restype = vp
        L.or_volume_from_leaves.argtypes = [C.POINTER(C.c_int32), C.POINTER(C.c_uint64), C.POINTER(C.c_float),
                                            C.c_size_t, C.c_float]
        L.or_volume_csg.argtypes = [vp, vp, C.c_int]
        L.or_volume_flood_fill.argtypes = [vp]
        L.or_volume_clone.restype = vp
        L.or_volume_clone.argtypes = [vp]
        L.or_volume_signs_at.argtypes = [vp, i32p, C.c_size_t, u8p]
        L.or_reftest_leaf_csg.argtypes = [C.c_int, f32p, f32p, f32p, C.POINTER(C.c_int)]
        L.or_reftest_leaf_flood_fill.argtypes = [C.c_int, u8p]
        L.or_reftest_internal_flood_fill.argtypes = [C.c_int, u8p]
        L.or_reftest_root_flood_fill.argtypes = [u8p]
        _lib = L
    return _lib


def _f32(a):
    a = np.ascontiguousarray(a, dtype=np.float32)
    return a, a.ctypes.data_as(C.POINTER(C.c_float))


@dataclass
class VolumeExport:
    """Canonical dump: leaves sorted by origin; mask bit i of word w = voxel w*64+i."""
    voxel_size: float
    origins: np.ndarray  # (L,3) int32
    masks: np.ndarray    # (L,8) uint64
    values: np.ndarray   # (L,512) float32
    tile_origins: np.ndarray  # (T,3) int32
    tile_sizes: np.ndarray    # (T,) int32
    tile_values: np.ndarray   # (T,) float32

    def active(self) -> np.ndarray:
        """(L,512) bool active flags."""
        bits = np.unpackbits(self.masks.view(np.uint8).reshape(-1, 8, 8), axis=2, bitorder="little")
        return bits.reshape(-1, 512).astype(bool)

    @property
    def n_active(self) -> int:
        return int(self.active().sum())


class Volume:
    def __init__(self, handle, voxel_size):
        self.h = handle
        self.voxel_size = voxel_size

    def _take(self):
        h, self.h = self.h, None
        return h

    def clone(self) -> "Volume":
        return Volume(lib().or_volume_clone(self.h), self.voxel_size)

    def _csg(self, other: "Volume", op: int) -> "Volume":
        lib().or_volume_csg(self.h, other._take(), op)
        return self

    def union(self, other):      # volume/mod.rs:74-79
        return self._csg(other, 0)

    def subtract(self, other):   # volume/mod.rs:88-93
        return self._csg(other, 1)

    def intersect(self, other):  # volume/mod.rs:81-86
        return self._csg(other, 2)

    def offset(self, distance: float) -> "Volume":   # volume/mod.rs:95-108
        if lib().or_volume_offset(self.h, distance) != 0:
            raise RuntimeError("offset: the reference panics (tile origin in the leaf queue)")
        return self

    def densify_tiles(self) -> "Volume":
        """Active tiles become full leaves of the tile value (an extension)."""
        lib().or_volume_densify_tiles(self.h)
        return self

    def marching_cubes(self, voxel_size: float) -> np.ndarray:
        """MarchingCubesMesher (marching_cubes.rs:43-63): (n_tris, 3, 3) float32."""
        out = C.POINTER(C.c_float)()
        n = C.c_size_t(0)
        lib().or_marching_cubes(self.h, voxel_size, C.byref(out), C.byref(n))
        try:
            return np.ctypeslib.as_array(out, shape=(n.value * 9,)).copy().reshape(-1, 3, 3) if n.value else \
                np.zeros((0, 3, 3), np.float32)
        finally:
            lib().or_free(C.cast(out, C.c_void_p))

    def dual_contour(self, voxel_size: float) -> np.ndarray:
        """DualContouringMesher (dual_contouring.rs:23-83): (n_tris, 3, 3) float32."""
        out = C.POINTER(C.c_float)()
        n = C.c_size_t(0)
        rc = lib().or_dual_contour(self.h, voxel_size, C.byref(out), C.byref(n))
        if rc != 0:
            raise RuntimeError({1: "dual contouring: tile (reference todo!)",
                                2: "dual contouring: gradient without neighbours (reference unreachable!)"}[rc])
        try:
            return np.ctypeslib.as_array(out, shape=(n.value * 9,)).copy().reshape(-1, 3, 3) if n.value else \
                np.zeros((0, 3, 3), np.float32)
        finally:
            lib().or_free(C.cast(out, C.c_void_p))

    def flood_fill(self) -> "Volume":
        lib().or_volume_flood_fill(self.h)
        return self

    def signs_at(self, idx) -> np.ndarray:
        a = np.ascontiguousarray(idx, dtype=np.int32)
        out = np.zeros(a.shape[0], np.uint8)
        lib().or_volume_signs_at(self.h, a.ctypes.data_as(C.POINTER(C.c_int32)), a.shape[0],
                                 out.ctypes.data_as(C.POINTER(C.c_uint8)))
        return out.astype(bool)

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_volume_free(self.h)
            self.h = None

    def export(self) -> VolumeExport:
        L = lib()
        n = L.or_volume_leaf_count(self.h)
        o = np.zeros((n, 3), np.int32)
        m = np.zeros((n, 8), np.uint64)
        v = np.zeros((n, 512), np.float32)
        L.or_volume_export(self.h, o.ctypes.data_as(C.POINTER(C.c_int32)), m.ctypes.data_as(C.POINTER(C.c_uint64)),
                           v.ctypes.data_as(C.POINTER(C.c_float)))
        nt = L.or_volume_tile_count(self.h)
        to = np.zeros((nt, 3), np.int32)
        ts = np.zeros(nt, np.int32)
        tv = np.zeros(nt, np.float32)
        if nt:
            L.or_volume_export_tiles(self.h, to.ctypes.data_as(C.POINTER(C.c_int32)),
                                     ts.ctypes.data_as(C.POINTER(C.c_int32)), tv.ctypes.data_as(C.POINTER(C.c_float)))
        return VolumeExport(self.voxel_size, o, m, v, to, ts, tv)


def mc_cube_tiling(corner_values) -> list:
    """One cube's triangles as crossing-edge triples (12 = centre vertex)."""
    c = np.ascontiguousarray(corner_values, np.float32)
    out = (C.c_int * 48)()
    n = lib().or_mc_cube_tiling(c.ctypes.data_as(C.POINTER(C.c_float)), out)
    return [tuple(out[3 * i:3 * i + 3]) for i in range(n)]


def mesh_to_volume(soup, voxel_size: float, band: int, threads: int = 0, timing: bool = False):
    """MeshToVolume::default().with_voxel_size(h).with_narrow_band_width(band).convert(soup)."""
    a, p = _f32(soup)
    sec = C.c_double(0.0)
    h = lib().or_mesh_to_volume(p, a.size // 9, voxel_size, band, threads, C.byref(sec))
    vol = Volume(h, voxel_size) if h else None
    return (vol, sec.value) if timing else vol


def unsigned_distance(soup, voxel_size: float, band: int, threads: int = 0):
    a, p = _f32(soup)
    h = lib().or_unsigned_distance(p, a.size // 9, voxel_size, band, threads)
    return Volume(h, voxel_size) if h else None


def subdivide(soup, voxel_size: float) -> np.ndarray:
    a, p = _f32(soup)
    n = lib().or_subdivide(p, a.size // 9, voxel_size, None)
    out = np.zeros((n, 3, 3), np.float32)
    lib().or_subdivide(p, a.size // 9, voxel_size, out.ctypes.data_as(C.POINTER(C.c_float)))
    return out


def pair_count(soup, voxel_size: float, band: int) -> int:
    a, p = _f32(soup)
    return int(lib().or_pair_count(p, a.size // 9, voxel_size, band))


def winding_numbers(soup, points, accuracy: float = 2.0, exact: bool = False, threads: int = 0):
    a, p = _f32(soup)
    q, qp = _f32(points)
    m = q.size // 3
    o32 = np.zeros(m, np.float32)
    o64 = np.zeros(m, np.float64)
    ex = np.zeros(m, np.float32) if exact else None
    lib().or_winding_numbers(p, a.size // 9, qp, m, accuracy, o32.ctypes.data_as(C.POINTER(C.c_float)),
                             o64.ctypes.data_as(C.POINTER(C.c_double)),
                             ex.ctypes.data_as(C.POINTER(C.c_float)) if exact else None, threads)
    return (o32, o64, ex) if exact else (o32, o64)


def solid_angle(tri, q) -> float:
    t, tp = _f32(tri)
    qq, qp = _f32(q)
    return float(lib().or_solid_angle(tp, qp))


def closest_point(tri, q) -> np.ndarray:
    t, tp = _f32(tri)
    qq, qp = _f32(q)
    out = np.zeros(3, np.float32)
    lib().or_closest_point(tp, qp, out.ctypes.data_as(C.POINTER(C.c_float)))
    return out


@dataclass
class Bvh:
    """The reference-tree BVH (aabb_tree.rs:636-644): nodes in push order, root last."""
    leaf: np.ndarray     # (N,) bool
    left: np.ndarray     # (N,) object start (leaf) or left child (branch)
    right: np.ndarray    # (N,) object end (leaf) or right child (branch)
    bbox: np.ndarray     # (N,6) min xyz, max xyz
    objects: np.ndarray  # (T,3,3) triangles in tree order
    centre: np.ndarray   # (N,3) dipole centre
    radius: np.ndarray   # (N,) dipole radius


def bvh_export(soup) -> Bvh:
    a, p = _f32(soup)
    n = a.size // 9
    L = lib()
    nn = L.or_bvh_export(p, n, None, None, None, None, None, None, None)
    leaf = np.zeros(nn, np.uint8)
    left = np.zeros(nn, np.uint64)
    right = np.zeros(nn, np.uint64)
    bbox = np.zeros((nn, 6), np.float32)
    objs = np.zeros((n, 3, 3), np.float32)
    cen = np.zeros((nn, 3), np.float32)
    rad = np.zeros(nn, np.float32)
    fp = lambda x: x.ctypes.data_as(C.POINTER(C.c_float))  # noqa: E731
    up = lambda x: x.ctypes.data_as(C.POINTER(C.c_uint64))  # noqa: E731
    L.or_bvh_export(p, n, leaf.ctypes.data_as(C.POINTER(C.c_uint8)), up(left), up(right), fp(bbox), fp(objs),
                    fp(cen), fp(rad))
    return Bvh(leaf.astype(bool), left.astype(np.int64), right.astype(np.int64), bbox, objs, cen, rad)


def dc_feature_point(points, normals) -> np.ndarray:
    """find_feature_point (dual_contouring.rs:429-458) in voxel units."""
    pa, pp = _f32(points)
    na, npp = _f32(normals)
    assert pa.shape == na.shape and pa.size % 3 == 0
    out = np.zeros(3, np.float32)
    lib().or_dc_feature_point(pp, npp, pa.size // 3, out.ctypes.data_as(C.POINTER(C.c_float)))
    return out


def unsigned_distance_f64(vol: Volume, soup, band: int, threads: int = 0) -> np.ndarray:
    """(L,512) float64 min over the reference pair set, canonical leaf order."""
    a, p = _f32(soup)
    n = lib().or_volume_leaf_count(vol.h)
    out = np.zeros((n, 512), np.float64)
    lib().or_unsigned_distance_f64(vol.h, p, a.size // 9, band, out.ctypes.data_as(C.POINTER(C.c_double)), threads)
    return out


def active_set(soup, voxel_size: float, band: int, threads: int = 0):
    """(origins (L,3) int32, masks (L,8) uint64): union of dilated boxes, canonical order."""
    a, p = _f32(soup)
    n = lib().or_active_set(p, a.size // 9, voxel_size, band, threads)
    o = np.zeros((n, 3), np.int32)
    m = np.zeros((n, 8), np.uint64)
    lib().or_active_set_export(o.ctypes.data_as(C.POINTER(C.c_int32)), m.ctypes.data_as(C.POINTER(C.c_uint64)))
    return o, m


def sample_unsigned(soup, voxel_size: float, band: int, voxels, threads: int = 0):
    """Reference f32 unsigned distance and its FP64 counterpart at voxel indices (m,3)."""
    a, p = _f32(soup)
    v = np.ascontiguousarray(voxels, dtype=np.int32)
    m = v.shape[0]
    o32 = np.zeros(m, np.float32)
    o64 = np.zeros(m, np.float64)
    lib().or_sample_unsigned(p, a.size // 9, voxel_size, band, v.ctypes.data_as(C.POINTER(C.c_int32)), m,
                             o32.ctypes.data_as(C.POINTER(C.c_float)), o64.ctypes.data_as(C.POINTER(C.c_double)),
                             threads)
    return o32, o64


def sphere(radius: float, center, voxel_size: float) -> Volume:
    """VolumeBuilder::default().with_voxel_size(h).sphere(radius, center)."""
    return Volume(lib().or_volume_sphere(radius, center[0], center[1], center[2], voxel_size), voxel_size)


def cuboid(mn, mx, voxel_size: float) -> Volume:
    """VolumeBuilder::default().with_voxel_size(h).cuboid(min, max) (builder.rs:32-51)."""
    a, pa = _f32(mn)
    b, pb = _f32(mx)
    return Volume(lib().or_volume_cuboid(pa, pb, voxel_size), voxel_size)


def iwp(mn, mx, cell_size: float, voxel_size: float) -> Volume:
    """VolumeBuilder::default().with_voxel_size(h).iwp(min, max, cell_size) (builder.rs:53-76)."""
    a, pa = _f32(mn)
    b, pb = _f32(mx)
    return Volume(lib().or_volume_iwp(pa, pb, cell_size, voxel_size), voxel_size)


def fs_compute_distance(a1, a2, a3, h) -> float:
    return lib().or_fs_compute_distance(a1, a2, a3, h)


def from_voxels(idx, values, voxel_size: float) -> Volume:
    """An empty grid receiving insert(idx[i], values[i]) in order."""
    a = np.ascontiguousarray(np.asarray(idx, np.int32).reshape(-1, 3))
    v = np.ascontiguousarray(np.asarray(values, np.float32).reshape(-1))
    assert a.shape[0] == v.shape[0]
    return Volume(lib().or_volume_from_voxels(a.ctypes.data_as(C.POINTER(C.c_int32)),
                                              v.ctypes.data_as(C.POINTER(C.c_float)), a.shape[0], voxel_size),
                  voxel_size)


def from_leaves(exp, voxel_size: float) -> Volume:
    """A grid holding the leaves of a canonical export as they are (masks and
    all 512 values per leaf); the export's tiles are not carried."""
    o = np.ascontiguousarray(exp.origins, np.int32)
    m = np.ascontiguousarray(exp.masks, np.uint64)
    v = np.ascontiguousarray(exp.values, np.float32)
    assert o.shape[0] == m.shape[0] == v.shape[0] and exp.tile_origins.shape[0] == 0
    return Volume(lib().or_volume_from_leaves(o.ctypes.data_as(C.POINTER(C.c_int32)),
                                              m.ctypes.data_as(C.POINTER(C.c_uint64)),
                                              v.ctypes.data_as(C.POINTER(C.c_float)), o.shape[0], voxel_size),
                  voxel_size)


def reftest_leaf_csg(op: int, a, b):
    a = np.ascontiguousarray(a, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    out = np.zeros(8, np.float32)
    full = C.c_int(0)
    lib().or_reftest_leaf_csg(op, a.ctypes.data_as(C.POINTER(C.c_float)), b.ctypes.data_as(C.POINTER(C.c_float)),
                              out.ctypes.data_as(C.POINTER(C.c_float)), C.byref(full))
    return out, bool(full.value)


def _signs(fn, n, *args):
    out = np.zeros(n, np.uint8)
    fn(*args, out.ctypes.data_as(C.POINTER(C.c_uint8)))
    return out.astype(bool)


def reftest_leaf_flood_fill(which: int):
    return _signs(lib().or_reftest_leaf_flood_fill, 64, which)


def reftest_internal_flood_fill(which: int):
    return _signs(lib().or_reftest_internal_flood_fill, 512, which)


def reftest_root_flood_fill():
    return _signs(lib().or_reftest_root_flood_fill, 2 * 2 * 8)
