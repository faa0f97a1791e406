"""Host-side mirror of baby_shark's `voxel::prelude` (src/voxel/prelude.rs:1-4).

Same names, argument meaning and error behaviour as the Rust API, backed by
the HIP kernels of libbs_voxel.so through its C ABI:

* ``MeshToVolume``  -- mesh_to_volume.rs:17-73 (``convert`` returns ``None``
  for an empty mesh, like the reference);
* ``Volume``        -- volume/mod.rs:10-123 (CSG and offset consume their
  operands: the Rust methods take ``self``/``other`` by value).

A mesh is a triangle soup: an ``(n, 3, 3)`` float32 array (triangle, vertex,
xyz), or a float32 tensor of that shape on a GPU, which is used in place.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import Optional

import numpy as np

from . import _lib
from ._lib import BS_DC_VERTEX_PARTICLE, BS_DC_VERTEX_QEF_F64, BSError, check, load


class Context:
    """One HIP stream on one device (``bs_ctx``)."""

    def __init__(self, device: int = 0):
        L = load()
        h = C.c_void_p()
        check(L.bs_ctx_create(int(device), C.byref(h)))
        self._h = h
        self.device = device

    @property
    def handle(self):
        return self._h

    def stream(self) -> int:
        return int(load().bs_ctx_stream(self._h) or 0)

    def set_profiling(self, on: bool, counters: bool = True) -> None:
        """Phase timing with HIP events; with `counters`, winding visit counts
        too (a slower kernel variant: time with counters=False)."""
        check(load().bs_ctx_set_profiling(self._h, (1 | (2 if counters else 0)) if on else 0))

    def set_paths(self, flags: int) -> None:
        """Force alternate code paths (BS_PATH_* flags of bs_voxel.h); 0 = by size."""
        check(load().bs_ctx_set_paths(self._h, int(flags)))

    def set_surrogate(self, classifier, threshold_vox: float) -> None:
        """The tile classifier (TileClassifier, or None for the perfect one
        that reads each tile's true smallest |SDF|) and its cull threshold in
        voxels, used by the surrogate sign mode on this context.  The
        classifier must outlive its use."""
        self._surrogate = classifier
        check(load().bs_ctx_set_surrogate(self._h, classifier.handle if classifier is not None else None,
                                          float(threshold_vox)))

    def last_offset_stats(self) -> dict:
        s = _lib.OffsetStats()
        check(load().bs_ctx_last_offset_stats(self.handle, C.byref(s)))
        return {k: getattr(s, k) for k, _ in s._fields_}

    def last_csg_stats(self) -> dict:
        s = _lib.CsgStats()
        check(load().bs_ctx_last_csg_stats(self.handle, C.byref(s)))
        return {k: getattr(s, k) for k, _ in s._fields_}

    def last_dc_stats(self) -> dict:
        s = _lib.DcStats()
        check(load().bs_ctx_last_dc_stats(self.handle, C.byref(s)))
        return {k: getattr(s, k) for k, _ in s._fields_}

    def last_mc_stats(self) -> dict:
        s = _lib.DcStats()
        check(load().bs_ctx_last_mc_stats(self.handle, C.byref(s)))
        return {k: getattr(s, k) for k, _ in s._fields_}

    def last_stats(self) -> dict:
        s = _lib.M2VStats()
        check(load().bs_ctx_last_m2v_stats(self._h, C.byref(s)))
        return {k: getattr(s, k) for k, _ in s._fields_}

    def close(self) -> None:
        if getattr(self, "_h", None):
            load().bs_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_default_ctx: dict[int, Context] = {}


def default_context(device: int = 0) -> Context:
    ctx = _default_ctx.get(device)
    if ctx is None:
        ctx = _default_ctx[device] = Context(device)
    return ctx


@dataclass
class VolumeExport:
    """Canonical dump: leaves sorted by origin (x, y, z); mask bit i of word w
    is voxel offset w*64+i with offset = (x&7)<<6 | (y&7)<<3 | (z&7)."""
    voxel_size: float
    origins: np.ndarray       # (L, 3) int32
    masks: np.ndarray         # (L, 8) uint64
    values: np.ndarray        # (L, 512) float32
    tile_origins: np.ndarray  # (T, 3) int32
    tile_sizes: np.ndarray    # (T,) int32
    tile_values: np.ndarray   # (T,) float32

    def active(self) -> np.ndarray:
        bits = np.unpackbits(self.masks.view(np.uint8).reshape(-1, 8, 8), axis=2, bitorder="little")
        return bits.reshape(-1, 512).astype(bool)

    @property
    def n_active(self) -> int:
        return int(self.active().sum())


class Volume:
    """A device-resident narrow-band volume (``bs_volume``)."""

    def __init__(self, handle: C.c_void_p, ctx: Context):
        self._h = handle
        self._ctx = ctx

    @staticmethod
    def with_voxel_size(voxel_size: float, ctx: Optional[Context] = None) -> "Volume":
        """An empty volume (Volume::with_voxel_size, volume/mod.rs:16-24)."""
        return volume_from_voxels(np.zeros((0, 3), np.int32), np.zeros(0, np.float32), voxel_size, ctx)

    # ---- ownership ------------------------------------------------------
    def _take(self) -> C.c_void_p:
        if not self._h:
            raise ValueError("volume was consumed by a previous operation")
        h, self._h = self._h, None
        return h

    def free(self) -> None:
        if getattr(self, "_h", None):
            load().bs_volume_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    # ---- reference API --------------------------------------------------
    def info(self) -> dict:
        i = _lib.VolumeInfo()
        check(load().bs_volume_info_get(self._h, C.byref(i)))
        return {"voxel_size": i.voxel_size, "n_leaves": i.n_leaves, "n_active": i.n_active, "n_tiles": i.n_tiles}

    def voxel_size(self) -> float:  # volume/mod.rs:32-34
        return self.info()["voxel_size"]

    def clone(self) -> "Volume":  # volume/mod.rs:116-123
        h = C.c_void_p()
        check(load().bs_volume_clone(self._h, C.byref(h)))
        return Volume(h, self._ctx)

    # ---- CSG (volume/mod.rs:74-93): both operands are consumed ----------
    def _csg(self, other: "Volume", fn: str) -> "Volume":
        if other is self:
            raise ValueError("CSG operands must be two distinct volumes")
        if not self._h or not other._h:
            raise ValueError("volume was consumed by a previous operation")
        check(getattr(load(), fn)(self._ctx.handle, self._h, other._h))
        other._h = None                      # freed by the library
        return Volume(self._take(), self._ctx)

    def union(self, other: "Volume") -> "Volume":
        return self._csg(other, "bs_volume_union")

    def subtract(self, other: "Volume") -> "Volume":
        return self._csg(other, "bs_volume_subtract")

    def intersect(self, other: "Volume") -> "Volume":
        return self._csg(other, "bs_volume_intersect")

    def offset(self, distance: float) -> "Volume":
        """Volume::offset (volume/mod.rs:95-108); consumes this handle."""
        if not self._h:
            raise ValueError("volume was consumed by a previous operation")
        check(load().bs_volume_offset(self._ctx.handle, self._h, float(distance)))
        return Volume(self._take(), self._ctx)

    def flood_fill(self) -> "Volume":
        """FloodFill of the grid in place (root_node/flood_fill.rs:8-42)."""
        check(load().bs_volume_flood_fill(self._ctx.handle, self._h))
        return self

    def densify_tiles(self) -> "Volume":
        """Active tiles become full leaves of the tile value (an extension:
        lets dual contouring run after a union that left tiles)."""
        check(load().bs_volume_densify_tiles(self._ctx.handle, self._h))
        return self

    def signs_at(self, idx) -> np.ndarray:
        """FloodFill::sign_at for (n, 3) voxel indices: True where negative."""
        a = np.ascontiguousarray(np.asarray(idx, dtype=np.int32).reshape(-1, 3))
        out = np.zeros(a.shape[0], np.uint8)
        check(load().bs_volume_signs_at(self._ctx.handle, self._h, a.ctypes.data_as(C.POINTER(C.c_int32)),
                                        a.shape[0], out.ctypes.data_as(C.POINTER(C.c_uint8))))
        return out.astype(bool)

    def device_arrays(self, copy: bool = True):
        """The leaf arrays as device tensors (origins (L, 4) int32, masks (L, 8)
        int64 holding the uint64 mask words' bits, values (L, 512) float32);
        copies unless copy=False, whose views live only as long as this volume."""
        import torch
        L = load()
        o, m, v = C.POINTER(C.c_int32)(), C.POINTER(C.c_uint64)(), C.POINTER(C.c_float)()
        n = C.c_uint64(0)
        check(L.bs_volume_device_leaves(self._h, C.byref(o), C.byref(m), C.byref(v), C.byref(n)))
        nl = n.value

        class _View:
            def __init__(self, ptr, shape, typestr):
                self.__cuda_array_interface__ = {"shape": shape, "typestr": typestr, "version": 2, "strides": None,
                                                 "data": (C.cast(ptr, C.c_void_p).value or 0, False)}
        dev = torch.device("cuda", self._ctx.device)
        out = []
        for ptr, shape, ts, dt in ((o, (nl, 4), "<i4", torch.int32), (m, (nl, 8), "<i8", torch.int64),
                                   (v, (nl, 512), "<f4", torch.float32)):
            if nl == 0:
                out.append(torch.zeros(shape, dtype=dt, device=dev))
                continue
            t = torch.as_tensor(_View(ptr, shape, ts), device=dev)
            out.append(t.clone() if copy else t)
        return tuple(out)

    def export(self) -> VolumeExport:
        L = load()
        inf = self.info()
        n, nt = inf["n_leaves"], inf["n_tiles"]
        o = np.zeros((n, 3), np.int32)
        m = np.zeros((n, 8), np.uint64)
        v = np.zeros((n, 512), np.float32)
        check(L.bs_volume_export(self._h, o.ctypes.data_as(C.POINTER(C.c_int32)),
                                 m.ctypes.data_as(C.POINTER(C.c_uint64)), v.ctypes.data_as(C.POINTER(C.c_float))))
        to = np.zeros((nt, 3), np.int32)
        ts = np.zeros(nt, np.int32)
        tv = np.zeros(nt, np.float32)
        if nt:
            check(L.bs_volume_export_tiles(self._h, to.ctypes.data_as(C.POINTER(C.c_int32)),
                                           ts.ctypes.data_as(C.POINTER(C.c_int32)),
                                           tv.ctypes.data_as(C.POINTER(C.c_float))))
        return VolumeExport(inf["voxel_size"], o, m, v, to, ts, tv)


class VolumeBuilder:
    """builder.rs:5-30: primitives sampled on the grid of one voxel size."""

    def __init__(self, ctx: Optional[Context] = None, voxel_size: float = 1.0):
        self._ctx = ctx
        self.voxel_size = float(voxel_size)

    @property
    def ctx(self) -> Context:
        if self._ctx is None:
            self._ctx = default_context()
        return self._ctx

    def with_voxel_size(self, size: float) -> "VolumeBuilder":
        self.set_voxel_size(size)
        return self

    def set_voxel_size(self, size: float) -> None:
        self.voxel_size = float(size)

    def sphere(self, radius: float, origin) -> Volume:
        """|p - origin| - radius, kept within 2 voxels (builder.rs:21-30)."""
        o = [float(c) for c in origin]
        h = C.c_void_p()
        check(load().bs_volume_sphere(self.ctx.handle, float(radius), o[0], o[1], o[2], self.voxel_size,
                                      C.byref(h)))
        return Volume(h, self.ctx)

    def cuboid(self, min_corner, max_corner) -> Volume:
        """The box's signed distance, kept within 2 voxels (builder.rs:32-51)."""
        a = [float(c) for c in min_corner]
        b = [float(c) for c in max_corner]
        h = C.c_void_p()
        check(load().bs_volume_cuboid(self.ctx.handle, *a, *b, self.voxel_size, C.byref(h)))
        return Volume(h, self.ctx)

    def iwp(self, min_corner, max_corner, cell_size: float) -> Volume:
        """The IWP triply periodic surface in a box, band width 2 (builder.rs:53-76)."""
        a = [float(c) for c in min_corner]
        b = [float(c) for c in max_corner]
        h = C.c_void_p()
        check(load().bs_volume_iwp(self.ctx.handle, *a, *b, float(cell_size), self.voxel_size, C.byref(h)))
        return Volume(h, self.ctx)


def release_cached_memory() -> None:
    """Frees the leaf arrays that dropped volumes left in the library's
    per-device pool (at most 4 sets and 8 GiB per device), for callers that
    need that device memory for other work."""
    load().bs_release_cached_memory()


def volume_from_voxels(idx, values, voxel_size: float, ctx: Optional[Context] = None) -> Volume:
    """A grid built by inserting values[i] at voxel idx[i] in order."""
    ctx = ctx or default_context()
    a = np.ascontiguousarray(np.asarray(idx, dtype=np.int32).reshape(-1, 3))
    v = np.ascontiguousarray(np.asarray(values, dtype=np.float32).reshape(-1))
    if a.shape[0] != v.shape[0]:
        raise ValueError("one value per voxel index")
    h = C.c_void_p()
    check(load().bs_volume_from_voxels(ctx.handle, a.ctypes.data_as(C.POINTER(C.c_int32)),
                                       v.ctypes.data_as(C.POINTER(C.c_float)), a.shape[0], float(voxel_size),
                                       C.byref(h)))
    return Volume(h, ctx)


def batch_mesh_to_volume(meshes, voxel_sizes, band: int, ctx: Optional[Context] = None, workers=0,
                         shared_signs: bool = False):
    """MeshToVolume::convert over many meshes in one library call.

    `meshes`: CUDA tensors (device input) or arrays (host input), all of one
    kind; `voxel_sizes`: one per mesh; `workers`: worker streams (0: the
    library's default of 4; "auto": the count tune.batch_workers measured or
    cached for this device and batch shape).  Returns a list of Volume or
    None (empty mesh), in input order."""
    ctx = ctx or default_context()
    if workers == "auto":
        from .tune import batch_workers
        workers = batch_workers(meshes, voxel_sizes, band, ctx=ctx, shared_signs=shared_signs)
    L = load()
    n = len(meshes)
    dev = n > 0 and hasattr(meshes[0], "is_cuda") and meshes[0].is_cuda
    keep, ptrs, counts = [], (C.c_void_p * n)(), (C.c_size_t * n)()
    params = (_lib.M2VParams * n)()
    for i, m in enumerate(meshes):
        if dev:
            import torch
            t = m.detach().reshape(-1, 9)
            if t.dtype != torch.float32 or not t.is_contiguous():
                t = t.to(torch.float32).contiguous()
            keep.append(t)
            ptrs[i] = t.data_ptr()
            counts[i] = t.shape[0]
        else:
            a = np.ascontiguousarray(np.asarray(m, dtype=np.float32).reshape(-1, 9))
            keep.append(a)
            ptrs[i] = a.ctypes.data
            counts[i] = a.shape[0]
        L.bs_m2v_params_default(C.byref(params[i]))
        params[i].voxel_size = float(voxel_sizes[i])
        params[i].band = int(band)
        params[i].sign_mode = _lib.BS_SIGN_SHARED if shared_signs else _lib.BS_SIGN_PER_VOXEL
    if dev:
        import torch
        torch.cuda.synchronize()
    out = (C.c_void_p * n)()
    status = (C.c_int * n)()
    check(L.bs_batch_mesh_to_volume(ctx.handle, n, ptrs, counts, params, 0 if dev else 1, int(workers), out,
                                    status))
    return [Volume(C.c_void_p(out[i]), ctx) if out[i] else None for i in range(n)]


class DualContouringMesher:
    """meshing/dual_contouring.rs:12-83.  Default voxel size 1.0."""

    def __init__(self, ctx: Optional[Context] = None):
        self._ctx = ctx
        self.voxel_size = 1.0
        self.vertex_mode = BS_DC_VERTEX_PARTICLE

    @property
    def ctx(self) -> Context:
        if self._ctx is None:
            self._ctx = default_context()
        return self._ctx

    def with_voxel_size(self, size: float) -> "DualContouringMesher":
        self.voxel_size = float(size)
        return self

    def with_qef_vertices(self) -> "DualContouringMesher":
        """Vertices at each cell's FP64 quadric-error minimiser (sharp
        features kept) instead of the reference's particle iteration; no
        counterpart in the reference (BS_DC_VERTEX_QEF_F64)."""
        self.vertex_mode = BS_DC_VERTEX_QEF_F64
        return self

    def _call(self, fn, *args) -> None:
        # the mode is a context setting: set for this call, the default after
        L = load()
        check(L.bs_ctx_set_dc_vertex(self.ctx.handle, self.vertex_mode))
        try:
            check(fn(*args))
        finally:
            L.bs_ctx_set_dc_vertex(self.ctx.handle, BS_DC_VERTEX_PARTICLE)

    def mesh(self, volume: "Volume") -> np.ndarray:
        """Triangle soup (n, 3, 3) float32 on the host."""
        L = load()
        out = C.POINTER(C.c_float)()
        n = C.c_size_t(0)
        self._call(L.bs_dual_contour, self.ctx.handle, volume._h, self.voxel_size, C.byref(out), C.byref(n))
        if n.value == 0:
            return np.zeros((0, 3, 3), np.float32)
        try:
            return np.ctypeslib.as_array(out, shape=(n.value * 9,)).copy().reshape(-1, 3, 3)
        finally:
            L.bs_free(C.cast(out, C.c_void_p))

    def mesh_device(self, volume: "Volume") -> "DeviceTriangles":
        """Triangle soup left in device memory (see DeviceTriangles)."""
        ptr = C.c_void_p()
        n = C.c_size_t(0)
        self._call(load().bs_dual_contour_device, self.ctx.handle, volume._h, self.voxel_size, C.byref(ptr),
                   C.byref(n))
        return DeviceTriangles(ptr.value or 0, n.value)


class MarchingCubesMesher:
    """meshing/marching_cubes.rs:17-63 (Marching Cubes 33).  Default voxel
    size 1.0.  Cube by cube the tiling the reference's tables select, in its
    vertex order."""

    def __init__(self, ctx: Optional[Context] = None):
        self._ctx = ctx
        self.voxel_size = 1.0

    @property
    def ctx(self) -> Context:
        if self._ctx is None:
            self._ctx = default_context()
        return self._ctx

    def with_voxel_size(self, size: float) -> "MarchingCubesMesher":
        self.voxel_size = float(size)
        return self

    def set_voxel_size(self, size: float) -> "MarchingCubesMesher":
        return self.with_voxel_size(size)

    def mesh(self, volume: "Volume") -> np.ndarray:
        """Triangle soup (n, 3, 3) float32 on the host."""
        L = load()
        out = C.POINTER(C.c_float)()
        n = C.c_size_t(0)
        check(L.bs_marching_cubes(self.ctx.handle, volume._h, self.voxel_size, C.byref(out), C.byref(n)))
        if n.value == 0:
            return np.zeros((0, 3, 3), np.float32)
        try:
            return np.ctypeslib.as_array(out, shape=(n.value * 9,)).copy().reshape(-1, 3, 3)
        finally:
            L.bs_free(C.cast(out, C.c_void_p))

    def mesh_device(self, volume: "Volume") -> "DeviceTriangles":
        """Triangle soup left in device memory (see DeviceTriangles)."""
        ptr = C.c_void_p()
        n = C.c_size_t(0)
        check(load().bs_marching_cubes_device(self.ctx.handle, volume._h, self.voxel_size, C.byref(ptr),
                                              C.byref(n)))
        return DeviceTriangles(ptr.value or 0, n.value)


class DeviceTriangles:
    """A device-resident (n, 3, 3) float32 soup owned by the library's
    allocator; exposes __cuda_array_interface__ so torch.as_tensor can view it
    (clone before the object is dropped)."""

    def __init__(self, ptr: int, n: int):
        self.ptr, self.n = ptr, n

    @property
    def __cuda_array_interface__(self):
        return {"shape": (self.n, 3, 3), "typestr": "<f4", "data": (self.ptr, False), "version": 2,
                "strides": None}

    def __len__(self):
        return self.n

    def free(self) -> None:
        if self.ptr:
            load().bs_device_free(C.c_void_p(self.ptr))
            self.ptr = 0

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class MeshToVolume:
    """mesh_to_volume.rs:17-73.  Defaults: voxel size 1.0, band 0."""

    def __init__(self, ctx: Optional[Context] = None):
        self._ctx = ctx
        self.params = _lib.M2VParams()
        load().bs_m2v_params_default(C.byref(self.params))

    @property
    def ctx(self) -> Context:
        if self._ctx is None:
            self._ctx = default_context()
        return self._ctx

    def with_narrow_band_width(self, width: int) -> "MeshToVolume":
        self.set_narrow_band_width(width)
        return self

    def set_narrow_band_width(self, width: int) -> "MeshToVolume":
        self.params.band = int(width)
        return self

    def with_voxel_size(self, size: float) -> "MeshToVolume":
        self.set_voxel_size(size)
        return self

    def set_voxel_size(self, size: float) -> "MeshToVolume":
        self.params.voxel_size = float(size)
        return self

    def with_surrogate_signs(self) -> "MeshToVolume":
        """The tile-gated sign mode (BS_SIGN_SURROGATE, closed meshes only):
        the context's classifier (Context.set_surrogate) picks the tiles whose
        voxels share one winding-number evaluation where the exact distances
        prove it; the others are evaluated voxel by voxel."""
        self.params.sign_mode = _lib.BS_SIGN_SURROGATE
        return self

    def with_shared_signs(self, on: bool = True) -> "MeshToVolume":
        """One winding-number evaluation per group of voxels that provably
        share a sign (closed meshes only); off = the reference's per-voxel
        signs."""
        self.params.sign_mode = _lib.BS_SIGN_SHARED if on else _lib.BS_SIGN_PER_VOXEL
        return self

    def convert(self, mesh) -> Optional[Volume]:
        """Narrow-band SDF of a triangle soup; ``None`` for an empty mesh."""
        L = load()
        h = C.c_void_p()
        if hasattr(mesh, "is_cuda") and mesh.is_cuda:
            import torch
            t = mesh.detach().reshape(-1, 9)
            if t.dtype != torch.float32 or not t.is_contiguous():
                t = t.to(torch.float32).contiguous()
            torch.cuda.current_stream(t.device).synchronize()
            st = L.bs_mesh_to_volume_device(self.ctx.handle, C.c_void_p(t.data_ptr()), t.shape[0],
                                            C.byref(self.params), C.byref(h))
        else:
            a = np.ascontiguousarray(np.asarray(mesh, dtype=np.float32).reshape(-1, 9))
            st = L.bs_mesh_to_volume(self.ctx.handle, a.ctypes.data_as(C.POINTER(C.c_float)), a.shape[0],
                                     C.byref(self.params), C.byref(h))
        if st == _lib.BS_ERR_EMPTY:
            return None
        check(st)
        return Volume(h, self.ctx)


def read_stl(path) -> np.ndarray:
    """Binary STL (StlReader, io/stl.rs:64-92) as an (n, 3, 3) float32 soup."""
    L = load()
    out = C.POINTER(C.c_float)()
    n = C.c_size_t(0)
    check(L.bs_stl_read(str(path).encode(), C.byref(out), C.byref(n)))
    if n.value == 0:
        return np.zeros((0, 3, 3), np.float32)
    try:
        return np.ctypeslib.as_array(out, shape=(n.value * 9,)).copy().reshape(-1, 3, 3)
    finally:
        L.bs_free(C.cast(out, C.c_void_p))


def write_stl(path, soup) -> None:
    """Binary STL (StlWriter, io/stl.rs:150-199) of an (n, 3, 3) soup."""
    a = np.ascontiguousarray(np.asarray(soup, dtype=np.float32).reshape(-1, 9))
    check(load().bs_stl_write(str(path).encode(), a.ctypes.data_as(C.c_void_p), a.shape[0]))


class MeshingMethod:
    """remeshing/voxel.rs:10-15."""
    FEATURE_PRESERVING = _lib.BS_MESHING_FEATURE_PRESERVING  # dual contouring
    MANIFOLD = _lib.BS_MESHING_MANIFOLD                      # marching cubes (the default)


class VoxelRemesher:
    """remeshing/voxel.rs:44-95: mesh -> band-0 SDF -> mesh, in one library
    call.  The default method is MANIFOLD (marching cubes), as in the
    reference (:87-95)."""

    def __init__(self, ctx: Optional[Context] = None):
        self._ctx = ctx
        self.voxel_size = 1.0
        self.method = MeshingMethod.MANIFOLD

    @property
    def ctx(self) -> Context:
        if self._ctx is None:
            self._ctx = default_context()
        return self._ctx

    def with_voxel_size(self, size: float) -> "VoxelRemesher":
        self.voxel_size = float(size)
        return self

    def with_meshing_method(self, method: int) -> "VoxelRemesher":
        self.method = int(method)
        return self

    def remesh(self, mesh) -> Optional[np.ndarray]:
        """(m, 3, 3) float32 soup; None for an empty mesh."""
        L = load()
        out = C.POINTER(C.c_float)()
        n = C.c_size_t(0)
        if hasattr(mesh, "is_cuda") and mesh.is_cuda:
            import torch
            t = mesh.detach().reshape(-1, 9)
            if t.dtype != torch.float32 or not t.is_contiguous():
                t = t.to(torch.float32).contiguous()
            torch.cuda.current_stream(t.device).synchronize()
            st = L.bs_voxel_remesh(self.ctx.handle, C.c_void_p(t.data_ptr()), t.shape[0], 1, self.voxel_size,
                                   self.method, C.byref(out), C.byref(n))
        else:
            a = np.ascontiguousarray(np.asarray(mesh, dtype=np.float32).reshape(-1, 9))
            st = L.bs_voxel_remesh(self.ctx.handle, a.ctypes.data_as(C.c_void_p), a.shape[0], 0, self.voxel_size,
                                   self.method, C.byref(out), C.byref(n))
        if st == _lib.BS_ERR_EMPTY:
            return None
        check(st)
        if n.value == 0:
            return np.zeros((0, 3, 3), np.float32)
        try:
            return np.ctypeslib.as_array(out, shape=(n.value * 9,)).copy().reshape(-1, 3, 3)
        finally:
            L.bs_free(C.cast(out, C.c_void_p))


class TileClassifier:
    """The tile classifier of BS_SIGN_SURROGATE (``bs_mlp``): 3 -> 64 -> 64 ->
    64 -> 1 ReLU network, BF16 weights and activations, FP32 accumulation,
    trained on the device.  Predicts a 4^3 tile's smallest |SDF| in voxels
    from its centre.  No counterpart in the reference."""

    N_PARAMS = _lib.BS_MLP_PARAMS

    def __init__(self, ctx: Optional[Context] = None, seed: int = 0):
        self.ctx = ctx or default_context()
        h = C.c_void_p()
        check(load().bs_mlp_create(self.ctx.handle, int(seed), C.byref(h)))
        self._h = h

    def close(self) -> None:
        if getattr(self, "_h", None):
            load().bs_mlp_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    @property
    def params(self) -> np.ndarray:
        out = np.zeros(self.N_PARAMS, np.float32)
        check(load().bs_mlp_get_params(self.ctx.handle, self._h, out.ctypes.data_as(C.POINTER(C.c_float))))
        return out

    @params.setter
    def params(self, p) -> None:
        a = np.ascontiguousarray(np.asarray(p, np.float32).reshape(-1))
        assert a.size == self.N_PARAMS
        check(load().bs_mlp_set_params(self.ctx.handle, self._h, a.ctypes.data_as(C.POINTER(C.c_float))))

    def set_input_map(self, shift, scale: float) -> "TileClassifier":
        check(load().bs_mlp_set_input_map(self._h, float(shift[0]), float(shift[1]), float(shift[2]), float(scale)))
        return self

    @staticmethod
    def _dev(t):
        import torch
        t = t.detach().to(dtype=torch.float32).contiguous()
        assert t.is_cuda
        return t

    def forward(self, x):
        """Predictions for the (n, 3) device tensor x (a device tensor back)."""
        import torch
        x = self._dev(x).reshape(-1, 3)
        out = torch.empty(x.shape[0], dtype=torch.float32, device=x.device)
        torch.cuda.current_stream(x.device).synchronize()
        check(load().bs_mlp_forward_device(self.ctx.handle, self._h, C.c_void_p(x.data_ptr()), x.shape[0],
                                           C.c_void_p(out.data_ptr())))
        return out

    def gradient(self, x, y):
        """(gradient of mean((net(x) - y)^2) as BS_MLP_PARAMS floats, that loss)."""
        import torch
        x = self._dev(x).reshape(-1, 3)
        y = self._dev(y).reshape(-1)
        g = np.zeros(self.N_PARAMS, np.float32)
        loss = C.c_double(0.0)
        torch.cuda.current_stream(x.device).synchronize()
        check(load().bs_mlp_gradient_device(self.ctx.handle, self._h, C.c_void_p(x.data_ptr()),
                                            C.c_void_p(y.data_ptr()), x.shape[0],
                                            g.ctypes.data_as(C.POINTER(C.c_float)), C.byref(loss)))
        return g, loss.value

    def train(self, x, y, epochs: int = 20, batch: int = 8192, lr: float = 3e-3, lr_final: float = 1e-4,
              seed: int = 0) -> np.ndarray:
        """Mini-batch Adam on the squared error; returns each epoch's mean loss."""
        import torch
        x = self._dev(x).reshape(-1, 3)
        y = self._dev(y).reshape(-1)
        tp = _lib.MlpTrainParams(epochs, batch, lr, lr_final, 0.9, 0.999, 1e-8, seed)
        losses = np.zeros(epochs, np.float32)
        torch.cuda.current_stream(x.device).synchronize()
        check(load().bs_mlp_train_device(self.ctx.handle, self._h, C.c_void_p(x.data_ptr()),
                                         C.c_void_p(y.data_ptr()), x.shape[0], C.byref(tp),
                                         losses.ctypes.data_as(C.POINTER(C.c_float))))
        return losses


def tile_samples(volume: "Volume"):
    """(centres (n, 3), smallest |value| / voxel size (n,)) of the tiles of a
    volume that hold an active voxel, as device tensors."""
    import torch
    n_cap = 8 * volume.info()["n_leaves"]
    x = torch.empty((max(n_cap, 1), 3), dtype=torch.float32, device="cuda")
    y = torch.empty(max(n_cap, 1), dtype=torch.float32, device="cuda")
    n = C.c_size_t(0)
    torch.cuda.synchronize()
    check(load().bs_volume_tile_samples(volume._ctx.handle, volume._h, C.c_void_p(x.data_ptr()),
                                        C.c_void_p(y.data_ptr()), n_cap, C.byref(n)))
    keep = y[:n.value] >= 0
    return x[:n.value][keep], y[:n.value][keep]
