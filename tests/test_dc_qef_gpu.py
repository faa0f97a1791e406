"""Dual contouring with FP64 quadric-error vertices (BS_DC_VERTEX_QEF_F64, an
opt-in mode with no counterpart in the reference, whose particle iteration
stays the default and the parity path).

The mode moves vertices only, so the soup keeps the particle soup's
triangles in the same order.  Each vertex must be the truncated
pseudo-inverse solution of its cell's quadric: the test rebuilds the cells'
Hermite data (edge crossings and normals, by the reference's rules) from the
same sampled field in numpy and solves each quadric with LAPACK's symmetric
eigensolver (numpy.linalg.eigh), an independent route to the same minimiser."""
import numpy as np
import pytest

import baby_shark_amd as bs

pytestmark = pytest.mark.gpu

H = 0.05
TRUNC = 0.1  # eigenvalues below this fraction of the largest are dropped


def _rotated_box_field(n=14):
    """max over six planes (a rotated box, 6.3 x 4.7 x 5.2 voxels half-extent):
    linear on each face region, so faces are exact planes and the edges and
    corners are sharp.  Voxels within 3 voxels of the surface, in voxel units."""
    a, b = 0.43, -0.61
    Rz = np.array([[np.cos(a), -np.sin(a), 0], [np.sin(a), np.cos(a), 0], [0, 0, 1]])
    Rx = np.array([[1, 0, 0], [0, np.cos(b), -np.sin(b)], [0, np.sin(b), np.cos(b)]])
    R = Rz @ Rx
    o = np.array([0.31, -0.17, 0.23])
    e = np.array([6.3, 4.7, 5.2])
    g = np.stack(np.meshgrid(*[np.arange(-n, n + 1)] * 3, indexing="ij"), -1).reshape(-1, 3)
    f = (np.abs((g - o) @ R) - e).max(axis=1)
    keep = np.abs(f) < 3.0
    return g[keep], f[keep].astype(np.float32)


_EDGES = [((0, 0, 0), 0), ((0, 0, 0), 1), ((0, 0, 0), 2), ((1, 0, 0), 1), ((1, 0, 0), 2), ((0, 0, 1), 0),
          ((0, 0, 1), 1), ((0, 1, 0), 0), ((0, 1, 0), 2), ((1, 0, 1), 1), ((0, 1, 1), 0), ((1, 1, 0), 2)]


def _qef_vertices(idx, vals):
    """Every surface cell's quadric minimiser (voxel units): crossings and
    normals as dual_contouring.rs:257-349 defines them, in f32; the quadric in
    f64 through numpy.linalg.eigh, truncated and clamped to the cell."""
    lo = idx.min(axis=0) - 2
    dims = idx.max(axis=0) - lo + 3
    V = np.zeros(dims, np.float32)
    A = np.zeros(dims, bool)
    V[tuple((idx - lo).T)] = vals
    A[tuple((idx - lo).T)] = True
    half = np.float32(0.5)

    def grad(c):
        out = np.zeros(3, np.float32)
        for ax in range(3):
            d = np.zeros(3, int)
            d[ax] = 1
            out[ax] = (V[tuple(c + d)] - V[tuple(c - d)]) * half  # both neighbours active in this band
        return out

    verts = []
    for c0 in np.argwhere(A):
        corners = [c0 + np.array(k) for k in np.ndindex(2, 2, 2)]
        if not all(A[tuple(c)] for c in corners):
            continue
        neg = [np.signbit(V[tuple(c)]) for c in corners]
        if all(neg) or not any(neg):
            continue
        P, N = [], []
        for off, d in _EDGES:
            a = c0 + np.array(off)
            b = a.copy()
            b[d] += 1
            v1, v2 = V[tuple(a)], V[tuple(b)]
            if np.signbit(v1) == np.signbit(v2):
                continue
            t = np.float32(0.5) if v1 == v2 else np.float32(v1 / (v1 - v2))
            p = (a + lo).astype(np.float32)
            p[d] = p[d] + t
            nv = (np.float32(1.0) - t) * grad(a) + t * grad(b)
            nv = nv / np.float32(np.sqrt((nv[0] * nv[0] + nv[1] * nv[1]) + nv[2] * nv[2]))
            P.append(p.astype(np.float64))
            N.append(nv.astype(np.float64))
        P, N = np.array(P), np.array(N)
        m = P.mean(axis=0)
        Aq = N.T @ N
        bq = N.T @ np.einsum("ij,ij->i", N, P - m)
        w, E = np.linalg.eigh(Aq)
        x = m.copy()
        for k in range(3):
            if w[k] > TRUNC * w.max():
                x += E[:, k] * (E[:, k] @ bq) / w[k]
        cell = (c0 + lo).astype(np.float64)
        verts.append(np.clip(x, cell, cell + 1.0))
    return np.array(verts)


def _mesh(ctx, qef):
    idx, vals = _rotated_box_field()
    vol = bs.volume_from_voxels(idx, vals * np.float32(H), H, ctx)
    m = bs.DualContouringMesher(ctx).with_voxel_size(H)
    if qef:
        m = m.with_qef_vertices()
    soup = m.mesh(vol)
    vol.free()
    return soup


def test_qef_keeps_the_particle_triangles(gpu_ctx):
    part = _mesh(gpu_ctx, False)
    qef = _mesh(gpu_ctx, True)
    again = _mesh(gpu_ctx, False)
    assert part.shape == qef.shape and part.shape[0] > 0
    # the mode is per call: a default mesher afterwards is the particle one
    assert np.array_equal(part.view(np.uint32), again.view(np.uint32))
    # the same cells: matching vertices stay within a cell of each other
    assert np.max(np.abs(part.astype(np.float64) - qef)) < 2.0 * H
    assert not np.array_equal(part, qef)


def test_qef_vertices_are_the_quadric_minimisers(gpu_ctx):
    idx, vals = _rotated_box_field()
    want = _qef_vertices(idx, vals)
    got = np.unique(_mesh(gpu_ctx, True).reshape(-1, 3).astype(np.float64), axis=0) / H
    d = np.sqrt(((got[:, None, :] - want[None, :, :]) ** 2).sum(-1))
    nearest = d.min(axis=1)
    print(f"{got.shape[0]} soup vertices, {want.shape[0]} surface cells; largest distance to the "
          f"numpy quadric minimiser {nearest.max():.2e} voxel", flush=True)
    assert nearest.max() < 1e-4
    # and the minimisers are the soup's vertices (a cell at the band's edge
    # may have no crossing edge with four cells, so a few can go unused)
    used = np.mean(d.min(axis=0) < 1e-4)
    print(f"minimisers used by the soup: {used:.4f}", flush=True)
    assert used > 0.95
