"""Pins the oracle's VolumeBuilder::cuboid and ::iwp (builder.rs:32-76)
against closed forms evaluated in float64.  CPU only."""
import numpy as np


def _points(e, h):
    li, off = np.nonzero(e.active())
    idx = e.origins[li] + np.stack([off >> 6, (off >> 3) & 7, off & 7], 1)
    return idx, e.values[li, off]


def _box_sdf(p, mn, mx):
    out = np.maximum(np.maximum(mn - p, p - mx), 0.0)
    inside = np.all(out == 0.0, axis=1)
    depth = np.min(np.minimum(p - mn, mx - p), axis=1)
    return np.where(inside, -depth, np.linalg.norm(out, axis=1))


def test_cuboid_is_the_box_sdf(oracle_lib):
    h = 0.05
    mn, mx = np.array([-0.33, -0.21, 0.07]), np.array([0.41, 0.18, 0.52])
    e = oracle_lib.cuboid(mn, mx, h).export()
    idx, v = _points(e, h)
    p = idx.astype(np.float32).astype(np.float64) * np.float32(h)
    want = _box_sdf(p, mn.astype(np.float32).astype(np.float64), mx.astype(np.float32).astype(np.float64))
    assert np.max(np.abs(v - want)) < 1e-6
    # every grid point of the index box with |f| <= 2h is present, no other
    lo = np.floor((mn - h) / h).astype(int)
    hi = np.ceil((mx + h) / h).astype(int)
    g = np.stack(np.meshgrid(*[np.arange(lo[k], hi[k] + 1) for k in range(3)], indexing="ij"), -1).reshape(-1, 3)
    f = _box_sdf(g * h, mn, mx)
    keep = np.abs(f) <= 2 * h - 1e-6
    border = np.abs(np.abs(f) - 2 * h) <= 1e-6
    got = {tuple(x) for x in idx.tolist()}
    assert {tuple(x) for x in g[keep].tolist()} <= got
    assert got <= {tuple(x) for x in g[keep | border].tolist()}


def test_iwp_formula_and_shell(oracle_lib):
    h, cell = 0.05, 0.5
    mn, mx = np.array([-0.6, -0.6, 0.0]), np.array([0.6, 0.6, 0.9])
    e = oracle_lib.iwp(mn, mx, cell, h).export()
    idx, v = _points(e, h)
    p = idx.astype(np.float32).astype(np.float64) * np.float32(h)
    mn32, mx32 = mn.astype(np.float32).astype(np.float64), mx.astype(np.float32).astype(np.float64)
    inside = np.all((p >= mn32) & (p <= mx32), axis=1)
    c = np.cos(p[inside] / cell)
    w = -(c.sum(1) - 0.51 * (c[:, 0] * c[:, 1] + c[:, 1] * c[:, 2] + c[:, 2] * c[:, 0]) - 1.0) * cell
    assert np.max(np.abs(v[inside] - w)) < 2e-5 * cell
    # outside the sampling box: the distance to the box grown by one voxel,
    # which is exactly zero in the one-voxel shell around it
    d = _box_sdf(p[~inside], mn32 - h, mx32 + h)
    assert np.max(np.abs(v[~inside] - np.maximum(d, 0.0))) < 1e-6
    shell = d <= 0.0
    assert shell.sum() > 0 and np.all(v[~inside][shell] == 0.0)
    assert np.all(np.abs(v) <= 3 * h * (1 + 1e-6))
