"""Pins the oracle's Volume::offset (fast sweeping) against closed forms and
invariants.  CPU only."""
import itertools

import numpy as np
import pytest

H = 0.05


def test_compute_distance_closed_forms(oracle_lib):
    f = oracle_lib.fs_compute_distance
    h = np.float32(0.1)
    # one term: a1 + h when the next smallest is at least that far
    assert f(0.3, 0.5, 0.7, h) == np.float32(np.float32(0.3) + h)
    assert f(0.3, 3.4e38, 3.4e38, h) == np.float32(np.float32(0.3) + h)
    # two equal terms: a + h / sqrt(2); three equal terms: a + h / sqrt(3)
    a = 0.25
    assert abs(f(a, a, 3.4e38, h) - (a + h / np.sqrt(2))) < 1e-6
    assert abs(f(a, a, a, h) - (a + h / np.sqrt(3))) < 1e-6
    # sort3: the argument order does not matter
    vals = [(0.31, 0.33, 0.36), (0.2, 0.27, 0.9), (1.0, 1.02, 1.05)]
    for v in vals:
        outs = {f(*p, h) for p in itertools.permutations(v)}
        assert len(outs) == 1
        # the Godunov solution satisfies sum((d - a_i)^+)^2 = h^2 over the terms it uses
        d = outs.pop()
        s = sum(max(d - x, 0.0) ** 2 for x in v)
        assert abs(s - h * h) < 1e-6


def _plane(x0_vox=0.3, n=48, yz=40):
    """f = (x - x0) h on a slab of voxels x in [-n/2, n/2), y, z in [0, yz)."""
    xs = np.arange(-n // 2, n // 2)
    g = np.stack(np.meshgrid(xs, np.arange(yz), np.arange(yz), indexing="ij"), -1).reshape(-1, 3)
    v = ((g[:, 0].astype(np.float64) - x0_vox) * H).astype(np.float32)
    return g, v


def _dense(e, lo, shape):
    val = np.full(shape, np.nan, np.float32)
    act = e.active()
    for li in range(e.origins.shape[0]):
        o = e.origins[li]
        for off in np.nonzero(act[li])[0]:
            p = o + np.array([off >> 6, (off >> 3) & 7, off & 7]) - lo
            if np.all(p >= 0) and np.all(p < shape):
                val[tuple(p)] = e.values[li, off]
    return val


@pytest.mark.parametrize("D", [0.12, -0.12, 0.0])
def test_offset_of_plane_is_shifted_plane(oracle_lib, D):
    g, v = _plane()
    vol = oracle_lib.from_voxels(g, v, H)
    before = vol.clone().export()
    out = vol.offset(D).export()
    # frozen voxels (|v| <= 2h before) are only shifted by -D, exactly
    act_b = before.active()
    keep = act_b & (np.abs(before.values) <= np.float32(2 * H))
    ob = {tuple(o): i for i, o in enumerate(out.origins)}
    for li in range(before.origins.shape[0]):
        m = keep[li]
        if not m.any():
            assert tuple(before.origins[li]) not in ob or out.active()[ob[tuple(before.origins[li])]].any()
            continue
        lo = ob[tuple(before.origins[li])]
        assert np.array_equal(out.values[lo][m], (before.values[li][m] - np.float32(D)).astype(np.float32))
    # away from the slab's sides the extension is the 1-D distance, shifted
    lo = np.array([-24, 0, 0])
    dense = _dense(out, lo, (48, 40, 40))
    core = dense[:, 12:28, 12:28]
    xs = np.arange(-24, 24)
    want = ((xs - 0.3) * H - D)[:, None, None]
    have = ~np.isnan(core)
    err = np.abs(np.where(have, core - want, 0.0))
    assert err.max() < 1e-5
    # the extension reaches |D| + 2h on the sweep side and stops there
    lim = abs(D) + 2 * H
    f = (xs - 0.3) * H
    side = f > 0 if D >= 0 else f < 0
    reach = side & (np.abs(f) <= lim - 1e-4) | (np.abs(f) <= 2 * H)
    assert np.array_equal(have.all(axis=(1, 2)), reach)
    assert not have[~reach].any()


def test_offset_of_sphere(oracle_lib):
    # a sphere offset by D is close to the sphere of radius r + D near its new surface
    r, D = 0.6, 0.1
    out = oracle_lib.sphere(r, (0.0, 0.0, 0.0), H).offset(D).export()
    act = out.active()
    li, off = np.nonzero(act)
    p = (out.origins[li] + np.stack([off >> 6, (off >> 3) & 7, off & 7], 1)) * H
    f = np.linalg.norm(p, axis=1) - (r + D)
    vals = out.values[li, off]
    near = np.abs(f) < 1.5 * H
    assert near.sum() > 1000
    assert np.max(np.abs(vals[near] - f[near])) < 0.35 * H
    assert np.all(np.signbit(vals[np.abs(f) > 0.5 * H]) == np.signbit(f[np.abs(f) > 0.5 * H]))
