"""Seeded synthetic closed meshes (f32 triangle soups, outward oriented).

Every mesh is built in float64 and rounded once to float32.  Randomness comes
from splitmix64 so a (kind, seed) pair names the same soup on every machine.

* ``torus``              -- R, r, (major x minor) segments, radial noise along
                            the tube normal, uniform in [-amp, +amp].
* ``displaced_icosphere`` -- icosahedron subdivided ``level`` times (20*4^level
                            faces) on the unit sphere, vertices moved radially
                            to ``radius * (1 + amp * fbm(p))`` with fbm a
                            3-octave hashed value noise (base frequency 4,
                            lacunarity 2, gain 0.5), normalised to [-1, 1].

A soup is a ``(n, 3, 3)`` float32 array: triangle, vertex, xyz.
"""
from __future__ import annotations

import functools

import numpy as np

_MASK = np.uint64(0xFFFFFFFFFFFFFFFF)


def splitmix64(x: np.ndarray) -> np.ndarray:
    """One splitmix64 output for each uint64 state in ``x`` (vectorised)."""
    with np.errstate(over="ignore"):
        z = (x.astype(np.uint64) + np.uint64(0x9E3779B97F4A7C15)) & _MASK
        z = ((z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)) & _MASK
        z = ((z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)) & _MASK
        return z ^ (z >> np.uint64(31))


def uniform01(seed: int, count: int, stream: int = 0) -> np.ndarray:
    """``count`` doubles in [0, 1) from the splitmix64 sequence of ``seed``."""
    with np.errstate(over="ignore"):
        base = np.uint64((seed * 0x9E3779B97F4A7C15 + stream * 0xD1B54A32D192ED03) & 0xFFFFFFFFFFFFFFFF)
        idx = np.arange(count, dtype=np.uint64) * np.uint64(0x9E3779B97F4A7C15)
        bits = splitmix64(base + idx)
    return (bits >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def _orient_outward(soup: np.ndarray) -> np.ndarray:
    a, b, c = soup[:, 0], soup[:, 1], soup[:, 2]
    vol = np.einsum("ij,ij->i", a, np.cross(b, c)).sum() / 6.0
    if vol < 0:
        soup = soup[:, [0, 2, 1]]
    return soup


def torus(R: float = 1.0, r: float = 0.35, major: int = 128, minor: int = 64, amp: float = 0.01,
          seed: int = 1, center=(0.0, 0.0, 0.0)) -> np.ndarray:
    """Closed torus about the z axis, 2*major*minor triangles."""
    i = np.arange(major)
    j = np.arange(minor)
    u = 2.0 * np.pi * i / major
    v = 2.0 * np.pi * j / minor
    uu, vv = np.meshgrid(u, v, indexing="ij")
    n = np.stack([np.cos(vv) * np.cos(uu), np.cos(vv) * np.sin(uu), np.sin(vv)], axis=-1)
    c = np.stack([R * np.cos(uu), R * np.sin(uu), np.zeros_like(uu)], axis=-1)
    delta = (uniform01(seed, major * minor).reshape(major, minor) * 2.0 - 1.0) * amp
    verts = c + (r + delta)[..., None] * n + np.asarray(center, dtype=np.float64)
    verts = verts.reshape(-1, 3)
    ii, jj = np.meshgrid(i, j, indexing="ij")
    i1 = (ii + 1) % major
    j1 = (jj + 1) % minor
    v00 = (ii * minor + jj).ravel()
    v10 = (i1 * minor + jj).ravel()
    v11 = (i1 * minor + j1).ravel()
    v01 = (ii * minor + j1).ravel()
    faces = np.concatenate([np.stack([v00, v10, v11], 1), np.stack([v00, v11, v01], 1)], 0)
    soup = _orient_outward(verts[faces])
    return soup.astype(np.float32)


def _icosahedron():
    t = (1.0 + 5.0 ** 0.5) / 2.0
    v = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t],
                  [0, -1, -t], [0, 1, -t], [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], dtype=np.float64)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    f = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9], [5, 11, 4],
                  [11, 10, 2], [10, 7, 6], [7, 1, 8], [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8],
                  [3, 8, 9], [4, 9, 5], [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]], dtype=np.int64)
    return v, f


@functools.lru_cache(maxsize=2)
def icosphere(level: int):
    """Unit icosphere: (vertices float64 (V,3), faces int64 (F,3)), 20*4^level faces.
    Cached: callers must not modify the arrays."""
    v, f = _icosahedron()
    for _ in range(level):
        e = np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]], 0)
        e.sort(axis=1)
        uniq, inv = np.unique(e, axis=0, return_inverse=True)
        inv = inv.ravel()
        mid = v[uniq[:, 0]] + v[uniq[:, 1]]
        mid /= np.linalg.norm(mid, axis=1, keepdims=True)
        base = v.shape[0]
        v = np.concatenate([v, mid], 0)
        nf = f.shape[0]
        ab = base + inv[:nf]
        bc = base + inv[nf:2 * nf]
        ca = base + inv[2 * nf:]
        a, b, c = f[:, 0], f[:, 1], f[:, 2]
        f = np.concatenate([np.stack([a, ab, ca], 1), np.stack([b, bc, ab], 1),
                            np.stack([c, ca, bc], 1), np.stack([ab, bc, ca], 1)], 0)
    return v, f


def _lattice_value(ix, iy, iz, seed: int, octave: int) -> np.ndarray:
    with np.errstate(over="ignore"):
        h = (ix.astype(np.int64).astype(np.uint64) * np.uint64(0x8DA6B343)) \
            ^ (iy.astype(np.int64).astype(np.uint64) * np.uint64(0xD8163841)) \
            ^ (iz.astype(np.int64).astype(np.uint64) * np.uint64(0xCB1AB31F))
        h = splitmix64(h ^ np.uint64((seed * 0x9E3779B97F4A7C15 + octave * 0x632BE59BD9B4E019) & 0xFFFFFFFFFFFFFFFF))
    return (h >> np.uint64(11)).astype(np.float64) * (2.0 / 9007199254740992.0) - 1.0


def value_noise(p: np.ndarray, seed: int, octave: int) -> np.ndarray:
    """Hashed-lattice value noise in [-1, 1] with smoothstep interpolation."""
    f = np.floor(p)
    t = p - f
    s = t * t * (3.0 - 2.0 * t)
    i0 = f.astype(np.int64)
    out = np.zeros(p.shape[0])
    for dx in (0, 1):
        wx = s[:, 0] if dx else 1.0 - s[:, 0]
        for dy in (0, 1):
            wy = s[:, 1] if dy else 1.0 - s[:, 1]
            for dz in (0, 1):
                wz = s[:, 2] if dz else 1.0 - s[:, 2]
                out += wx * wy * wz * _lattice_value(i0[:, 0] + dx, i0[:, 1] + dy, i0[:, 2] + dz, seed, octave)
    return out


def fbm(p: np.ndarray, seed: int, octaves: int = 3, base_freq: float = 4.0, lacunarity: float = 2.0,
        gain: float = 0.5) -> np.ndarray:
    total = np.zeros(p.shape[0])
    norm = 0.0
    amp = 1.0
    freq = base_freq
    for o in range(octaves):
        total += amp * value_noise(p * freq, seed, o)
        norm += amp
        amp *= gain
        freq *= lacunarity
    return total / norm


def displaced_icosphere(level: int = 8, radius: float = 1.0, amp: float = 0.05, seed: int = 2,
                        center=(0.0, 0.0, 0.0)) -> np.ndarray:
    """Icosphere of 20*4^level faces, radius * (1 + amp * fbm(p)), outward CCW."""
    v, f = icosphere(level)
    r = radius * (1.0 + amp * fbm(v, seed))
    verts = v * r[:, None] + np.asarray(center, dtype=np.float64)
    soup = _orient_outward(verts[f])
    return soup.astype(np.float32)


def sphere_soup(level: int, radius: float = 1.0, center=(0.0, 0.0, 0.0)) -> np.ndarray:
    """Undisplaced icosphere soup."""
    return displaced_icosphere(level, radius, 0.0, 0, center)


def signed_volume(soup: np.ndarray) -> float:
    s = soup.astype(np.float64)
    return float(np.einsum("ij,ij->i", s[:, 0], np.cross(s[:, 1], s[:, 2])).sum() / 6.0)


def batch_meshes(count: int = 256, seed: int = 5, lo: float = 1e3, hi: float = 2e5, span_voxels: int = 256):
    """The uneven batch: `count` closed meshes whose triangle counts are drawn
    log-uniform in [lo, hi], alternating tori (2 * major * minor triangles,
    major = 2 * minor, noise 0.01) and displaced icospheres (the level nearest
    the drawn count, 3..6), each with the voxel size that makes its largest
    bounding-box extent span `span_voxels` voxels.  Returns [(soup, h)]."""
    u = uniform01(seed, 2 * count)
    out = []
    for i in range(count):
        target = lo * (hi / lo) ** u[2 * i]
        s = seed * 1000 + i
        if i % 2 == 0:
            minor = max(3, int(round((target / 4.0) ** 0.5)))
            soup = torus(1.0, 0.35, 2 * minor, minor, 0.01, seed=s)
        else:
            level = int(min(6, max(3, round(np.log(target / 20.0) / np.log(4.0)))))
            soup = displaced_icosphere(level, 1.0, 0.05, seed=s)
        ext = float((soup.reshape(-1, 3).max(0) - soup.reshape(-1, 3).min(0)).max())
        out.append((soup, ext / span_voxels))
    return out


def _displaced_icosphere_args(args):
    return displaced_icosphere(*args)


def displaced_icospheres(level: int, seeds, radius: float = 1.0, amp: float = 0.05, processes: int = 0):
    """displaced_icosphere for each seed, built in worker processes (started
    fresh, not forked, so a caller that holds a GPU context may use it)."""
    import concurrent.futures
    import multiprocessing
    import os
    seeds = list(seeds)
    if not processes:
        processes = max(1, min(len(seeds), 16, os.cpu_count() or 1))
    if processes == 1:
        return [displaced_icosphere(level, radius, amp, s) for s in seeds]
    with concurrent.futures.ProcessPoolExecutor(max_workers=processes,
                                                mp_context=multiprocessing.get_context("spawn")) as ex:
        return list(ex.map(_displaced_icosphere_args, [(level, radius, amp, s) for s in seeds]))
