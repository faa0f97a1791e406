"""Pins the oracle's marching cubes (MarchingCubesMesher,
meshing/marching_cubes.rs).  CPU only.

* the reference's known answer (volume/mod.rs:134-152): box2.stl through
  MeshToVolume::default() at voxel 0.2 (band 0), offset 0.5 and marching
  cubes gives 7944 vertices;
* cube by cube, the topology of the emitted surface (components, their
  crossing edges, Euler characteristics, centre vertex, triangle counts, and
  the directed boundary) equals that of the reference's table-driven tiling
  on 5822 cubes covering all 41 subcases (mc_topology.npz, written by
  golden/make_mc_fixtures.py from the reference's tables);
* closed smooth surfaces come out as closed, consistently oriented 2-manifolds
  of the right Euler characteristic and enclosed volume."""
import os

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
EDGE = [(0, 1), (1, 2), (3, 2), (0, 3), (4, 5), (5, 6), (7, 6), (4, 7), (0, 4), (1, 5), (2, 6), (3, 7)]


def _stl(path):
    b = open(path, "rb").read()
    n = int.from_bytes(b[80:84], "little")
    rec = np.frombuffer(b[84:84 + 50 * n], dtype=np.dtype([("n", "<f4", 3), ("v", "<f4", 9), ("a", "<u2")]))
    return rec["v"].reshape(-1, 9).astype(np.float32)


def test_known_answer_box2_offset_7944(oracle_lib):
    v = oracle_lib.mesh_to_volume(_stl(os.path.join(GOLDEN, "box2.stl")), 0.2, 0)
    v.offset(0.5)
    assert v.marching_cubes(0.2).shape[0] * 3 == 7944


def _signature(tris):
    par = {}

    def find(x):
        while par.setdefault(x, x) != x:
            x = par[x]
        return x
    for t in tris:
        for v in t[1:]:
            a, b = find(t[0]), find(v)
            if a != b:
                par[b] = a
    comps = {}
    for t in tris:
        comps.setdefault(find(t[0]), []).append(t)
    parts = []
    for ts in comps.values():
        V = {v for t in ts for v in t}
        E = {tuple(sorted((t[i], t[(i + 1) % 3]))) for t in ts for i in range(3)}
        parts.append("%s/%d/%d/%d" % (".".join(str(v) for v in sorted(V - {12})), len(V) - len(E) + len(ts),
                                       int(12 in V), len(ts)))
    de = {}
    for t in tris:
        for i in range(3):
            de[(t[i], t[(i + 1) % 3])] = de.get((t[i], t[(i + 1) % 3]), 0) + 1
    bd = sorted((a, b) for (a, b) in de if (b, a) not in de)
    return "|".join(sorted(parts)) + "#" + ",".join("%d>%d" % e for e in bd)


def test_cube_topology_matches_reference_tables(oracle_lib):
    f = np.load(os.path.join(GOLDEN, "mc_topology.npz"))
    vals, sigs, subs = f["values"], f["signature"], f["subcase"]
    assert len(set(subs.tolist())) == 41
    bad = {}
    for v, s, sub in zip(vals, sigs, subs):
        if _signature(oracle_lib.mc_cube_tiling(v)) != str(s):
            bad[str(sub)] = bad.get(str(sub), 0) + 1
    assert not bad, bad


def test_planted_mistake_is_caught(oracle_lib):
    """The fixture check fails on a wrong topology: a case-3 cube whose
    surface is reported as the other subcase (two triangles instead of one
    band, or the reverse) has a different signature."""
    f = np.load(os.path.join(GOLDEN, "mc_topology.npz"))
    i = [k for k, s in enumerate(f["subcase"]) if s == "3.2"][0]
    j = [k for k, s in enumerate(f["subcase"]) if s == "3.1" and
         np.array_equal(f["values"][k] < 0, f["values"][i] < 0)]
    t = oracle_lib.mc_cube_tiling(f["values"][i])
    flipped = [(a, c, b) for a, b, c in t]
    assert _signature(t) == str(f["signature"][i])
    assert _signature(flipped) != str(f["signature"][i])
    if j:
        assert str(f["signature"][j[0]]) != str(f["signature"][i])


def test_every_crossing_is_a_vertex(oracle_lib):
    rng = np.random.default_rng(3)
    for _ in range(3000):
        v = (np.exp(rng.uniform(-5, 2, 8)) * np.where(rng.random(8) < 0.5, -1, 1)).astype(np.float32)
        used = {e for t in oracle_lib.mc_cube_tiling(v) for e in t} - {12}
        cross = {k for k, (a, b) in enumerate(EDGE) if (v[a] < 0) != (v[b] < 0)}
        assert used == cross


def _closed_oriented(tris):
    """Welds vertices by exact position; returns (V, E, F) after checking that
    every directed edge is matched by exactly one reversed edge."""
    pts = tris.reshape(-1, 3)
    _, vid = np.unique(pts.view(np.dtype((np.void, 12))), return_inverse=True)
    vid = vid.reshape(-1, 3)
    d = {}
    for a, b, c in vid:
        for u, w in ((a, b), (b, c), (c, a)):
            d[(u, w)] = d.get((u, w), 0) + 1
    assert all(n == 1 for n in d.values())
    assert all((w, u) in d for (u, w) in d)
    return int(vid.max()) + 1, len(d) // 2, vid.shape[0]


def _signed_volume(tris):
    a, b, c = tris[:, 0].astype(np.float64), tris[:, 1].astype(np.float64), tris[:, 2].astype(np.float64)
    return float(np.sum(np.einsum("ij,ij->i", a, np.cross(b, c))) / 6.0)


@pytest.mark.parametrize("r,c,h", [(0.5, (0.01, 0.02, -0.03), 0.02), (0.31, (0.3, -0.2, 0.1), 0.013)])
def test_sphere_is_closed_oriented_genus0(oracle_lib, r, c, h):
    t = oracle_lib.sphere(r, c, h).marching_cubes(h)
    V, E, F = _closed_oriented(t)
    assert V - E + F == 2
    assert abs(_signed_volume(t) - 4.0 / 3.0 * np.pi * r ** 3) < 0.01 * r ** 3


def test_torus_is_closed_oriented_genus1(oracle_lib):
    h, R, r = 0.02, 0.6, 0.2
    n = int(1.0 / h)
    g = np.stack(np.meshgrid(*(np.arange(-n, n),) * 3, indexing="ij"), -1).reshape(-1, 3)
    p = g * h
    d = np.sqrt((np.sqrt(p[:, 0] ** 2 + p[:, 1] ** 2) - R) ** 2 + p[:, 2] ** 2) - r
    keep = np.abs(d) < 3 * h
    t = oracle_lib.from_voxels(g[keep].astype(np.int32), d[keep].astype(np.float32), h).marching_cubes(h)
    V, E, F = _closed_oriented(t)
    assert V - E + F == 0
    assert abs(_signed_volume(t) - 2 * np.pi ** 2 * R * r ** 2) < 0.02 * (2 * np.pi ** 2 * R * r ** 2)


def test_cube_triangles_match_reference_tables(oracle_lib):
    """Triangle for triangle, in the emitted vertex order: the oracle's tiling
    of every fixture cube equals the one the reference's tables select."""
    f = np.load(os.path.join(GOLDEN, "mc_topology.npz"))
    bad = {}
    for v, t, n, sub in zip(f["values"], f["triangles"], f["n_triangles"], f["subcase"]):
        want = [tuple(int(x) for x in r) for r in t[:n]]
        if oracle_lib.mc_cube_tiling(v) != want:
            bad[str(sub)] = bad.get(str(sub), 0) + 1
    assert not bad, bad


def test_planted_triangle_order_mistakes_are_caught(oracle_lib):
    """The exact check fails when a tiling is emitted with its vertices
    rotated, its triangles reordered or a quad split along the other diagonal."""
    f = np.load(os.path.join(GOLDEN, "mc_topology.npz"))
    i = [k for k, s in enumerate(f["subcase"]) if s == "2"][0]
    want = [tuple(int(x) for x in r) for r in f["triangles"][i][:f["n_triangles"][i]]]
    got = oracle_lib.mc_cube_tiling(f["values"][i])
    assert got == want and len(got) == 2
    assert [(b, c, a) for a, b, c in got] != want
    assert got[::-1] != want
    (a, b, c), (d, e, g) = got
    quad = [a, b, c, d, e, g]
    shared = [x for x in (a, b, c) if x in (d, e, g)]
    other = [x for x in quad if x not in shared]
    assert len(shared) == 2 and len(set(other)) == 2
    assert [(other[0], shared[0], other[1]), (other[1], shared[1], other[0])] != want
