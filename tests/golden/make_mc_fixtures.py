"""Writes mc_topology.npz: corner values of single cubes, the triangles the
reference's MarchingCubesMesher emits for each of them (crossing-edge
triples, 12 = centre vertex, in the emitted vertex order) and their topology
(src/voxel/meshing/marching_cubes.rs:72-243 with the tables of
src/voxel/meshing/lookup_table.rs, read here as data).  Run in a container
that holds the reference at /root/reference; the output is a fixture of
inputs and expected outputs only.

The topology of a cube's triangle list is recorded without the
triangulation: per connected component (triangles sharing a vertex) the
crossing edges it uses, its Euler characteristic, whether it holds the
centre vertex and its triangle count, plus the directed boundary edges of
the whole list (its orientation)."""
import os
import ast
import re
import sys

import numpy as np

TABLES = "/root/reference/src/voxel/meshing/lookup_table.rs"
f32 = np.float32
EPS = f32(1.1920929e-07)
FACES = {1: (0, 4, 5, 1), 2: (1, 5, 6, 2), 3: (2, 6, 7, 3), 4: (3, 7, 4, 0), 5: (0, 3, 2, 1), 6: (4, 7, 6, 5)}
VER = [((0, 1), (4, 5), (7, 6), (3, 2)), ((3, 2), (0, 1), (4, 5), (7, 6)), ((2, 3), (6, 7), (5, 4), (1, 0)),
       ((1, 0), (2, 3), (6, 7), (5, 4)), ((2, 1), (3, 0), (7, 4), (6, 5)), ((3, 0), (2, 1), (6, 5), (7, 4)),
       ((0, 3), (4, 7), (5, 6), (1, 2)), ((1, 2), (0, 3), (4, 7), (5, 6)), ((4, 0), (7, 3), (6, 2), (5, 1)),
       ((5, 1), (4, 0), (7, 3), (6, 2)), ((6, 2), (5, 1), (4, 0), (7, 3)), ((7, 3), (6, 2), (5, 1), (4, 0))]
AMB = {1: [(1, 7, 4), (0, 6, 5), (3, 5, 6), (2, 4, 7)], 2: [(1, 7, 0), (2, 4, 1), (3, 5, 2), (0, 6, 3)],
       5: [(0, 6, 8), (1, 7, 9), (2, 4, 10), (3, 5, 11)]}
FAM = {1: 1, 3: 1, 2: 2, 4: 2, 5: 5, 6: 5, 0: 5}


def load_tables():
    s = open(TABLES).read()
    s = re.sub(r"/\*.*?\*/", "", s, flags=re.S)
    s = re.sub(r"//[^\n]*", "", s)
    out = {}
    for m in re.finditer(r"pub const (\w+)\s*:\s*[^=]+=\s*(\[.*?\]);", s, flags=re.S):
        if "convert" not in m.group(2):
            out[m.group(1)] = ast.literal_eval(m.group(2))
    return out


class Cube:
    """handle_cube's tests on one cube's values (f32 arithmetic)."""

    def __init__(self, c, T):
        self.c = [f32(x) for x in c]
        self.T = T

    def test_face(self, face):
        a, b, cc, d = (self.c[i] for i in FACES[abs(face)])
        val = f32(f32(a * cc) - f32(b * d))
        if abs(val) < EPS:
            return face >= 0
        return f32(f32(f32(face) * a) * val) >= 0

    def amb(self, amb_face, face):
        s, e = f32(face), 0
        for i, j, ed in AMB[FAM[amb_face]]:
            if self.c[i] * s > 0 and self.c[j] * s > 0:
                e = ed
        return e

    def ver(self, edge):
        c = self.c
        (pa, qa), (pb, qb), (pc, qc), (pd, qd) = VER[edge]
        dA, dB, dC, dD = f32(c[pa] - c[qa]), f32(c[pb] - c[qb]), f32(c[pc] - c[qc]), f32(c[pd] - c[qd])
        a = f32(f32(dA * dC) - f32(dB * dD))
        b = f32(f32(f32(f32(c[qc] * dA) + f32(c[qa] * dC)) - f32(c[qd] * dB)) - f32(c[qb] * dD))
        if a > 0:
            return 1
        with np.errstate(all="ignore"):
            t = f32(-b / f32(f32(2.0) * a))
        if t < 0 or t > 1:
            return 1
        at, bt = f32(c[qa] + f32(dA * t)), f32(c[qb] + f32(dB * t))
        ct, dt = f32(c[qc] + f32(dC * t)), f32(c[qd] + f32(dD * t))
        v = f32(f32(at * ct) - f32(bt * dt))
        return 1 if v < 0 else 0

    def interior(self, case, cfg, face):
        T = self.T
        if case in (4, 7):
            s = face if case == 4 else -face
            return (self.ver(self.amb(1, s)) + self.ver(self.amb(2, s)) + self.ver(self.amb(5, s))) != 0
        if case == 6:
            return self.ver(self.amb(abs(T["TEST_6"][cfg][0]), face)) != 0
        if case == 10:
            return self.ver(self.amb(abs(T["TEST_10"][cfg][0]), face)) != 0
        return (self.ver(self.amb(abs(T["TEST_12"][cfg][0]), face)) +
                self.ver(self.amb(abs(T["TEST_12"][cfg][1]), face))) != 0

    def case13(self):
        c = self.c
        a = f32(f32(f32(c[0] - c[1]) * f32(c[7] - c[6])) - f32(f32(c[4] - c[5]) * f32(c[3] - c[2])))
        b = f32(f32(f32(f32(c[6] * f32(c[0] - c[1])) + f32(c[1] * f32(c[7] - c[6]))) - f32(c[2] * f32(c[4] - c[5])))
                - f32(c[5] * f32(c[3] - c[2])))
        cc = f32(f32(c[1] * c[6]) - f32(c[5] * c[2]))
        with np.errstate(all="ignore"):
            delta = f32(f32(b * b) - f32(f32(f32(4.0) * a) * cc))
            sq = f32(np.sqrt(delta))
            t1, t2 = f32(f32(-b + sq) / f32(a + a)), f32(f32(-b - sq) / f32(a + a))
            if 0 < t1 < 1 and 0 < t2 < 1:
                def xy(t):
                    at, bt = f32(c[1] + f32(f32(c[0] - c[1]) * t)), f32(c[5] + f32(f32(c[4] - c[5]) * t))
                    ct, dt = f32(c[6] + f32(f32(c[7] - c[6]) * t)), f32(c[2] + f32(f32(c[3] - c[2]) * t))
                    den = f32(f32(f32(at + ct) - bt) - dt)
                    return f32(f32(at - dt) / den), f32(f32(at - bt) / den)
                (x1, y1), (x2, y2) = xy(t1), xy(t2)
                return not (0 < x1 < 1 and 0 < x2 < 1 and 0 < y1 < 1 and 0 < y2 < 1)
        return True


def reference_tiling(values, T):
    """The triangles (edge numbers, 12 = centre) handle_cube emits, in the
    emitted vertex order (add_faces reads entries i, i+2, i+1)."""
    R = Cube(values, T)
    cid = sum(1 << i for i in range(8) if R.c[i] < 0)
    case, cfg = T["CASES"][cid]
    tf = R.test_face
    ti = lambda f: R.interior(case, cfg, f)  # noqa: E731
    tl, sub = [], str(case)
    if case in (1, 2, 5, 8, 9, 11, 14):
        tl = T["TILING_%d" % case][cfg]
    elif case == 3:
        tl, sub = (T["TILING_3_2"][cfg], "3.2") if tf(T["TEST_3"][cfg]) else (T["TILING_3_1"][cfg], "3.1")
    elif case == 4:
        tl, sub = (T["TILING_4_1"][cfg], "4.1") if ti(T["TEST_4"][cfg]) else (T["TILING_4_2"][cfg], "4.2")
    elif case == 6:
        if tf(T["TEST_6"][cfg][0]):
            tl, sub = T["TILING_6_2"][cfg], "6.2"
        elif ti(T["TEST_6"][cfg][1]):
            tl, sub = T["TILING_6_1_1"][cfg], "6.1.1"
        else:
            tl, sub = T["TILING_6_1_2"][cfg], "6.1.2"
    elif case == 7:
        t7 = T["TEST_7"][cfg]
        s = int(tf(t7[0])) + 2 * int(tf(t7[1])) + 4 * int(tf(t7[2]))
        if s == 0:
            tl, sub = T["TILING_7_1"][cfg], "7.1"
        elif s in (1, 2, 4):
            tl, sub = T["TILING_7_2"][cfg][{1: 0, 2: 1, 4: 2}[s]], "7.2"
        elif s in (3, 5, 6):
            tl, sub = T["TILING_7_3"][cfg][{3: 0, 5: 1, 6: 2}[s]], "7.3"
        elif ti(t7[3]):
            tl, sub = T["TILING_7_4_1"][cfg], "7.4.1"
        else:
            tl, sub = T["TILING_7_4_2"][cfg], "7.4.2"
    elif case in (10, 12):
        tt, n, k = T["TEST_%d" % case][cfg], (5 if case == 10 else 23), str(case)
        if tf(tt[0]):
            if tf(tt[1]):
                tl, sub = ((T["TILING_%s_1_1_" % k][cfg], k + ".1.1_") if ti(-tt[2])
                           else (T["TILING_%s_1_2" % k][n - cfg], k + ".1.2_"))
            else:
                tl, sub = T["TILING_%s_2" % k][cfg], k + ".2"
        elif tf(tt[1]):
            tl, sub = T["TILING_%s_2_" % k][cfg], k + ".2_"
        elif ti(tt[2]):
            tl, sub = T["TILING_%s_1_1" % k][cfg], k + ".1.1"
        else:
            tl, sub = T["TILING_%s_1_2" % k][cfg], k + ".1.2"
    elif case == 13:
        t13 = T["TEST_13"][cfg]
        o = T["SUB_CONFIG_13"][sum(1 << k for k in range(6) if tf(t13[k]))]
        if o == 0:
            tl, sub = T["TILING_13_1"][cfg], "13.1"
        elif 1 <= o <= 6:
            tl, sub = T["TILING_13_2"][cfg][o - 1], "13.2"
        elif 7 <= o <= 18:
            tl, sub = T["TILING_13_3"][cfg][o - 7], "13.3"
        elif 19 <= o <= 22:
            tl, sub = T["TILING_13_4"][cfg][o - 19], "13.4"
        elif 23 <= o <= 26:
            if R.case13():
                tl, sub = T["TILING_13_5_1"][cfg][o - 23], "13.5.1"
            else:
                tl, sub = T["TILING_13_5_2"][1 - cfg][o - 23], "13.5.2"
        elif 27 <= o <= 38:
            tl, sub = T["TILING_13_3_"][cfg][o - 27], "13.3_"
        elif 39 <= o <= 44:
            tl, sub = T["TILING_13_2_"][cfg][o - 39], "13.2_"
        elif o == 45:
            tl, sub = T["TILING_13_1_"][cfg], "13.1_"
    return [(tl[i], tl[i + 2], tl[i + 1]) for i in range(0, len(tl), 3)], sub


def signature(tris):
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


def main(out=os.path.join(os.path.dirname(os.path.abspath(__file__)), "mc_topology.npz")):
    T = load_tables()
    rng = np.random.default_rng(20261202)
    vals, sigs, subs, tri_lists, seen = [], [], [], [], {}
    special = [90, 165, 60, 85, 170, 105, 150, 195, 26, 88, 218, 130, 125]
    k = 0
    while len(vals) < 3000 or min(seen.values()) < 40:
        cid = int(rng.choice(special)) if k % 3 == 0 else int(rng.integers(0, 256))
        mode = k % 5
        k += 1
        if mode < 2:
            m = rng.uniform(0.01, 1, 8)
        elif mode < 4:
            m = np.exp(rng.uniform(-8, 3, 8))
        else:
            m = rng.choice([0.25, 0.5, 1.0, 2.0], 8)
        v = np.array([-x if (cid >> i) & 1 else x for i, x in enumerate(m)], np.float32)
        tris, sub = reference_tiling(v, T)
        if seen.get(sub, 0) >= 150:
            continue
        seen[sub] = seen.get(sub, 0) + 1
        vals.append(v)
        sigs.append(signature(tris))
        subs.append(sub)
        tri_lists.append(tris)
        if k > 2000000:
            sys.exit("a subcase was not reached")
    tri = np.full((len(vals), 12, 3), -1, np.int8)
    for i, ts in enumerate(tri_lists):
        if ts:
            tri[i, :len(ts)] = ts
    np.savez_compressed(out, values=np.array(vals, np.float32), signature=np.array(sigs), subcase=np.array(subs),
                        triangles=tri, n_triangles=np.array([len(t) for t in tri_lists], np.int8))
    print(len(vals), "cubes,", len(seen), "subcases:", dict(sorted(seen.items())))


if __name__ == "__main__":
    main()
