"""Binary STL reading and writing (io/stl.rs) through the C ABI, on the CPU:
the reference's known answer for assets/box.stl (io/mod.rs:247-270: 12 faces,
36 soup vertices), the bunny's 13,000 faces, an independent numpy parse of the
same bytes, write/read round trips and the error paths."""
import os

import numpy as np
import pytest

import baby_shark_amd as bs
from baby_shark_amd import _lib

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
FACE = np.dtype([("n", "<f4", 3), ("v", "<f4", (3, 3)), ("a", "<u2")])


def parse(path):
    data = open(path, "rb").read()
    n = int(np.frombuffer(data[80:84], "<u4")[0])
    return n, np.frombuffer(data[84:84 + 50 * n], FACE)


def test_box_known_answer_and_numpy_parse():
    soup = bs.read_stl(os.path.join(GOLDEN, "box.stl"))
    assert soup.shape == (12, 3, 3) and soup.reshape(-1, 3).shape[0] == 36
    n, f = parse(os.path.join(GOLDEN, "box.stl"))
    assert n == 12 and np.array_equal(soup.view(np.uint32), f["v"].view(np.uint32))


def test_bunny_face_count():
    soup = bs.read_stl(os.path.join(GOLDEN, "bunny.stl"))
    n, f = parse(os.path.join(GOLDEN, "bunny.stl"))
    assert soup.shape == (13000, 3, 3) and n == 13000
    assert np.array_equal(soup.view(np.uint32), f["v"].view(np.uint32))


def test_write_read_round_trip(tmp_path):
    rng = np.random.default_rng(4)
    soup = rng.normal(size=(1000, 3, 3)).astype(np.float32)
    soup[7] = soup[7, 0]              # a point: degenerate
    soup[8, 2] = soup[8, 1]           # two equal vertices: degenerate
    p = tmp_path / "r.stl"
    bs.write_stl(p, soup)
    data = open(p, "rb").read()
    assert len(data) == 84 + 50 * 1000 and data[:80] == bytes(80)
    n, f = parse(p)
    assert n == 1000 and np.all(f["a"] == 0)
    assert np.array_equal(bs.read_stl(p).view(np.uint32), soup.view(np.uint32))
    # unit normals of (b - a) x (c - a), zeros for degenerate faces
    s = soup.astype(np.float64)
    c = np.cross(s[:, 1] - s[:, 0], s[:, 2] - s[:, 0])
    ln = np.linalg.norm(c, axis=1)
    ok = ln > 0
    assert np.all(f["n"][~ok] == 0) and set(np.nonzero(~ok)[0]) == {7, 8}
    assert np.max(np.abs(f["n"][ok] - c[ok] / ln[ok, None])) < 1e-5
    # empty soups
    bs.write_stl(tmp_path / "e.stl", np.zeros((0, 3, 3), np.float32))
    assert bs.read_stl(tmp_path / "e.stl").shape == (0, 3, 3)


def test_io_errors(tmp_path):
    with pytest.raises(bs.BSError) as e:
        bs.read_stl(tmp_path / "missing.stl")
    assert e.value.status == _lib.BS_ERR_IO
    data = open(os.path.join(GOLDEN, "box.stl"), "rb").read()
    for cut in (50, 84 + 50 * 5 + 20):
        q = tmp_path / f"short{cut}.stl"
        q.write_bytes(data[:cut])
        with pytest.raises(bs.BSError) as e:
            bs.read_stl(q)
        assert e.value.status == _lib.BS_ERR_IO
    with pytest.raises(bs.BSError) as e:
        bs.write_stl(tmp_path / "no_dir" / "x.stl", np.zeros((1, 3, 3), np.float32))
    assert e.value.status == _lib.BS_ERR_IO
