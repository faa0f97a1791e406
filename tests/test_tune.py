"""The batch autotuner's selection and JSON cache (host logic, no GPU)."""
import json

from baby_shark_amd import tune


def test_pick_fastest_takes_the_minimum_and_breaks_ties_low():
    t = {1: 3.0, 2: 1.0, 4: 1.0, 8: 2.0}
    r = tune.pick_fastest(lambda w: t[w], candidates=(1, 2, 4, 8), reps=2)
    assert r["workers"] == 2
    assert r["ms"] == {"1": 3000.0, "2": 1000.0, "4": 1000.0, "8": 2000.0}


def test_batch_key_rounds_the_shape():
    a = tune.batch_key("MI355X", [1000, 3000, 190000], 3, False)
    b = tune.batch_key("MI355X", [1100, 2900, 200000], 3, False)
    assert a == b
    assert a != tune.batch_key("MI355X", [1000, 3000, 190000], 3, True)
    assert a != tune.batch_key("MI355X", [1000] * 300, 3, False)


def test_cache_round_trip(tmp_path):
    path = str(tmp_path / "sub" / "tune.json")
    assert tune.load_cache(path) == {}
    tune.store("k1", {"workers": 4}, path)
    tune.store("k2", {"workers": 8}, path)
    assert tune.load_cache(path) == {"k1": {"workers": 4}, "k2": {"workers": 8}}
    with open(path, "w") as f:
        f.write("not json")
    assert tune.load_cache(path) == {}
    json.dumps(tune.load_cache(path))
