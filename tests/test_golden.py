"""The oracle reproduces the committed golden fixtures (tests/golden/
fixtures.npz, written by tests/golden/make_fixtures.py) exactly.  CPU only."""
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def test_oracle_reproduces_fixtures(oracle_lib):
    import importlib.util
    spec = importlib.util.spec_from_file_location("make_fixtures", os.path.join(GOLDEN, "make_fixtures.py"))
    mf = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mf)
    want = np.load(os.path.join(GOLDEN, "fixtures.npz"))
    got = mf.build()
    assert set(got) == set(want.files)
    for k in want.files:
        a, b = np.asarray(got[k]), want[k]
        assert a.shape == b.shape and a.dtype == b.dtype, k
        assert np.array_equal(a.view(np.uint8), b.view(np.uint8)), k
