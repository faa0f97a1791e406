"""Pins the CPU reference of the tile classifier (tests/mlp_reference.py):
BF16 rounding known answers, and its analytic gradient against central
finite differences in float64.  CPU only."""
import numpy as np

import mlp_reference as M


def _params(seed):
    rng = np.random.default_rng(seed)
    p = np.zeros(M.N_PARAMS)
    d = M.unpack(p)
    for name, fan in (("W1", 3), ("W2", 64), ("W3", 64), ("W4", 64)):
        d[name][...] = rng.uniform(-1, 1, d[name].shape) * np.sqrt(6.0 / fan)
    for name in ("b1", "b2", "b3"):
        d[name][...] = rng.normal(0, 0.1, 64)
    q = M.pack(d)
    q[-1] = 0.3
    return q


def test_bf16_rounding_known_answers():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -9, 1.0 + 2 ** -8 + 2 ** -20, -2.5, 0.0,
                  3.0e38], np.float32)
    want = np.array([1.0, 1.0, 1.0 + 2 ** -6, 1.0, 1.0 + 2 ** -7, -2.5, 0.0], np.float32)
    got = M.bf16(x)
    assert np.array_equal(got[:7], want)              # ties to even, exact values kept
    assert np.isfinite(got[7]) and abs(got[7] / 3.0e38 - 1) < 2 ** -8
    assert np.all((M.bf16(np.random.default_rng(0).normal(size=1000).astype(np.float32)).view(np.uint32)
                   & 0xFFFF) == 0)


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(4)
    p = _params(1)
    x = rng.uniform(-1, 1, (64, 3))
    y = rng.uniform(0, 3, 64)
    g, loss = M.gradient(p, x, y, rounded=False)
    idx = np.concatenate([rng.choice(192, 20, replace=False), 192 + rng.choice(64, 8, replace=False),
                          256 + rng.choice(4096, 40, replace=False), 4352 + rng.choice(64, 8, replace=False),
                          4416 + rng.choice(4096, 40, replace=False), 8512 + rng.choice(64, 8, replace=False),
                          8576 + rng.choice(64, 16, replace=False), [8640]])
    eps = 1e-6
    for i in idx:
        a, b = p.copy(), p.copy()
        a[i] += eps
        b[i] -= eps
        la = np.mean((M.forward(a, x, rounded=False)[0] - y) ** 2)
        lb = np.mean((M.forward(b, x, rounded=False)[0] - y) ** 2)
        fd = (la - lb) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-6 + 1e-5 * abs(fd), (i, fd, g[i])
    assert abs(loss - np.mean((M.forward(p, x, rounded=False)[0] - y) ** 2)) < 1e-12


def test_rounded_forward_stays_close_to_float64():
    rng = np.random.default_rng(5)
    p = _params(2)
    x = rng.uniform(-1, 1, (256, 3))
    a = M.forward(p, x, rounded=False)[0]
    b = M.forward(p, x, rounded=True)[0]
    assert np.max(np.abs(a - b)) < 0.05 * (1 + np.max(np.abs(a)))   # BF16 activations: ~3 significant digits
    y = rng.uniform(0, 3, 256)
    g64, _ = M.gradient(p, x, y, rounded=False)
    g16, _ = M.gradient(p, x, y, rounded=True)
    assert np.corrcoef(g64, g16)[0, 1] > 0.9
