"""The surrogate sign mode and its tile classifier (no counterpart in the
reference).  The classifier's forward pass and gradient match the CPU
reference (tests/mlp_reference.py, itself pinned by finite differences), a
planted transposition in that reference is caught, training lowers the loss,
and on unseen meshes of the headline family the surrogate-signed volume is
bit-identical to the per-voxel (reference) volume, with the gain over the
per-voxel and shared-sign paths reported."""

import numpy as np
import pytest

import baby_shark_amd as bs
from baby_shark_amd import meshgen as mg
from baby_shark_amd.surrogate import compare_sign_modes, train_tile_classifier

import mlp_reference as M
from test_mlp_reference import _params

pytestmark = pytest.mark.gpu
SHIFT, SCALE = (0.1, -0.2, 0.05), 1.7


def _net(ctx, seed):
    net = bs.TileClassifier(ctx, seed=0)
    net.params = _params(seed).astype(np.float32)
    net.set_input_map(SHIFT, SCALE)
    return net


def test_forward_matches_cpu_reference(gpu_ctx):
    import torch
    net = _net(gpu_ctx, 3)
    x = np.random.default_rng(7).uniform(-1.2, 1.2, (4099, 3)).astype(np.float32)  # not a multiple of 16
    got = net.forward(torch.from_numpy(x).cuda()).cpu().numpy().astype(np.float64)
    want, _ = M.forward(net.params, x, SHIFT, SCALE)
    err = np.abs(got - want)
    print(f"forward: max err {err.max():.3g}, median {np.median(err):.3g}, |out| max {np.abs(want).max():.3g}",
          flush=True)
    assert err.max() <= 5e-3 * (1.0 + np.abs(want).max())
    assert np.median(err) <= 1e-5 * (1.0 + np.abs(want).max())


def _close(g, r, what):
    scale = np.max(np.abs(r))
    bad = np.abs(g - r) > 2e-3 * scale + 1e-3 * np.abs(r)
    assert not bad.any(), f"{what}: {bad.sum()} entries differ, worst {np.max(np.abs(g - r)) / scale:.3g} of max"


def test_gradient_matches_cpu_reference(gpu_ctx):
    import torch
    net = _net(gpu_ctx, 4)
    rng = np.random.default_rng(8)
    x = rng.uniform(-1.2, 1.2, (3001, 3)).astype(np.float32)
    y = rng.uniform(0.0, 4.0, 3001).astype(np.float32)
    g, loss = net.gradient(torch.from_numpy(x).cuda(), torch.from_numpy(y).cuda())
    r, rloss = M.gradient(net.params, x, y, SHIFT, SCALE)
    assert abs(loss - rloss) <= 1e-4 * rloss
    d = M.unpack(g.astype(np.float64))
    e = M.unpack(r)
    for name in ("W1", "b1", "W2", "b2", "W3", "b3", "W4"):
        _close(d[name], e[name], name)
    assert abs(d["b4"] - e["b4"]) <= 1e-3 * (abs(e["b4"]) + 1e-3)
    # a mistake planted in the reference (W2 used transposed) is caught
    q = M.unpack(net.params.astype(np.float32).copy())
    q["W2"] = q["W2"].T.copy()
    bad, _ = M.gradient(M.pack(q).astype(np.float32), x, y, SHIFT, SCALE)
    with pytest.raises(AssertionError):
        _close(d["W3"], M.unpack(bad)["W3"], "W3 vs transposed reference")


def test_training_lowers_the_loss(gpu_ctx):
    import torch
    rng = np.random.default_rng(9)
    x = rng.uniform(-1.0, 1.0, (65536, 3)).astype(np.float32)
    y = (np.abs(np.linalg.norm(x, axis=1) - 0.6) * 8.0).astype(np.float32)
    net = bs.TileClassifier(gpu_ctx, seed=5)
    xd, yd = torch.from_numpy(x).cuda(), torch.from_numpy(y).cuda()
    losses = net.train(xd, yd, epochs=60, batch=4096, lr=3e-3, lr_final=1e-4, seed=1)
    print("training losses:", np.round(losses, 4).tolist(), flush=True)
    assert losses[-1] < 0.05 * losses[0]
    _, final = net.gradient(xd, yd)
    assert final < 0.05 * losses[0]
    # the trained network's outputs still match the reference arithmetic
    got = net.forward(xd[:2048]).cpu().numpy()
    want, _ = M.forward(net.params, x[:2048])
    assert np.max(np.abs(got - want)) <= 5e-3 * (1.0 + np.abs(want).max())


@pytest.mark.timeout(900)
def test_surrogate_signs_on_unseen_meshes(gpu_ctx):
    """The headline family (displaced icospheres of 8 subdivisions, voxel
    0.002, band 3): the classifier is trained on 2M tile samples of 32 seeded
    meshes and tried on 8 unseen seeds.  Every surrogate-signed volume equals
    the per-voxel volume bit for bit (so no sign differs); culled tiles whose
    voxels failed the exact check are counted, and the step time is compared
    with the per-voxel and shared-sign paths."""
    ctx = gpu_ctx
    h, band = 0.002, 3
    train = mg.displaced_icospheres(8, list(range(100, 132)))
    net, rep = train_tile_classifier(ctx, train, h, band, samples=2_000_000, epochs=12, seed=3)
    del train
    print("classifier:", rep, flush=True)
    thr = 2.0 + max(rep["over_prediction_q999_vox"], 0.0)
    tests = mg.displaced_icospheres(8, list(range(200, 208)))
    rows = compare_sign_modes(ctx, tests, h, band, net, thr)
    for row in rows:
        print("unseen mesh:", row, flush=True)
        for name in ("shared", "surrogate", "exact_gate"):
            r = row[name]
            assert r["masks_equal"] and r["values_differing"] == 0 and r["sign_mismatches"] == 0, (name, r)
    for name in ("shared", "surrogate", "exact_gate"):
        gain = np.mean([r["per_voxel"]["ms"] / r[name]["ms"] for r in rows])
        vs_shared = np.mean([r["shared"]["ms"] / r[name]["ms"] for r in rows])
        print(f"{name}: mean gain over per-voxel {gain:.3f}x, over shared {vs_shared:.3f}x", flush=True)
