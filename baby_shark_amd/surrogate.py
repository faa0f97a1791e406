"""Training of the tile classifier used by the surrogate sign mode.

The classifier (TileClassifier: 3 -> 64 -> 64 -> 64 -> 1, BF16 weights, FP32
accumulation, hand-written HIP kernels) predicts the smallest |SDF| of a 4^3
tile, in voxels, from the tile centre.  Its samples come from the exact path:
each training mesh is converted with per-voxel signs and every tile that
holds an active voxel gives (centre, smallest |value| / voxel size).  No
counterpart in the reference: a tile it culls only changes how many winding
numbers are evaluated, never a sign (the exact distances check every culled
voxel, see BS_SIGN_SURROGATE in include/bs_voxel.h)."""
from __future__ import annotations

import time
from typing import Optional

import numpy as np

from .voxel import Context, MeshToVolume, TileClassifier, default_context, tile_samples


def collect_samples(ctx: Context, meshes, voxel_size: float, band: int, total: int, seed: int = 0):
    """(x, y) device tensors: `total` tile samples drawn evenly from the meshes."""
    import torch
    gen = torch.Generator(device="cuda")
    gen.manual_seed(seed)
    m2v = MeshToVolume(ctx).with_voxel_size(voxel_size).with_narrow_band_width(band)
    per = -(-total // len(meshes))
    xs, ys = [], []
    for soup in meshes:
        vol = m2v.convert(soup)
        x, y = tile_samples(vol)
        vol.free()
        pick = torch.randperm(x.shape[0], generator=gen, device="cuda")[:per]
        xs.append(x[pick])
        ys.append(y[pick])
    return torch.cat(xs)[:total].contiguous(), torch.cat(ys)[:total].contiguous()


def train_tile_classifier(ctx: Optional[Context], meshes, voxel_size: float, band: int = 3,
                          samples: int = 2_000_000, epochs: int = 20, batch: int = 8192, lr: float = 3e-3,
                          lr_final: float = 1e-4, seed: int = 0):
    """A classifier trained on the meshes' tile samples; returns it with a
    report (sample count, seconds, loss per epoch, over-prediction quantiles
    on the training samples)."""
    import torch
    ctx = ctx or default_context()
    t0 = time.perf_counter()
    x, y = collect_samples(ctx, meshes, voxel_size, band, samples, seed)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    lo, hi = x.min(0).values, x.max(0).values
    centre = ((lo + hi) * 0.5).cpu().numpy()
    scale = float(2.0 / (hi - lo).max().item())
    net = TileClassifier(ctx, seed=seed)
    p = net.params
    p[-1] = float(y.mean().item())  # output bias starts at the mean target
    net.params = p
    net.set_input_map(centre, scale)
    losses = net.train(x, y, epochs=epochs, batch=batch, lr=lr, lr_final=lr_final, seed=seed)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    over = (net.forward(x) - y).cpu().numpy()
    report = {"samples": int(x.shape[0]), "meshes": len(meshes), "collect_s": round(t1 - t0, 3),
              "train_s": round(t2 - t1, 3), "epochs": epochs, "batch": batch,
              "loss_per_epoch": [round(float(v), 5) for v in losses],
              "rmse_vox": round(float(np.sqrt(np.mean(over ** 2))), 4),
              "over_prediction_q999_vox": round(float(np.quantile(over, 0.999)), 4),
              "target_mean_vox": round(float(y.mean().item()), 4)}
    return net, report


def compare_sign_modes(ctx: Context, meshes, voxel_size: float, band: int, net, threshold: float,
                       exact_threshold: float = 2.0, reps: int = 3):
    """Per mesh: the step time (best of reps), winding evaluations and tile
    gate counts of the per-voxel, shared-sign, surrogate (network gate at
    `threshold`) and exact-gate (each tile's true smallest |SDF| against
    `exact_threshold`) modes, and for each of the last three the number of
    values and of signs that differ from the per-voxel volume."""
    import torch
    modes = (("per_voxel", None), ("shared", None), ("surrogate", (net, threshold)),
             ("exact_gate", (None, exact_threshold)))
    rows = []
    try:
        for soup in meshes:
            dev = soup if getattr(soup, "is_cuda", False) else torch.from_numpy(soup).cuda()
            row, ref = {}, None
            for name, gate in modes:
                m2v = MeshToVolume(ctx).with_voxel_size(voxel_size).with_narrow_band_width(band)
                if name == "shared":
                    m2v.with_shared_signs()
                elif gate is not None:
                    m2v.with_surrogate_signs()
                    ctx.set_surrogate(*gate)
                best = None
                for _ in range(reps):
                    torch.cuda.synchronize()
                    t0 = time.perf_counter()
                    vol = m2v.convert(dev)
                    dt = (time.perf_counter() - t0) * 1e3
                    best = dt if best is None else min(best, dt)
                    st = ctx.last_stats()
                    _, masks, values = vol.device_arrays()
                    vol.free()
                r = {"ms": round(best, 3), "evals": int(st["n_sign_evals"]), "active": int(st["n_active"])}
                if gate is not None:
                    r.update({"tiles": int(st["n_sg_tiles"]), "culled": int(st["n_sg_culled"]),
                              "misclassified": int(st["n_sg_misclassified"]),
                              "fallback_voxels": int(st["n_sg_fallback"])})
                if ref is None:
                    ref = (masks, values)
                else:
                    r["masks_equal"] = bool(torch.equal(masks, ref[0]))
                    r["values_differing"] = int((values.view(torch.int32) != ref[1].view(torch.int32)).sum().item())
                    r["sign_mismatches"] = int((torch.signbit(values) != torch.signbit(ref[1])).sum().item())
                row[name] = r
            rows.append(row)
    finally:
        ctx.set_surrogate(None, 0.0)
    return rows
