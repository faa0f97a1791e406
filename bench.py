"""Narrow-band SDF throughput on one MI355X.

Workload (default): displaced icosphere, 8 subdivisions (1,310,720 f32
triangles, radius 1, radial 3-octave value noise amplitude 0.05, base
frequency 4, seed 2), voxel 0.002, 3-voxel band, sign by winding number.
A step is one MeshToVolume::convert of the device-resident soup into a
device-resident volume; value = active voxels / second.

Prints one JSON line.  `roofline` is for the dominant kernel (FP32 VALU work
from the kernels' own counters / its HIP-event time on the library stream);
`cpu_baseline` times the CPU oracle (the reference algorithm restated in C++,
same parallel decomposition) on a bounded sample of the same mesh.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP32_TFLOPS = 157.3  # MI355X dense FP32 vector, published
PEAK_HBM_GBS = 8000.0


def distance_flop(stats):
    # 82 FLOP per (voxel, sub-triangle) pair of the reference pair set
    return 82.0 * stats["n_pairs"]


def winding_flop(stats):
    # 11 per near internal-node visit, 89 per far (order-2) evaluation,
    # 11 + 77 per triangle for each near leaf visit (atan2 counted as 1)
    leaf_visits = stats.get("wn_leaf_visits", 0)
    return 11.0 * stats["wn_node_visits"] + 89.0 * stats["wn_far_evals"] + 11.0 * leaf_visits + \
        77.0 * stats["wn_leaf_tris"]


def cpu_baseline(soup, h, band, max_threads=16):
    """The oracle on a polar cap of the same mesh (triangles with centroid z > 0.8,
    about a tenth of the surface)."""
    import numpy as np
    from oracle import oracle as O
    if not os.path.exists(O.LIB_PATH):
        O.build()
    cent = soup.mean(axis=1)
    cap = np.ascontiguousarray(soup[cent[:, 2] > 0.8])
    threads = max(1, min(max_threads, os.cpu_count() or 1))
    vol, sec = O.mesh_to_volume(cap, h, band, threads=threads, timing=True)
    n_act = vol.export().n_active
    return {"value": n_act / sec, "unit": "voxels/s", "cores": threads, "kind": "port",
            "sample": f"polar cap z>0.8 of the same mesh: {cap.shape[0]} triangles, {n_act} active voxels, "
                      f"{sec:.2f} s",
            "seconds": round(sec, 3)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--level", type=int, default=8)
    ap.add_argument("--voxel", type=float, default=0.002)
    ap.add_argument("--band", type=int, default=3)
    ap.add_argument("--profile-steps", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the multi-GPU layout is a design "
              "note, not a measured path.", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import baby_shark_amd as bs
    from baby_shark_amd import meshgen as mg

    soup = mg.displaced_icosphere(args.level, 1.0, 0.05, seed=2)
    dev = torch.from_numpy(soup).cuda()
    torch.cuda.synchronize()
    ctx = bs.default_context(0)
    m2v = bs.MeshToVolume(ctx).with_voxel_size(args.voxel).with_narrow_band_width(args.band)

    for _ in range(args.warmup):
        v = m2v.convert(dev)
        v.free()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    n_active = 0
    for _ in range(args.steps):
        v = m2v.convert(dev)
        v.free()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    stats = ctx.last_stats()
    n_active = stats["n_active"]
    ms_step = (t1 - t0) * 1e3 / args.steps

    # kernel durations: HIP events on the library stream around each phase
    ctx.set_profiling(True)
    prof = []
    for _ in range(max(1, args.profile_steps)):
        v = m2v.convert(dev)
        v.free()
        prof.append(ctx.last_stats())
    ctx.set_profiling(False)
    ms_dist = float(np.mean([p["ms_distance"] for p in prof]))
    ms_wn = float(np.mean([p["ms_winding"] for p in prof]))
    st = prof[-1]
    kernels = {
        "k_distance": {"ms": ms_dist, "flop": distance_flop(st)},
        "k_winding": {"ms": ms_wn, "flop": winding_flop(st)},
    }
    for k in kernels.values():
        k["tflops"] = k["flop"] / (k["ms"] * 1e-3) / 1e12 if k["ms"] > 0 else 0.0
    dom = max(kernels, key=lambda k: kernels[k]["ms"])
    roof = {"bound": "valu_fp32", "kernel": dom, "achieved": round(kernels[dom]["tflops"], 3),
            "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s", "frac": round(kernels[dom]["tflops"] / PEAK_FP32_TFLOPS, 4),
            "traffic": None}

    out = {
        "metric": "narrow-band SDF voxels/s at 1.3M tris (plain), err<=1e-4 vox",
        "value": n_active / (ms_step * 1e-3),
        "unit": "voxels/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"mesh_to_volume displaced icosphere L{args.level} ({soup.shape[0]} tris), "
                               f"voxel {args.voxel}, band {args.band}, winding-number sign",
                   "n_tris": int(soup.shape[0]), "voxel_size": args.voxel, "band": args.band,
                   "active_voxels": int(n_active), "leaves": int(st["n_leaves"]),
                   "sub_triangles": int(st["n_sub"]), "pairs": int(st["n_pairs"])},
        "roofline": roof,
        "kernels": {k: {"ms": round(v["ms"], 4), "tflops": round(v["tflops"], 3),
                        "frac": round(v["tflops"] / PEAK_FP32_TFLOPS, 4)} for k, v in kernels.items()},
        "phases_ms": {k: round(float(np.mean([p[k] for p in prof])), 4)
                      for k in ("ms_subdivide", "ms_leaves", "ms_distance", "ms_bvh", "ms_winding")},
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(soup, args.voxel, args.band)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
