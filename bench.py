"""Narrow-band SDF throughput on one MI355X.

Workload (default): displaced icosphere, 8 subdivisions (1,310,720 f32
triangles, radius 1, radial 3-octave value noise amplitude 0.05, base
frequency 4, seed 2), voxel 0.002, 3-voxel band, sign by winding number.
A step is one MeshToVolume::convert of the device-resident soup into a
device-resident volume; value = active voxels / second.

Prints one JSON line.  `roofline` is for the dominant kernel (FP32 VALU work
from the kernels' own counters / its HIP-event time on the library stream);
`cpu_baseline` times the CPU oracle (the reference algorithm restated in C++,
same parallel decomposition) on the whole mesh, and `error` checks the GPU's
volume against that oracle volume (and against FP64 distances on a sample).
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
# HBM bytes per k_winding launch at the default workload, from the FETCH_SIZE
# (x2.0 calibration) and WRITE_SIZE counters of `bench.py --traffic-probe`
# under rocprofv3 --pmc
WINDING_TRAFFIC_BYTES = 1.015e9


def traffic_probe():
    """For rocprofv3 --pmc runs: a 1 GiB fill and a 1 GiB sum (counter
    calibration), then one mesh-to-volume step of the default workload."""
    import torch
    import baby_shark_amd as bs
    from baby_shark_amd import meshgen as mg
    n = 256 * 2 ** 20
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    total = float(src.sum())
    torch.cuda.synchronize()
    print("calibration: 1 GiB written, 1 GiB read", total, flush=True)
    soup = torch.from_numpy(mg.displaced_icosphere(8, 1.0, 0.05, seed=2)).cuda()
    torch.cuda.synchronize()
    ctx = bs.default_context(0)
    v = bs.MeshToVolume(ctx).with_voxel_size(0.002).with_narrow_band_width(3).convert(soup)
    print("mesh_to_volume active voxels", ctx.last_stats()["n_active"], flush=True)
    v.free()


def distance_flop(stats):
    # 82 FLOP per (voxel, sub-triangle) pair of the reference pair set
    return 82.0 * stats["n_pairs"]


def winding_flop(stats):
    # 11 per near internal-node visit, 89 per far (order-2) evaluation,
    # 11 + 77 per triangle for each near leaf visit (atan2 counted as 1)
    leaf_visits = stats.get("wn_leaf_visits", 0)
    return 11.0 * stats["wn_node_visits"] + 89.0 * stats["wn_far_evals"] + 11.0 * leaf_visits + \
        77.0 * stats["wn_leaf_tris"]


def surrogate_leg(ctx, n_train=32, n_test=8, h=0.002, band=3):
    """The learned tile gate on the headline family: a 3-64-64-64-1 network
    (BF16 weights, FP32 accumulation, trained on the GPU by the library's own
    kernels) learns a 4^3 tile's smallest |SDF| from 2M tile samples of
    n_train seeded displaced icospheres (8 subdivisions), then gates the sign
    work on n_test unseen seeds.  Reported beside the per-voxel and shared-sign
    paths and the same gate fed each tile's true smallest |SDF| (a perfect
    classifier); every volume is compared with the per-voxel one."""
    import numpy as np
    from baby_shark_amd import meshgen as mg
    from baby_shark_amd.surrogate import compare_sign_modes, train_tile_classifier
    t0 = time.perf_counter()
    train = mg.displaced_icospheres(8, list(range(100, 100 + n_train)))
    tests = mg.displaced_icospheres(8, list(range(200, 200 + n_test)))
    gen_s = time.perf_counter() - t0
    net, rep = train_tile_classifier(ctx, train, h, band, samples=2_000_000, epochs=12, seed=3)
    del train
    thr = 2.0 + max(rep["over_prediction_q999_vox"], 0.0)
    rows = compare_sign_modes(ctx, tests, h, band, net, thr)
    net.close()
    out = {"workload": f"displaced icospheres L8, voxel {h}, band {band}: trained on {n_train} seeds, "
                       f"tested on {n_test} unseen seeds", "mesh_generation_s": round(gen_s, 1),
           "classifier": rep, "threshold_vox": round(thr, 4)}
    for name in ("per_voxel", "shared", "surrogate", "exact_gate"):
        r = {"ms_mean": round(float(np.mean([x[name]["ms"] for x in rows])), 3),
             "evaluated_fraction": round(float(np.mean([x[name]["evals"] / x[name]["active"] for x in rows])), 5),
             "speedup_vs_per_voxel": round(float(np.mean([x["per_voxel"]["ms"] / x[name]["ms"] for x in rows])), 3),
             "speedup_vs_shared": round(float(np.mean([x["shared"]["ms"] / x[name]["ms"] for x in rows])), 3)}
        if name != "per_voxel":
            r["values_differing"] = int(sum(x[name]["values_differing"] for x in rows))
            r["sign_mismatches"] = int(sum(x[name]["sign_mismatches"] for x in rows))
        if name in ("surrogate", "exact_gate"):
            for k in ("tiles", "culled", "misclassified", "fallback_voxels"):
                r[k] = int(sum(x[name][k] for x in rows))
        out[name] = r
    return out


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline_and_error(soup, h, band, gpu_export, max_threads=16, n_sample=20000):
    """The CPU oracle (the reference's algorithm and parallel decomposition,
    restated in C++) converts the whole mesh on the host cores -- the CPU
    baseline -- and its volume is the checker of the GPU's: leaf keys and
    masks, every value (max and RMS error in voxels), every sign off the
    surface; distances at a fixed sample of voxels against FP64."""
    import numpy as np
    from oracle import oracle as O
    if not os.path.exists(O.LIB_PATH):
        O.build()
    nproc = os.cpu_count() or 1
    threads = max(1, min(max_threads, nproc))
    vol, sec = O.mesh_to_volume(soup, h, band, threads=threads, timing=True)
    ref = vol.export()
    vol = None
    n_act = ref.n_active
    base = {"value": n_act / sec, "unit": "voxels/s", "cores": threads, "kind": "port",
            "sample": f"the whole mesh: {soup.shape[0]} triangles, {n_act} active voxels, {sec:.2f} s, one run",
            "seconds": round(sec, 3), "nproc": nproc, "cpu_model": _cpu_model()}
    g = gpu_export
    err = {"checked_against": "oracle whole volume (reference algorithm, f32); FP64 distances on a fixed sample"}
    same_keys = g.origins.shape == ref.origins.shape and np.array_equal(g.origins, ref.origins)
    same_masks = same_keys and np.array_equal(g.masks, ref.masks)
    err["keys_equal"], err["masks_equal"] = bool(same_keys), bool(same_masks)
    if same_masks:
        act = ref.active()
        gv = g.values[act].astype(np.float64)
        rv = ref.values[act].astype(np.float64)
        e = np.abs(gv - rv) / h
        err["max_vox"] = float(e.max())
        err["rms_vox"] = float(np.sqrt(np.mean(e ** 2)))
        err["bit_exact_frac"] = float(np.mean(g.values[act] == ref.values[act]))
        li, off = np.nonzero(act)
        vox = ref.origins[li] + np.stack([off >> 6, (off >> 3) & 7, off & 7], 1)
        pts = vox.astype(np.float32) * np.float32(h)
        ulp = np.spacing(np.max(np.abs(pts), axis=1).astype(np.float32)).astype(np.float64)
        off_surface = np.abs(rv) > np.maximum(1e-6 * h, 16.0 * ulp)
        err["sign_mismatches"] = int(np.sum(np.signbit(gv[off_surface]) != np.signbit(rv[off_surface])))
        err["signs_checked"] = int(off_surface.sum())
        rng = np.random.default_rng(12345)
        pick = rng.choice(li.size, min(n_sample, li.size), replace=False)
        _, r64 = O.sample_unsigned(soup, h, band, vox[pick])
        e64 = np.abs(np.abs(gv[pick]) - r64) / h
        err["max_vs_fp64_vox"] = float(e64.max())
        err["rms_vs_fp64_vox"] = float(np.sqrt(np.mean(e64 ** 2)))
        err["fp64_sample"] = int(pick.size)
    return base, err


def pipeline(ctx, level=7, h=0.001, band=3, reps=2):
    """CSG, offset and dual contouring on the two-icosphere workload: displaced
    icospheres of `level` subdivisions (seeds 3 and 4, the second moved by
    (0.3, 0, 0)) voxelized at h; S = A - B, I = A n B, S.offset(+0.02),
    I.offset(-0.01), dual contouring of S.  Each operation is timed on
    fresh clones (clone time excluded); the last of `reps` rounds is kept.
    Algorithmic bytes: flood fill 4160 B per operand leaf (read 2048 + 64,
    write 2048), assembly 6336 B per paired result leaf (two leaves read, one
    written) and 4224 B per copied one; offset 2176 B per input leaf
    (remove_if), 4224 B per leaf sweep (read + write), 4096 B per output leaf
    (shift)."""
    import torch
    import baby_shark_amd as bs
    from baby_shark_amd import meshgen as mg
    a = torch.from_numpy(mg.displaced_icosphere(level, 1.0, 0.05, seed=3)).cuda()
    b = torch.from_numpy(mg.displaced_icosphere(level, 1.0, 0.05, seed=4, center=(0.3, 0.0, 0.0))).cuda()
    torch.cuda.synchronize()
    m2v = bs.MeshToVolume(ctx).with_voxel_size(h).with_narrow_band_width(band)
    A, B = m2v.convert(a), m2v.convert(b)
    res = {}
    for _ in range(reps):
        r = {}
        for name, op in (("subtract", "subtract"), ("intersect", "intersect")):
            x, y = A.clone(), B.clone()
            t0 = time.perf_counter()
            out = getattr(x, op)(y)
            wall = (time.perf_counter() - t0) * 1e3
            st = ctx.last_csg_stats()
            flood_b = 4160.0 * (st["n_a"] + st["n_b"])
            asm_b = 6336.0 * st["n_paired"] + 4224.0 * (st["n_out"] - st["n_paired"])
            dev_ms = st["ms_flood"] + st["ms_assemble"]
            r[name] = {"wall_ms": round(wall, 3), "flood_ms": round(st["ms_flood"], 4),
                       "assemble_ms": round(st["ms_assemble"], 4), "leaves_a": st["n_a"], "leaves_b": st["n_b"],
                       "leaves_out": st["n_out"], "paired": st["n_paired"],
                       "flood_gbs": round(flood_b / (st["ms_flood"] * 1e6), 1) if st["ms_flood"] > 0 else None,
                       "assemble_gbs": round(asm_b / (st["ms_assemble"] * 1e6), 1) if st["ms_assemble"] > 0 else None,
                       "kernels_gbs": round((flood_b + asm_b) / (dev_ms * 1e6), 1) if dev_ms > 0 else None,
                       "wall_gbs": round((flood_b + asm_b) / (wall * 1e6), 1),
                       "host_ms": {k: round(st[k], 2) for k in ("ms_topology", "ms_plan", "ms_plan_fill", "ms_plan_tree",
                                                                 "ms_plan_order", "ms_memory", "ms_compact")}}
            r["_" + name] = out
        for name, src, dist in (("offset_S_+0.02", "_subtract", 0.02), ("offset_I_-0.01", "_intersect", -0.01)):
            x = r[src].clone()
            n_in = x.info()["n_leaves"]
            t0 = time.perf_counter()
            out = x.offset(dist)
            wall = (time.perf_counter() - t0) * 1e3
            st = ctx.last_offset_stats()
            byts = 2176.0 * n_in + 4224.0 * st["n_swept"] + 4096.0 * st["n_out"]
            r[name] = {"wall_ms": round(wall, 3), "device_ms": round(st["ms_device"], 3), "leaves_in": n_in,
                       "leaves_swept": st["n_swept"], "leaves_out": st["n_out"], "planes": st["n_planes"],
                       "launches": st["n_launches"],
                       "device_gbs": round(byts / (st["ms_device"] * 1e6), 1) if st["ms_device"] > 0 else None}
            r["_" + name] = out
        t0 = time.perf_counter()
        dm = bs.DualContouringMesher(ctx).with_voxel_size(h).mesh_device(r["_offset_S_+0.02"])
        wall = (time.perf_counter() - t0) * 1e3
        st = ctx.last_dc_stats()
        r["dual_contour_S"] = {"wall_ms": round(wall, 3), "device_ms": round(st["ms_total"], 3),
                               "leaves": st["n_leaves"], "triangles": st["n_tris"],
                               "tris_per_s": round(st["n_tris"] / (st["ms_total"] * 1e-3)) if st["ms_total"] else None}
        dm.free()
        # marching cubes of the same volume (algorithmic bytes: 2112 B per leaf
        # read, 36 B per triangle written)
        t0 = time.perf_counter()
        mm = bs.MarchingCubesMesher(ctx).with_voxel_size(h).mesh_device(r["_offset_S_+0.02"])
        wall = (time.perf_counter() - t0) * 1e3
        st = ctx.last_mc_stats()
        r["marching_cubes_S"] = {"wall_ms": round(wall, 3), "device_ms": round(st["ms_total"], 3),
                                 "leaves": st["n_leaves"], "triangles": st["n_tris"],
                                 "tris_per_s": round(st["n_tris"] / (st["ms_total"] * 1e-3)) if st["ms_total"] else None,
                                 "device_gbs": round((2112.0 * st["n_leaves"] + 36.0 * st["n_tris"]) /
                                                     (st["ms_total"] * 1e6), 1) if st["ms_total"] else None}
        mm.free()
        for k in [k for k in r if k.startswith("_")]:
            r[k].free()
        res = r
    A.free()
    B.free()
    res = {k: v for k, v in res.items() if not k.startswith("_")}
    res["workload"] = (f"two displaced icospheres L{level} ({a.shape[0]} tris each, second moved by 0.3 in x), "
                       f"voxel {h}, band {band}")
    return res


def shared_signs(ctx, dev, args, ms_plain):
    """The same convert with shared signs (one winding-number evaluation per
    group of voxels that no surface point separates): K timed steps after W
    warm-ups, phase times from HIP events, and the whole volume compared bit
    for bit with the per-voxel-sign volume."""
    import numpy as np
    import baby_shark_amd as bs
    import torch
    m2v = bs.MeshToVolume(ctx).with_voxel_size(args.voxel).with_narrow_band_width(args.band)
    m2s = bs.MeshToVolume(ctx).with_voxel_size(args.voxel).with_narrow_band_width(args.band).with_shared_signs()
    for _ in range(args.warmup):
        m2s.convert(dev).free()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        m2s.convert(dev).free()
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / args.steps
    ctx.set_profiling(True, counters=False)
    v = m2s.convert(dev)
    st = ctx.last_stats()
    ctx.set_profiling(False)
    a = v.export()
    v.free()
    p = m2v.convert(dev)
    b = p.export()
    p.free()
    diff = int(np.count_nonzero(a.values.view(np.uint32) != b.values.view(np.uint32)))
    sign_diff = int(np.count_nonzero(np.signbit(a.values[a.active()]) != np.signbit(b.values[b.active()])))
    return {"ms_per_step": round(ms, 3), "voxels_per_s": st["n_active"] / (ms * 1e-3),
            "speedup_vs_plain": round(ms_plain / ms, 3), "winding_evaluations": st["n_sign_evals"],
            "evaluated_fraction": round(st["n_sign_evals"] / st["n_active"], 5),
            "ms_groups": round(st["ms_sign_share"], 4), "ms_winding": round(st["ms_winding"], 4),
            "values_differing_from_plain": diff, "sign_mismatches_vs_plain": sign_diff}


def torus_and_sphere(ctx, reps=5):
    """The small end-to-end workload: torus (R 1, r 0.35, 128 x 64 segments)
    to a volume at h = 0.02, band 3, united with the sphere r = 0.5 at
    (1, 0, 0), its active tiles densified, then dual contoured.  Wall time of
    each stage from a resident soup; the best of `reps` runs is kept."""
    import torch
    import baby_shark_amd as bs
    from baby_shark_amd import meshgen as mg
    soup = torch.from_numpy(mg.torus(1.0, 0.35, 128, 64, 0.01, seed=1)).cuda()
    torch.cuda.synchronize()
    best = None
    for _ in range(reps):
        t = [time.perf_counter()]
        v = bs.MeshToVolume(ctx).with_voxel_size(0.02).with_narrow_band_width(3).convert(soup)
        t.append(time.perf_counter())
        v = v.union(bs.VolumeBuilder(ctx, 0.02).sphere(0.5, (1.0, 0.0, 0.0)))
        t.append(time.perf_counter())
        n_tiles = v.info()["n_tiles"]
        t.append(time.perf_counter())
        mc = bs.MarchingCubesMesher(ctx).with_voxel_size(0.02).mesh_device(v)   # tiles met directly
        t.append(time.perf_counter())
        n_mc = ctx.last_mc_stats()["n_tris"]
        mc.free()
        t.append(time.perf_counter())
        v.densify_tiles()
        t.append(time.perf_counter())
        tris = bs.DualContouringMesher(ctx).with_voxel_size(0.02).mesh_device(v)
        t.append(time.perf_counter())
        n_tris = ctx.last_dc_stats()["n_tris"]
        tris.free()
        v.free()
        ms = [(b - a) * 1e3 for a, b in zip(t, t[1:])]
        ms = [ms[0], ms[1], ms[3], ms[5], ms[6]]   # m2v, union, marching cubes, densify, dual contour
        if best is None or sum(ms) < sum(best[0]):
            best = (ms, n_tiles, n_tris, n_mc)
    ms, n_tiles, n_tris, n_mc = best
    return {"workload": f"torus {soup.shape[0]} tris, voxel 0.02, band 3, union sphere r 0.5; marching cubes of "
                        f"the union (tiles included); densify, dual contour",
            "mesh_to_volume_ms": round(ms[0], 3), "union_ms": round(ms[1], 3), "marching_cubes_ms": round(ms[2], 3),
            "mc_triangles": n_mc, "densify_ms": round(ms[3], 3), "dual_contour_ms": round(ms[4], 3),
            "total_ms": round(ms[0] + ms[1] + ms[3] + ms[4], 3), "tiles_densified": n_tiles, "triangles": n_tris}


def lightweighting(ctx, reps=3):
    """The reference's examples/lightweighting.rs on the GPU: the bunny
    (13,000 triangles, read from binary STL) to a volume at voxel 0.2 (band
    0), an IWP lattice (cell 0.5) over [-26, 26]^2 x [0, 50], offset -1,
    intersect, subtract, union, subtract of a clip box, dual contouring.  Wall
    time per stage from a resident soup; the best of `reps` runs."""
    import os
    import torch
    import baby_shark_amd as bs
    soup = torch.from_numpy(bs.read_stl(os.path.join(os.path.dirname(os.path.abspath(__file__)), "tests", "golden",
                                                     "bunny.stl"))).cuda()
    torch.cuda.synchronize()
    h = 0.2
    B = bs.VolumeBuilder(ctx, h)
    best = None
    for _ in range(reps):
        t = [time.perf_counter()]
        bunny = bs.MeshToVolume(ctx).with_voxel_size(h).convert(soup)
        t.append(time.perf_counter())
        iwp = B.iwp((-26.0, -26.0, 0.0), (26.0, 26.0, 50.0), 0.5)
        t.append(time.perf_counter())
        bunny_offset = bunny.clone().offset(-1.0)
        t.append(time.perf_counter())
        bunny_iwp = bunny_offset.clone().intersect(iwp)
        clip = B.cuboid((-30.0, -30.0, -2.0), (30.0, -5.0, 50.0))
        res = bunny.subtract(bunny_offset).union(bunny_iwp).subtract(clip)
        t.append(time.perf_counter())
        tris = bs.DualContouringMesher(ctx).with_voxel_size(h).mesh_device(res)
        t.append(time.perf_counter())
        n_tris = ctx.last_dc_stats()["n_tris"]
        tris.free()
        res.free()
        ms = [(b - a) * 1e3 for a, b in zip(t, t[1:])]
        if best is None or sum(ms) < sum(best[0]):
            best = (ms, n_tris)
    ms, n_tris = best
    return {"workload": "examples/lightweighting.rs: bunny 13000 tris, voxel 0.2, IWP cell 0.5, offset -1, "
                        "4 CSG ops, dual contour",
            "mesh_to_volume_ms": round(ms[0], 3), "iwp_ms": round(ms[1], 3), "offset_ms": round(ms[2], 3),
            "csg_ms": round(ms[3], 3), "dual_contour_ms": round(ms[4], 3), "total_ms": round(sum(ms), 3),
            "triangles": n_tris}


def remesh(ctx, reps=3):
    """VoxelRemesher (band-0 conversion and meshing in one library call) on
    the bunny at voxel 0.1 and at 0.05, with the default manifold method
    (marching cubes) and the feature-preserving one (dual contouring), from a
    resident soup to a host soup; the best of `reps` runs."""
    import os
    import torch
    import baby_shark_amd as bs
    soup = torch.from_numpy(bs.read_stl(os.path.join(os.path.dirname(os.path.abspath(__file__)), "tests", "golden",
                                                     "bunny.stl"))).cuda()
    torch.cuda.synchronize()
    res = {"workload": "VoxelRemesher of the bunny, 13000 tris, to a host soup (manifold = marching cubes, "
                       "feature preserving = dual contouring)"}
    for h in (0.1, 0.05):
        for name, method in (("manifold", bs.MeshingMethod.MANIFOLD),
                             ("feature_preserving", bs.MeshingMethod.FEATURE_PRESERVING)):
            r = bs.VoxelRemesher(ctx).with_voxel_size(h).with_meshing_method(method)
            best, n = None, 0
            for _ in range(reps):
                t0 = time.perf_counter()
                out = r.remesh(soup)
                dt = (time.perf_counter() - t0) * 1e3
                n = out.shape[0]
                best = dt if best is None else min(best, dt)
            res[f"{name}_voxel_{h}"] = {"ms": round(best, 3), "triangles": n}
    return res


def batch(ctx, count=256, seed=5):
    """The uneven batch: 256 closed meshes (tori and displaced icospheres)
    with triangle counts log-uniform in [1e3, 2e5], each voxelized so its
    bounding box spans 256 voxels, band 3, soups resident in HBM.  Timed as
    one bs_batch_mesh_to_volume call over 8 and 4 streams, and as the same
    call on 1 stream (the per-mesh loop a caller would write)."""
    import torch
    import baby_shark_amd as bs
    from baby_shark_amd import meshgen as mg
    ms = mg.batch_meshes(count=count, seed=seed)
    soups = [torch.from_numpy(s).cuda() for s, _ in ms]
    hs = [h for _, h in ms]
    torch.cuda.synchronize()
    res = {"workload": f"{count} meshes (tori / displaced icospheres), {sum(s.shape[0] for s, _ in ms)} triangles "
                       f"in all, {min(s.shape[0] for s, _ in ms)}-{max(s.shape[0] for s, _ in ms)} each, "
                       "bounding box 256 voxels, band 3"}
    # the autotuned worker count (timed once per device and batch shape, then cached)
    from baby_shark_amd import tune
    res["autotuned_workers"] = tune.batch_workers(soups, hs, 3, ctx=ctx)
    res["autotune_cache"] = tune.load_cache().get(
        tune.batch_key(torch.cuda.get_device_name(0), [s.shape[0] for s in soups], 3, False))
    runs = [("streams_8", 8), ("streams_4", 4), ("streams_1", 1)]
    runs.insert(0, ("autotuned", res["autotuned_workers"]))
    for name, workers in runs:
        best = None
        for _ in range(2):
            t0 = time.perf_counter()
            vols = bs.batch_mesh_to_volume(soups, hs, 3, ctx=ctx, workers=workers)
            dt = time.perf_counter() - t0
            n_act = sum(v.info()["n_active"] for v in vols if v is not None)
            for v in vols:
                if v is not None:
                    v.free()
            best = dt if best is None else min(best, dt)
        res[name] = {"ms": round(best * 1e3, 2), "active_voxels": n_act, "voxels_per_s": round(n_act / best)}
    return res


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
    ap.add_argument("--no-pipeline", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-shared", action="store_true")
    ap.add_argument("--no-surrogate", action="store_true")
    ap.add_argument("--traffic-probe", action="store_true")
    args = ap.parse_args()
    if args.traffic_probe:
        traffic_probe()
        return
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

    # work: the winding kernel's visit counters (one step with the counting
    # variant); kernel durations: HIP events on the library stream around each
    # phase of the production (non-counting) kernels
    ctx.set_profiling(True, counters=True)
    v = m2v.convert(dev)
    v.free()
    counted = ctx.last_stats()
    ctx.set_profiling(True, counters=False)
    prof = []
    for _ in range(max(1, args.profile_steps)):
        v = m2v.convert(dev)
        v.free()
        prof.append(ctx.last_stats())
    ctx.set_profiling(False)
    ms_dist = float(np.mean([p["ms_distance"] for p in prof]))
    ms_wn = float(np.mean([p["ms_winding"] for p in prof]))
    st = counted
    kernels = {
        "k_distance": {"ms": ms_dist, "flop": distance_flop(st)},
        "k_winding": {"ms": ms_wn, "flop": winding_flop(st)},
    }
    for k in kernels.values():
        k["tflops"] = k["flop"] / (k["ms"] * 1e-3) / 1e12 if k["ms"] > 0 else 0.0
    dom = max(kernels, key=lambda k: kernels[k]["ms"])
    # compute-bound: the contract names that "mfma"; the work here is vector
    # FP32 (VALU), priced against MI355X's published FP32 peak
    roof = {"bound": "mfma", "bound_detail": "compute, vector FP32 (VALU); no matrix-shaped work", "kernel": dom, "achieved": round(kernels[dom]["tflops"], 3),
            "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s", "frac": round(kernels[dom]["tflops"] / PEAK_FP32_TFLOPS, 4),
            "traffic": WINDING_TRAFFIC_BYTES if dom == "k_winding" else None,
            "traffic_source": "profiles/m2v_icosphere_L8_traffic.md (rocprofv3 --pmc FETCH_SIZE, WRITE_SIZE)"}

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
        # the winding-number tree (top-down binned build) and its visits per active voxel
        "winding_tree": {"build": "topdown" if st["n_bvh_iterations"] else "lbvh",
                         "level_passes": int(st["n_bvh_iterations"]), "height": int(st["bvh_height"]),
                         "near_visits_per_voxel": round(st["wn_node_visits"] / n_active, 2),
                         "far_evals_per_voxel": round(st["wn_far_evals"] / n_active, 2),
                         "leaf_visits_per_voxel": round(st["wn_leaf_visits"] / n_active, 2),
                         "leaf_tris_per_voxel": round(st["wn_leaf_tris"] / n_active, 2)},
    }
    if not args.no_shared:
        out["shared_signs"] = shared_signs(ctx, dev, args, ms_step)
    if not args.no_pipeline:
        out["pipeline"] = pipeline(ctx)
        out["torus_and_sphere"] = torus_and_sphere(ctx)
        out["lightweighting"] = lightweighting(ctx)
        out["remesh"] = remesh(ctx)
    if not args.no_batch:
        out["batch"] = batch(ctx)
    if not args.no_surrogate:
        out["surrogate"] = surrogate_leg(ctx)
    if not args.no_cpu:
        v = m2v.convert(dev)
        ge = v.export()
        v.free()
        out["cpu_baseline"], out["error"] = cpu_baseline_and_error(soup, args.voxel, args.band, ge)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
