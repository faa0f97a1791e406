"""Autotuning of the batched mesh-to-SDF call.

`bs_batch_mesh_to_volume` spreads meshes over worker streams, each with its
own host thread and scratch arena.  How many pay off depends on the device,
the host and the batch (a few large meshes want few workers, many small ones
want more to overlap their short kernels and host round trips).
`batch_workers` times the caller's own batch once per candidate count and
keeps the fastest in a JSON cache keyed by device and batch shape (mesh count,
total triangles and the largest mesh, both rounded to a power of two), so
later batches of that shape skip the timing.  The result of a batch does not
depend on the worker count (tests/test_batch_gpu.py), only its speed.

The cache file is `$BS_TUNE_CACHE`, else ~/.cache/baby_shark_amd/tune.json.
"""
from __future__ import annotations

import json
import math
import os
import time
from typing import Callable, Optional, Sequence

CANDIDATES = (1, 2, 4, 8, 16)


def cache_path() -> str:
    return os.environ.get("BS_TUNE_CACHE") or os.path.join(os.path.expanduser("~"), ".cache", "baby_shark_amd",
                                                           "tune.json")


def _pow2(x: int) -> int:
    return 1 << max(0, int(math.ceil(math.log2(max(1, x)))))


def batch_key(device: str, tri_counts: Sequence[int], band: int, shared_signs: bool) -> str:
    """The batch shape a tuned worker count applies to."""
    return (f"batch_workers|{device}|meshes={_pow2(len(tri_counts))}|tris={_pow2(sum(tri_counts))}"
            f"|largest={_pow2(max(tri_counts, default=0))}|band={band}|shared={int(bool(shared_signs))}")


def load_cache(path: Optional[str] = None) -> dict:
    path = path or cache_path()
    try:
        with open(path, "r", encoding="utf-8") as f:
            data = json.load(f)
        return data if isinstance(data, dict) else {}
    except (OSError, ValueError):
        return {}


def store(key: str, value: dict, path: Optional[str] = None) -> None:
    """Adds one entry; a cache that cannot be written is skipped."""
    path = path or cache_path()
    data = load_cache(path)
    data[key] = value
    try:
        os.makedirs(os.path.dirname(path), exist_ok=True)
        tmp = f"{path}.{os.getpid()}.tmp"
        with open(tmp, "w", encoding="utf-8") as f:
            json.dump(data, f, indent=1, sort_keys=True)
        os.replace(tmp, path)
    except OSError:
        pass


def pick_fastest(run: Callable[[int], float], candidates: Sequence[int] = CANDIDATES, reps: int = 1) -> dict:
    """run(workers) -> seconds.  Each candidate is timed `reps` times (the
    best kept); returns {"workers": best, "ms": {workers: ms}}."""
    times = {}
    for w in candidates:
        times[int(w)] = min(run(int(w)) for _ in range(max(1, reps))) * 1e3
    best = min(times, key=lambda w: (times[w], w))
    return {"workers": best, "ms": {str(k): round(v, 3) for k, v in times.items()}}


def batch_workers(meshes, voxel_sizes, band: int, ctx=None, shared_signs: bool = False,
                  candidates: Sequence[int] = CANDIDATES, reps: int = 1, path: Optional[str] = None,
                  retune: bool = False) -> int:
    """The worker count to pass to batch_mesh_to_volume for this batch: from
    the cache, else timed on the batch itself (every candidate runs the whole
    batch `reps` times) and stored."""
    import torch

    from .voxel import batch_mesh_to_volume, default_context

    ctx = ctx or default_context()
    counts = [int(getattr(m, "shape", (len(m),))[0]) for m in meshes]
    key = batch_key(torch.cuda.get_device_name(ctx.device), counts, band, shared_signs)
    if not retune:
        hit = load_cache(path).get(key)
        if isinstance(hit, dict) and int(hit.get("workers", 0)) in candidates:
            return int(hit["workers"])

    def run(workers: int) -> float:
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        vols = batch_mesh_to_volume(meshes, voxel_sizes, band, ctx=ctx, workers=workers, shared_signs=shared_signs)
        dt = time.perf_counter() - t0
        for v in vols:
            if v is not None:
                v.free()
        return dt

    run(4)  # warm: arenas and pools at their size before anything is timed
    result = pick_fastest(run, candidates, reps)
    store(key, result, path)
    return int(result["workers"])
