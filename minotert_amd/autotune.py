"""Autotuner for the traversal kernels' launch shape.

Every variant in the library's table (block size x LDS stack depth x lane-refill policy x
XCD partitioning) returns identical results; the autotuner times each one on the caller's own
rays (HIP events on the library stream) in interleaved rounds, one launch of every variant per
round after a warm-up launch, so that clock and thermal drift during the measurement falls on
all variants alike; it keeps the fastest median and records the choice in a JSON cache keyed by
device ("gfx950...-256CU"), kernel and workload name. A cached choice is reused as is, so
re-running on the same machine gives the same choice without re-measuring.
"""
from __future__ import annotations

import json
import os
import statistics

from .runtime import Context, DeviceArray, Scene, trace_variants

DEFAULT_CACHE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning.json")


def measure(ctx: Context, scene: Scene, rays: DeviceArray, out: DeviceArray, any_hit: bool = False,
            reps: int = 9, warmup: int = 1) -> dict:
    """Median kernel milliseconds of every variant, by name, over `reps` interleaved rounds."""
    name = "trace_any" if any_hit else "trace_closest"
    run = scene.trace_any if any_hit else scene.trace_closest
    names = trace_variants()
    before = ctx.trace_variant(any_hit)
    times = {v: [] for v in names}
    try:
        for _ in range(reps):
            for i, v in enumerate(names):
                ctx.set_trace_variant(i, any_hit)
                for _ in range(warmup):
                    run(rays, out)
                ctx.profile(True)
                ctx.profile_reset()
                run(rays, out)
                ms, _ = ctx.profile_query(name)
                ctx.profile(False)
                times[v].append(ms)
    finally:
        ctx.set_trace_variant(before, any_hit)
    return {v: statistics.median(t) for v, t in times.items()}


def _key(ctx: Context, any_hit: bool, workload: str) -> str:
    return f"{ctx.device_name()}|{'trace_any' if any_hit else 'trace_closest'}|{workload}"


def load_cache(path: str = DEFAULT_CACHE) -> dict:
    if not os.path.exists(path):
        return {}
    with open(path, "r", encoding="utf-8") as f:
        return json.load(f)


def tune(ctx: Context, scene: Scene, rays: DeviceArray, out: DeviceArray, workload: str, any_hit: bool = False,
         cache_path: str | None = DEFAULT_CACHE, force: bool = False) -> dict:
    """Select the traversal variant for this workload: the cached choice if there is one,
    otherwise the fastest measured variant (written back to the cache). Returns a record
    {"variant": name, "index": i, "source": "cache"|"measured", "ms": {...}}."""
    names = trace_variants()
    key = _key(ctx, any_hit, workload)
    cache = load_cache(cache_path) if cache_path else {}
    if not force and key in cache and cache[key]["variant"] in names:
        idx = names.index(cache[key]["variant"])
        ctx.set_trace_variant(idx, any_hit)
        return {"variant": names[idx], "index": idx, "source": "cache", "ms": cache[key].get("ms", {})}
    ms = measure(ctx, scene, rays, out, any_hit)
    best = min(ms, key=ms.get)
    idx = names.index(best)
    ctx.set_trace_variant(idx, any_hit)
    if cache_path:
        cache[key] = {"variant": best, "ms": {k: round(v, 5) for k, v in ms.items()}}
        with open(cache_path, "w", encoding="utf-8") as f:
            json.dump(cache, f, indent=1, sort_keys=True)
    return {"variant": best, "index": idx, "source": "measured", "ms": ms}


def apply_cached(ctx: Context, workload: str, cache_path: str = DEFAULT_CACHE) -> dict:
    """Apply cached choices (closest and any) for a workload without measuring."""
    names = trace_variants()
    cache = load_cache(cache_path)
    out = {}
    for any_hit in (False, True):
        key = _key(ctx, any_hit, workload)
        if key in cache and cache[key]["variant"] in names:
            ctx.set_trace_variant(names.index(cache[key]["variant"]), any_hit)
            out["trace_any" if any_hit else "trace_closest"] = cache[key]["variant"]
    return out
