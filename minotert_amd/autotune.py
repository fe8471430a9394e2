"""Autotuner for the traversal kernels' launch shape and the path tracer's frame launch shapes.

Every variant in the library's tables returns identical results:
  * traversal (mrt_trace_variant_*): block size x LDS stack depth x lane-refill policy x XCD
    partitioning; `tune` times each on the caller's own rays (HIP events on the library stream);
  * frame (mrt_frame_variant_*): the shading block, which is also the queue-append chunk, and the
    megakernel's block and LDS stack depth; `tune_frame` times whole frames of the caller's own
    workload (a callable, synchronized wall time), first over the traversal variants with the
    default frame shape, then over the frame variants with the traversal variant chosen.
Measurements run in interleaved rounds (one run of every candidate per round after a warm-up
run), so that clock and thermal drift falls on all candidates alike; the fastest median wins and
the choice is recorded in a JSON cache keyed by device ("gfx950...-256CU"), kind and workload
name. A cached choice is reused as is, so re-running on the same machine gives the same choice
without re-measuring.
"""
from __future__ import annotations

import json
import os
import statistics
import time

from .runtime import Context, DeviceArray, Scene, frame_variants, trace_variants

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


def _interleaved(ctx: Context, run, candidates: list, apply, reps: int, warmup: int = 1) -> dict:
    """Median synchronized wall milliseconds of run() under each candidate (applied by apply)."""
    times = {c: [] for c in candidates}
    for _ in range(reps):
        for c in candidates:
            apply(c)
            for _ in range(warmup):
                run()
            ctx.sync()
            t = time.perf_counter()
            run()
            ctx.sync()
            times[c].append((time.perf_counter() - t) * 1e3)
    return {c: statistics.median(t) for c, t in times.items()}


def tune_frame(ctx: Context, run, workload: str, cache_path: str | None = DEFAULT_CACHE, force: bool = False,
               reps: int = 5) -> dict:
    """Select the traversal variant (closest and any-hit kernels alike) and the frame variant for
    a whole-frame workload `run()` (e.g. one batch frame): the cached choice if there is one,
    otherwise measured. Returns {"trace_variant", "frame_variant", "source", "ms"}."""
    tnames, fnames = trace_variants(), frame_variants()
    key = f"{ctx.device_name()}|frame|{workload}"
    cache = load_cache(cache_path) if cache_path else {}
    if not force and key in cache and cache[key].get("trace_variant") in tnames and \
            cache[key].get("frame_variant") in fnames:
        rec = cache[key]
        ti, fi = tnames.index(rec["trace_variant"]), fnames.index(rec["frame_variant"])
        for any_hit in (False, True):
            ctx.set_trace_variant(ti, any_hit)
        ctx.set_frame_variant(fi)
        return {"trace_variant": rec["trace_variant"], "frame_variant": rec["frame_variant"], "source": "cache",
                "ms": rec.get("ms", {})}

    def set_trace(i):
        for any_hit in (False, True):
            ctx.set_trace_variant(i, any_hit)

    ctx.set_frame_variant(0)
    ms_t = _interleaved(ctx, run, list(range(len(tnames))), set_trace, reps)
    ti = min(ms_t, key=ms_t.get)
    set_trace(ti)
    ms_f = _interleaved(ctx, run, list(range(len(fnames))), ctx.set_frame_variant, reps)
    fi = min(ms_f, key=ms_f.get)
    ctx.set_frame_variant(fi)
    ms = {f"trace:{tnames[i]}": round(v, 4) for i, v in ms_t.items()}
    ms.update({f"frame:{fnames[i]}": round(v, 4) for i, v in ms_f.items()})
    if cache_path:
        cache = load_cache(cache_path)
        cache[key] = {"trace_variant": tnames[ti], "frame_variant": fnames[fi], "ms": ms}
        with open(cache_path, "w", encoding="utf-8") as f:
            json.dump(cache, f, indent=1, sort_keys=True)
    return {"trace_variant": tnames[ti], "frame_variant": fnames[fi], "source": "measured", "ms": ms}
