#!/usr/bin/env python3
"""bench.py — incoherent closest-hit throughput of the BVH4 traversal kernel on one MI355X.

Workload (one step = one closest-hit launch over one resident batch):
  scene   800 level-3 icospheres, 1,024,000 FP32 triangles, centres in [-40, 40]^3
  rays    the bounce-1 extension rays (cosine-weighted, from the primary hits) of a
          1920x1080 frame, samples 0..3 of frame 0, produced on the GPU by the library's
          own primary + shade kernels, ordered by path id, resident in HBM
Prints ONE JSON line: Mrays/s = rays / (wall time of K launches / K), plus the roofline of
the trace kernel (HBM bytes per launch from the committed rocprofv3 counter summary over the
device time per step of the timed region, from HIP events on the library stream around the K
launches, with the modelled fetch bytes of the oracle's
binary-BVH counts and the VALU issue figures beside it), the CPU oracle on the host cores, and
("cache") the second half of the metric: frame time to relMSE <= 0.01 of the radiance-cache
tracer against plain path tracing on the same scene at 1920x1080.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Mrays/s incoherent closest-hit; ms/frame to relMSE<=0.01, cache vs plain PT"
HBM_PEAK_GBS = 8000.0


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def read_profile(kernel: str, workload: str) -> dict:
    """`key: number` lines of the committed rocprofv3 PMC summary of one kernel and workload
    (profiles/<kernel>_<workload>_counters.txt): HBM bytes per launch, L2 hit rate, VALU issue and
    lane utilization, ... An empty dict if the summary is absent."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_{workload}_counters.txt")
    out = {}
    if not os.path.exists(path):
        return out
    with open(path, "r", encoding="utf-8") as f:
        for line in f:
            m = re.match(r"([a-z_0-9]+):\s*([-0-9.eE+]+)", line)
            if m:
                out[m.group(1)] = float(m.group(2))
    return out


def timed_calls(ctx, fn, k: int, region: str):
    """Calls fn() k times, timed by the host clock and by one pair of HIP events on the library
    stream around all k calls (mrt_profile_region: no events between the launches). Returns
    (wall ms per call, device ms per call)."""
    ctx.sync()
    ctx.profile_reset()
    t0 = time.perf_counter()
    ctx.profile_region(region, True)
    for _ in range(k):
        fn()
    ctx.profile_region(region, False)
    ctx.sync()
    wall = (time.perf_counter() - t0) * 1e3 / k
    dev_ms, _ = ctx.profile_query(region)
    return wall, dev_ms / k


def traversal_roofline(kernel_ms: float, rays: int, workload: str, b_ray_model: float | None) -> dict:
    """Roofline of the traversal kernel against the HBM peak. `achieved` = HBM bytes per launch
    measured by rocprofv3 (FETCH_SIZE and WRITE_SIZE, calibrated on a copy kernel; committed
    summary) over the launch duration of this run's timed steps (HIP events on the library stream
    around them, divided by the step count), so `frac` is a fraction of a bound
    the kernel cannot exceed. The modelled fetch bytes (oracle binary-BVH node and triangle
    counts, 48 + 32 N_node + 40 N_tri per ray) are reported beside it: most of those fetches hit
    the L2 and the Infinity Cache, so that rate is not an HBM figure. The kernel's binding limit
    is VALU issue; the committed counters give its issue fraction and lane utilization."""
    prof = read_profile("trace_closest", workload)
    traffic = prof.get("hbm_bytes_per_launch")
    roof = {"bound": "hbm", "achieved": None, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": None, "traffic": traffic,
            "kernel": "trace_closest", "kernel_ms": round(kernel_ms, 4),
            "basis": f"rocprofv3 HBM bytes per launch (profiles/trace_closest_{workload}_counters.txt) / "
                     f"device time per step of this run's timed region (HIP events on the library stream; a "
                     f"step is the traversal launch and its queue-clear kernel)"}
    if traffic:
        gbs = traffic / (kernel_ms * 1e-3) / 1e9
        roof["achieved"] = round(gbs, 1)
        roof["frac"] = round(gbs / HBM_PEAK_GBS, 4)
        roof["hbm_bytes_per_ray"] = round(traffic / rays, 1)
    if b_ray_model is not None:
        roof["modelled_fetch"] = {"bytes_per_ray": round(b_ray_model, 1),
                                  "GBps": round(b_ray_model * rays / (kernel_ms * 1e-3) / 1e9, 1),
                                  "model": "48 + 32 N_node + 40 N_tri, N from the oracle's binary BVH on the same rays; "
                                           "cache-served fetches included, not an HBM rate"}
    if "valu_issue_frac" in prof:
        roof["valu"] = {"issue_frac": prof["valu_issue_frac"], "lane_util": prof.get("valu_lane_util"),
                        "insts_per_ray": prof.get("valu_insts_per_ray"), "l2_hit_rate": prof.get("l2_hit_rate"),
                        "source": "same committed summary (its own PMC-run duration)"}
    return roof


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo", "r", encoding="utf-8") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def make_batch(mrt, cfg, ctx, scene, samples: int):
    cam = cfg.camera(1)
    p = cfg.path_params(1)
    pt = mrt.Pathtracer(ctx, scene)
    gb = pt.primary_rays((cam.width, cam.height), cam, planes=False)
    env = mrt.Environment(sky=tuple(cfg.CONFIG["sky"]), sun_dir=tuple(cfg.CONFIG["sun_dir"]),
                          sun_e=tuple(cfg.CONFIG["sun_e"]))
    opts = mrt.RenderOptions(frame_seed=p.frame_seed, sample_base=0, spp=samples, bounces=p.bounces,
                             rr_start=p.rr_start)
    d_rays, d_pid, n = pt.capture_rays(gb, cam, env, opts, bounce=1)
    rays = d_rays.download()[:n]
    pid = d_pid.download()[:n]
    rays = rays[np.argsort(pid, kind="stable")]
    return rays


def cpu_baseline(rays: np.ndarray, verts: np.ndarray, budget_s: float = 20.0):
    """Oracle (scalar C++ port) traversal of the same rays through its binary SAH BVH on all
    host threads; also yields the mean node and triangle tests per ray."""
    from oracle import oracle as orc
    threads = orc.hardware_threads()
    t0 = time.perf_counter()
    bvh = orc.BVH(verts)
    build_s = time.perf_counter() - t0
    flat = rays.view(np.float32).reshape(-1, 8)
    t0 = time.perf_counter()
    _, nbox, ntri = bvh.trace_closest(flat, threads=threads, with_stats=True)
    dt = time.perf_counter() - t0
    reps = 1
    total = dt
    while total < min(budget_s, 10.0) and reps < 50:
        t0 = time.perf_counter()
        bvh.trace_closest(flat, threads=threads)
        total += time.perf_counter() - t0
        reps += 1
    n = flat.shape[0]
    return {
        "value": round(n * reps / total / 1e6, 3), "unit": "Mrays/s", "cores": threads, "kind": "port",
        "cpu": cpu_model(),
        "sample": f"all {n} benchmark rays x{reps} passes through the oracle's binary SAH BVH "
                  f"(16 bins, leaf<=4, build {build_s:.2f}s)",
    }, nbox / n, ntri / n


def cache_section(mrt, cfg, ctx, scene, verts) -> dict:
    """The cache-vs-plain protocol of the radiance-cache workload (config-1 scene, 1920x1080,
    4 bounces): a plain reference of reference_spp samples, then 1-spp frames accumulated until
    relMSE <= target (frame time summed, training included), for plain path tracing and for
    the cache-terminated tracer ending paths at vertex 1 (the workload's setting) and at vertex 0.
    Also the steady state after training: ms/frame and single-frame relMSE."""
    from minotert_amd import convergence as cv
    from minotert_amd.pathtracer import CachedPathtracer, rel_mse
    c3 = cfg.config(3)
    cam = cfg.camera(3)
    env = cv.environment()
    npix = cam.width * cam.height
    target = float(c3["target_relmse"])
    pt = mrt.Pathtracer(ctx, scene)
    t0 = time.perf_counter()
    ref = cv.render_reference(pt, cam, env, int(c3["reference_spp"]))
    log(f"cache protocol: {c3['reference_spp']}-spp plain reference in {time.perf_counter() - t0:.2f} s")
    rgb = ctx.empty((npix, 3), np.float32)

    def summary(c):
        return {"frames": c.frames, "ms": round(c.ms, 2), "ms_per_frame": round(c.ms_per_frame, 3),
                "reached": c.reached, "relmse_last": round(c.relmse[-1], 5) if c.relmse else None}

    def throughput(render, k=20):
        render(0, rgb)
        ctx.sync()
        t = time.perf_counter()
        for f in range(k):
            render(1 + f, rgb)
        ctx.sync()
        return (time.perf_counter() - t) * 1e3 / k

    def single_frame_relmse(render, first, k=8):
        vals = []
        for f in range(k):
            render(first + f, rgb)
            vals.append(rel_mse(ctx, rgb, ref))
        return float(np.mean(vals))

    out = {"protocol": "1-spp frames accumulated until relMSE <= target vs a plain reference; frame time "
                       "summed (render + accumulate, synchronized, training included)",
           "resolution": [cam.width, cam.height], "reference_spp": int(c3["reference_spp"]), "target": target,
           "eps": float(cfg.CONFIG["relmse_eps"])}
    plain = cv.converge(ctx, cv.plain_frames(pt, cam, env), ref, npix, target, max_frames=256)
    out["plain"] = summary(plain)
    log(f"plain: {out['plain']}")
    steady_plain = {}
    for spp in (1, 2, 4, 8):
        render = (lambda s: (lambda f, o: pt.render(cam, env, cv.frame_options(3, 5000 + f, spp=s), out=o)))(spp)
        steady_plain[spp] = {"ms_per_frame": round(throughput(render), 3),
                             "frame_relmse": round(single_frame_relmse(render, 9000), 5)}
    out["plain_per_frame"] = steady_plain
    need = next((s for s in (1, 2, 4, 8) if steady_plain[s]["frame_relmse"] <= target), None)
    for vertex in (1, 0):
        cache = cv.make_cache(ctx, cv.scene_bounds(verts))
        cpt = CachedPathtracer(ctx, scene, cache)
        co = cv.cache_options(vertex=vertex)
        render = cv.cached_frames(cpt, cam, env, co)
        conv = cv.converge(ctx, render, ref, npix, target, max_frames=400)
        rec = summary(conv)
        trained = conv.frames
        while trained < 300:  # keep training to a steady state
            render(trained, rgb)
            trained += 1
        rec["steady_after_frames"] = trained
        rec["steady_ms_per_frame"] = round(throughput(lambda f, o: render(trained + f, o)), 3)
        # the same trained cache with each frame's steps at its own end instead of beside the next
        # frame's traversal
        seq = cv.cached_frames(cpt, cam, env, cv.cache_options(vertex=vertex, overlap_training=False))
        rec["overlap_training"] = bool(co.overlap_training)
        rec["steady_ms_per_frame_steps_at_frame_end"] = round(throughput(lambda f, o: seq(trained + 50 + f, o)), 3)
        rec["steady_frame_relmse"] = round(single_frame_relmse(render, trained + 100), 5)
        rec["ratio_progressive"] = round(plain.ms / conv.ms, 3) if conv.reached and plain.reached else None
        if need is not None and rec["steady_frame_relmse"] <= target:
            rec["ratio_per_frame"] = round(steady_plain[need]["ms_per_frame"] / rec["steady_ms_per_frame"], 3)
        else:
            rec["ratio_per_frame"] = None
        # equal time: plain gets at least the time of one steady cache frame (the smallest number
        # of 1-spp frames whose time reaches it), and, as a second reading, its single-frame error
        # at exactly that time under the 1/spp law (error x time constant)
        ms1 = steady_plain[1]["ms_per_frame"]
        k = max(1, math.ceil(rec["steady_ms_per_frame"] / ms1))
        eq = cv.converge(ctx, lambda f, o: pt.render(cam, env, cv.frame_options(3, 7000 + f), out=o), ref, npix,
                         0.0, max_frames=k)
        rec["equal_time_plain_frames"] = k
        rec["equal_time_plain_ms"] = round(k * ms1, 3)
        rec["equal_time_plain_relmse"] = round(eq.relmse[-1], 5)
        rec["equal_time_plain_relmse_1_over_spp"] = round(steady_plain[1]["frame_relmse"] * ms1
                                                          / rec["steady_ms_per_frame"], 5)
        rec["beats_plain_at_equal_time"] = bool(rec["steady_frame_relmse"] < min(rec["equal_time_plain_relmse"],
                                                                             rec["equal_time_plain_relmse_1_over_spp"]))
        # training samples per frame (the exported batch of steady frames)
        ex = cpt.export_buffers(int(co.train_samples + co.train_samples // 8 + 256))
        counts = []
        for f in range(4):
            cpt.render(cam, env, cv.frame_options(3, trained + 400 + f), out=rgb, cache_opts=co, export=ex)
            counts.append(int(ex["count"].download()[0]))
        rec["train_samples_target"] = int(co.train_samples)
        rec["train_samples_per_frame"] = int(round(float(np.mean(counts))))
        for a in ex.values():
            a.free()
        st = cache.stats()
        rec["nonfinite"] = st["nonfinite"]
        out[f"cache_vertex{vertex}"] = rec
        log(f"cache vertex {vertex}: {rec}")
        cache.close()
    out["plain_spp_for_target_per_frame"] = need
    w = out.get(f"cache_vertex{int(c3['cache_vertex'])}", {})
    out["target_ratio"] = 3.0
    out["ratio_workload_vertex"] = w.get("ratio_progressive")
    out["target_met"] = bool(w.get("ratio_progressive") is not None and w["ratio_progressive"] >= 3.0)
    rgb.free()
    ref.free()
    out["mlp"] = mlp_microbench(mrt, ctx, verts)
    return out


def cache_frames(mrt, cfg, ctx, scene, verts, k: int) -> dict:
    """k trained 1080p frames of the cache workload at vertex 1, each frame synchronized, with its
    training steps at its own end (no second stream): the training kernels one after another."""
    from minotert_amd import convergence as cv
    from minotert_amd.pathtracer import CachedPathtracer
    cam = cfg.camera(3)
    rgb = ctx.empty((cam.width * cam.height, 3), np.float32)
    cache = cv.make_cache(ctx, cv.scene_bounds(verts))
    cpt = CachedPathtracer(ctx, scene, cache)
    co = cv.cache_options(vertex=1, overlap_training=False)
    ms = []
    for f in range(k):
        t0 = time.perf_counter()
        cpt.render(cam, cv.environment(), cv.frame_options(3, 100 + f), out=rgb, cache_opts=co)
        ctx.sync()
        ms.append((time.perf_counter() - t0) * 1e3)
    st = cache.stats()
    cache.close()
    rgb.free()
    return {"frames": k, "ms_per_frame_last_half": round(float(np.mean(ms[k // 2:])), 3), "steps": st["steps"],
            "nonfinite": st["nonfinite"]}


def ground_plane_row(mrt, cfg, ctx) -> dict:
    """Sensitivity of the cache-vs-plain protocol to the ground plane: the cache workload's scene
    with the 2-triangle plane of the small workload added (z = -box - 2, +-4 box, albedo 0.5), the
    same reference, frames and target. The headline rows keep the scene without it."""
    from minotert_amd import convergence as cv
    from minotert_amd.pathtracer import CachedPathtracer
    c3 = cfg.config(3)
    d = cfg.scene_desc(3)
    d.ground, g = 1, cfg.CONFIG["ground"]
    d.ground_z, d.ground_half, d.ground_albedo = -d.box - g["z_offset_below_box"], g["half_extent_over_box"] * d.box, g["albedo"]
    verts, mat, alb = mrt.generate_scene(d)
    scene = mrt.Scene(ctx, verts, mat, alb, builder="device")
    cam = cfg.camera(3)
    env = cv.environment()
    npix = cam.width * cam.height
    target = float(c3["target_relmse"])
    pt = mrt.Pathtracer(ctx, scene)
    ref = cv.render_reference(pt, cam, env, int(c3["reference_spp"]))
    plain = cv.converge(ctx, cv.plain_frames(pt, cam, env), ref, npix, target, max_frames=256)
    cache = cv.make_cache(ctx, cv.scene_bounds(verts))
    cpt = CachedPathtracer(ctx, scene, cache)
    conv = cv.converge(ctx, cv.cached_frames(cpt, cam, env, cv.cache_options(vertex=int(c3["cache_vertex"]))), ref,
                       npix, target, max_frames=400)
    out = {"scene": "cache workload scene + 2-triangle ground plane", "triangles": int(verts.shape[0]),
           "plain": {"frames": plain.frames, "ms": round(plain.ms, 2), "reached": plain.reached,
                     "relmse_first": round(plain.relmse[0], 5)},
           "cache_vertex1": {"frames": conv.frames, "ms": round(conv.ms, 2), "reached": conv.reached,
                             "relmse_first": round(conv.relmse[0], 5)},
           "ratio_progressive": round(plain.ms / conv.ms, 3) if conv.reached and plain.reached else None}
    cache.close()
    ref.free()
    scene.close()
    log(f"ground-plane sensitivity: {out}")
    return out


def scale_section(mrt, cfg, log_prefix="scale", frames: bool = True, cpu: bool = True, tune: str = "cache",
                  tune_cache: str | None = None) -> dict:
    """The scale workload: 3200 level-4 icospheres (16,384,000 triangles) in [-100, 100]^3 at
    3840x2160 with 8 bounces. BVH footprint and build time, incoherent closest-hit throughput on
    the bounce-1 rays of one frame (HBM traffic per launch from the committed rocprof summary),
    frame times of the wavefront, the megakernel and the cache-terminated tracer."""
    from minotert_amd import convergence as cv
    from minotert_amd.pathtracer import CachedPathtracer
    c4 = cfg.config(4)
    verts, mat, alb = mrt.generate_scene(cfg.scene_desc(4))
    ctx = mrt.Context(0)
    # device build (the first one loads the builder's code), a second warm one, and the host
    # build for its time and a byte comparison
    mrt.Scene(ctx, verts, mat, alb, builder="device").close()
    scene = mrt.Scene(ctx, verts, mat, alb, builder="device")
    st = scene.stats()
    host = mrt.Scene(ctx, verts, mat, alb, builder="host")
    hst = host.stats()
    dn, dt = scene.export()
    hn, ht = host.export()
    identical = bool(np.array_equal(dn, hn) and np.array_equal(dt, ht))
    host.close()
    del dn, dt, hn, ht
    log(f"{log_prefix}: {verts.shape[0]} triangles, BVH4 {st['n_nodes']} nodes depth {st['depth']}, "
        f"built on the device in {st['build_ms']:.1f} ms (host {hst['build_ms']:.0f} ms, identical {identical}), "
        f"{st['device_bytes'] / 2**30:.2f} GiB on the device")
    cam = cfg.camera(4)
    env = cv.environment()
    pt = mrt.Pathtracer(ctx, scene)
    bounces, rr = int(c4["bounces"]), int(c4["rr_start"])
    opts = mrt.RenderOptions(frame_seed=cfg.frame_seed(0), spp=1, bounces=bounces, rr_start=rr)
    gb = pt.primary_rays((cam.width, cam.height), cam, planes=False)
    d_rays, d_pid, n = pt.capture_rays(gb, cam, env, opts, bounce=1)
    rays = d_rays.download()[:n]
    pid = d_pid.download()[:n]
    d_rays.free()
    d_pid.free()
    rays = rays[np.argsort(pid, kind="stable")]
    dr = ctx.to_device(rays)
    dh = ctx.empty(n, mrt.HIT_DTYPE)
    from minotert_amd import autotune
    tuned = autotune.tune(ctx, scene, dr, dh, "scale4k_bounce1", force=tune == "force",
                          cache_path=tune_cache or autotune.DEFAULT_CACHE) if tune != "off" else None
    scene.trace_closest(dr, dh)
    ms, kernel_ms = timed_calls(ctx, lambda: scene.trace_closest(dr, dh), 20, "scale_trace")
    out = {"workload": "3200 level-4 icospheres, 3840x2160, 8 bounces, roulette from bounce 3",
           "triangles": int(verts.shape[0]), "bvh_nodes": st["n_nodes"], "bvh_depth": st["depth"],
           "bvh_build_ms_device": round(st["build_ms"], 1), "bvh_build_ms_host": round(hst["build_ms"], 1),
           "bvh_device_equals_host": identical,
           "device_bytes": {"bvh_nodes": st["n_nodes"] * 128, "leaf_triangles": int(verts.shape[0]) * 48,
                            "soup": int(verts.shape[0]) * 36, "total_scene": st["device_bytes"]},
           "bounce1_rays": int(n), "bounce1_ms_per_launch": round(ms, 4),
           "bounce1_mrays_per_s": round(n / (ms * 1e-3) / 1e6, 1), "kernel_ms": round(kernel_ms, 4),
           "trace_variant": tuned["variant"] if tuned else mrt.runtime.trace_variants()[ctx.trace_variant()]}
    dr.free()
    dh.free()
    if cpu:
        # the oracle on every 64th ray of the pixel-ordered set (all host threads)
        from oracle import oracle as orc
        sub = np.ascontiguousarray(rays[::64]).view(np.float32).reshape(-1, 8)
        t0 = time.perf_counter()
        bvh = orc.BVH(verts)
        build_s = time.perf_counter() - t0
        threads = orc.hardware_threads()
        _, nbox, ntri = bvh.trace_closest(sub, threads=threads, with_stats=True)
        t0 = time.perf_counter()
        reps = 0
        while reps < 1 or (time.perf_counter() - t0 < 5.0 and reps < 20):
            bvh.trace_closest(sub, threads=threads)
            reps += 1
        dt = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": round(sub.shape[0] * reps / dt / 1e6, 3), "unit": "Mrays/s", "cores": threads,
                               "kind": "port", "cpu": cpu_model(),
                               "sample": f"every 64th bounce-1 ray ({sub.shape[0]} rays) x{reps} through the "
                                         f"oracle's binary SAH BVH (build {build_s:.1f} s)"}
        b_model = 48.0 + 32.0 * nbox / sub.shape[0] + 40.0 * ntri / sub.shape[0]
        out["oracle_nodes_per_ray"] = round(nbox / sub.shape[0], 2)
        out["oracle_tris_per_ray"] = round(ntri / sub.shape[0], 2)
        log(f"{log_prefix}: CPU {out['cpu_baseline']}, nodes/ray {out['oracle_nodes_per_ray']} "
            f"tris/ray {out['oracle_tris_per_ray']}")
        del bvh
    else:
        b_model = None
    out["roofline"] = traversal_roofline(kernel_ms, int(n), "scale4k_bounce1", b_model)
    if not frames:
        log(f"{log_prefix}: {out}")
        scene.close()
        ctx.close()
        return out
    rgb = ctx.empty((cam.width * cam.height, 3), np.float32)

    def frame_ms(render, k=5):
        render(0)
        ctx.sync()
        t = time.perf_counter()
        for f in range(k):
            render(1 + f)
        ctx.sync()
        return round((time.perf_counter() - t) * 1e3 / k, 3)

    def fo(f, mega=False):
        return mrt.RenderOptions(frame_seed=cfg.frame_seed(f), spp=1, bounces=bounces, rr_start=rr, megakernel=mega)

    # The frames run with the launch shapes tuned on whole frames of this workload (their
    # traversals are mixed and camera launches, not the bounce-1 closest-hit launch above).
    if tune != "off":
        from minotert_amd import autotune
        ft = autotune.tune_frame(ctx, lambda: pt.render(cam, env, fo(0), out=rgb), "scale4k_frame",
                                 cache_path=tune_cache or autotune.DEFAULT_CACHE, force=tune == "force")
        out["frame_trace_variant"], out["frame_variant"] = ft["trace_variant"], ft["frame_variant"]
        out["frame_variant_source"] = ft["source"]
        log(f"{log_prefix}: frame shapes {ft['trace_variant']}, {ft['frame_variant']} ({ft['source']})")
    out["frame_ms_wavefront"] = frame_ms(lambda f: pt.render(cam, env, fo(f), out=rgb))
    out["frame_ms_megakernel"] = frame_ms(lambda f: pt.render(cam, env, fo(f, True), out=rgb))
    cache = cv.make_cache(ctx, cv.scene_bounds(verts))
    cpt = CachedPathtracer(ctx, scene, cache)
    co = cv.cache_options(vertex=1)
    for f in range(20):  # train before timing
        cpt.render(cam, env, fo(100 + f), out=rgb, cache_opts=co)
    out["frame_ms_cached_vertex1"] = frame_ms(lambda f: cpt.render(cam, env, fo(200 + f), out=rgb, cache_opts=co))
    log(f"{log_prefix}: {out}")
    cache.close()
    rgb.free()
    scene.close()
    ctx.close()
    return out


def mlp_microbench(mrt, ctx, verts, n: int = 2073600, steps: int = 10) -> dict:
    """Cache inference kernels on one 1080p frame's worth of queries (random surface samples
    inside the scene bounds): event-timed encoding (gather-bound) and BF16 MFMA forward, with the
    forward's rate against the dense BF16 MFMA peak (43,008 FLOP per query: 41 inputs padded to
    64, five 64x64 hidden layers, RGB padded to 16); and the training MLP kernels on one step's
    16,384 samples (~3 x 43,008 FLOP per sample)."""
    from minotert_amd import convergence as cv
    from minotert_amd.cache import SAMPLE_DTYPE
    lo, hi = cv.scene_bounds(verts)
    rng = np.random.default_rng(1)
    smp = np.zeros(n, SAMPLE_DTYPE)
    smp["pos"] = rng.uniform(lo, hi, (n, 3))
    d = rng.normal(size=(n, 3))
    smp["dir"] = d / np.linalg.norm(d, axis=1)[:, None]
    smp["nrm"] = -smp["dir"]
    smp["alb"] = rng.uniform(0.2, 0.9, (n, 3))
    cache = cv.make_cache(ctx, (lo, hi))
    d_s = ctx.to_device(smp)
    d_y = ctx.empty((n, 4), np.float32)
    cache.query(d_s, d_y)
    ctx.sync()
    ctx.profile(True)
    ctx.profile_reset()
    for _ in range(steps):
        cache.query(d_s, d_y)
    enc_ms, el = ctx.profile_query("cache_encode")
    fwd_ms, fl = ctx.profile_query("cache_mlp_forward")
    ctx.profile(False)
    enc_ms /= max(el, 1)
    fwd_ms /= max(fl, 1)
    flop = 43008.0 * n
    tflops = flop / (fwd_ms * 1e-3) / 1e12
    # encoding: 16 levels x 8 corners x 8 B of table reads + 64 B sample + 128 B BF16 inputs per query
    enc_bytes = n * (16 * 8 * 8 + 64 + 128)
    res = {"queries": n, "encode_ms": round(enc_ms, 4), "encode_GBps_algorithmic": round(enc_bytes / (enc_ms * 1e-3) / 1e9, 1),
           "forward_ms": round(fwd_ms, 4), "forward_TFLOPs": round(tflops, 1),
           "forward_frac_bf16_peak": round(tflops / 2500.0, 4)}
    # one training step of the workload's size (65,536 samples / 4 steps): the MLP's forward, data
    # and weight gradients (k_mlp_train + k_bwd_weight) at ~3 x 43,008 FLOP per sample
    m = 16384
    tgt = np.zeros((m, 4), np.float32)
    tgt[:, :3] = rng.uniform(0.0, 2.0, (m, 3))
    d_ts, d_tt = ctx.to_device(smp[:m].copy()), ctx.to_device(tgt)
    cache.train(d_ts, d_tt)
    ctx.sync()
    ctx.profile(True)
    ctx.profile_reset()
    for _ in range(steps):
        cache.train(d_ts, d_tt)
    tr_ms, tl = ctx.profile_query("cache_mlp_train")
    ctx.profile(False)
    tr_ms /= max(tl, 1)
    tr_tflops = 3.0 * 43008.0 * m / (tr_ms * 1e-3) / 1e12
    res.update({"train_samples_per_step": m, "train_mlp_ms": round(tr_ms, 4), "train_mlp_TFLOPs": round(tr_tflops, 1),
                "train_mlp_frac_bf16_peak": round(tr_tflops / 2500.0, 4)})
    log(f"cache inference and training kernels: {res}")
    cache.close()
    d_s.free()
    d_y.free()
    d_ts.free()
    d_tt.free()
    return res


def extra_trace_sections(mrt, cfg, ctx, scene, d_rays, d_hits, n: int, steps: int):
    """The same batch traversed in coherence order (octant + origin Morton radix sort, sort
    included), and the secondary incoherent set."""
    from minotert_amd import runtime as rtm
    d_order = ctx.empty(n, np.uint32)
    rtm.ray_order(scene, d_rays, d_order)
    rtm.trace_closest_ordered(scene, d_rays, d_order, d_hits)
    ctx.sync()
    ctx.profile(True)
    ctx.profile_reset()
    for _ in range(steps):
        rtm.ray_order(scene, d_rays, d_order)
        rtm.trace_closest_ordered(scene, d_rays, d_order, d_hits)
    sort_ms, sl = ctx.profile_query("ray_sort")
    otr_ms, ol = ctx.profile_query("trace_closest")
    ctx.profile(False)
    sort_ms /= max(sl, 1)
    otr_ms /= max(ol, 1)
    sorted_rec = {"sort_ms": round(sort_ms, 4), "trace_ms": round(otr_ms, 4),
                  "mrays_per_s_incl_sort": round(n / ((sort_ms + otr_ms) * 1e-3) / 1e6, 1)}
    log(f"sorted order: sort {sort_ms:.4f} ms + trace {otr_ms:.4f} ms -> {sorted_rec['mrays_per_s_incl_sort']} Mrays/s")
    secondary = secondary_set(mrt, cfg, ctx, scene, steps)
    return sorted_rec, secondary


def secondary_set(mrt, cfg, ctx, scene, steps: int) -> dict:
    """The secondary incoherent set: 2^21 rays with origins uniform in the scene box
    [-40, 40]^3 and directions uniform on the sphere (seeded), traced closest-hit; kernel time
    from HIP events."""
    n = 1 << 21
    box = float(cfg.config(1)["box"])
    rng = np.random.default_rng(int(cfg.CONFIG["scene_seed_base"]) + 1)
    rays = np.zeros(n, mrt.RAY_DTYPE)
    rays["o"] = rng.uniform(-box, box, (n, 3))
    d = rng.normal(size=(n, 3))
    rays["d"] = d / np.linalg.norm(d, axis=1)[:, None]
    rays["tmin"] = 0.0
    rays["tmax"] = np.inf
    d_rays = ctx.to_device(rays)
    d_hits = ctx.empty(n, mrt.HIT_DTYPE)
    for _ in range(3):
        scene.trace_closest(d_rays, d_hits)
    _, kernel_ms = timed_calls(ctx, lambda: scene.trace_closest(d_rays, d_hits), steps, "secondary_trace")
    hit = float(np.mean(d_hits.download()["tri"] != mrt.MISS))
    res = {"rays": n, "kernel_ms": round(kernel_ms, 4), "mrays_per_s": round(n / (kernel_ms * 1e-3) / 1e6, 1),
           "hit_fraction": round(hit, 4), "origins": f"uniform in [-{box:g}, {box:g}]^3", "directions": "uniform on S2"}
    log(f"secondary set: {res}")
    d_rays.free()
    d_hits.free()
    return res


def batch_section(mrt, cfg, ctx, steps: int = 10, tune: str = "cache", tune_cache: str | None = None) -> dict:
    """The batched workload: the 16 scenes of the batch layout (level-3 icospheres, triangle
    counts log-uniform in [1e4, 2e6], resolutions from 256^2 to 1920x1080, 2 bounces). One frame
    of all 16 views rendered (a) scene by scene, (b) as one batch (concatenated BVHs, one
    traversal launch per bounce over a shared queue), (c) by replaying the batch frame captured
    as a HIP graph; wall time per frame, synchronized."""
    from minotert_amd.pathtracer import Batch
    layout = cfg.batch_layout(2)
    c2 = cfg.config(2)
    scenes, cams, outs = [], [], []
    for e in layout:
        v, m, a = mrt.generate_scene(e.scene)
        scenes.append(mrt.Scene(ctx, v, m, a, builder="device"))
        cams.append(cfg.camera(1, width=e.width, height=e.height, box=e.scene.box))
        outs.append(ctx.empty((e.width * e.height, 3), np.float32))
    env = mrt.Environment(sky=tuple(cfg.CONFIG["sky"]), sun_dir=tuple(cfg.CONFIG["sun_dir"]),
                          sun_e=tuple(cfg.CONFIG["sun_e"]))

    def opts(f):
        return mrt.RenderOptions(frame_seed=cfg.frame_seed(f), spp=1, bounces=int(c2["bounces"]), rr_start=int(c2["rr_start"]))

    def timed(fn):
        fn(0)
        ctx.sync()
        t = time.perf_counter()
        for f in range(steps):
            fn(1 + f)
        ctx.sync()
        return (time.perf_counter() - t) * 1e3 / steps

    pts = [mrt.Pathtracer(ctx, sc) for sc in scenes]
    b = Batch(ctx, scenes)
    views = list(zip(cams, outs))
    # the batch's traversal, shading-block and megakernel shapes from the autotuner (cached
    # choice, measured on whole batch frames when absent)
    from minotert_amd import autotune
    before = (ctx.trace_variant(False), ctx.trace_variant(True), ctx.frame_variant())
    tuned = {"trace_variant": mrt.runtime.trace_variants()[before[0]],
             "frame_variant": mrt.runtime.frame_variants()[before[2]], "source": "default"}
    if tune != "off":
        tuned = autotune.tune_frame(ctx, lambda: b.render(views, env, opts(0)), "batched16", force=tune == "force",
                                    cache_path=tune_cache or autotune.DEFAULT_CACHE)
        log(f"batch: traversal {tuned['trace_variant']}, frame {tuned['frame_variant']} ({tuned['source']})")
    separate = timed(lambda f: [pt.render(cam, env, opts(f), out=o) for pt, cam, o in zip(pts, cams, outs)])
    batched = timed(lambda f: b.render(views, env, opts(f)))
    g = b.graph(views, env, opts(0))
    g.set_frame(0)  # each launch renders the next frame index
    graphed = timed(lambda f: g.launch())
    info = b.info()
    pixels = sum(c.width * c.height for c in cams)
    res = {"scenes": len(scenes), "triangles": int(info["triangles"]), "pixels": int(pixels),
           "bounces": int(c2["bounces"]), "ms_separate": round(separate, 3), "ms_batch": round(batched, 3),
           "ms_graph": round(graphed, 3), "mpaths_per_s_graph": round(pixels / (graphed * 1e-3) / 1e6, 1),
           "trace_variant": tuned["trace_variant"], "frame_variant": tuned["frame_variant"],
           "variant_source": tuned["source"]}
    # the context's previous shapes for the sections after this one
    ctx.set_trace_variant(before[0], False)
    ctx.set_trace_variant(before[1], True)
    ctx.set_frame_variant(before[2])
    log(f"batch: {res}")
    g.close()
    b.close()
    for o in outs:
        o.free()
    for sc in scenes:
        sc.close()
    return res


def lambert_section(mrt, cfg, ctx, steps: int = 20) -> dict:
    """The Lambert workload: 20,482 triangles at 256x256, primary rays (G-buffer) then Lambert
    shading with one shadow ray per hit pixel; wall time per frame and the two traversal
    launches' kernel time."""
    verts, mat, alb = mrt.generate_scene(cfg.scene_desc(0))
    scene = mrt.Scene(ctx, verts, mat, alb, builder="device")
    cam = cfg.camera(0)
    pt = mrt.Pathtracer(ctx, scene)
    light = tuple(cfg.CONFIG["sun_dir"]), tuple(cfg.CONFIG["lambert_light_e"])

    def frame():
        gb = pt.primary_rays((cam.width, cam.height), cam, planes=False)
        rgb = pt.shade_lambert(gb, cam, *light)
        return gb, rgb
    for _ in range(3):
        frame()
    ctx.sync()
    t = time.perf_counter()
    for _ in range(steps):
        frame()
    ctx.sync()
    ms = (time.perf_counter() - t) * 1e3 / steps
    ctx.profile(True)
    ctx.profile_reset()
    for _ in range(steps):
        frame()
    c_ms, cl = ctx.profile_query("trace_closest")
    a_ms, al = ctx.profile_query("trace_any")
    ctx.profile(False)
    res = {"triangles": int(verts.shape[0]), "pixels": cam.width * cam.height, "ms_per_frame": round(ms, 4),
           "primary_kernel_ms": round(c_ms / max(cl, 1), 4), "shadow_kernel_ms": round(a_ms / max(al, 1), 4)}
    log(f"lambert: {res}")
    scene.close()
    return res


def sky_section(mrt, ctx, steps: int = 10) -> dict:
    """The physical sky: the transmittance (256x64) and multiple-scattering (32x32) LUTs, the
    sky-view LUT (192x108) and the miss radiance of 2,073,600 random directions; kernel times
    from HIP events."""
    from minotert_amd.sky import PhysicalSky
    ctx.profile(True)
    ctx.profile_reset()
    skies = [PhysicalSky(ctx) for _ in range(steps)]
    luts_ms, ll = ctx.profile_query("sky_luts")
    for sk in skies:
        sk.update_view()
    view_ms, vl = ctx.profile_query("sky_view")
    ctx.profile(False)
    n = 2073600
    rng = np.random.default_rng(3)
    d = rng.normal(size=(n, 3))
    d = (d / np.linalg.norm(d, axis=1)[:, None]).astype(np.float32)
    d_dirs = ctx.to_device(d)
    d_rgb = ctx.empty((n, 3), np.float32)
    sk = skies[0]
    sk.radiance(d_dirs, d_rgb)
    ctx.sync()
    t = time.perf_counter()
    for _ in range(steps):
        sk.radiance(d_dirs, d_rgb)
    ctx.sync()
    rad_ms = (time.perf_counter() - t) * 1e3 / steps
    res = {"transmittance_and_multiscattering_ms": round(luts_ms / max(ll, 1), 4),
           "sky_view_ms": round(view_ms / max(vl, 1), 4), "radiance_directions": n,
           "radiance_ms": round(rad_ms, 4)}
    log(f"sky: {res}")
    for s_ in skies:
        s_.close()
    d_dirs.free()
    d_rgb.free()
    return res


def post_section(mrt, cfg, ctx, scene, steps: int = 10, images: str | None = None) -> dict:
    """Post-processing of one 1920x1080 frame of the 1,024,000-triangle scene (1-spp colour and
    its G-buffer): the bilateral denoiser with the reference's defaults (sigma 5, kSigma 2:
    radius 10, 325 taps per pixel with a primary hit) and the default AMD tonemapper writing
    float sRGB and RGBA8 (28 B of HBM traffic per pixel). Event-timed on the library stream."""
    from minotert_amd import post
    cam = cfg.camera(1)
    w, h = cam.width, cam.height
    pt = mrt.Pathtracer(ctx, scene)
    env = mrt.Environment(sky=tuple(cfg.CONFIG["sky"]), sun_dir=tuple(cfg.CONFIG["sun_dir"]),
                          sun_e=tuple(cfg.CONFIG["sun_e"]))
    gb = pt.primary_rays((w, h), cam)
    rgb = pt.secondary_rays(gb, cam, env, mrt.RenderOptions(frame_seed=1))
    den = ctx.empty((w * h, 3), np.float32)
    srgb = ctx.empty((w * h, 3), np.float32)
    q = ctx.empty((w * h, 4), np.uint8)
    dn, tm = post.Denoiser(ctx), post.Tonemapper(ctx)
    tp = post.TonemapParams()
    km = 1e-3  # scene units read as metres (the filter's depth term is in kilometres)
    dn.bilateral(rgb, gb.depth, gb.normal, w, h, cam.near_plane, den, km_per_unit=km)
    tm.apply(den, tp, srgb, q)
    ctx.sync()
    ctx.profile(True)
    ctx.profile_reset()
    for f in range(steps):
        dn.bilateral(rgb, gb.depth, gb.normal, w, h, cam.near_plane, den, frame=f, km_per_unit=km)
        tm.apply(den, tp, srgb, q)
    d_ms, dl = ctx.profile_query("denoise_bilateral")
    t_ms, tl = ctx.profile_query("tonemap")
    ctx.profile(False)
    d_ms /= max(dl, 1)
    t_ms /= max(tl, 1)
    hit_px = int(np.sum(gb.depth.download() > 0.0))
    taps = sum(int(np.floor(2.0 * np.sqrt(100.0 - dx * dx))) + 1 for dx in range(-10, 11))  # dy from -pt by 1 to pt
    res = {"width": w, "height": h, "hit_pixels": hit_px, "denoise_ms": round(d_ms, 4),
           "denoise_gtaps_per_s": round(hit_px * taps / (d_ms * 1e-3) / 1e9, 1), "taps_per_pixel": taps,
           "tonemap_ms": round(t_ms, 4), "tonemap_GBps": round(28.0 * w * h / (t_ms * 1e-3) / 1e9, 1),
           "tonemap_frac_hbm": round(28.0 * w * h / (t_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
    log(f"post-processing: {res}")
    if images:
        os.makedirs(images, exist_ok=True)
        for name, img in (("incoherent_1080p_1spp.png", rgb), ("incoherent_1080p_1spp_denoised.png", den)):
            tm.apply(img, tp, None, q)
            post.write_png(os.path.join(images, name), q.download(), w, h)
        log(f"images written to {images}")
    for a in (den, srgb, q, rgb, gb.hits, gb.depth, gb.normal, gb.motion):
        a.free()
    return res


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--samples", type=int, default=4, help="samples per pixel whose bounce-1 rays form the batch")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baseline")
    ap.add_argument("--no-cache", action="store_true", help="skip the radiance-cache vs plain protocol")
    ap.add_argument("--no-scale", action="store_true", help="skip the 16.4M-triangle 4K workload")
    ap.add_argument("--no-post", action="store_true",
                    help="skip the denoiser / tonemapper, Lambert, sky and batched-scene timings")
    ap.add_argument("--metric-only", action="store_true",
                    help="only the headline workload (no sorted order, secondary set, cache, post or scale "
                         "sections), so that a profiler's per-kernel averages cover its launches alone")
    ap.add_argument("--images", default=None, metavar="DIR",
                    help="write the post-processing frames (AMD tonemap, sRGB) as PNG files into DIR")
    ap.add_argument("--scale-only", action="store_true",
                    help="run only the 16.4M-triangle 4K workload (for profiling its traversal)")
    ap.add_argument("--cache-frames-only", type=int, default=0, metavar="K",
                    help="only K trained cache frames of the radiance-cache workload (vertex 1, each frame's "
                         "training steps at its own end), for profiling the training kernels")
    ap.add_argument("--tune", choices=["cache", "force", "off"], default="cache",
                    help="traversal launch shape: cached autotuner choice (measured if absent), re-measure, or default")
    ap.add_argument("--tune-cache", default=None, help="autotuner cache file (default minotert_amd/tuning.json)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="also run the plain copy kernel over MIB MiB three times (reference point for "
                         "rocprof FETCH_SIZE/WRITE_SIZE) and report its bandwidth")
    args = ap.parse_args()
    if args.metric_only:
        args.no_cache = args.no_post = args.no_scale = True
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); the path shards by pixel tiles "
              f"with a replicated BVH, see the design notes", file=sys.stderr)
        return 2

    import minotert_amd as mrt
    from minotert_amd import configs as cfg

    if args.scale_only:
        print(json.dumps({"scale": scale_section(mrt, cfg, frames=False, cpu=not args.no_cpu, tune=args.tune,
                                                 tune_cache=args.tune_cache)}), flush=True)
        return 0
    log("generating the 1,024,000-triangle scene")
    verts, mat, alb = mrt.generate_scene(cfg.scene_desc(1))
    ctx = mrt.Context(0)
    if args.cache_frames_only:
        scene = mrt.Scene(ctx, verts, mat, alb, builder="device")
        print(json.dumps({"cache_frames": cache_frames(mrt, cfg, ctx, scene, verts, args.cache_frames_only)}),
              flush=True)
        return 0
    t0 = time.perf_counter()
    scene = mrt.Scene(ctx, verts, mat, alb, builder="device")
    st = scene.stats()
    log(f"BVH4 built on the device: {st['n_nodes']} nodes, depth {st['depth']}, {st['build_ms']:.1f} ms "
        f"(+upload {1e3 * (time.perf_counter() - t0) - st['build_ms']:.0f} ms)")
    rays = make_batch(mrt, cfg, ctx, scene, args.samples)
    n = rays.shape[0]
    log(f"batch: {n} bounce-1 rays")
    d_rays = ctx.to_device(rays)
    d_hits = ctx.empty(n, mrt.HIT_DTYPE)
    tuned = {"variant": mrt.runtime.trace_variants()[ctx.trace_variant()], "source": "default"}
    if args.tune != "off":
        from minotert_amd import autotune
        tuned = autotune.tune(ctx, scene, d_rays, d_hits, "incoherent_bounce1", force=args.tune == "force",
                              cache_path=args.tune_cache or autotune.DEFAULT_CACHE)
        log(f"traversal variant {tuned['variant']} ({tuned['source']})")
        if tuned["source"] == "measured":
            for k, v in sorted(tuned["ms"].items(), key=lambda kv: kv[1]):
                log(f"  {k:22s} {v:.4f} ms  {n / v / 1e3:.0f} Mrays/s")

    for _ in range(args.warmup):
        scene.trace_closest(d_rays, d_hits)
    # the timed steps, bracketed by host synchronizations (wall clock: the metric) and by one
    # pair of HIP events on the library stream (device time per step: the roofline's duration)
    ms_per_step, kernel_ms = timed_calls(ctx, lambda: scene.trace_closest(d_rays, d_hits), args.steps,
                                         "timed_steps")
    mrays = n / (ms_per_step * 1e-3) / 1e6
    log(f"timed: {args.steps} steps, {ms_per_step:.4f} ms/step ({kernel_ms:.4f} ms on the device), "
        f"{mrays:.1f} Mrays/s")
    hits = d_hits.download()
    hit_frac = float(np.mean(hits["tri"] != mrt.MISS))

    sorted_rec = secondary = None
    if not args.metric_only:
        sorted_rec, secondary = extra_trace_sections(mrt, cfg, ctx, scene, d_rays, d_hits, n, args.steps)

    copy = None
    if args.calibrate_copy > 0:
        nbytes = args.calibrate_copy << 20
        src = ctx.empty(nbytes, np.uint8)
        dst = ctx.empty(nbytes, np.uint8)
        ctx.copy_device(dst, src)
        ctx.sync()
        ctx.profile(True)
        ctx.profile_reset()
        for _ in range(3):
            ctx.copy_device(dst, src)
        cms, cl = ctx.profile_query("copy_bytes")
        ctx.profile(False)
        copy_ms = cms / max(cl, 1)
        copy = {"bytes_read": nbytes, "bytes_written": nbytes, "kernel_ms": round(copy_ms, 4),
                "GBps": round(2 * nbytes / (copy_ms * 1e-3) / 1e9, 1)}
        log(f"copy kernel: {copy}")
        src.free()
        dst.free()

    cpu = None
    b_ray = None
    if not args.no_cpu:
        log("CPU oracle baseline")
        cpu, n_node, n_tri = cpu_baseline(rays, verts)
        b_ray = 48.0 + 32.0 * n_node + 40.0 * n_tri
        log(f"CPU {cpu['value']} Mrays/s on {cpu['cores']} threads; nodes/ray {n_node:.2f} tris/ray {n_tri:.2f}")
    roof = traversal_roofline(kernel_ms, int(n), "incoherent_bounce1", b_ray)

    out = {
        "metric": METRIC, "value": round(mrays, 2), "unit": "Mrays/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms_per_step, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "incoherent closest-hit, bounce-1 extension rays of the 1,024,000-triangle "
                               "icosphere scene at 1920x1080", "triangles": int(verts.shape[0]),
                   "rays_per_step": int(n), "samples_per_pixel": args.samples, "hit_fraction": round(hit_frac, 4),
                   "bvh": "BVH4, 128-B nodes, binned SAH 16 bins, leaf<=4", "bvh_nodes": st["n_nodes"],
                   "bvh_build_ms": round(st["build_ms"], 1), "bvh_builder": "device", "trace_variant": tuned["variant"],
                   "variant_source": tuned["source"]},
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    if sorted_rec is not None:
        out["sorted_order"] = sorted_rec
        out["secondary_set"] = secondary
    if not args.no_cache:
        out["cache"] = cache_section(mrt, cfg, ctx, scene, verts)
        out["cache"]["ground_plane_sensitivity"] = ground_plane_row(mrt, cfg, ctx)
    if not args.no_post:
        out["post"] = post_section(mrt, cfg, ctx, scene, images=args.images)
        out["lambert"] = lambert_section(mrt, cfg, ctx)
        out["sky"] = sky_section(mrt, ctx)
        out["batch"] = batch_section(mrt, cfg, ctx, tune=args.tune, tune_cache=args.tune_cache)
    if copy is not None:
        out["copy_calibration"] = copy
    scene.close()
    ctx.close()
    if not args.no_scale:
        out["scale"] = scale_section(mrt, cfg, cpu=not args.no_cpu, tune=args.tune, tune_cache=args.tune_cache)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
