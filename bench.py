#!/usr/bin/env python3
"""Throughput of the gfx950 PatchMatch engine on one MI355X.

Workload (one step): a 1920x1080 reference view with 10 source views of a
seeded heightfield scene (64x64 mesh, Gaussian bumps, depth ~3..30), window
15x15 step 1: the photometric pass (8 iterations, random init) followed by
the geometric-consistency pass (2 iterations) whose initial state is the
photometric result and whose source depth maps are the sources' own
photometric results (computed once before timing; everything is resident in
HBM when the timed region starts, and the photometric -> geometric hand-off
stays on the device).

Metric: Mpixel.iter/s = W x H x (8 + 2) per step / time; also views/s, the
median relative depth error against the rendered ground truth, and p50/p99
step latency.

roofline: the checkerboard sweep kernel, algorithmic FP32 FLOPs per launch
(23 per bilateral-NCC window sample + 75 per (hypothesis, source) evaluation
+ 61 per geometric term, counted in a calibration run) over its average
launch duration measured with hipEvents on the engine's stream during the
timed steps.  cpu_baseline: the C/OpenMP transcription on a band of rows of
the same view, photometric pass.

after_path: the stages that follow the path on the same 11 views (cross
filter over the device-resident photometric problems, deduplication, depth
map fusion), each timed and compared bit for bit with its transcription.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP32_PEAK_TFLOPS = 157.3        # MI355X published FP32 vector peak
FLOP_PER_SAMPLE = 23
FLOP_PER_EVAL = 75
FLOP_PER_GEO = 61
METRIC = "Mpixel·iter/s + views/s at fixed depth error; p50/p99 view latency"


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--rcp", choices=["exact", "fast"], default="exact")
    ap.add_argument("--storage", choices=["u8", "f16"], default="f16",
                    help="source image storage in HBM (f16 holds the 8-bit values exactly: same maps)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--skip-forms", action="store_true",
                    help="view workload: leave out the batched and served forms' summaries")
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="target CPU baseline duration")
    ap.add_argument("--only-photometric", action="store_true",
                    help="profiling mode: the step is the photometric pass alone, no source passes")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="profiling mode: one device-to-device copy of MIB MiB, then exit")
    ap.add_argument("--workload", choices=["view", "batch", "serve", "select", "filters"], default="view",
                    help="view: the 1920x1080 view (headline); batch: 64 orbit views, plain vs batched; "
                         "serve: Poisson arrivals into the serving loop; select: source-view selection "
                         "on a 500-camera sparse model; filters: RunFilter, cross filter, normal "
                         "encoding, the steps after the geometric pass, deduplication and depth-map fusion "
                         "after the path")
    ap.add_argument("--views", type=int, default=64, help="batch workload: reference views")
    ap.add_argument("--jobs", type=int, default=24, help="serve workload: jobs per arrival rate")
    ap.add_argument("--rates", default="5,20,50,200", help="serve workload: arrival rates, requests/s")
    ap.add_argument("--budget-gb", type=float, default=200.0, help="serve workload: HBM admission budget")
    ap.add_argument("--max-active", type=int, default=4, help="serve workload: jobs on the device at once")
    return ap.parse_args()


def copy_calibration(mib: int) -> int:
    """One hipMemcpy device-to-device of a known size: the reference point
    that scales FETCH_SIZE / WRITE_SIZE counter readings to bytes."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    n = mib << 20
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), C.c_size_t(n)) == 0 and hip.hipMalloc(C.byref(b), C.c_size_t(n)) == 0
    assert hip.hipMemset(a, 1, C.c_size_t(n)) == 0
    assert hip.hipDeviceSynchronize() == 0
    assert hip.hipMemcpy(b, a, C.c_size_t(n), 3) == 0
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    print(json.dumps({"copy_calibration_bytes": n}))
    return 0


def c2_options(pmm, geom: bool, rcp: int, storage: int = 1):  # 1: PM_STORAGE_F16, the engine's default
    return pmm.PatchMatchOptions(window_radius=7, window_step=1, sigma_spatial=3.0, sigma_color=0.2,
                                 num_iterations=2 if geom else 8, geom_consistency=geom,
                                 init_depth_random=not geom, thk=2, seed=2, rcp_mode=rcp, image_storage=storage)


PMC_SUMMARY = os.path.join(ROOT, "profiles", "k_sweep_1920x1080_pmc.md")


def pmc_traffic():
    """HBM bytes per photometric k_sweep launch of this workload, from the
    committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary (scaled by the copy
    calibration; the file gives the commands), or None."""
    try:
        with open(PMC_SUMMARY) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch:"):
                    return int(float(line.split(":", 1)[1]))
    except OSError:
        return None
    return None


def sweep_flops(stats) -> float:
    return (stats.sweep_ncc_samples * FLOP_PER_SAMPLE + stats.sweep_ncc_evaluations * FLOP_PER_EVAL
            + stats.sweep_geo_evaluations * FLOP_PER_GEO)


def cpu_baseline(pmm, scene_mod, sc, target_s: float) -> dict:
    from oracle import binding
    prob = scene_mod.to_problem(sc)
    opts = c2_options(pmm, False, 0)
    view = pmm.ViewDesc(prob, opts)
    threads, cores = cpu_threads()
    W, H = view.width, view.height
    mid = H // 2
    # size the band for ~target_s: the cost of one iteration over 2 rows is
    # the difference of a 2-iteration and a 1-iteration probe (same init)
    probe = c2_options(pmm, False, 0)
    probe_t = []
    for it in (1, 2):
        probe.num_iterations = it
        t = time.perf_counter()
        binding.run_view(probe, view, mid, mid + 2, threads)
        probe_t.append(time.perf_counter() - t)
    per_row_iter = max(probe_t[1] - probe_t[0], 1e-4) / 2.0
    rows = int(max(2, min(64, target_s / (per_row_iter * opts.num_iterations))))
    r0 = max(0, mid - rows // 2)
    r1 = min(H, r0 + rows)
    t = time.perf_counter()
    binding.run_view(opts, view, r0, r1, threads)
    dt = time.perf_counter() - t
    mpix_iter = (r1 - r0) * W * opts.num_iterations / 1e6
    # one thread on a 2-row band, 1 iteration (scaling reference)
    one = c2_options(pmm, False, 0)
    one.num_iterations = 1
    t = time.perf_counter()
    binding.run_view(one, view, mid, mid + 2, 1)
    d1 = time.perf_counter() - t
    t = time.perf_counter()
    binding.run_view(one, view, mid, mid + 2, threads)
    dn = time.perf_counter() - t
    return {"value": round(mpix_iter / dt, 4), "unit": "Mpixel·iter/s", "cores": threads, "kind": "port",
            "host_cpus": cores,
            "sample": f"C/OpenMP transcription, photometric pass of the 1920x1080 view, rows [{r0},{r1}) "
                      f"x {W} px, {opts.num_iterations} iterations, 10 sources, 15x15 step 1; {dt:.1f} s",
            "cpu_model": cpu_model(),
            "single_thread": {"value": round(2 * W / 1e6 / d1, 5), "unit": "Mpixel·iter/s",
                              "same_sample_all_threads": round(2 * W / 1e6 / dn, 5),
                              "sample": f"rows [{mid},{mid + 2}) x {W} px, 1 iteration (with the init); "
                                        f"{d1:.1f} s on 1 thread"}}


def cpu_threads():
    """(threads the CPU baseline uses, what the host shows).  The threads are
    the CPUs this process may run on (its affinity mask), capped by
    OMP_NUM_THREADS where set: a GPU's share of a shared host is given that
    way, and the rest of the machine belongs to other GPUs' jobs."""
    try:
        allowed = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        allowed = os.cpu_count() or 1
    env = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    threads = min(allowed, env) if env > 0 else allowed
    phys = set()
    try:
        with open("/proc/cpuinfo") as f:
            pid = cid = None
            for line in f:
                if line.startswith("physical id"):
                    pid = line.split(":", 1)[1].strip()
                elif line.startswith("core id"):
                    cid = line.split(":", 1)[1].strip()
                elif not line.strip():
                    if pid is not None and cid is not None:
                        phys.add((pid, cid))
                    pid = cid = None
    except OSError:
        pass
    return threads, {"affinity": allowed, "logical": os.cpu_count(), "physical_cores": len(phys) or None,
                     "omp_num_threads": env or None}


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def progress(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def rel_depth_errors(depth: np.ndarray, gt: np.ndarray) -> np.ndarray:
    m = gt > 0
    return np.abs(depth[m] - gt[m]) / gt[m]


def run_batch_workload(args) -> dict:
    """64 reference views of a 200-camera orbit (sizes 640x480..3000x2000,
    5..20 sources, window 11x11 step 2 or 15x15 step 1, 6 iterations), run
    (a) plain: one view at a time, one launch per phase of the view,
    default launch shapes; (b) batched: all views advanced together in
    shared persistent launches replayed from a hipGraph, launch shapes
    autotuned per class.  Both use f16 image storage.  Inputs are resident on the
    device before either timed region."""
    from sensemap_amd import abi, scene
    from sensemap_amd import patch_match as pmm

    t0 = time.perf_counter()
    orbit = scene.orbit_scene(seed=3)
    render_s = time.perf_counter() - t0
    progress(f"orbit rendered in {render_s:.1f} s")
    plain_items = scene.batch_problems(orbit, num_views=args.views, seed=3, image_storage=abi.PM_STORAGE_F16)
    batch_items = scene.batch_problems(orbit, num_views=args.views, seed=3, image_storage=abi.PM_STORAGE_F16)
    gts = [orbit.views[p.ref_image_idx].depth for _, p in plain_items]
    pix_iter = sum(g.size * o.num_iterations for g, (o, _) in zip(gts, plain_items))

    eng_plain = pmm.default_engine()
    plain = [pmm.PatchMatchACMMCuda(o, p, eng_plain) for o, p in plain_items]
    plain[0].Run()  # warm-up
    lat = []
    t = time.perf_counter()
    for v in plain:
        t1 = time.perf_counter()
        v.Run()
        lat.append((time.perf_counter() - t1) * 1000.0)
    plain_s = time.perf_counter() - t
    progress(f"plain path: {len(plain)} views in {plain_s:.2f} s")
    plain_depth = [v.GetDepthMap() for v in plain]
    for v in plain:
        v.close()
    del plain

    eng_batch = pmm.Engine(0)
    b = pmm.PatchMatchBatch(batch_items, engine=eng_batch, graph=True, autotune=True)
    t = time.perf_counter()
    b.Run()  # warm-up: autotune + graph capture + first replay
    warm_s = time.perf_counter() - t
    progress(f"batched warm-up (autotune + capture + replay) {warm_s:.2f} s")
    steps = max(1, args.steps)
    t = time.perf_counter()
    for _ in range(steps):
        b.Run()
    batch_s = (time.perf_counter() - t) / steps
    batch_depth = [v.GetDepthMap() for v in b.views]
    identical = all(np.array_equal(a, c) for a, c in zip(plain_depth, batch_depth))
    e_plain = np.concatenate([rel_depth_errors(d, g) for d, g in zip(plain_depth, gts)])
    e_batch = np.concatenate([rel_depth_errors(d, g) for d, g in zip(batch_depth, gts)])
    shapes = {}
    for o, p in batch_items:
        n = len(p.src_image_idxs)
        key = f"r{o.window_radius}s{o.window_step}_n{'<=8' if n <= 8 else ('<=16' if n <= 16 else '>16')}"
        shapes[key] = eng_batch.get_tuning(abi.PM_PHASE_SWEEP, o, n)
    b.close()
    n = len(plain_items)
    return {
        "metric": METRIC,
        "value": round(n / batch_s, 4),
        "unit": "views/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": 1,
        "ms_per_step": round(batch_s * 1000.0, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (f16 image storage)",
        "data": "synthetic",
        "config": {"workload": f"{n} orbit views 640x480..3000x2000, 5-20 sources, window 11x11 s2 or 15x15 s1, "
                               "6 iterations; batched persistent launches + hipGraph + autotuned shapes, f16 storage",
                   "views": n, "mpixel_iter": round(pix_iter / 1e6, 2)},
        "batched_views_per_s": round(n / batch_s, 4),
        "plain_views_per_s": round(n / plain_s, 4),
        "speedup_vs_plain": round(plain_s / batch_s, 3),
        "batched_mpixel_iter_per_s": round(pix_iter / 1e6 / batch_s, 3),
        "plain_mpixel_iter_per_s": round(pix_iter / 1e6 / plain_s, 3),
        "median_rel_depth_error_plain": float(np.median(e_plain)),
        "median_rel_depth_error_batched": float(np.median(e_batch)),
        "median_error_ratio": float(np.median(e_batch) / np.median(e_plain)),
        "depth_maps_identical": bool(identical),
        "plain_view_p50_ms": float(np.percentile(lat, 50)),
        "plain_view_p99_ms": float(np.percentile(lat, 99)),
        "batch_warmup_s": round(warm_s, 2),
        "orbit_render_s": round(render_s, 1),
        "sweep_shapes": shapes,
    }


def run_serve_workload(args) -> dict:
    """Depth-map requests arriving as a Poisson process into the serving loop
    (HBM admission budget, device LRU source cache, continuous batching),
    over a 500-image orbit; job sizes from the 64-view batch distribution
    (640x480..3000x2000, 5-20 sources, window 11x11 s2 or 15x15 s1, 6
    iterations, f16 storage).  Per arrival rate: views/s, p50/p99 latency
    (submit -> depth map in host memory) and the cache hit rate.  The plain
    path serves the same jobs one at a time (create + run + read back the
    depth map)."""
    from sensemap_amd import abi, scene
    from sensemap_amd import patch_match as pmm

    t0 = time.perf_counter()
    orbit = scene.orbit_scene(seed=5, num_cams=500)
    progress(f"orbit of 500 images rendered in {time.perf_counter() - t0:.1f} s")
    rates = [float(r) for r in args.rates.split(",")]
    J = args.jobs
    pool = scene.batch_problems(orbit, num_views=J * len(rates), seed=5, image_storage=abi.PM_STORAGE_F16)
    eng = pmm.default_engine()
    # launch shapes for every kernel class, tuned once before serving
    classes = {}
    for o, p in pool:
        n = len(p.src_image_idxs)
        key = (o.window_radius, o.window_step, 0 if n <= 8 else (1 if n <= 16 else 2))
        if key not in classes:
            classes[key] = (o, p)
    t = time.perf_counter()
    for o, p in classes.values():
        v = pmm.PatchMatchACMMCuda(o, p, eng)
        v.Autotune()
        v.close()
    progress(f"autotuned {len(classes)} classes in {time.perf_counter() - t:.1f} s")

    # plain path on the first rate's jobs: one view at a time
    plain_jobs = pool[:J]
    t = time.perf_counter()
    plain_err = []
    plain_lat = []
    for o, p in plain_jobs:
        t1 = time.perf_counter()
        v = pmm.PatchMatchACMMCuda(o, p, eng)
        v.Run()
        depth = v.GetDepthMap()
        v.close()
        plain_lat.append((time.perf_counter() - t1) * 1000.0)
        plain_err.append(rel_depth_errors(depth, orbit.views[p.ref_image_idx].depth))
    plain_s = time.perf_counter() - t
    progress(f"plain path: {J} jobs in {plain_s:.1f} s")

    server = pmm.PatchMatchServer(eng, hbm_budget_bytes=int(args.budget_gb * (1 << 30)),
                                  max_active_views=args.max_active)
    rng = np.random.default_rng(5)
    per_rate = []
    served_err = []
    prev = server.stats()
    for ri, lam in enumerate(rates):
        jobs_spec = pool[ri * J:(ri + 1) * J]
        gaps = rng.exponential(1.0 / lam, size=J)
        handles = [None] * J
        t_start = time.perf_counter()
        t_next = t_start
        for k, (o, p) in enumerate(jobs_spec):
            t_next += gaps[k]
            d = t_next - time.perf_counter()
            if d > 0:
                time.sleep(d)
            handles[k] = server.submit(o, p, maps=("depth",))  # what the plain path reads back too
        lat = []
        for k, h in enumerate(handles):
            st = h.wait()
            lat.append(st.latency_ms)
            if ri == 0:
                served_err.append(rel_depth_errors(h.depth, orbit.views[jobs_spec[k][1].ref_image_idx].depth))
        span = time.perf_counter() - t_start
        cur = server.stats()
        hits = cur.cache_hits - prev.cache_hits
        misses = cur.cache_misses - prev.cache_misses
        prev = cur
        per_rate.append({"rate_req_s": lam, "jobs": J, "views_per_s": round(J / span, 4),
                         "p50_ms": round(float(np.percentile(lat, 50)), 1),
                         "p99_ms": round(float(np.percentile(lat, 99)), 1),
                         "cache_hit_rate": round(hits / max(1, hits + misses), 4),
                         "max_active": cur.max_active_seen})
        progress(f"rate {lam}/s: {per_rate[-1]}")
    st = server.stats()
    server.close()
    e_plain = np.concatenate(plain_err)
    e_served = np.concatenate(served_err)
    best = max(per_rate, key=lambda r: r["views_per_s"])
    return {
        "metric": METRIC,
        "value": best["views_per_s"],
        "unit": "views/s",
        "n_gpus": 1,
        "steps": len(rates),
        "warmup": 0,
        "ms_per_step": None,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (f16 image storage)",
        "data": "synthetic",
        "config": {"workload": f"Poisson arrivals at {args.rates} req/s, {J} jobs per rate, 500-image orbit, "
                               "sizes 640x480..3000x2000, 5-20 sources, 6 iterations; serving loop with "
                               f"{args.budget_gb:g} GB HBM budget, device LRU source cache, continuous batching",
                   "jobs_per_rate": J, "budget_gb": args.budget_gb, "max_active_views": args.max_active},
        "per_rate": per_rate,
        "plain_views_per_s": round(J / plain_s, 4),
        "plain_job_p50_ms": round(float(np.percentile(plain_lat, 50)), 1),
        "plain_job_p99_ms": round(float(np.percentile(plain_lat, 99)), 1),
        "speedup_vs_plain": round(best["views_per_s"] / (J / plain_s), 3),
        "median_rel_depth_error_plain": float(np.median(e_plain)),
        "median_rel_depth_error_served": float(np.median(e_served)),
        "median_error_ratio": float(np.median(e_served) / np.median(e_plain)),
        "cache_hits": int(st.cache_hits), "cache_misses": int(st.cache_misses),
        "cache_evictions": int(st.cache_evictions), "steps_issued": int(st.steps), "launches": int(st.launches),
    }


def run_plane_view(pmm, scene_mod, steps: int, with_cpu: bool) -> dict:
    """The 640x480 three-plane view (4 sources, window 11x11 step 2, sigma_s 3,
    sigma_c 0.2, 6 iterations, best of 8 random inits): GPU Mpixel.iter/s
    and, when asked, the CPU transcription on the whole view, whose maps
    must equal the GPU's bit for bit."""
    sc = scene_mod.planes_scene(seed=1)
    prob = scene_mod.to_problem(sc)
    opts = pmm.PatchMatchOptions(window_radius=5, window_step=2, sigma_spatial=3.0, sigma_color=0.2,
                                 num_iterations=6, geom_consistency=False, thk=2, seed=1, init_samples=8)
    v = pmm.PatchMatchACMMCuda(opts, prob)
    v.Run()
    t = time.perf_counter()
    for _ in range(steps):
        v.Run()
    ms = (time.perf_counter() - t) * 1000.0 / steps
    H, W = sc.ref.image.shape
    mpix_iter = W * H * opts.num_iterations / 1e6
    out = {"workload": "640x480 three planes, 4 sources, window 11x11 step 2, 6 iterations, best of 8 inits",
           "gpu_mpixel_iter_per_s": round(mpix_iter / (ms / 1000.0), 3), "gpu_ms_per_view": round(ms, 2),
           "median_rel_depth_error": float(np.median(rel_depth_errors(v.GetDepthMap(), sc.ref.depth)))}
    if with_cpu:
        from oracle import binding
        threads = cpu_threads()[0]
        t = time.perf_counter()
        cpu = binding.run_view(opts, v.view, 0, 0, threads)
        cs = time.perf_counter() - t
        gpu = {"depth": v.GetDepthMap(), "normal": v.GetNormalMap(), "conf": v.GetConfMap(),
               "cost": v.GetCostMap(), "view_sel": v.GetViewSelMap()}
        same = all(np.array_equal(np.asarray(gpu[k]).view(np.uint32), np.asarray(cpu[k]).view(np.uint32))
                   for k in gpu)
        out["cpu_transcription"] = {"value": round(mpix_iter / cs, 4), "unit": "Mpixel·iter/s", "cores": threads,
                                    "seconds": round(cs, 2), "maps_bitwise_equal_to_gpu": bool(same)}
    v.close()
    return out


def run_source_selection(scene_mod, steps: int, with_cpu: bool) -> dict:
    """Model::GetMaxOverlappingImages on a 500-camera orbit model with 300k
    points (mean track 6, 4 blind cameras), 40 sources at 1 degree, plus the
    depth ranges: device time per phase (hipEvents), the call's wall time
    (host checks, copies, fallback) and the transcription's single-thread
    time on the same call, with the lists compared."""
    s = scene_mod.sparse_scene(seed=5, num_cams=500, num_points=300_000, num_blind=4)
    m = s.model()
    m.GetMaxOverlappingImages(40, 1.0)
    walls, dev, phases = [], [], []
    got = None
    for _ in range(max(steps, 3)):
        t = time.perf_counter()
        got = m.GetMaxOverlappingImages(40, 1.0)
        walls.append((time.perf_counter() - t) * 1000.0)
        st = m.last_overlap_stats
        dev.append(st.total_ms)
        phases.append((st.emit_ms, st.sort_ms, st.reduce_ms, st.rank_ms))
    k = int(np.argsort(dev)[len(dev) // 2])
    t = time.perf_counter()
    ranges = m.ComputeDepthRanges()
    range_ms = (time.perf_counter() - t) * 1000.0
    st = m.last_overlap_stats
    out = {"workload": "500-camera orbit, 300k points (mean track 6), 40 sources, 1 degree",
           "track_entries": int(m.track_offsets[-1]), "pair_records": int(st.pair_records),
           "image_pairs": int(st.image_pairs), "fallback_images": int(st.fallback_images),
           "gpu_device_ms": round(dev[k], 3), "gpu_call_ms": round(float(np.median(walls)), 3),
           "phase_ms": dict(zip(("emit", "sort", "reduce", "rank"), (round(x, 3) for x in phases[k]))),
           "mrecords_per_s": round(st.pair_records / dev[k] / 1e3, 1),
           "depth_ranges_call_ms": round(range_ms, 3)}
    if with_cpu:
        from oracle import binding
        t = time.perf_counter()
        want = binding.max_overlapping_images(m, 40, 1.0)
        cs = time.perf_counter() - t
        want_r = binding.depth_ranges(m)
        out["cpu_transcription"] = {"ms": round(cs * 1000.0, 1), "cores": 1,
                                    "lists_equal_to_gpu": bool(want == got),
                                    "depth_ranges_equal_to_gpu": bool(np.array_equal(
                                        np.asarray(ranges, np.float32), want_r))}
    return out


def run_filters_workload(args) -> dict:
    """The steps after the path on the 1920x1080 heightfield scene: each of the
    11 views' photometric maps (untimed), then RunFilter of the reference
    view against its 10 sources (geometric consistency), the cross-view
    filter of all 11 views against the other 10, the NormalMap encoding of
    the reference's normals on the device, the steps after the geometric
    pass, deduplication and depth-map fusion of the 11 refined maps.  Each is timed as the call a
    caller makes (device-resident inputs for RunFilter and the encoding; the
    cross filter takes host maps, so its call includes their upload) and
    beside its CPU transcription on the same inputs, with the results
    compared bit for bit."""
    from oracle import binding, io_oracle
    from sensemap_amd import abi, scene
    from sensemap_amd import patch_match as pmm
    from sensemap_amd.cross_filter import CrossImage, cross_filter

    threads = cpu_threads()[0]
    sc = scene.heightfield_scene(seed=2)
    views = [sc.ref] + sc.srcs
    n = len(views)
    images = [pmm.Image(v.image, v.camera.K, v.camera.R, v.camera.T) for v in views]
    engine = pmm.default_engine()
    photo_opts = c2_options(pmm, False, abi.PM_RCP_EXACT)
    progress("photometric maps of the 11 views")
    depths, normals, photos = [], [], []
    for k in range(n):
        prob = pmm.Problem(k, [j for j in range(n) if j != k], images, sc.depth_min, sc.depth_max)
        e = pmm.PatchMatchACMMCuda(photo_opts, prob, engine)
        e.Run()
        depths.append(e.GetDepthMap())
        normals.append(e.GetNormalMap())
        photos.append(e)
    ref_photo = photos[0]
    H, W = depths[0].shape
    px = W * H
    reps = max(args.steps, 3)

    # RunFilter (k_run_filter): a fresh problem per repetition (the filter
    # rewrites the depth map it is given)
    fo = c2_options(pmm, True, abi.PM_RCP_EXACT)
    fo.init_depth_random = False
    fprob = pmm.Problem(0, list(range(1, n)), images, sc.depth_min, sc.depth_max, depth_maps=depths,
                        normal_maps=normals)
    times, got = [], None
    for _ in range(reps):
        e = pmm.PatchMatchACMMCuda(fo, fprob, engine)
        t = time.perf_counter()
        e.RunFilter(False)
        times.append((time.perf_counter() - t) * 1000.0)
        got = e.GetDepthMap()
        e.close()
    t = time.perf_counter()
    want = binding.run_filter(fo, pmm.ViewDesc(fprob, fo), False, threads)
    cpu_ms = (time.perf_counter() - t) * 1000.0
    run_filter = {"call_ms": round(float(np.median(times)), 3), "cpu_ms": round(cpu_ms, 1), "cpu_cores": threads,
                  "mpixel_per_s": round(px / 1e6 / (np.median(times) / 1000.0), 1),
                  "kept": int((got > 0).sum()), "bitwise_equal": bool(np.array_equal(got.view(np.uint32),
                                                                                     want.view(np.uint32)))}

    progress("cross-view filter")
    # cross-view filter (k_cross_filter): all 11 views against the other 10
    ims = [CrossImage(v.camera.K, v.camera.R, v.camera.T, depths[k], normals[k]) for k, v in enumerate(views)]
    cviews = [(k, [j for j in range(n) if j != k]) for k in range(n)]
    cross_filter(ims, cviews)
    times = []
    for _ in range(reps):
        t = time.perf_counter()
        outs = cross_filter(ims, cviews)
        times.append((time.perf_counter() - t) * 1000.0)
    t = time.perf_counter()
    couts = binding.cross_filter(ims, cviews, threads=threads)
    cpu_ms = (time.perf_counter() - t) * 1000.0
    # the same over the problems, maps already on the device
    from sensemap_amd.cross_filter import cross_filter_problems
    cross_filter_problems(photos, cviews)
    dtimes = []
    for _ in range(reps):
        t = time.perf_counter()
        douts = cross_filter_problems(photos, cviews)
        dtimes.append((time.perf_counter() - t) * 1000.0)
    cross = {"views": n, "sources_per_view": n - 1, "call_ms": round(float(np.median(times)), 3),
             "device_maps_call_ms": round(float(np.median(dtimes)), 3),
             "device_maps_equal": bool(all(np.array_equal(a.view(np.uint32), b.view(np.uint32))
                                           for a, b in zip(douts, outs))),
             "cpu_ms": round(cpu_ms, 1), "cpu_cores": threads,
             "mpixel_per_s": round(n * px / 1e6 / (np.median(times) / 1000.0), 1),
             "bitwise_equal": bool(all(np.array_equal(a.view(np.uint32), b.view(np.uint32))
                                       for a, b in zip(outs, couts)))}

    progress("normal encoding")
    # NormalMap encoding on the device (k_encode_normals + read back)
    ref_photo.GetNormalCodes()
    times = []
    for _ in range(reps):
        t = time.perf_counter()
        codes = ref_photo.GetNormalCodes()
        times.append((time.perf_counter() - t) * 1000.0)
    t = time.perf_counter()
    want_codes = io_oracle.encode(ref_photo.GetNormalMap())
    cpu_ms = (time.perf_counter() - t) * 1000.0
    enc = {"call_ms": round(float(np.median(times)), 3), "cpu_ms": round(cpu_ms, 1), "cpu_cores": 1,
           "mpixel_per_s": round(px / 1e6 / (np.median(times) / 1000.0), 1),
           "bitwise_equal": bool(np.array_equal(codes, want_codes))}
    # the steps after the iterations of the reference's geometric pass: the
    # curvature map, the median filter and a mask (k_curvature, k_geom_grad,
    # k_median, k_mask), timed on the device (pm_run_stats.post_ms) and
    # checked against the transcription's chain applied to the same pass's
    # maps without those steps
    progress("steps after the geometric pass")
    po = c2_options(pmm, True, abi.PM_RCP_EXACT)
    rng = np.random.default_rng(3)
    mask = (rng.random((H, W)) > 0.1).astype(np.uint8) * 255
    gprob = pmm.Problem(0, list(range(1, n)), images, sc.depth_min, sc.depth_max, depth_maps=depths,
                        normal_maps=normals, mask_maps=[mask] + [None] * (n - 1))
    plain_prob = pmm.Problem(0, list(range(1, n)), images, sc.depth_min, sc.depth_max, depth_maps=depths,
                             normal_maps=normals)
    base = pmm.PatchMatchACMMCuda(po, plain_prob, engine)
    base.Run()
    d0, n0 = base.GetDepthMap(), base.GetNormalMap()
    base.close()
    po_post = c2_options(pmm, True, abi.PM_RCP_EXACT)
    po_post.est_curvature = True
    po_post.median_filter = True
    e = pmm.PatchMatchACMMCuda(po_post, gprob, engine)
    post_times = []
    for _ in range(reps + 1):
        e.Run()
        post_times.append(e.GetStats().post_ms)
    gd, gn, gc = e.GetDepthMap(), e.GetNormalMap(), e.GetCurvatureMap()
    e.close()
    t = time.perf_counter()
    want_c = binding.curvature_map(views[0].image, n0)
    want_d, want_n = binding.median_filter(d0, n0, po_post.depth_diff_threshold)
    want_d = binding.mask_depth(mask, want_d)
    cpu_ms = (time.perf_counter() - t) * 1000.0
    post_ms = float(np.median(post_times[1:]))
    post = {"post_ms": round(post_ms, 3), "mpixel_per_s": round(px / 1e6 / (post_ms / 1000.0), 1),
            "algorithmic_bytes_per_px": 66,
            "achieved_gb_s": round(66 * px / (post_ms / 1000.0) / 1e9, 1),
            "cpu_ms": round(cpu_ms, 1), "cpu_cores": threads,
            "bitwise_equal": bool(np.array_equal(gc, want_c) and np.array_equal(gd.view(np.uint32), want_d.view(np.uint32))
                                  and np.array_equal(gn.view(np.uint32), want_n.view(np.uint32)))}
    progress("deduplication")
    # deduplication after the cross filter (k_dedup): the 11 refined maps,
    # every view a filtered reference, in view order
    from sensemap_amd.cross_filter import dedup
    dims = [(CrossImage(v.camera.K, v.camera.R, v.camera.T, outs[k], normals[k]), True) for k, v in enumerate(views)]
    dedup(dims, cviews)
    times = []
    for _ in range(reps):
        t = time.perf_counter()
        dd = dedup(dims, cviews)
        times.append((time.perf_counter() - t) * 1000.0)
    t = time.perf_counter()
    dd_cpu = binding.cross_dedup(dims, cviews, threads=threads)
    cpu_ms = (time.perf_counter() - t) * 1000.0
    dedup_res = {"views": n, "call_ms": round(float(np.median(times)), 3), "cpu_ms": round(cpu_ms, 1),
                 "cpu_cores": threads, "mpixel_per_s": round(n * px / 1e6 / (np.median(times) / 1000.0), 1),
                 "kept": int(sum((v != 0).sum() for v in dd["visibility"])),
                 "bitwise_equal": bool(all(np.array_equal(a, b) for a, b in zip(dd["visibility"], dd_cpu["visibility"]))
                                       and all(np.array_equal(a.view(np.uint32), b.view(np.uint32))
                                               for key in ("dedup_depth", "delt")
                                               for a, b in zip(dd[key], dd_cpu[key])))}

    progress("depth-map fusion")
    # depth-map fusion (pm_fusion_run) of the 11 refined maps with the
    # photometric normals and the images' intensities; every view overlaps
    # the other 10
    from sensemap_amd.fusion import FusionImage, StereoFusionOptions, fuse, fusion_stats
    fims = [FusionImage(v.camera.K, v.camera.R, v.camera.T, outs[k], normals[k], v.image,
                        [j for j in range(n) if j != k]) for k, v in enumerate(views)]
    fo_ = StereoFusionOptions()
    t = time.perf_counter()
    fuse(fims, fo_)
    progress(f"fusion first call {time.perf_counter() - t:.2f} s")
    times = []
    for _ in range(max(2, reps // 2)):
        t = time.perf_counter()
        fz = fuse(fims, fo_)
        times.append((time.perf_counter() - t) * 1000.0)
        progress(f"fusion call {times[-1]:.1f} ms")
    t = time.perf_counter()
    fz_cpu = binding.fusion(fims, fo_)
    cpu_ms = (time.perf_counter() - t) * 1000.0
    progress(f"fusion transcription {cpu_ms:.1f} ms")
    fst = fusion_stats(fims, fo_)
    fusion_res = {"views": n, "points": int(len(fz["score"])), "call_ms": round(float(np.median(times)), 2),
                  "walks": int(fst.walks), "walks_with_pixels": int(fst.kept), "dataflow_launches": int(fst.rounds),
                  "walks_run_alone": int(fst.big_walks), "longest_chains": int(fst.chain),
                  "dataflow_ms": round(float(fst.flow_ms), 1),
                  "cpu_ms": round(cpu_ms, 1), "cpu_cores": 1,
                  "mpixel_per_s": round(n * px / 1e6 / (np.median(times) / 1000.0), 2),
                  "bitwise_equal": bool(len(fz["score"]) == len(fz_cpu["score"]) and all(
                      np.array_equal(fz[k].view(np.uint32) if fz[k].dtype == np.float32 else fz[k],
                                     fz_cpu[k].view(np.uint32) if fz_cpu[k].dtype == np.float32 else fz_cpu[k])
                      for k in ("xyz", "normal", "rgb", "score", "vis_offsets", "vis_images")))}
    for e in photos:
        e.close()
    return {"metric": "filter and encoding throughput (Mpixel/s) after the path", "value": cross["mpixel_per_s"],
            "unit": "Mpixel/s", "n_gpus": 1, "steps": reps, "warmup": 1, "ms_per_step": cross["call_ms"],
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
            "config": {"workload": "1920x1080 heightfield scene, 11 views: RunFilter of the reference (10 sources), "
                                   "cross-view filter of all 11 views (10 sources each), normal encoding, "
                                   "curvature map + median filter + mask after the geometric pass, deduplication "
                                   "and depth-map fusion of the 11 refined maps"},
            "run_filter": run_filter, "cross_filter": cross, "normal_encoding": enc, "post_steps": post,
            "dedup": dedup_res, "fusion": fusion_res}


def run_after_path(views, problems, with_cpu: bool) -> dict:
    """The stages after the path on the step's own views: the cross-view
    filter over the 11 photometric problems whose maps are already on the
    device (every view against the other 10), deduplication of the refined
    maps and depth-map fusion of the refined maps with the photometric
    normals and the images.  Each stage is timed as one call, host copies
    included, and compared bit for bit with its CPU transcription on the
    same inputs (skipped with with_cpu False)."""
    from oracle import binding
    from sensemap_amd.cross_filter import CrossImage, cross_filter_problems, dedup
    from sensemap_amd.fusion import FusionImage, StereoFusionOptions, fuse

    threads = cpu_threads()[0]
    n = len(views)
    cviews = [(k, [j for j in range(n) if j != k]) for k in range(n)]
    normals = [p.GetNormalMap() for p in problems]
    depths = [p.GetDepthMap() for p in problems]

    def timed(fn, reps=3):
        fn()
        ts, out = [], None
        for _ in range(reps):
            t = time.perf_counter()
            out = fn()
            ts.append((time.perf_counter() - t) * 1000.0)
        return out, round(float(np.median(ts)), 2)

    def same_f32(a, b):
        return bool(np.array_equal(a.view(np.uint32), b.view(np.uint32)))

    progress("after the path: cross filter")
    refined, cross_ms = timed(lambda: cross_filter_problems(problems, cviews))
    progress("after the path: deduplication")
    dims = [(CrossImage(v.camera.K, v.camera.R, v.camera.T, refined[k], normals[k]), True)
            for k, v in enumerate(views)]
    dd, dedup_ms = timed(lambda: dedup(dims, cviews))
    progress("after the path: fusion")
    fims = [FusionImage(v.camera.K, v.camera.R, v.camera.T, refined[k], normals[k], v.image,
                        [j for j in range(n) if j != k]) for k, v in enumerate(views)]
    fo = StereoFusionOptions()
    fz, fusion_ms = timed(lambda: fuse(fims, fo), reps=2)
    res = {"workload": "cross filter of the 11 photometric maps on the device (10 sources each), "
                       "deduplication of the refined maps, depth-map fusion of the refined maps with the "
                       "photometric normals and the images (default StereoFusionOptions)",
           "cross_filter_ms": cross_ms, "dedup_ms": dedup_ms, "fusion_ms": fusion_ms,
           "seconds": round((cross_ms + dedup_ms + fusion_ms) / 1000.0, 3),
           "kept_after_filter": int(sum((r > 0).sum() for r in refined)), "points": int(len(fz["score"])),
           "bitwise_equal": None}
    if not with_cpu:
        return res
    progress("after the path: transcriptions")
    hims = [CrossImage(v.camera.K, v.camera.R, v.camera.T, depths[k], normals[k]) for k, v in enumerate(views)]
    t = time.perf_counter()
    want_refined = binding.cross_filter(hims, cviews, threads=threads)
    cross_cpu = (time.perf_counter() - t) * 1000.0
    t = time.perf_counter()
    want_dd = binding.cross_dedup(dims, cviews, threads=threads)
    dedup_cpu = (time.perf_counter() - t) * 1000.0
    t = time.perf_counter()
    want_fz = binding.fusion(fims, fo)
    fusion_cpu = (time.perf_counter() - t) * 1000.0
    eq = {"cross_filter": all(same_f32(a, b) for a, b in zip(refined, want_refined)),
          "dedup": all(np.array_equal(a, b) for a, b in zip(dd["visibility"], want_dd["visibility"]))
          and all(same_f32(a, b) for key in ("dedup_depth", "delt") for a, b in zip(dd[key], want_dd[key])),
          "fusion": len(fz["score"]) == len(want_fz["score"]) and all(
              same_f32(fz[k], want_fz[k]) if fz[k].dtype == np.float32 else np.array_equal(fz[k], want_fz[k])
              for k in ("xyz", "normal", "rgb", "score", "vis_offsets", "vis_images"))}
    res["bitwise_equal"] = {k: bool(v) for k, v in eq.items()}
    res["cpu_ms"] = {"cross_filter": round(cross_cpu, 1), "dedup": round(dedup_cpu, 1),
                     "fusion": round(fusion_cpu, 1)}
    res["cpu_cores"] = {"cross_filter": threads, "dedup": threads, "fusion": 1}
    return res


def main() -> int:
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU: bench.py --gpus must be 1 "
                                   "(multi-GPU sharding by reference view is design-only)"}))
        return 2
    if args.copy_calibration:
        return copy_calibration(args.copy_calibration)
    if args.workload == "batch":
        print(json.dumps(run_batch_workload(args)))
        return 0
    if args.workload == "serve":
        print(json.dumps(run_serve_workload(args)))
        return 0
    if args.workload == "filters":
        print(json.dumps(run_filters_workload(args)))
        return 0
    if args.workload == "select":
        from sensemap_amd import scene
        sel = run_source_selection(scene, args.steps, not args.no_cpu_baseline)
        print(json.dumps({"metric": "pair records/s (source-view selection)", "value": sel["mrecords_per_s"] * 1e6,
                          "unit": "records/s", "n_gpus": 1, "steps": max(args.steps, 3), "warmup": 1,
                          "ms_per_step": sel["gpu_device_ms"], "higher_is_better": True, "scaling": "weak",
                          "vs_baseline": None, "dtype": "u64", "data": "synthetic",
                          "config": {"workload": sel["workload"]}, "source_selection": sel}))
        return 0
    from sensemap_amd import abi, scene
    from sensemap_amd import patch_match as pmm

    rcp = abi.PM_RCP_FAST if args.rcp == "fast" else abi.PM_RCP_EXACT
    sc = scene.heightfield_scene(seed=2)
    views = [sc.ref] + sc.srcs
    n_views = len(views)
    images = [pmm.Image(v.image, v.camera.K, v.camera.R, v.camera.T) for v in views]
    engine = pmm.default_engine()

    # sources' photometric passes (inputs of the geometric pass), untimed
    storage = abi.PM_STORAGE_F16 if args.storage == "f16" else abi.PM_STORAGE_U8
    photo_opts = c2_options(pmm, False, rcp, storage)
    src_photo = []
    for k in range(1, 1 if args.only_photometric else n_views):
        prob = pmm.Problem(k, [j for j in range(n_views) if j != k], images, sc.depth_min, sc.depth_max)
        p = pmm.PatchMatchACMMCuda(photo_opts, prob, engine)
        p.Run()
        src_photo.append(p)
    ref_prob = pmm.Problem(0, list(range(1, n_views)), images, sc.depth_min, sc.depth_max,
                           depth_maps=[v.depth for v in views], normal_maps=[v.normal for v in views])
    photo = pmm.PatchMatchACMMCuda(photo_opts, ref_prob, engine)
    geo = pmm.PatchMatchACMMCuda(c2_options(pmm, True, rcp, storage), ref_prob, engine)
    for i, p in enumerate(src_photo):
        geo.SetSrcDepthFrom(i, p)

    def step():
        photo.Run()
        if not args.only_photometric:
            geo.InitFrom(photo)
            geo.Run()

    # calibration step with counting on (deterministic: same evaluations as
    # every timed step); the atomics it adds are off for the timed steps
    engine.set_counting(True)
    step()
    s1 = photo.GetStats()
    s2 = None if args.only_photometric else geo.GetStats()
    engine.set_counting(False)
    flops_sweep = sweep_flops(s1) + (sweep_flops(s2) if s2 else 0)
    launches_per_step = s1.sweep_launches + (s2.sweep_launches if s2 else 0)

    for _ in range(args.warmup):
        step()
    lat = []
    sweep_ms = 0.0
    t_all = time.perf_counter()
    for _ in range(args.steps):
        t = time.perf_counter()
        step()
        lat.append((time.perf_counter() - t) * 1000.0)
        sweep_ms += photo.GetStats().sweep_ms + (0.0 if args.only_photometric else geo.GetStats().sweep_ms)
    total_s = time.perf_counter() - t_all
    ms_per_step = total_s * 1000.0 / args.steps
    W, H = sc.ref.image.shape[1], sc.ref.image.shape[0]
    iters = photo_opts.num_iterations + (0 if args.only_photometric else 2)
    mpix_iter = W * H * iters / 1e6
    value = mpix_iter / (ms_per_step / 1000.0)
    final = photo if args.only_photometric else geo
    step_depth = final.GetDepthMap()
    err = np.abs(step_depth - sc.ref.depth)[sc.ref.depth > 0] / sc.ref.depth[sc.ref.depth > 0]
    err_photo = np.abs(photo.GetDepthMap() - sc.ref.depth)[sc.ref.depth > 0] / sc.ref.depth[sc.ref.depth > 0]
    avg_launch_ms = sweep_ms / (args.steps * launches_per_step)
    achieved = (flops_sweep / launches_per_step) / (avg_launch_ms / 1000.0) / 1e12
    # the same step with the other image storage (f16 holds the 8-bit values
    # exactly: the maps must come out identical), timed the same way
    other = None
    if not args.only_photometric:
        alt = abi.PM_STORAGE_U8 if storage == abi.PM_STORAGE_F16 else abi.PM_STORAGE_F16
        o16, g16 = c2_options(pmm, False, rcp, alt), c2_options(pmm, True, rcp, alt)
        photo16 = pmm.PatchMatchACMMCuda(o16, ref_prob, engine)
        geo16 = pmm.PatchMatchACMMCuda(g16, ref_prob, engine)
        for i, p in enumerate(src_photo):
            geo16.SetSrcDepthFrom(i, p)

        def step16():
            photo16.Run()
            geo16.InitFrom(photo16)
            geo16.Run()

        for _ in range(args.warmup):
            step16()
        t = time.perf_counter()
        for _ in range(args.steps):
            step16()
        ms16 = (time.perf_counter() - t) * 1000.0 / args.steps
        other = {"image_storage": "u8" if alt == abi.PM_STORAGE_U8 else "f16",
                 "value": round(mpix_iter / (ms16 / 1000.0), 3), "ms_per_step": round(ms16, 2),
                 "depth_maps_identical": bool(np.array_equal(geo16.GetDepthMap(), geo.GetDepthMap()))}
    # the same step with the fast reciprocal (one Newton step, within one ulp
    # of 1/z, not bitwise): what the exact reciprocal costs
    fast = None
    if rcp == abi.PM_RCP_EXACT and not args.only_photometric:
        of, gf = c2_options(pmm, False, abi.PM_RCP_FAST, storage), c2_options(pmm, True, abi.PM_RCP_FAST, storage)
        photof = pmm.PatchMatchACMMCuda(of, ref_prob, engine)
        geof = pmm.PatchMatchACMMCuda(gf, ref_prob, engine)
        for i, p in enumerate(src_photo):
            geof.SetSrcDepthFrom(i, p)

        def stepf():
            photof.Run()
            geof.InitFrom(photof)
            geof.Run()

        for _ in range(args.warmup):
            stepf()
        t = time.perf_counter()
        for _ in range(args.steps):
            stepf()
        msf = (time.perf_counter() - t) * 1000.0 / args.steps
        df, de = geof.GetDepthMap(), geo.GetDepthMap()
        m = (df > 0) & (de > 0)
        errf = np.abs(df - sc.ref.depth)[sc.ref.depth > 0] / sc.ref.depth[sc.ref.depth > 0]
        fast = {"value": round(mpix_iter / (msf / 1000.0), 3), "ms_per_step": round(msf, 2),
                "median_rel_depth_error": float(np.median(errf)),
                "median_rel_diff_vs_exact": float(np.median(np.abs(df[m] - de[m]) / de[m]))}
    # the whole C2 pipeline once: the 10 sources' photometric passes, whose
    # depth maps feed the geometric pass, then the reference's photometric and
    # geometric passes (11 views, 8 photometric iterations each, plus 2)
    full = None
    if not args.only_photometric:
        full_runs = []
        for _ in range(3):
            t = time.perf_counter()
            for p in src_photo:
                p.Run()
            for i, p in enumerate(src_photo):
                geo.SetSrcDepthFrom(i, p)
            step()
            full_runs.append(time.perf_counter() - t)
        full_s = float(np.median(full_runs))
        full_mpix = (n_views * photo_opts.num_iterations + 2) * W * H / 1e6
        full = {"workload": "10 source photometric passes + the reference's photometric and geometric passes",
                "views": n_views, "mpixel_iter": round(full_mpix, 2), "seconds": round(full_s, 3),
                "runs_s": [round(x, 3) for x in full_runs],
                "mpixel_iter_per_s": round(full_mpix / full_s, 3), "views_per_s": round(n_views / full_s, 4),
                "depth_equal_to_step": bool(np.array_equal(geo.GetDepthMap().view(np.uint32),
                                                           step_depth.view(np.uint32)))}
    # the same pipeline with the 11 photometric passes advanced together in
    # shared launches (pm_batch, hipGraph replay), the geometric pass fed from
    # the batch's views on the device
    full_batched = None
    if not args.only_photometric:
        items = [(photo_opts, pmm.Problem(k, [j for j in range(n_views) if j != k], images, sc.depth_min,
                                          sc.depth_max)) for k in range(n_views)]
        fb = pmm.PatchMatchBatch(items, engine=engine, graph=True)

        def batched_pipeline():
            fb.Run()
            for i in range(1, n_views):
                geo.SetSrcDepthFrom(i - 1, fb.views[i])
            geo.InitFrom(fb.views[0])
            geo.Run()

        batched_pipeline()  # graph capture
        runs = []
        for _ in range(3):
            t = time.perf_counter()
            batched_pipeline()
            runs.append(time.perf_counter() - t)
        fb_s = float(np.median(runs))
        full_batched = {"workload": "the same 11 photometric passes in one batch (shared persistent launches, "
                                    "hipGraph replay), then the reference's geometric pass",
                        "seconds": round(fb_s, 3), "runs_s": [round(x, 3) for x in runs],
                        "mpixel_iter_per_s": round(full_mpix / fb_s, 3), "views_per_s": round(n_views / fb_s, 4),
                        "speedup_vs_one_view_per_launch": round(full_s / fb_s, 3),
                        "depth_equal_to_step": bool(np.array_equal(geo.GetDepthMap().view(np.uint32),
                                                                   step_depth.view(np.uint32)))}
        fb.close()
        for i, p in enumerate(src_photo):  # restore the geometric pass's inputs
            geo.SetSrcDepthFrom(i, p)
    # the reference's flow for the same step: each pass's problem created from
    # host buffers (images; for the geometric pass the sources' depth maps and
    # the photometric maps as its initial state), run, maps read back
    host = None
    if not args.only_photometric:
        src_depths = [p.GetDepthMap() for p in src_photo]

        def host_step():
            t0 = time.perf_counter()
            ph = pmm.PatchMatchACMMCuda(photo_opts, ref_prob, engine)
            ph.Run()
            d, nm = ph.GetDepthMap(), ph.GetNormalMap()
            ph.close()
            gprob = pmm.Problem(0, list(range(1, n_views)), images, sc.depth_min, sc.depth_max,
                                depth_maps=[d] + src_depths, normal_maps=[nm] + [v.normal for v in views[1:]])
            g = pmm.PatchMatchACMMCuda(c2_options(pmm, True, rcp, storage), gprob, engine)
            g.Run()
            gd = g.GetDepthMap()
            g.GetNormalMap()
            g.close()
            return (time.perf_counter() - t0) * 1000.0, gd

        host_step()
        runs = [host_step() for _ in range(2)]
        hms = float(np.median([r[0] for r in runs]))
        host = {"workload": "the same step with each pass's problem created from host buffers and its depth and "
                            "normal maps read back (images, source depth maps and initial maps cross PCIe)",
                "ms_per_step": round(hms, 2), "value": round(mpix_iter / (hms / 1000.0), 3),
                "maps_equal_to_device_flow": bool(np.array_equal(runs[-1][1].view(np.uint32),
                                                                 geo.GetDepthMap().view(np.uint32)))}
    after = None
    if not args.only_photometric:
        after = run_after_path(views, [photo] + src_photo, not args.no_cpu_baseline)
    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "Mpixel·iter/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "1920x1080 reference + 10 sources, heightfield scene, window 15x15 step 1, "
                               "8 photometric + 2 geometric iterations, one view per launch",
                   "width": W, "height": H, "num_src": n_views - 1, "window_radius": 7, "window_step": 1,
                   "iterations": iters, "rcp_mode": args.rcp, "image_storage": args.storage},
        "views_per_s": round(1000.0 / ms_per_step, 4),
        "median_rel_depth_error": float(np.median(err)),
        "frac_within_1pct": float(np.mean(err < 0.01)),
        "median_rel_depth_error_photometric": float(np.median(err_photo)),
        "other_storage": other,
        "fast_reciprocal": fast,
        "full_pipeline": full,
        "full_pipeline_batched": full_batched,
        "host_buffers": host,
        "after_path": after,
        "p50_ms": float(np.percentile(lat, 50)),
        "p99_ms": float(np.percentile(lat, 99)),
        "roofline": {
            "kernel": "k_sweep (checkerboard propagation + view selection + refinement)",
            "bound": "valu",
            "achieved": round(achieved, 3),
            "peak": FP32_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": round(achieved / FP32_PEAK_TFLOPS, 4),
            "traffic": pmc_traffic(),
            "avg_launch_ms": round(avg_launch_ms, 3),
            "flop_per_launch": int(flops_sweep / launches_per_step),
        },
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(pmm, scene, sc, args.cpu_seconds)
    if not args.only_photometric:
        out["plane_view"] = run_plane_view(pmm, scene, args.steps, not args.no_cpu_baseline)
        progress("source selection")
        out["source_selection"] = run_source_selection(scene, args.steps, not args.no_cpu_baseline)
    if not args.only_photometric and not args.skip_forms:
        # the batched (64 orbit views) and served (Poisson arrivals) forms
        # against the plain path, at reduced repetitions so the default run
        # stays within minutes; --workload batch / serve run them in full
        progress("batched form")
        b = run_batch_workload(argparse.Namespace(views=64, steps=1))
        out["batch"] = {k: b[k] for k in ("config", "batched_views_per_s", "plain_views_per_s", "speedup_vs_plain",
                                          "batched_mpixel_iter_per_s", "plain_mpixel_iter_per_s",
                                          "median_rel_depth_error_plain", "median_rel_depth_error_batched",
                                          "median_error_ratio", "depth_maps_identical", "plain_view_p50_ms",
                                          "plain_view_p99_ms", "batch_warmup_s")}
        progress("serving loop")
        s = run_serve_workload(argparse.Namespace(rates="20,200", jobs=16, budget_gb=200.0, max_active=4))
        out["serve"] = {k: s[k] for k in ("config", "per_rate", "plain_views_per_s", "plain_job_p50_ms",
                                          "plain_job_p99_ms", "speedup_vs_plain", "median_rel_depth_error_plain",
                                          "median_rel_depth_error_served", "median_error_ratio", "cache_hits",
                                          "cache_misses", "cache_evictions")}
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
