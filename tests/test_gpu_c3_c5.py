"""The batched and served forms at the sizes of their workloads:

- batched: 64 reference views of a 200-camera orbit, sizes log-uniform from
  640x480 to 3000x2000, 5-20 sources, window 11x11 step 2 or 15x15 step 1,
  6 iterations, f16 storage, hipGraph replay and autotuned launch shapes.
  Every view's maps equal, bit for bit, the view run alone on the plain
  path (so the median depth error is the plain path's), and the two
  cheapest views equal the CPU transcription bit for bit;
- served: Poisson arrivals at two rates over a 500-image orbit, jobs drawn
  from the same size distribution, a 200 GB budget.  Every served job's maps
  equal the job run alone, and the cheapest job equals the transcription.

These replace the controller's loop over problems, one engine per host
thread (patch_match_controller.cc:964-995, :3620-3632)."""
import functools
import time

import numpy as np
import pytest

from oracle import binding
from sensemap_amd import abi, scene
from sensemap_amd import patch_match as pmm
from tests import helpers

pytestmark = pytest.mark.gpu

MAPS = ("depth", "normal", "conf", "cost", "view_sel")


def grab(v):
    return {"depth": v.GetDepthMap(), "normal": v.GetNormalMap(), "conf": v.GetConfMap(), "cost": v.GetCostMap(),
            "view_sel": v.GetViewSelMap()}


def work(item):
    o, p = item
    H, W = p.images[p.ref_image_idx].bitmap.shape
    S = (2 * (o.window_radius // o.window_step) + 1) ** 2
    return W * H * len(p.src_image_idxs) * S


@functools.lru_cache(maxsize=None)
def c3_items():
    orbit = scene.orbit_scene(seed=3)
    return orbit, scene.batch_problems(orbit, num_views=64, seed=3, image_storage=abi.PM_STORAGE_F16)


def test_c3_batch_of_64_orbit_views():
    orbit, items = c3_items()
    sizes = [p.images[p.ref_image_idx].bitmap.shape for _, p in items]
    print("[c3] 64 views,", sum(h * w for h, w in sizes) / 1e6, "Mpixel; batch with graph + autotune", flush=True)
    b = pmm.PatchMatchBatch(items, engine=pmm.Engine(0), graph=True, autotune=True)
    b.Run()
    eng = pmm.Engine(0)
    e_plain, e_batch = [], []
    for k, (o, p) in enumerate(items):
        e = pmm.PatchMatchACMMCuda(o, p, eng)
        e.Run()
        alone = grab(e)
        e.close()
        got = grab(b.views[k])
        diff = helpers.maps_equal(got, alone, MAPS)
        assert all(v == 0 for v in diff.values()), (k, sizes[k], diff)
        gt = orbit.views[p.ref_image_idx].depth
        e_plain.append(helpers.depth_errors(alone["depth"], gt))
        e_batch.append(helpers.depth_errors(got["depth"], gt))
    med_p, med_b = float(np.median(np.concatenate(e_plain))), float(np.median(np.concatenate(e_batch)))
    print("[c3] every view equals its plain run; median rel. depth error", med_p, med_b, flush=True)
    assert abs(med_b - med_p) <= 0.01 * med_p
    cheapest = sorted(range(len(items)), key=lambda k: work(items[k]))[:2]
    for k in cheapest:
        o, p = items[k]
        print("[c3] view", k, sizes[k], "sources", len(p.src_image_idxs), "r", o.window_radius,
              "against the transcription", flush=True)
        cpu = binding.run_view(o, pmm.ViewDesc(p, o))
        diff = helpers.maps_equal(grab(b.views[k]), cpu, MAPS)
        assert all(v == 0 for v in diff.values()), (k, diff)
    b.close()


def test_c5_serving_at_two_rates_over_500_images():
    orbit = scene.orbit_scene(seed=5, num_cams=500)
    J = 12
    pool = scene.batch_problems(orbit, num_views=2 * J, seed=5, image_storage=abi.PM_STORAGE_F16)
    eng = pmm.Engine(0)
    sv = pmm.PatchMatchServer(eng, hbm_budget_bytes=200 << 30, max_active_views=4)
    rng = np.random.default_rng(5)
    jobs = []
    for ri, lam in enumerate((20.0, 200.0)):
        gaps = rng.exponential(1.0 / lam, size=J)
        t_next = time.perf_counter()
        for k in range(J):
            t_next += gaps[k]
            d = t_next - time.perf_counter()
            if d > 0:
                time.sleep(d)
            o, p = pool[ri * J + k]
            jobs.append(sv.submit(o, p))
        print(f"[c5] rate {lam}/s: {J} jobs submitted", flush=True)
    lat = [j.wait().latency_ms for j in jobs]
    st = sv.stats()
    print("[c5] p50/p99 latency ms", np.percentile(lat, 50), np.percentile(lat, 99), "hits", st.cache_hits,
          "misses", st.cache_misses, "max active", st.max_active_seen, flush=True)
    assert st.completed == 2 * J and st.failed == 0 and st.cache_hits > 0
    for k, ((o, p), j) in enumerate(zip(pool, jobs)):
        e = pmm.PatchMatchACMMCuda(o, p, eng)
        e.Run()
        diff = helpers.maps_equal(j.maps(), grab(e), MAPS)
        e.close()
        assert all(v == 0 for v in diff.values()), (k, diff)
    print("[c5] every served job equals the job alone", flush=True)
    k = min(range(len(pool)), key=lambda i: work(pool[i]))
    o, p = pool[k]
    cpu = binding.run_view(o, pmm.ViewDesc(p, o))
    diff = helpers.maps_equal(jobs[k].maps(), cpu, MAPS)
    print("[c5] job", k, "against the transcription", diff, flush=True)
    assert all(v == 0 for v in diff.values()), diff
    sv.close()
