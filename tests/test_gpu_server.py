"""Serving loop: jobs that arrive while others run are admitted as memory and
slots free up, their phases share launches, and each job's maps equal, bit
for bit, the maps of the same job run alone.  Budget admission, the source
cache (hits, LRU eviction) and the over-budget rejection are checked through
the server's counters."""
import time

import numpy as np
import pytest

from sensemap_amd import abi, scene
from sensemap_amd import patch_match as pmm
from tests import helpers

pytestmark = pytest.mark.gpu

MAPS = ("depth", "normal", "conf", "cost", "view_sel")


def grab(v):
    return {"depth": v.GetDepthMap(), "normal": v.GetNormalMap(), "conf": v.GetConfMap(),
            "cost": v.GetCostMap(), "view_sel": v.GetViewSelMap()}


def orbit_jobs():
    """Jobs over one small orbit: neighbouring views share sources."""
    orbit = scene.orbit_scene(seed=5, num_cams=24, min_px=96 * 72, max_px=200 * 150, grid=32, extent=20.0)
    items = scene.batch_problems(orbit, num_views=10, seed=5, n_min=2, n_max=6, iterations=4)
    rng = np.random.default_rng(9)
    out = []
    for k, (o, p) in enumerate(items):
        if k % 3 == 1:
            o.image_storage = abi.PM_STORAGE_F16
        if k % 4 == 2:
            o.window_radius, o.window_step = 4, 2  # generic kernel shape
        out.append((o, p))
    rng.shuffle(out)
    return out


def alone_maps(items, engine):
    res = []
    for o, p in items:
        e = pmm.PatchMatchACMMCuda(o, p, engine)
        e.Run()
        res.append(grab(e))
        e.close()
    return res


def workspace_reserve(engine):
    sv = pmm.PatchMatchServer(engine, hbm_budget_bytes=4 << 30, max_active_views=4)
    cap = sv.stats().arena_capacity
    sv.close()
    return (4 << 30) - cap


def run_staggered(server, items, gap_s):
    jobs = []
    for o, p in items:
        jobs.append(server.submit(o, p))
        time.sleep(gap_s)
    return jobs


def test_jobs_in_the_mix_equal_jobs_alone():
    eng = pmm.Engine(0)
    items = orbit_jobs()
    want = alone_maps(items, eng)
    sv = pmm.PatchMatchServer(eng, hbm_budget_bytes=8 << 30, max_active_views=6)
    jobs = run_staggered(sv, items, 0.003)
    for k, j in enumerate(jobs):
        st = j.wait()
        diff = helpers.maps_equal(j.maps(), want[k], MAPS)
        print("[server] job", k, "latency_ms", round(st.latency_ms, 2), diff, flush=True)
        assert all(d == 0 for d in diff.values()), (k, diff)
    s = sv.stats()
    print("[server] stats steps", s.steps, "launches", s.launches, "max_active", s.max_active_seen,
          "hits", s.cache_hits, "misses", s.cache_misses, flush=True)
    assert s.completed == len(items) and s.failed == 0
    assert s.max_active_seen > 1            # jobs overlapped on the device
    assert s.launches < s.steps * 4         # phases of several jobs shared launches
    assert s.cache_hits > 0                 # shared sources came from the cache
    sv.close()
    eng.close()


def test_small_budget_admits_few_and_evicts():
    eng = pmm.Engine(0)
    items = orbit_jobs()
    want = alone_maps(items, eng)
    reserve = workspace_reserve(eng)
    sv = pmm.PatchMatchServer(eng, hbm_budget_bytes=reserve + (3 << 20), max_active_views=8)
    jobs = [sv.submit(o, p) for o, p in items]
    for k, j in enumerate(jobs):
        j.wait()
        diff = helpers.maps_equal(j.maps(), want[k], MAPS)
        assert all(d == 0 for d in diff.values()), (k, diff)
    s = sv.stats()
    print("[server] small budget: max_active", s.max_active_seen, "evictions", s.cache_evictions,
          "capacity", s.arena_capacity, flush=True)
    assert s.completed == len(items)
    assert s.max_active_seen < len(items)   # the budget held jobs back
    assert s.cache_evictions > 0
    assert s.arena_used <= s.arena_capacity
    sv.close()
    eng.close()


def test_jobs_dropped_by_the_caller_still_complete():
    # the server borrows a job's buffers until it is waited for; the Python
    # server holds the jobs, so dropping every reference to them is safe
    import gc
    eng = pmm.Engine(0)
    items = orbit_jobs()[:4]
    sv = pmm.PatchMatchServer(eng, hbm_budget_bytes=8 << 30, max_active_views=2)
    for o, p in items:
        sv.submit(o, p)
    gc.collect()
    for _ in range(20):
        bytearray(1 << 20)  # churn the host heap while the jobs run
    sv.close()
    assert not sv._inflight
    eng.close()


def test_one_slot_runs_jobs_one_by_one():
    eng = pmm.Engine(0)
    items = orbit_jobs()[:4]
    want = alone_maps(items, eng)
    sv = pmm.PatchMatchServer(eng, hbm_budget_bytes=8 << 30, max_active_views=1)
    jobs = [sv.submit(o, p) for o, p in items]
    for k, j in enumerate(jobs):
        j.wait()
        diff = helpers.maps_equal(j.maps(), want[k], MAPS)
        assert all(d == 0 for d in diff.values()), (k, diff)
    assert sv.stats().max_active_seen == 1
    sv.close()
    eng.close()


def test_oversized_job_is_rejected_at_submit():
    eng = pmm.Engine(0)
    reserve = workspace_reserve(eng)
    sv = pmm.PatchMatchServer(eng, hbm_budget_bytes=reserve + (1 << 20), max_active_views=2)
    o = helpers.c1_options(num_iterations=2)
    p = scene.to_problem(helpers.small_planes(seed=1, width=640, height=480, f=500.0))
    with pytest.raises(abi.PMError) as ei:
        sv.submit(o, p)
    assert "PM_ERR_OVER_BUDGET" in str(ei.value)
    sv.close()
    eng.close()


def test_geometric_job_through_the_server():
    eng = pmm.Engine(0)
    d = helpers.small_planes(seed=4)
    pd = scene.to_problem(d, with_gt_maps=True)
    rng = np.random.default_rng(1)
    pd.depth_maps = [x * (1 + 0.01 * rng.standard_normal(x.shape)).astype(np.float32) for x in pd.depth_maps]
    o = helpers.c1_options(num_iterations=3, geom_consistency=True)
    want = alone_maps([(o, pd)], eng)[0]
    sv = pmm.PatchMatchServer(eng, hbm_budget_bytes=4 << 30, max_active_views=4)
    j = sv.submit(o, pd)
    j.wait()
    diff = helpers.maps_equal(j.maps(), want, MAPS)
    assert all(x == 0 for x in diff.values()), diff
    sv.close()
    eng.close()


def test_cache_key_of_another_size_is_rejected():
    eng = pmm.Engine(0)
    items = orbit_jobs()
    sv = pmm.PatchMatchServer(eng, hbm_budget_bytes=4 << 30, max_active_views=2)
    o, p = items[0]
    a = sv.submit(o, p, maps=("depth",), cache_ids=[1000 + k for k in range(len(p.src_image_idxs))])
    a.wait()
    # the same key for a source of another size: an error, not a cache hit
    o2, p2 = next((o2, p2) for o2, p2 in items[1:]
                  if p2.images[p2.src_image_idxs[0]].bitmap.shape != p.images[p.src_image_idxs[0]].bitmap.shape)
    with pytest.raises(abi.PMError) as ei:
        sv.submit(o2, p2, cache_ids=[1000] + [0] * (len(p2.src_image_idxs) - 1))
    assert ei.value.status == abi.PM_ERR_INVALID_ARG
    print("[server] rejected:", ei.value, flush=True)
    # one key for two sources of different sizes in one job
    srcs = [(k, p2.images[i].bitmap.shape) for k, i in enumerate(p2.src_image_idxs)]
    pair = next(((k1, k2) for k1, s1 in srcs for k2, s2 in srcs if k1 < k2 and s1 != s2), None)
    if pair is not None:
        ids = [0] * len(srcs)
        ids[pair[0]] = ids[pair[1]] = 2000
        with pytest.raises(abi.PMError) as ei:
            sv.submit(o2, p2, cache_ids=ids)
        assert ei.value.status == abi.PM_ERR_INVALID_ARG
    # the server still serves
    b = sv.submit(o2, p2, maps=("depth",))
    b.wait()
    sv.close()


def test_jobs_keep_their_results_after_close():
    eng = pmm.Engine(0)
    items = orbit_jobs()[:4]
    want = alone_maps(items, eng)
    sv = pmm.PatchMatchServer(eng, hbm_budget_bytes=4 << 30, max_active_views=2)
    jobs = [sv.submit(o, p) for o, p in items]
    sv.close()  # finishes every job; each keeps its maps and status
    for k, j in enumerate(jobs):
        assert j.done()
        st = j.wait()
        assert st.status == abi.PM_OK and j.wait() is st
        diff = helpers.maps_equal(j.maps(), want[k], MAPS)
        assert all(d == 0 for d in diff.values()), (k, diff)
