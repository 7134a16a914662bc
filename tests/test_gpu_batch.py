"""Batched form: a mixed batch (uneven sizes, source counts, window shapes,
photometric and geometric views) gives every view the same maps, bit for
bit, as running that view alone; hipGraph replays equal the eager run."""
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


def mixed_items():
    items = []
    a = helpers.small_planes(seed=1)
    items.append((helpers.c1_options(num_iterations=5), scene.to_problem(a)))
    b = helpers.small_planes(seed=2, width=200, height=90, f=140.0, num_src=3)
    items.append((helpers.c1_options(num_iterations=5, window_radius=7, window_step=1), scene.to_problem(b)))
    c = helpers.small_heightfield(seed=3, width=150, height=130, f=120.0, num_src=6)
    items.append((helpers.c1_options(num_iterations=3, thk=4, image_storage=abi.PM_STORAGE_F16), scene.to_problem(c)))
    d = helpers.small_planes(seed=4)
    pd = scene.to_problem(d, with_gt_maps=True)
    rng = np.random.default_rng(1)
    pd.depth_maps = [x * (1 + 0.01 * rng.standard_normal(x.shape)).astype(np.float32) for x in pd.depth_maps]
    items.append((helpers.c1_options(num_iterations=4, geom_consistency=True), pd))
    e = helpers.small_planes(seed=5, width=64, height=40, f=60.0, num_src=2)
    items.append((helpers.c1_options(num_iterations=6, window_radius=4, window_step=2), scene.to_problem(e)))
    return items


def test_batch_equals_views_alone():
    items = mixed_items()
    print("[batch] views", [p.images[0].bitmap.shape for _, p in items], flush=True)
    alone = []
    for opt, prob in items:
        e = pmm.PatchMatchACMMCuda(opt, prob)
        e.Run()
        alone.append(grab(e))
        e.close()
    b = pmm.PatchMatchBatch(items)
    b.Run()
    for k, v in enumerate(b.views):
        diff = helpers.maps_equal(grab(v), alone[k], MAPS)
        print("[batch] view", k, diff, flush=True)
        assert all(x == 0 for x in diff.values()), (k, diff)
    st = b.GetStats()
    assert st.sweep_launches > 0 and st.total_ms > 0


def test_batch_graph_replays_equal_eager():
    items = mixed_items()
    eager = pmm.PatchMatchBatch(items)
    eager.Run()
    want = [grab(v) for v in eager.views]
    g = pmm.PatchMatchBatch(items, graph=True)
    for rep in range(3):
        g.Run()
        # the program's own eager work between replays
        x = pmm.PatchMatchACMMCuda(items[0][0], items[0][1])
        x.Run()
        x.close()
        for k, v in enumerate(g.views):
            diff = helpers.maps_equal(grab(v), want[k], MAPS)
            assert all(d == 0 for d in diff.values()), (rep, k, diff)


def shape_engine(items, lw, bpc):
    """A fresh engine whose every kernel class of `items` uses one launch shape."""
    eng = pmm.Engine(0)
    for opt, prob in items:
        for ph in (abi.PM_PHASE_INIT, abi.PM_PHASE_SWEEP, abi.PM_PHASE_LOCAL, abi.PM_PHASE_FILTER):
            eng.set_tuning(ph, opt, len(prob.src_image_idxs), lw, bpc)
    return eng


def test_launch_shapes_give_identical_maps():
    opt = helpers.c1_options(num_iterations=4)
    prob = scene.to_problem(helpers.small_planes(seed=7, width=150, height=100, f=120.0))
    base = pmm.PatchMatchACMMCuda(opt, prob)
    base.Run()
    want = grab(base)
    base.close()
    for lw in (4, 5, 6):
        for bpc in (0, 1):
            print("[shape] tile_lw", lw, "blocks_per_cu", bpc, flush=True)
            eng = shape_engine([(opt, prob)], lw, bpc)
            e = pmm.PatchMatchACMMCuda(opt, prob, eng)
            e.Run()
            diff = helpers.maps_equal(grab(e), want, MAPS)
            e.close()
            eng.close()
            assert all(d == 0 for d in diff.values()), (lw, bpc, diff)


def test_batch_with_other_shapes_equals_views_alone():
    items = mixed_items()
    alone = []
    for opt, prob in items:
        e = pmm.PatchMatchACMMCuda(opt, prob)
        e.Run()
        alone.append(grab(e))
        e.close()
    for lw, bpc in ((6, 0), (4, 1)):
        print("[shape] batch tile_lw", lw, "blocks_per_cu", bpc, flush=True)
        eng = shape_engine(items, lw, bpc)
        b = pmm.PatchMatchBatch(items, engine=eng, graph=True)
        b.Run()
        for k, v in enumerate(b.views):
            diff = helpers.maps_equal(grab(v), alone[k], MAPS)
            assert all(d == 0 for d in diff.values()), (lw, bpc, k, diff)
        b.close()
        eng.close()


def test_autotune_keeps_maps_and_round_trips(tmp_path):
    opt = helpers.c1_options(num_iterations=5)
    prob = scene.to_problem(helpers.small_heightfield(seed=3, width=256, height=160, f=200.0, num_src=5))
    n = len(prob.src_image_idxs)
    base = pmm.PatchMatchACMMCuda(opt, prob)
    base.Run()
    want = grab(base)
    base.close()
    eng = pmm.Engine(0)
    e = pmm.PatchMatchACMMCuda(opt, prob, eng)
    r = e.Autotune()
    print("[autotune] sweep", r.sweep.tile_lw, r.sweep.blocks_per_cu, r.sweep_ms_default, r.sweep_ms_best,
          "local", r.local_opt.tile_lw, r.local_opt.blocks_per_cu, flush=True)
    assert r.candidates >= 3
    assert 4 <= r.sweep.tile_lw <= 6 and r.sweep.blocks_per_cu >= 0
    assert 0 < r.sweep_ms_best <= r.sweep_ms_default
    assert eng.get_tuning(abi.PM_PHASE_SWEEP, opt, n) == (r.sweep.tile_lw, r.sweep.blocks_per_cu)
    e.Run()
    diff = helpers.maps_equal(grab(e), want, MAPS)
    assert all(d == 0 for d in diff.values()), diff
    e.close()
    path = str(tmp_path / "shapes.txt")
    eng.save_tuning(path)
    eng2 = pmm.Engine(0)
    eng2.load_tuning(path)
    assert eng2.get_tuning(abi.PM_PHASE_SWEEP, opt, n) == (r.sweep.tile_lw, r.sweep.blocks_per_cu)
    assert eng2.get_tuning(abi.PM_PHASE_LOCAL, opt, n) == (r.local_opt.tile_lw, r.local_opt.blocks_per_cu)
    eng.close()
    eng2.close()


def test_batch_autotune_equals_views_alone():
    items = mixed_items()
    alone = []
    for opt, prob in items:
        e = pmm.PatchMatchACMMCuda(opt, prob)
        e.Run()
        alone.append(grab(e))
        e.close()
    eng = pmm.Engine(0)
    b = pmm.PatchMatchBatch(items, engine=eng, graph=True, autotune=True)
    for rep in range(2):
        b.Run()
        for k, v in enumerate(b.views):
            diff = helpers.maps_equal(grab(v), alone[k], MAPS)
            assert all(d == 0 for d in diff.values()), (rep, k, diff)
    b.close()
    eng.close()


def test_view_keeps_its_batch_alive():
    import gc
    items = mixed_items()[:2]
    alone = []
    for opt, prob in items:
        e = pmm.PatchMatchACMMCuda(opt, prob)
        e.Run()
        alone.append(grab(e))
        e.close()
    b = pmm.PatchMatchBatch(items)
    b.Run()
    v = b.views[1]
    del b
    gc.collect()
    diff = helpers.maps_equal(grab(v), alone[1], MAPS)
    assert all(x == 0 for x in diff.values()), diff
