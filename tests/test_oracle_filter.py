"""RunFilter's transcription (oracle/pm_oracle.c::pmo_run_filter,
FilterDepthMapsKernel / DeduplicDepthMapsKernel) pinned by the geometry:
ground-truth maps of a rendered scene are consistent, so the filter keeps
them and deduplication removes them; a depth map scaled by 30% is not
explained by its sources and is removed (the cost is a reprojection error in
pixels, so with this scene's small baselines a few percent still pass); the number kept never grows with
filter_min_num_consistent."""
import numpy as np

from oracle import binding
from sensemap_amd import patch_match as pmm
from sensemap_amd import scene
from tests import helpers


def problem(scale=1.0):
    sc = helpers.small_planes(seed=5)
    prob = scene.to_problem(sc, with_gt_maps=True)
    if scale != 1.0:
        prob.depth_maps = list(prob.depth_maps)
        prob.depth_maps[prob.ref_image_idx] = prob.depth_maps[prob.ref_image_idx] * np.float32(scale)
    return sc, prob


def opts(**kw):
    return helpers.c1_options(geom_consistency=True, init_depth_random=False, **kw)


def run(prob, o, dedup=False):
    return binding.run_filter(o, pmm.ViewDesc(prob, o), dedup)


def test_ground_truth_is_kept_and_deduplicated():
    sc, prob = problem()
    gt = sc.ref.depth
    kept = run(prob, opts())
    valid = gt > 0
    assert np.all(kept[kept > 0] == gt[kept > 0])
    assert (kept > 0).sum() > 0.9 * valid.sum()
    dedup = run(prob, opts(), dedup=True)
    assert (dedup > 0).sum() < 0.1 * valid.sum()
    assert np.all(dedup[(kept > 0)] == 0) or ((dedup > 0) & (kept > 0)).sum() == 0


def test_scaled_depth_is_removed():
    sc, prob = problem(scale=1.3)
    out = run(prob, opts())
    assert (out > 0).sum() < 0.05 * (sc.ref.depth > 0).sum()


def test_kept_count_is_monotone_in_min_consistent():
    sc, prob = problem()
    counts = [int((run(prob, opts(filter_min_num_consistent=m)) > 0).sum()) for m in (1, 2, 3, 4, 9)]
    assert all(a >= b for a, b in zip(counts, counts[1:]))
    assert counts[-1] == counts[-2]  # capped at the 4 sources
