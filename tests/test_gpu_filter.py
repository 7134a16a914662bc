"""RunFilter on the device equals its transcription bit for bit (filter and
deduplication, default and strict consistency counts), on ground-truth and
noisy depth maps."""
import numpy as np
import pytest

from oracle import binding
from sensemap_amd import patch_match as pmm
from sensemap_amd import scene
from tests import helpers

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("noise", [0.0, 0.02])
def test_run_filter_equals_transcription(noise):
    sc = helpers.small_planes(seed=5)
    prob = scene.to_problem(sc, with_gt_maps=True)
    rng = np.random.default_rng(2)
    prob.depth_maps = [(d * (1 + noise * rng.standard_normal(d.shape))).astype(np.float32) for d in prob.depth_maps]
    for m in (1, 3):
        o = helpers.c1_options(geom_consistency=True, init_depth_random=False, filter_min_num_consistent=m)
        for dedup in (False, True):
            e = pmm.PatchMatchACMMCuda(o, prob)
            e.RunFilter(dedup)
            got = e.GetDepthMap()
            e.close()
            want = binding.run_filter(o, pmm.ViewDesc(prob, o), dedup)
            diff = int(np.sum(got.view(np.uint32) != want.view(np.uint32)))
            print("[filter] noise", noise, "min", m, "dedup", dedup, "kept", int((got > 0).sum()), "diff", diff,
                  flush=True)
            assert diff == 0
