"""Cross-view filter on the device equals its transcription bit for bit, for
views of different sizes and source counts in one pass, with ground-truth
and with perturbed maps."""
import numpy as np
import pytest

from oracle import binding
from sensemap_amd import scene
from sensemap_amd.cross_filter import CrossFilterOptions, CrossImage, cross_filter
from tests import helpers

pytestmark = pytest.mark.gpu


def test_cross_filter_equals_transcription():
    o = scene.orbit_scene(seed=8, num_cams=16, min_px=100 * 75, max_px=240 * 180, grid=32, extent=20.0)
    rng = np.random.default_rng(4)
    ims = []
    for k, v in enumerate(o.views):
        d = v.depth.copy()
        if k % 3 == 1:  # noisy depth on some views
            d = (d * (1 + 0.01 * rng.standard_normal(d.shape))).astype(np.float32)
        ims.append(CrossImage(v.camera.K, v.camera.R, v.camera.T, d, v.normal))
    views = [(k, [(k + d) % 16 for d in (-3, -2, -1, 1, 2, 3)][: 2 + k % 5]) for k in range(16)]
    for opts in (CrossFilterOptions(), CrossFilterOptions(filter_min_num_consistent=3, max_normal_error=30.0,
                                                          depth_diff_threshold=0.01)):
        got = cross_filter(ims, views, opts)
        want = binding.cross_filter(ims, views, opts)
        for k, (a, b) in enumerate(zip(got, want)):
            diff = int(np.sum(a.view(np.uint32) != b.view(np.uint32)))
            print("[cross] view", k, a.shape, "kept", int((a > 0).sum()), "diff", diff, flush=True)
            assert diff == 0, k


def test_cross_filter_over_device_problems_equals_host_maps():
    """pm_cross_filter_problems (maps already on the device) gives the same
    refined maps as pm_cross_filter on the read-back maps, bit for bit, and
    its write-back replaces each problem's depth map with them."""
    from sensemap_amd import patch_match as pmm
    from sensemap_amd.cross_filter import cross_filter_problems
    sc = scene.planes_scene(seed=6, width=96, height=72, f=80.0, num_src=4)
    views = [sc.ref] + sc.srcs
    n = len(views)
    images = [pmm.Image(v.image, v.camera.K, v.camera.R, v.camera.T) for v in views]
    opts = helpers.c1_options(num_iterations=3)
    probs = []
    for k in range(n):
        prob = pmm.Problem(k, [j for j in range(n) if j != k], images, sc.depth_min, sc.depth_max)
        e = pmm.PatchMatchACMMCuda(opts, prob)
        e.Run()
        probs.append(e)
    cviews = [(k, [j for j in range(n) if j != k][: 2 + k % 3]) for k in range(n)]
    ims = [CrossImage(v.camera.K, v.camera.R, v.camera.T, p.GetDepthMap(), p.GetNormalMap())
           for v, p in zip(views, probs)]
    want = cross_filter(ims, cviews)
    got = cross_filter_problems(probs, cviews)
    for k, (a, b) in enumerate(zip(got, want)):
        diff = int(np.sum(a.view(np.uint32) != b.view(np.uint32)))
        print("[cross-dev] view", k, "kept", int((a > 0).sum()), "diff", diff, flush=True)
        assert diff == 0
    cross_filter_problems(probs, cviews, write_back=True)
    for k, p in enumerate(probs):
        assert np.array_equal(p.GetDepthMap().view(np.uint32), want[k].view(np.uint32))
        p.close()
