"""Deduplication after the cross filter (CrossFilterCuda::DedupRun) on the
device equals its transcription bit for bit: visibility words, deduplicated
depth maps and delt maps, on the designed plane scenes and on an orbit whose
views overlap in chains (each view reads the words earlier views wrote) and
in independent sets (views that share one launch)."""
import numpy as np
import pytest

from oracle import binding
from sensemap_amd import scene
from sensemap_amd.cross_filter import CrossFilterOptions, CrossImage, cross_filter, dedup
from tests import test_oracle_dedup as designed

pytestmark = pytest.mark.gpu


def same(a, b):
    for key in ("visibility", "dedup_depth", "delt"):
        for k, (x, y) in enumerate(zip(a[key], b[key])):
            xb = x.view(np.uint32) if x.dtype == np.float32 else x
            yb = y.view(np.uint32) if y.dtype == np.float32 else y
            n = int(np.sum(xb != yb))
            assert n == 0, (key, k, n)


def test_designed_scenes_equal_transcription():
    im = designed.image
    cases = [
        ([(im(0.0), True), (im(0.5), False)], [(0, [1])], None),
        ([(im(0.0), True), (im(0.5), True)], [(0, [1]), (1, [0])], None),
        ([(im(0.0), True), (im(0.5), True)], [(1, [0]), (0, [1])], None),
        ([(im(0.0), True), (im(0.5), False), (im(-0.5, 1.5), False)], [(0, [1, 2])],
         CrossFilterOptions(filter_min_num_consistent=2)),
        ([(im(0.0), True), (im(0.5, 1.010), False), (im(-0.5, 1.001), False), (im(0.3, 1.005), False)],
         [(0, [1, 2, 3])], None),
    ]
    for imgs, views, opts in cases:
        got = dedup(imgs, views, opts)
        want = binding.cross_dedup(imgs, views, opts)
        same(got, want)
    print("[dedup] designed scenes equal", flush=True)


def test_orbit_dedup_equals_transcription():
    o = scene.orbit_scene(seed=8, num_cams=16, min_px=100 * 75, max_px=240 * 180, grid=32, extent=20.0)
    rng = np.random.default_rng(4)
    ims = []
    for k, v in enumerate(o.views):
        d = v.depth.copy()
        if k % 3 == 1:
            d = (d * (1 + 0.004 * rng.standard_normal(d.shape))).astype(np.float32)
        ims.append(CrossImage(v.camera.K, v.camera.R, v.camera.T, d, v.normal))
    # refs 0..11 with ring neighbours: consecutive views overlap (chains);
    # the last views are far apart on the ring (independent, one launch)
    views = [(k, [(k + d) % 16 for d in (-2, -1, 1, 2, 3)][: 2 + k % 4]) for k in range(12)]
    views += [(12, [13]), (14, [15])]
    refined = cross_filter(ims, views)
    ref_of = {r: w for (r, _), w in zip(views, refined)}
    dims = [(CrossImage(im.K, im.R, im.T, ref_of.get(k, im.depth), im.normal), k in ref_of)
            for k, im in enumerate(ims)]
    for opts in (CrossFilterOptions(), CrossFilterOptions(filter_min_num_consistent=2, depth_diff_threshold=0.01)):
        got = dedup(dims, views, opts)
        want = binding.cross_dedup(dims, views, opts)
        kept = [int((v != 0).sum()) for v in got["visibility"][:12]]
        print("[dedup] orbit kept per view", kept, flush=True)
        assert sum(kept) > 0
        same(got, want)
    # the order matters: reversed views give other words, both equal to the transcription
    rv = views[::-1]
    got = dedup(dims, rv)
    same(got, binding.cross_dedup(dims, rv))
    assert any(not np.array_equal(a, b) for a, b in zip(got["visibility"], dedup(dims, views)["visibility"]))
