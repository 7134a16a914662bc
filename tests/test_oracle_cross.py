"""The cross-view filter's transcription (oracle/pm_oracle.c::pmo_cross_filter,
CrossFilterKernel kernel_functions.cu:3561-3778) pinned by what the geometry
fixes: rendered ground-truth maps of one surface are mutually consistent
and are kept with their depth; a view whose depth is scaled by 5% or whose
normals face away is rejected; the border band the kernel skips is zero; a
source identical to the reference returns the reference depth itself (one
consistent source: the median of the single sorted value)."""
import functools

import numpy as np
import pytest

from oracle import binding
from sensemap_amd import scene
from sensemap_amd.cross_filter import CrossFilterOptions, CrossImage


@functools.lru_cache(maxsize=None)
def orbit():
    return scene.orbit_scene(seed=7, num_cams=20, min_px=120 * 90, max_px=200 * 150, grid=32, extent=20.0)


def images_of(views):
    return [CrossImage(v.camera.K, v.camera.R, v.camera.T, v.depth, v.normal) for v in views]


def neighbours(k, n, m=20):
    return [(k + d) % m for d in range(-n // 2, n // 2 + 1) if d != 0]


def test_consistent_ground_truth_is_kept():
    o = orbit()
    ims = images_of(o.views)
    views = [(k, neighbours(k, 4)) for k in range(0, 20, 5)]
    out = binding.cross_filter(ims, views)
    for (ref, _), r in zip(views, out):
        gt = o.views[ref].depth
        inner = np.zeros_like(gt, bool)
        inner[3:-2, 3:-2] = True
        valid = inner & (gt > 0)
        kept = valid & (r > 0)
        assert kept.sum() > 0.6 * valid.sum(), (ref, kept.sum(), valid.sum())
        rel = np.abs(r[kept] - gt[kept]) / gt[kept]
        # nearest-texel sampling of the source depth on sloped surfaces
        assert np.median(rel) < 2e-3 and rel.max() < 0.02
        assert np.all(r[~inner] == 0) and np.all(r[gt == 0] == 0)


def test_scaled_depth_and_flipped_normals_are_rejected():
    o = orbit()
    ims = images_of(o.views)
    ref = 3
    scaled = list(ims)
    scaled[ref] = CrossImage(ims[ref].K, ims[ref].R, ims[ref].T, ims[ref].depth * np.float32(1.05), ims[ref].normal)
    out = binding.cross_filter(scaled, [(ref, neighbours(ref, 4))])[0]
    base = binding.cross_filter(ims, [(ref, neighbours(ref, 4))])[0]
    assert (out > 0).sum() < 0.05 * (base > 0).sum()
    flipped = list(ims)
    flipped[ref] = CrossImage(ims[ref].K, ims[ref].R, ims[ref].T, ims[ref].depth, -ims[ref].normal)
    out = binding.cross_filter(flipped, [(ref, neighbours(ref, 4))])[0]
    assert np.all(out == 0)


def test_identical_source_returns_the_reference_depth():
    o = orbit()
    v = o.views[2]
    ims = images_of([v, v])
    out = binding.cross_filter(ims, [(0, [1])])[0]
    gt = v.depth
    kept = out > 0
    assert kept.sum() > 0.5 * (gt > 0).sum()
    assert np.array_equal(out[kept], gt[kept])


def test_min_consistent_is_capped_by_the_source_count():
    o = orbit()
    ims = images_of(o.views)
    views = [(6, neighbours(6, 2))]
    strict = binding.cross_filter(ims, views, CrossFilterOptions(filter_min_num_consistent=5))[0]
    both = binding.cross_filter(ims, views, CrossFilterOptions(filter_min_num_consistent=2))[0]
    assert np.array_equal(strict, both)
    loose = binding.cross_filter(ims, views, CrossFilterOptions(filter_min_num_consistent=1))[0]
    assert (loose > 0).sum() >= (both > 0).sum()
