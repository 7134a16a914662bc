"""Shared scene/problem builders for the tests."""
from __future__ import annotations

import functools

import numpy as np

from sensemap_amd import scene
from sensemap_amd import patch_match as pmm


@functools.lru_cache(maxsize=None)
def small_planes(seed=1, width=160, height=120, f=125.0, num_src=4, amp=40.0, noise=0.0):
    return scene.planes_scene(seed=seed, width=width, height=height, f=f, num_src=num_src,
                              tex_amplitude=amp, noise_sigma=noise)


@functools.lru_cache(maxsize=None)
def small_heightfield(seed=2, width=192, height=108, f=150.0, num_src=5):
    return scene.heightfield_scene(seed=seed, width=width, height=height, f=f, num_src=num_src)


def c1_options(**kw) -> pmm.PatchMatchOptions:
    """Matching parameters of the 640x480 three-plane workload (11x11 step 2, sigma_s 3, sigma_c 0.2, 6 iterations)."""
    base = dict(window_radius=5, window_step=2, sigma_spatial=3.0, sigma_color=0.2, num_iterations=6,
                geom_consistency=False, thk=2, seed=1)
    base.update(kw)
    return pmm.PatchMatchOptions(**base)


def depth_errors(depth, gt):
    m = gt > 0
    return np.abs(depth[m] - gt[m]) / gt[m]


def maps_equal(a: dict, b: dict, keys=("depth", "normal", "conf", "cost", "view_sel")) -> dict:
    """Per map: number of elements that differ bitwise (NaN == NaN)."""
    out = {}
    for k in keys:
        x, y = np.asarray(a[k]), np.asarray(b[k])
        if x.dtype == np.float32:
            xb, yb = x.view(np.uint32), y.view(np.uint32)
            both_nan = np.isnan(x) & np.isnan(y)
            out[k] = int(np.count_nonzero((xb != yb) & ~both_nan))
        else:
            out[k] = int(np.count_nonzero(x != y))
    return out
