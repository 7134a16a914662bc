"""Accuracy against seeded ground truth at the full size of the planes scene
(640x480, 4 sources, 11x11 step 2, 6 iterations) and failure handling of the
GPU entry points."""
import ctypes as C

import numpy as np
import pytest

from sensemap_amd import abi, scene
from sensemap_amd import patch_match as pmm
from tests import helpers

pytestmark = pytest.mark.gpu


def test_planes_640x480_accuracy():
    sc = scene.planes_scene(seed=1)
    print("[accuracy] planes 640x480 N=4", flush=True)
    e = pmm.PatchMatchACMMCuda(helpers.c1_options(init_samples=8), scene.to_problem(sc))
    e.Run()
    err = helpers.depth_errors(e.GetDepthMap(), sc.ref.depth)
    n = e.GetNormalMap()
    print("[accuracy] median", float(np.median(err)), "within1%", float(np.mean(err < 0.01)), flush=True)
    assert np.median(err) < 0.01
    assert np.mean(err < 0.01) > 0.6
    assert np.all(np.isfinite(n))
    st = e.GetStats()
    assert st.sweep_launches == 12 and st.local_opt_launches == 6 and st.sweep_ms > 0


def test_counting_matches_oracle_on_small_view():
    from oracle import binding
    sc = helpers.small_planes()
    opts = helpers.c1_options(num_iterations=4)
    eng = pmm.Engine(0)
    eng.set_counting(True)
    e = pmm.PatchMatchACMMCuda(opts, scene.to_problem(sc), engine=eng)
    e.Run()
    st = e.GetStats()
    cpu = binding.run_view(opts, e.view, count=True)
    assert (st.ncc_evaluations, st.ncc_samples, st.geo_evaluations) == cpu["counters"]
    assert st.ncc_samples == st.ncc_evaluations * 36 or st.ncc_samples < st.ncc_evaluations * 36


def test_invalid_views_are_rejected():
    sc = helpers.small_planes()
    prob = scene.to_problem(sc)
    opts = helpers.c1_options()
    lib = abi.load_engine()
    eng = pmm.default_engine()
    v = pmm.ViewDesc(prob, opts)
    h = C.c_void_p()
    bad = v.desc
    keep = (bad.num_src, bad.depth_min)
    bad.num_src = 0
    assert lib.pm_problem_create(eng.handle, C.byref(opts.to_c()), C.byref(bad), C.byref(h)) == 1
    bad.num_src = 40
    assert lib.pm_problem_create(eng.handle, C.byref(opts.to_c()), C.byref(bad), C.byref(h)) == 4
    bad.num_src = keep[0]
    bad.depth_min = -1.0
    assert lib.pm_problem_create(eng.handle, C.byref(opts.to_c()), C.byref(bad), C.byref(h)) == 1
    bad.depth_min = keep[1]
    bad.ref_cam.K[0] = float("nan")
    assert lib.pm_problem_create(eng.handle, C.byref(opts.to_c()), C.byref(bad), C.byref(h)) == 1
    g = helpers.c1_options(geom_consistency=True, init_depth_random=True)
    v2 = pmm.ViewDesc(prob, helpers.c1_options())
    assert lib.pm_problem_create(eng.handle, C.byref(g.to_c()), C.byref(v2.desc), C.byref(h)) == 1
    # the engine still works afterwards
    e = pmm.PatchMatchACMMCuda(helpers.c1_options(num_iterations=1), prob)
    e.Run()
    assert np.isfinite(e.GetDepthMap()).all()
