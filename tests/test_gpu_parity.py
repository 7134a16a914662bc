"""GPU parity: the gfx950 engine (called through the C-ABI) against the CPU
transcription on the same seeded inputs.  With PM_RCP_EXACT every output map
must be bitwise identical (depth, normal, conf, per-source cost, view_sel)."""
import sys

import numpy as np
import pytest

from oracle import binding
from sensemap_amd import abi, scene
from sensemap_amd import patch_match as pmm
from tests import helpers

pytestmark = pytest.mark.gpu


def run_both(sc, opts, problem=None):
    prob = problem or scene.to_problem(sc)
    print(f"[parity] {sc.ref.image.shape} N={len(prob.src_image_idxs)} r={opts.window_radius} "
          f"s={opts.window_step} it={opts.num_iterations} geom={opts.geom_consistency}", flush=True)
    eng = pmm.PatchMatchACMMCuda(opts, prob)
    eng.Run()
    gpu = {"depth": eng.GetDepthMap(), "normal": eng.GetNormalMap(), "conf": eng.GetConfMap(),
           "cost": eng.GetCostMap(), "view_sel": eng.GetViewSelMap()}
    cpu = binding.run_view(opts, eng.view)
    eng.close()
    return gpu, cpu


def assert_bitwise(gpu, cpu):
    diff = helpers.maps_equal(gpu, cpu)
    print("[parity] differing elements", diff, flush=True)
    assert all(v == 0 for v in diff.values()), diff


def test_poses_bitwise_equal_to_transcription():
    sc = helpers.small_planes()
    opts = helpers.c1_options(num_iterations=0)
    eng = pmm.PatchMatchACMMCuda(opts, scene.to_problem(sc))
    poses, k8 = eng.GetPoses()
    import ctypes as C
    p2 = np.zeros_like(poses)
    k2 = np.zeros_like(k8)
    binding.load().pmo_pack_poses(C.byref(eng.view.desc), p2.ctypes.data_as(C.POINTER(C.c_float)),
                                  k2.ctypes.data_as(C.POINTER(C.c_float)))
    assert np.array_equal(poses.view(np.uint32), p2.view(np.uint32))
    assert np.array_equal(k8.view(np.uint32), k2.view(np.uint32))


def test_init_only_bitwise():
    gpu, cpu = run_both(helpers.small_planes(), helpers.c1_options(num_iterations=0))
    assert_bitwise(gpu, cpu)


def test_c1_shape_bitwise():
    # 11x11 step 2, 6 iterations (local optimisation from iteration 3)
    gpu, cpu = run_both(helpers.small_planes(), helpers.c1_options())
    assert_bitwise(gpu, cpu)


def test_c2_window_bitwise():
    sc = helpers.small_heightfield()
    opts = helpers.c1_options(window_radius=7, window_step=1, sigma_spatial=3.0, num_iterations=4, thk=4)
    gpu, cpu = run_both(sc, opts)
    assert_bitwise(gpu, cpu)


def test_generic_window_shape_bitwise():
    # radius 4 step 2 has no specialised kernel: exercises the runtime-shape path
    opts = helpers.c1_options(window_radius=4, window_step=2, num_iterations=4)
    gpu, cpu = run_both(helpers.small_planes(seed=3), opts)
    assert_bitwise(gpu, cpu)


def test_best_of_k_init_bitwise():
    gpu, cpu = run_both(helpers.small_planes(), helpers.c1_options(init_samples=8, num_iterations=2))
    assert_bitwise(gpu, cpu)


def test_geometric_pass_bitwise():
    sc = helpers.small_planes(seed=4)
    prob = scene.to_problem(sc, with_gt_maps=True)
    # the sources' "photometric" depth maps: ground truth with a little noise
    rng = np.random.default_rng(0)
    prob.depth_maps = [d * (1 + 0.01 * rng.standard_normal(d.shape)).astype(np.float32) for d in prob.depth_maps]
    opts = helpers.c1_options(geom_consistency=True, num_iterations=4, geom_consistency_regularizer=0.3)
    gpu, cpu = run_both(sc, opts, problem=prob)
    assert_bitwise(gpu, cpu)


def test_compensated_moments_bitwise():
    sc = helpers.small_planes(seed=5, amp=2.0, noise=0.5)
    opts = helpers.c1_options(accum_mode=abi.PM_ACCUM_COMPENSATED, num_iterations=3)
    gpu, cpu = run_both(sc, opts)
    assert_bitwise(gpu, cpu)


def test_f16_storage_equals_u8_storage():
    sc = helpers.small_planes()
    prob = scene.to_problem(sc)
    outs = []
    for st in (abi.PM_STORAGE_U8, abi.PM_STORAGE_F16):
        e = pmm.PatchMatchACMMCuda(helpers.c1_options(image_storage=st), prob)
        e.Run()
        outs.append({"depth": e.GetDepthMap(), "normal": e.GetNormalMap(), "conf": e.GetConfMap(),
                     "cost": e.GetCostMap(), "view_sel": e.GetViewSelMap()})
        e.close()
    assert all(v == 0 for v in helpers.maps_equal(outs[0], outs[1]).values())


def test_run_is_deterministic():
    sc = helpers.small_planes()
    e = pmm.PatchMatchACMMCuda(helpers.c1_options(), scene.to_problem(sc))
    e.Run()
    a = e.GetDepthMap()
    e.Run()
    b = e.GetDepthMap()
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_fast_reciprocal_within_tolerance():
    # PM_RCP_FAST (v_rcp + one Newton step, <= 1 ulp from 1/z) is not bitwise;
    # it must give the same accuracy against ground truth.
    sc = helpers.small_planes()
    prob = scene.to_problem(sc)
    res = {}
    for mode in (abi.PM_RCP_EXACT, abi.PM_RCP_FAST):
        e = pmm.PatchMatchACMMCuda(helpers.c1_options(rcp_mode=mode), prob)
        e.Run()
        res[mode] = helpers.depth_errors(e.GetDepthMap(), sc.ref.depth)
        e.close()
    me, mf = np.median(res[abi.PM_RCP_EXACT]), np.median(res[abi.PM_RCP_FAST])
    print("[parity] median rel depth error exact", me, "fast", mf, flush=True)
    assert abs(mf - me) <= 0.01 * max(me, 1e-3) + 2e-4


@pytest.mark.parametrize("width,height,num_src", [(97, 61, 1), (45, 130, 17), (33, 17, 3)])
def test_ragged_shapes_and_source_counts_bitwise(width, height, num_src):
    # sizes that are no multiple of the 32x16 tile (partial tiles, windows
    # over the image border), one source, and more sources than the
    # reference's 8
    sc = scene.planes_scene(seed=5, width=width, height=height, f=0.8 * max(width, height), num_src=num_src)
    gpu, cpu = run_both(sc, helpers.c1_options(num_iterations=4))
    assert_bitwise(gpu, cpu)


def test_maximum_source_count_bitwise():
    # PM_MAX_NUM_SRC = 32 sources: every bit of view_sel is a source
    sc = scene.planes_scene(seed=7, width=64, height=48, f=50.0, num_src=32)
    gpu, cpu = run_both(sc, helpers.c1_options(num_iterations=3))
    assert_bitwise(gpu, cpu)
    assert (np.asarray(gpu["view_sel"]) >> 16).any()


def test_maximum_window_radius_bitwise():
    # PM_MAX_WINDOW_RADIUS = 32 (65x65 window, step 2: 33x33 samples) on the runtime-shape path
    sc = helpers.small_planes(seed=3, width=64, height=48, f=50.0)
    gpu, cpu = run_both(sc, helpers.c1_options(window_radius=32, window_step=2, sigma_spatial=12.0,
                                               num_iterations=2))
    assert_bitwise(gpu, cpu)
