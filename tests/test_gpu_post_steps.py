"""The steps of Run() after the iterations on the device -- ComputeCurvatureMap
(GetCurvatureMap), the median filter and MaskDepthMapKernel -- called through
the C-ABI, against the CPU transcription: bitwise on the engine's own maps,
on a given normal field whose float dot products round above 1 (NaN angles),
in the batched form and in the serving loop."""
import numpy as np
import pytest

from oracle import binding
from sensemap_amd import abi, scene
from sensemap_amd import patch_match as pmm
from tests import helpers
from tests.test_oracle_curvature import over_unit_normal

pytestmark = pytest.mark.gpu


def geom_problem(seed=4):
    sc = helpers.small_planes(seed=seed)
    prob = scene.to_problem(sc, with_gt_maps=True)
    rng = np.random.default_rng(0)
    prob.depth_maps = [d * (1 + 0.01 * rng.standard_normal(d.shape)).astype(np.float32) for d in prob.depth_maps]
    return sc, prob


def masked_problem(seed=6):
    sc = helpers.small_planes(seed=seed)
    prob = scene.to_problem(sc)
    rng = np.random.default_rng(seed)
    H, W = sc.ref.image.shape
    mask = (rng.random((H, W)) > 0.25).astype(np.uint8) * 255
    mask[30:60, 40:90] = 0
    prob.mask_maps = [mask] + [None] * len(prob.src_image_idxs)
    return sc, prob, mask


def maps_of(e, curvature=False):
    m = {"depth": e.GetDepthMap(), "normal": e.GetNormalMap(), "conf": e.GetConfMap(), "cost": e.GetCostMap(),
         "view_sel": e.GetViewSelMap()}
    if curvature:
        m["curvature"] = e.GetCurvatureMap()
    return m


def test_curvature_map_bitwise_on_a_geometric_run():
    sc, prob = geom_problem()
    opts = helpers.c1_options(geom_consistency=True, num_iterations=3, est_curvature=True)
    print("[curv] geometric run 160x120 N=4 it=3", flush=True)
    e = pmm.PatchMatchACMMCuda(opts, prob)
    e.Run()
    gpu = maps_of(e, curvature=True)
    cpu = binding.run_view(opts, e.view)
    e.close()
    diff = helpers.maps_equal(gpu, cpu, ("depth", "normal", "conf", "cost", "view_sel", "curvature"))
    print("[curv] differing elements", diff, flush=True)
    assert all(v == 0 for v in diff.values()), diff
    assert np.array_equal(gpu["curvature"], binding.curvature_map(sc.ref.image, gpu["normal"]))
    assert gpu["curvature"][1].max() > 0 and gpu["curvature"][0].max() > 0


def test_curvature_of_a_given_normal_field_with_nan_angles():
    sc, prob = geom_problem(seed=7)
    H, W = sc.ref.image.shape
    u = over_unit_normal()
    w = np.array([0.3, 0.1, -0.95])
    w = (w / np.linalg.norm(w)).astype(np.float32)
    rng = np.random.default_rng(2)
    field = np.empty((3, H, W), np.float32)
    field[:] = u[:, None, None]
    pick = rng.random((H, W)) < 0.4
    field[:, pick] = w[:, None]
    # around (60, 80): the first 12 neighbours in window order hold w, the
    # rest and the centre u; the median angle is NaN, stored as 0
    order = [(y, x) for y in range(-2, 3) for x in range(-2, 3) if (y, x) != (0, 0)]
    field[:, 58:63, 78:83] = u[:, None, None]
    for (dy, dx) in order[:12]:
        field[:, 60 + dy, 80 + dx] = w
    prob.normal_maps = list(prob.normal_maps)
    prob.normal_maps[0] = field
    prob.depth_maps = list(prob.depth_maps)
    prob.depth_maps[0] = np.maximum(prob.depth_maps[0], np.float32(1.0))
    opts = helpers.c1_options(geom_consistency=True, init_depth_random=False, num_iterations=0, est_curvature=True)
    print("[curv] given normal field, no iterations", flush=True)
    e = pmm.PatchMatchACMMCuda(opts, prob)
    e.Run()
    normal = e.GetNormalMap()
    curv = e.GetCurvatureMap()
    e.close()
    assert np.array_equal(normal.view(np.uint32), field.view(np.uint32))  # the init kept the given field
    want = binding.curvature_map(sc.ref.image, field)
    assert np.array_equal(curv, want)
    assert curv[1, 60, 80] == 0


def test_mask_map_bitwise():
    sc, prob, mask = masked_problem()
    opts = helpers.c1_options(num_iterations=3, conf_filter=True)
    print("[mask] 160x120 N=4 it=3 conf_filter", flush=True)
    e = pmm.PatchMatchACMMCuda(opts, prob)
    e.Run()
    gpu = maps_of(e)
    cpu = binding.run_view(opts, e.view)
    e.close()
    diff = helpers.maps_equal(gpu, cpu)
    print("[mask] differing elements", diff, flush=True)
    assert all(v == 0 for v in diff.values()), diff
    plain_prob = scene.to_problem(sc)
    e2 = pmm.PatchMatchACMMCuda(opts, plain_prob)
    e2.Run()
    plain = e2.GetDepthMap()
    e2.close()
    assert np.array_equal(gpu["depth"].view(np.uint32), binding.mask_depth(mask, plain).view(np.uint32))
    assert np.all(gpu["depth"][:-1, :-1][mask[:-1, :-1] == 0] == 0)


def test_median_filter_bitwise():
    sc = helpers.small_planes(seed=8)
    prob = scene.to_problem(sc)
    opts = helpers.c1_options(num_iterations=3, median_filter=True)
    print("[median] 160x120 N=4 it=3", flush=True)
    e = pmm.PatchMatchACMMCuda(opts, prob)
    e.Run()
    gpu = maps_of(e)
    cpu = binding.run_view(opts, e.view)
    e.close()
    diff = helpers.maps_equal(gpu, cpu)
    print("[median] differing elements", diff, flush=True)
    assert all(v == 0 for v in diff.values()), diff
    plain_opts = helpers.c1_options(num_iterations=3)
    e2 = pmm.PatchMatchACMMCuda(plain_opts, prob)
    e2.Run()
    d, n = binding.median_filter(e2.GetDepthMap(), e2.GetNormalMap(), opts.depth_diff_threshold)
    e2.close()
    assert np.array_equal(gpu["depth"].view(np.uint32), d.view(np.uint32))
    assert np.array_equal(gpu["normal"].view(np.uint32), n.view(np.uint32))


def test_curvature_getter_needs_the_option():
    sc = helpers.small_planes()
    e = pmm.PatchMatchACMMCuda(helpers.c1_options(num_iterations=0), scene.to_problem(sc))
    e.Run()
    with pytest.raises(abi.PMError):
        e.GetCurvatureMap()
    e.close()


def test_batch_and_server_carry_curvature_and_mask():
    _, gprob = geom_problem()
    gopts = helpers.c1_options(geom_consistency=True, num_iterations=3, est_curvature=True, median_filter=True)
    # one scene for both jobs: the serving loop caches source images by their
    # index, and images under one key must be identical
    _, mprob, _ = masked_problem(seed=4)
    mopts = helpers.c1_options(num_iterations=3, conf_filter=True)
    single = []
    for o, p in ((gopts, gprob), (mopts, mprob)):
        e = pmm.PatchMatchACMMCuda(o, p)
        e.Run()
        single.append(maps_of(e, curvature=o.est_curvature))
        e.close()
    print("[curv] batch with graph", flush=True)
    b = pmm.PatchMatchBatch([(gopts, gprob), (mopts, mprob)], graph=True)
    for rep in range(2):
        b.Run()
        got = [maps_of(b.views[0], curvature=True), maps_of(b.views[1])]
        for g, s in zip(got, single):
            diff = helpers.maps_equal(g, s, tuple(s.keys()))
            assert all(v == 0 for v in diff.values()), (rep, diff)
    b.close()
    print("[curv] server", flush=True)
    sv = pmm.PatchMatchServer(hbm_budget_bytes=2 << 30, max_active_views=2)
    jobs = [sv.submit(gopts, gprob), sv.submit(mopts, mprob)]
    for j, s in zip(jobs, single):
        diff = helpers.maps_equal(j.maps(), s, tuple(s.keys()))
        assert all(v == 0 for v in diff.values()), diff
    sv.close()
