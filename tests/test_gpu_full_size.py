"""The 1920x1080 heightfield view with 10 sources and a 15x15 window, the size
the throughput is quoted at, checked through properties that hold at any
size:
- f16 and u8 image storage give the same maps, bit for bit;
- two runs give the same maps, bit for bit;
- the per-source costs the engine stores equal the transcription's FP32 NCC
  at the engine's own final hypotheses (the same bar as the precision test:
  costs of sources the local optimisation did not select may be older);
- the geometric pass, fed the photometric result on the device, keeps the
  depth error against the rendered ground truth small;
- the curvature map, the median filter and a mask after the geometric pass
  equal the transcription's chain applied to the pass's maps without them."""
import functools

import numpy as np
import pytest

from oracle import binding
from sensemap_amd import abi, scene
from sensemap_amd import patch_match as pmm
from tests import helpers

pytestmark = pytest.mark.gpu


@functools.lru_cache(maxsize=None)
def view_scene():
    return scene.heightfield_scene(seed=2)


def opts(geom=False, **kw):
    base = dict(window_radius=7, window_step=1, sigma_spatial=3.0, sigma_color=0.2,
                num_iterations=2 if geom else 8, geom_consistency=geom, init_depth_random=not geom, thk=2, seed=2)
    base.update(kw)
    return pmm.PatchMatchOptions(**base)


def run(o, prob, engine=None):
    e = pmm.PatchMatchACMMCuda(o, prob, engine)
    e.Run()
    m = {"depth": e.GetDepthMap(), "normal": e.GetNormalMap(), "conf": e.GetConfMap(), "cost": e.GetCostMap(),
         "view_sel": e.GetViewSelMap()}
    return e, m


def test_full_size_storage_and_determinism():
    sc = view_scene()
    prob = scene.to_problem(sc)
    print("[full] 1920x1080 N=10 r=7 s=1 it=8: f16, f16 again, u8", flush=True)
    e1, a = run(opts(), prob)
    e1.close()
    e2, b = run(opts(), prob)
    e2.close()
    e3, c = run(opts(image_storage=abi.PM_STORAGE_U8), prob)
    e3.close()
    for other in (b, c):
        diff = helpers.maps_equal(a, other)
        print("[full] differing elements", diff, flush=True)
        assert all(v == 0 for v in diff.values()), diff
    err = helpers.depth_errors(a["depth"], sc.ref.depth)
    assert np.median(err) < 1e-3


def test_full_size_stored_costs_equal_the_transcription():
    sc = view_scene()
    prob = scene.to_problem(sc)
    o = opts()
    print("[full] 1920x1080 photometric pass, then the transcription's costs at its hypotheses", flush=True)
    e, m = run(o, prob)
    view = e.view
    c32, c64 = binding.cost_maps(o, view, m["depth"], m["normal"])
    e.close()
    valid = np.isfinite(c64)
    assert valid.mean() > 0.9
    same = np.mean(m["cost"][valid].view(np.uint32) == c32[valid].view(np.uint32))
    print("[full] stored cost == transcription FP32:", same, flush=True)
    assert same > 0.999


def test_full_size_geometric_pass_on_the_device():
    sc = view_scene()
    views = [sc.ref] + sc.srcs
    images = [pmm.Image(v.image, v.camera.K, v.camera.R, v.camera.T) for v in views]
    eng = pmm.default_engine()
    n = len(views)
    print("[full] 10 source photometric passes (3 iterations), then the reference's two passes", flush=True)
    srcs = []
    for k in range(1, n):
        p = pmm.Problem(k, [j for j in range(n) if j != k], images, sc.depth_min, sc.depth_max)
        s = pmm.PatchMatchACMMCuda(opts(num_iterations=3), p, eng)
        s.Run()
        srcs.append(s)
    ref = pmm.Problem(0, list(range(1, n)), images, sc.depth_min, sc.depth_max,
                      depth_maps=[v.depth for v in views], normal_maps=[v.normal for v in views])
    photo = pmm.PatchMatchACMMCuda(opts(), ref, eng)
    photo.Run()
    geo = pmm.PatchMatchACMMCuda(opts(geom=True), ref, eng)
    for i, s in enumerate(srcs):
        geo.SetSrcDepthFrom(i, s)
    geo.InitFrom(photo)
    geo.Run()
    gt = sc.ref.depth
    err_p = helpers.depth_errors(photo.GetDepthMap(), gt)
    err_g = helpers.depth_errors(geo.GetDepthMap(), gt)
    print("[full] median rel. depth error photometric", float(np.median(err_p)), "geometric",
          float(np.median(err_g)), flush=True)
    assert np.median(err_g) < 1e-3
    assert np.mean(err_g < 0.01) > 0.9
    for s in srcs + [photo, geo]:
        s.close()


def test_full_size_post_run_steps_equal_the_transcription():
    # the curvature map, the median filter and a mask after the geometric
    # pass of the 1920x1080 view, against the transcription's chain applied
    # to the same pass's maps without those steps
    sc = view_scene()
    views = [sc.ref] + sc.srcs
    n = len(views)
    images = [pmm.Image(v.image, v.camera.K, v.camera.R, v.camera.T) for v in views]
    depths = [v.depth for v in views]
    normals = [v.normal for v in views]
    H, W = sc.ref.image.shape
    mask = (np.random.default_rng(3).random((H, W)) > 0.1).astype(np.uint8) * 255
    print("[full] geometric pass with and without the post-run steps", flush=True)
    plain = pmm.Problem(0, list(range(1, n)), images, sc.depth_min, sc.depth_max, depth_maps=depths, normal_maps=normals)
    e, m = run(opts(geom=True), plain)
    e.close()
    post = pmm.Problem(0, list(range(1, n)), images, sc.depth_min, sc.depth_max, depth_maps=depths,
                       normal_maps=normals, mask_maps=[mask] + [None] * (n - 1))
    o = opts(geom=True, est_curvature=True, median_filter=True)
    e, g = run(o, post)
    curv = e.GetCurvatureMap()
    e.close()
    assert np.array_equal(curv, binding.curvature_map(sc.ref.image, m["normal"]))
    d, nrm = binding.median_filter(m["depth"], m["normal"], o.depth_diff_threshold)
    d = binding.mask_depth(mask, d)
    assert np.array_equal(g["depth"].view(np.uint32), d.view(np.uint32))
    assert np.array_equal(g["normal"].view(np.uint32), nrm.view(np.uint32))
    for k in ("conf", "cost", "view_sel"):
        assert np.array_equal(g[k], m[k])
