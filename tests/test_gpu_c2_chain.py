"""The 1920x1080 workload's exact kernel configuration, at 256x144 so that the
transcription finishes in seconds, compared with the transcription bit for
bit on every map of both passes:

- heightfield scene, 10 sources, window 15x15 step 1 (sigma_s 3, sigma_c
  0.2), thk 2, f16 image storage (the default), exact reciprocal;
- 8 photometric iterations with local optimisation (from the fourth);
- then 2 geometric iterations fed on the device, as the throughput step
  does: the reference's photometric maps through InitFrom and every
  source's photometric depth map through SetSrcDepthFrom
  (patch_match_cuda.cu:318-368, kernel_functions.cu:505-568).

The transcription gets the same inputs from the host: the device's source
depth maps read back and the photometric pass's maps (themselves checked
bitwise first)."""
import numpy as np
import pytest

from oracle import binding
from sensemap_amd import abi, scene
from sensemap_amd import patch_match as pmm
from tests import helpers

pytestmark = pytest.mark.gpu

MAPS = ("depth", "normal", "conf", "cost", "view_sel")


def opts(geom=False, **kw):
    base = dict(window_radius=7, window_step=1, sigma_spatial=3.0, sigma_color=0.2,
                num_iterations=2 if geom else 8, geom_consistency=geom, init_depth_random=not geom, thk=2, seed=2,
                image_storage=abi.PM_STORAGE_F16)
    base.update(kw)
    return pmm.PatchMatchOptions(**base)


def grab(e):
    return {"depth": e.GetDepthMap(), "normal": e.GetNormalMap(), "conf": e.GetConfMap(), "cost": e.GetCostMap(),
            "view_sel": e.GetViewSelMap()}


def test_c2_chain_bitwise_at_256x144():
    sc = scene.heightfield_scene(seed=2, width=256, height=144, f=200.0, num_src=10)
    views = [sc.ref] + sc.srcs
    n = len(views)
    images = [pmm.Image(v.image, v.camera.K, v.camera.R, v.camera.T) for v in views]
    eng = pmm.Engine(0)
    print("[c2] 10 source photometric passes on the device (2 iterations)", flush=True)
    srcs = []
    for k in range(1, n):
        p = pmm.Problem(k, [j for j in range(n) if j != k], images, sc.depth_min, sc.depth_max)
        s = pmm.PatchMatchACMMCuda(opts(num_iterations=2), p, eng)
        s.Run()
        srcs.append(s)
    ref = pmm.Problem(0, list(range(1, n)), images, sc.depth_min, sc.depth_max)
    photo = pmm.PatchMatchACMMCuda(opts(), ref, eng)
    photo.Run()
    pm_gpu = grab(photo)
    print("[c2] photometric pass: transcription", flush=True)
    pm_cpu = binding.run_view(opts(), photo.view)
    diff = helpers.maps_equal(pm_gpu, pm_cpu, MAPS)
    print("[c2] photometric differing elements", diff, flush=True)
    assert all(v == 0 for v in diff.values()), diff

    print("[c2] geometric pass fed on the device", flush=True)
    placeholder = [np.zeros_like(v.depth) for v in views]
    gprob = pmm.Problem(0, list(range(1, n)), images, sc.depth_min, sc.depth_max, depth_maps=placeholder,
                        normal_maps=[np.zeros((3,) + v.depth.shape, np.float32) for v in views])
    geo = pmm.PatchMatchACMMCuda(opts(geom=True), gprob, eng)
    for i, s in enumerate(srcs):
        geo.SetSrcDepthFrom(i, s)
    geo.InitFrom(photo)
    geo.Run()
    g_gpu = grab(geo)

    src_depths = [s.GetDepthMap() for s in srcs]
    cprob = pmm.Problem(0, list(range(1, n)), images, sc.depth_min, sc.depth_max,
                        depth_maps=[pm_gpu["depth"]] + src_depths,
                        normal_maps=[pm_gpu["normal"]] + [np.zeros((3,) + d.shape, np.float32) for d in src_depths])
    og = opts(geom=True)
    print("[c2] geometric pass: transcription", flush=True)
    g_cpu = binding.run_view(og, pmm.ViewDesc(cprob, og))
    diff = helpers.maps_equal(g_gpu, g_cpu, MAPS)
    print("[c2] geometric differing elements", diff, flush=True)
    assert all(v == 0 for v in diff.values()), diff
    # the pass did work: it moved depths, and stays close to the ground truth
    assert np.count_nonzero(g_gpu["depth"] != pm_gpu["depth"]) > 0.05 * g_gpu["depth"].size
    err = helpers.depth_errors(g_gpu["depth"], sc.ref.depth)
    print("[c2] median rel. depth error photometric", float(np.median(helpers.depth_errors(pm_gpu["depth"],
          sc.ref.depth))), "geometric", float(np.median(err)), flush=True)
    assert np.median(err) < 5e-3
    for s in srcs + [photo, geo]:
        s.close()
