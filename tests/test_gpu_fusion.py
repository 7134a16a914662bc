"""Depth-map fusion on the device (pm_fusion_run) equals its transcription
bit for bit -- positions, normals, colours, scores, visibility and the order
of the points -- on the designed plane scenes, on an orbit of rendered views
with noisy depths, on a scene whose walks outgrow the per-walk buffers
(rerun alone with buffers for max_num_pixels), and on the 11 views of the
1920x1080 heightfield scene."""
import numpy as np
import pytest

from oracle import binding
from sensemap_amd import abi, scene
from sensemap_amd import patch_match as pmm
from sensemap_amd.fusion import FusionImage, StereoFusionOptions, fuse
from tests import test_oracle_fusion as designed

pytestmark = pytest.mark.gpu


def same(a, b):
    assert len(a["score"]) == len(b["score"]), (len(a["score"]), len(b["score"]))
    for k in ("xyz", "normal", "score"):
        x, y = a[k].view(np.uint32), b[k].view(np.uint32)
        n = int(np.sum(x != y))
        assert n == 0, (k, n)
    for k in ("rgb", "vis_offsets", "vis_images"):
        assert np.array_equal(a[k], b[k]), k


def stats(eng, images, options):
    import ctypes as C
    from sensemap_amd.fusion import FusionInputs
    inp = FusionInputs(images)
    h = C.c_void_p()
    abi.check(eng.lib.pm_fusion_run(eng.handle, C.byref(options.to_c()), len(images), inp.images, C.byref(h)))
    s = abi.pm_fusion_stats()
    abi.check(eng.lib.pm_fusion_get_stats(h, C.byref(s)))
    eng.lib.pm_fusion_destroy(h)
    return s


@pytest.mark.parametrize("case", range(7))
def test_designed_scenes(case):
    d = designed
    O = StereoFusionOptions
    cases = [
        (d.scene(), O()),
        (d.scene(), O(max_num_pixels=3, min_num_pixels=3)),
        (d.scene(depth_scale=[1, 1, 1, 1.02, 1]), O(min_num_pixels=4)),
        (d.scene(depth_scale=[1, 1, 1, 1.005, 1.008]), O()),
        (d.scene(normal_tilt=[0, 0, 15.0, 0, 0]), O(min_num_pixels=4, with_normal=False)),
        (d.scene(shifts=(0, 2, -2, 4, -3.6)), O(min_num_pixels=4, max_reproj_error=0.5)),
        (d.scene(depth_scale=[1, 1.004, 1, 1, 1], rig=[1, 0, 0, 0, 0]), O()),
    ]
    ims, o = cases[case]
    got, want = fuse(ims, o), binding.fusion(ims, o)
    print("[fusion] designed case", case, "points", len(got["score"]), flush=True)
    same(got, want)


def orbit_images(seed=8, cams=12, noise=0.002, nbr=6):
    o = scene.orbit_scene(seed=seed, num_cams=cams, min_px=120 * 90, max_px=200 * 150, grid=32, extent=20.0)
    rng = np.random.default_rng(seed)
    ims = []
    n = len(o.views)
    for k, v in enumerate(o.views):
        d = (v.depth * (1 + noise * rng.standard_normal(v.depth.shape))).astype(np.float32)
        ov = [(k + s) % n for s in (1, -1, 2, -2, 3, -3)][:nbr]
        ims.append(FusionImage(v.camera.K, v.camera.R, v.camera.T, d, v.normal, v.image, ov))
    return ims


def test_orbit_equals_transcription():
    eng = pmm.Engine(0)
    ims = orbit_images()
    for o in (StereoFusionOptions(), StereoFusionOptions(min_num_pixels=3, max_depth_error=0.02, step_size=2)):
        got, want = fuse(ims, o, eng), binding.fusion(ims, o)
        s = stats(eng, ims, o)
        print("[fusion] orbit points", len(got["score"]), "walks", s.walks, "kept", s.kept, "rounds", s.rounds,
              "alone", s.big_walks, "ms", round(s.total_ms, 2), flush=True)
        assert len(got["score"]) > 100
        same(got, want)


def test_walks_that_outgrow_their_buffers():
    # 24 cameras of one plane, 0.35 px apart, and a 4-pixel reprojection
    # bound: a walk accepts a disc of pixels in every image, more than a walk
    # graph holds, so those walks run alone
    shifts = tuple(0.35 * k for k in range(24))
    ims = designed.scene(shifts=shifts, width=64)
    o = StereoFusionOptions(max_reproj_error=4.0)
    eng = pmm.Engine(0)
    got, want = fuse(ims, o, eng), binding.fusion(ims, o)
    s = stats(eng, ims, o)
    print("[fusion] dense walks: points", len(got["score"]), "alone", s.big_walks, "rounds", s.rounds, flush=True)
    assert s.big_walks > 0
    same(got, want)


def test_full_size_eleven_views_equal_transcription():
    # the 1920x1080 heightfield scene, all 11 views fused (each overlapping the
    # other 10), depths with 0.2% noise: about 2.7 M walks over 13 dataflow
    # segments, every point bitwise equal to the transcription's
    from sensemap_amd import scene as sc_mod
    sc = sc_mod.heightfield_scene(seed=2)
    views = [sc.ref] + sc.srcs
    n = len(views)
    rng = np.random.default_rng(2)
    ims = []
    for k, v in enumerate(views):
        d = (v.depth * (1 + 0.002 * rng.standard_normal(v.depth.shape))).astype(np.float32)
        ims.append(FusionImage(v.camera.K, v.camera.R, v.camera.T, d, v.normal, v.image,
                               [j for j in range(n) if j != k]))
    o = StereoFusionOptions()
    eng = pmm.Engine(0)
    got = fuse(ims, o, eng)
    s = stats(eng, ims, o)
    print("[fusion] full size: points", len(got["score"]), "walks", s.walks, "kept", s.kept, "rounds", s.rounds,
          "chain", s.chain, "ms", round(s.total_ms, 1), flush=True)
    want = binding.fusion(ims, o)
    assert len(got["score"]) > 100000
    same(got, want)


def test_grey_and_three_channel_images_equal_transcription():
    import dataclasses
    ims = orbit_images(cams=8)
    rgb = [dataclasses.replace(im, color=np.stack([im.color] * 3, axis=-1)) for im in ims]
    o = StereoFusionOptions()
    eng = pmm.Engine(0)
    grey, col = fuse(ims, o, eng), fuse(rgb, o, eng)
    print("[fusion] grey vs three channels: points", len(grey["score"]), flush=True)
    same(grey, col)
    same(grey, binding.fusion(ims, o))
