"""The transcription of StereoFusion's walk (oracle/pm_fusion_oracle.c,
fusion.cc:697-727, 1150-1215, 1730-2183) pinned on a designed scene whose
answers follow from geometry.

Five cameras (R = I, f = 100) look at the plane z = 10 from x = 0.1 s_k with
s = (0, 2, -2, 4, -4): a world point seen at column c of image 0 is at
column c - s_k of image k, exactly, so every walk reaches one pixel per
image and every projection lands on a pixel centre.  Depth maps are the
plane's depth (times a per-image factor where stated), normals face the
cameras, and colours are a function of the world point plus a per-image
offset.  Each overlapping list holds the other images in ascending order.

Checked, each against a closed form: which world points become points (all
five images see them within the 3-pixel margin) and their order, positions,
normals and visibility; the depth-first (last in, first out) order through
max_num_pixels; the depth, normal and reprojection thresholds; the median of
an even count (mean of the two middle values) for colours and the score's
depth errors; that the first non-rig image seeds the walks."""
import math

import numpy as np
import pytest

from oracle import binding as B
from sensemap_amd.fusion import FusionImage, StereoFusionOptions

W, H, F, Z = 48, 36, 100.0, 10.0
SHIFTS = (0, 2, -2, 4, -4)
CX, CY = (W - 1) / 2.0, (H - 1) / 2.0


def base_color(wc, r):
    return (wc * 7 + r * 13) % 200 + 20


def scene(depth_scale=None, normal_tilt=None, color_off=None, shifts=SHIFTS, rig=None, width=W):
    n = len(shifts)
    ims = []
    rows, cols = np.mgrid[0:H, 0:width]
    cx = (width - 1) / 2.0
    for k, s in enumerate(shifts):
        wc = cols + s  # world column (image-0 units) seen at this pixel
        off = 0 if color_off is None else color_off[k]
        grey = np.clip(base_color(np.floor(wc).astype(int), rows) + off, 0, 255).astype(np.uint8)
        d = np.full((H, width), Z * (1.0 if depth_scale is None else depth_scale[k]), np.float32)
        nm = np.zeros((3, H, width), np.float32)
        nm[2] = -1.0
        if normal_tilt is not None and normal_tilt[k]:
            a = math.radians(normal_tilt[k])
            nm[0], nm[2] = math.sin(a), -math.cos(a)
        T = np.array([-0.1 * s, 0.0, 0.0])
        ims.append(FusionImage((F, cx, F, CY), np.eye(3), T, d, nm, grey, [j for j in range(n) if j != k],
                               is_rig=bool(rig and rig[k])))
    return ims


def vis_lists(out):
    off = out["vis_offsets"]
    return [list(out["vis_images"][off[i]:off[i + 1]]) for i in range(len(off) - 1)]


def interior_points():
    # world columns every image sees within its margin: 3 <= c - s_k <= W - 4
    c0, c1 = 3 + max(SHIFTS), W - 4 + min(SHIFTS)
    return [(r, c) for r in range(3, H - 3) for c in range(c0, c1 + 1)]


def test_points_of_the_plane():
    out = B.fusion(scene())
    want = interior_points()
    assert len(out["xyz"]) == len(want)
    xyz = np.array([((c - CX) * Z / F, (r - CY) * Z / F, Z) for r, c in want], np.float32)
    assert np.allclose(out["xyz"], xyz, atol=1e-5)
    assert np.allclose(out["normal"], [0.0, 0.0, -1.0], atol=1e-6)
    assert all(v == [0, 1, 2, 3, 4] for v in vis_lists(out))
    assert np.array_equal(out["rgb"][:, 0], [base_color(c, r) for r, c in want])
    assert (out["score"] < 1e-3).all()


def test_walk_is_depth_first():
    # max_num_pixels 3: the seed, then the last image pushed (4), then the
    # last image 4 pushes that is not yet taken (3)
    o = StereoFusionOptions(max_num_pixels=3, min_num_pixels=3)
    out = B.fusion(scene(), o)
    vis = vis_lists(out)
    cols = {(round(float(p[1]) * F / Z + CY), round(float(p[0]) * F / Z + CX)) for p in out["xyz"]}
    inner = [(r, c) for r, c in interior_points()]
    idx = {(round(float(p[1]) * F / Z + CY), round(float(p[0]) * F / Z + CX)): i for i, p in enumerate(out["xyz"])}
    assert set(inner) <= cols
    assert all(vis[idx[rc]] == [0, 3, 4] for rc in inner)


def test_depth_error_threshold_and_even_median_score():
    # image 3's depth 2% off is rejected (error 0.0196 > 0.01)
    o = StereoFusionOptions(min_num_pixels=4)
    out = B.fusion(scene(depth_scale=[1, 1, 1, 1.02, 1]), o)
    assert len(out["xyz"]) > 0 and all(3 not in v for v in vis_lists(out))
    # 0.5% and 0.8% off pass; errors {0, 0, e3, e4}: the even median is e3 / 2
    out = B.fusion(scene(depth_scale=[1, 1, 1, 1.005, 1.008]))
    e3 = 0.005 / 1.005
    want = 0.6 * (e3 / 2) / np.float32(0.01)
    full = [i for i, v in enumerate(vis_lists(out)) if v == [0, 1, 2, 3, 4]]
    assert len(full) == len(interior_points())
    assert np.allclose(out["score"][full], want, rtol=2e-3)


def test_normal_threshold():
    tilt = [0, 0, 15.0, 0, 0]  # 15 degrees off: rejected at max_normal_error 10
    out = B.fusion(scene(normal_tilt=tilt), StereoFusionOptions(min_num_pixels=4))
    assert len(out["xyz"]) > 0 and all(2 not in v for v in vis_lists(out))
    out = B.fusion(scene(normal_tilt=tilt), StereoFusionOptions(min_num_pixels=4, with_normal=False))
    assert sum(v == [0, 1, 2, 3, 4] for v in vis_lists(out)) == len(interior_points())
    assert not out["normal"].any()


def test_reprojection_threshold():
    # image 4 shifted by -3.6 px: its pixel centre is 0.4 px from the projection
    shifts = (0, 2, -2, 4, -3.6)
    strict = B.fusion(scene(shifts=shifts), StereoFusionOptions(min_num_pixels=4, max_reproj_error=0.3))
    assert len(strict["xyz"]) > 0 and all(4 not in v for v in vis_lists(strict))
    loose = B.fusion(scene(shifts=shifts), StereoFusionOptions(min_num_pixels=4, max_reproj_error=0.5))
    assert any(4 in v for v in vis_lists(loose))


@pytest.mark.parametrize("off,want", [((0, 10, 20, 3, 7), 7), ((0, 10, 99, 3, 8), 5.5)])
def test_colour_median(off, want):
    # the second case rejects image 2 by its normal: offsets {0, 10, 3, 8},
    # even median (3 + 8) / 2 = 5.5, rounded half away from zero to 6
    tilt = [0, 0, 15.0, 0, 0] if want == 5.5 else None
    out = B.fusion(scene(color_off=off, normal_tilt=tilt), StereoFusionOptions(min_num_pixels=4))
    pts = interior_points()
    vis = vis_lists(out)
    full = [i for i, v in enumerate(vis) if len(v) == (5 if tilt is None else 4)]
    assert len(full) == len(pts)
    base = np.array([base_color(c, r) for r, c in pts])
    assert np.array_equal(out["rgb"][full, 0], base + math.floor(want + 0.5))


def test_first_non_rig_image_seeds():
    # image 1's depth is 0.4% off; seeded from image 0 the score is ~0 (a
    # single non-zero error among four), seeded from image 1 (image 0 a rig)
    # every error is 0.004 and the score 0.6 * 0.4
    sc = [1, 1.004, 1, 1, 1]
    out0 = B.fusion(scene(depth_scale=sc))
    assert np.median(out0["score"]) < 1e-3
    out1 = B.fusion(scene(depth_scale=sc, rig=[1, 0, 0, 0, 0]))
    assert np.allclose(np.median(out1["score"]), 0.6 * 0.4, rtol=5e-3)


def test_grey_images_equal_their_three_channel_copies():
    # a grey image (channels 1) reads its value as r, g and b: the points equal
    # those of the same images with the grey value stacked into three channels
    import dataclasses
    ims = scene(color_off=[0, 3, -2, 5, 1])
    rgb = [dataclasses.replace(im, color=np.stack([im.color] * 3, axis=-1)) for im in ims]
    o = StereoFusionOptions()
    a, b = B.fusion(ims, o), B.fusion(rgb, o)
    assert len(a["score"]) == len(b["score"]) > 0
    for k in ("xyz", "normal", "rgb", "score", "vis_offsets", "vis_images"):
        assert np.array_equal(a[k], b[k]), k
