"""Depth-map fusion (pm_fusion_run): StereoFusion's walk over the depth and
normal maps (src/mvs/fusion.cc:1730-2183) into fused points, with the image
order of StereoFusion::Run (fusion.cc:697-727, 1150-1215)."""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import List, Optional, Sequence

import numpy as np

from . import abi


@dataclass
class StereoFusionOptions:
    """The StereoFusion::Options fields the walk reads (fusion.h:47-90), same defaults."""
    step_size: int = 1
    min_num_pixels: int = 5
    max_num_pixels: int = 10000
    min_num_visible_images: int = 3
    max_traversal_depth: int = 100
    with_normal: bool = True
    max_reproj_error: float = 2.0
    max_depth_error: float = float(np.float32(0.01))   # 0.01f stored in a double
    max_normal_error: float = float(np.float32(10.0))

    def to_c(self) -> abi.pm_fusion_options:
        return abi.pm_fusion_options(int(self.step_size), int(self.min_num_pixels), int(self.max_num_pixels),
                                     int(self.min_num_visible_images), int(self.max_traversal_depth),
                                     int(bool(self.with_normal)), float(self.max_reproj_error),
                                     float(self.max_depth_error), float(self.max_normal_error))


@dataclass
class FusionImage:
    """One image: camera (fx, cx, fy, cy), world->camera R, T; depth (H, W);
    normal (3, H, W) in the camera frame or None; colour (H, W, 3) or grey
    (H, W) uint8; its overlapping images (Model::GetMaxOverlappingImages)."""
    K: Sequence[float]
    R: np.ndarray
    T: np.ndarray
    depth: np.ndarray
    normal: Optional[np.ndarray]
    color: np.ndarray
    overlapping: Sequence[int]
    is_rig: bool = False
    used: bool = True


def _camera(im) -> abi.pm_camera:
    c = abi.pm_camera()
    for k in range(4):
        c.K[k] = float(im.K[k])
    R = np.asarray(im.R, np.float64).reshape(9)
    T = np.asarray(im.T, np.float64).reshape(3)
    for k in range(9):
        c.R[k] = float(R[k])
    for k in range(3):
        c.T[k] = float(T[k])
    return c


class FusionInputs:
    """Owns the buffers behind one call."""

    def __init__(self, images: List[FusionImage]):
        self._keep = []
        self.images = (abi.pm_fusion_image * len(images))()
        for i, im in enumerate(images):
            d = np.ascontiguousarray(im.depth, np.float32)
            H, W = d.shape
            col = np.ascontiguousarray(np.asarray(im.color, np.uint8))  # (H, W, 3) or grey (H, W)
            ov = np.ascontiguousarray(np.asarray(list(im.overlapping), np.int32).reshape(-1))
            self._keep += [d, col, ov]
            c = self.images[i]
            c.cam = _camera(im)
            c.width, c.height = W, H
            c.depth = d.ctypes.data_as(C.POINTER(C.c_float))
            if im.normal is not None:
                n = np.ascontiguousarray(im.normal, np.float32)
                self._keep.append(n)
                c.normal = n.ctypes.data_as(C.POINTER(C.c_float))
            c.rgb = col.ctypes.data_as(C.POINTER(C.c_uint8))
            c.channels = 1 if col.ndim == 2 else 3
            c.overlapping = ov.ctypes.data_as(C.POINTER(C.c_int32))
            c.num_overlapping = len(ov)
            c.is_rig = int(bool(im.is_rig))
            c.used = int(bool(im.used))


def unpack(n: int, nv: int, get) -> dict:
    """Reads the points of a call through get(xyz, normal, rgb, score, offsets, images)."""
    xyz = np.zeros((n, 3), np.float32)
    nrm = np.zeros((n, 3), np.float32)
    rgb = np.zeros((n, 3), np.uint8)
    score = np.zeros(n, np.float32)
    off = np.zeros(n + 1, np.int64)
    vis = np.zeros(max(nv, 1), np.int32)
    get(xyz.ctypes.data, nrm.ctypes.data, rgb.ctypes.data, score.ctypes.data, off.ctypes.data, vis.ctypes.data)
    return {"xyz": xyz, "normal": nrm, "rgb": rgb, "score": score, "vis_offsets": off, "vis_images": vis[:nv]}


def fuse(images: List[FusionImage], options: Optional[StereoFusionOptions] = None, engine=None) -> dict:
    """Fused points: xyz (n, 3), normal (n, 3), rgb (n, 3), score (n,) and the
    visibility as CSR (vis_offsets (n + 1,), vis_images)."""
    from .patch_match import default_engine
    eng = engine or default_engine()
    opt = (options or StereoFusionOptions()).to_c()
    inp = FusionInputs(images)
    h = C.c_void_p()
    abi.check(eng.lib.pm_fusion_run(eng.handle, C.byref(opt), len(images), inp.images, C.byref(h)))
    try:
        n, nv = C.c_int64(), C.c_int64()
        abi.check(eng.lib.pm_fusion_get_counts(h, C.byref(n), C.byref(nv)))
        return unpack(n.value, nv.value, lambda *a: abi.check(eng.lib.pm_fusion_get_points(h, *a)))
    finally:
        eng.lib.pm_fusion_destroy(h)


def fusion_stats(images: List[FusionImage], options: Optional[StereoFusionOptions] = None, engine=None):
    """pm_fusion_stats of one fusion call (walks, walks that accepted pixels,
    dataflow launches, walks run alone, points, wall time)."""
    from .patch_match import default_engine
    eng = engine or default_engine()
    opt = (options or StereoFusionOptions()).to_c()
    inp = FusionInputs(images)
    h = C.c_void_p()
    abi.check(eng.lib.pm_fusion_run(eng.handle, C.byref(opt), len(images), inp.images, C.byref(h)))
    try:
        s = abi.pm_fusion_stats()
        abi.check(eng.lib.pm_fusion_get_stats(h, C.byref(s)))
        return s
    finally:
        eng.lib.pm_fusion_destroy(h)
