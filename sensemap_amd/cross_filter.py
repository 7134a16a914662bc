"""Cross-view depth filter (pm_cross_filter): CrossFilterCuda::Run
(src/mvs/patch_match_cuda.cu:1493-1551) over a set of views, each checked
against its extended source images' depth and normal maps."""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import List, Optional, Sequence, Tuple

import numpy as np

from . import abi


@dataclass
class CrossFilterOptions:
    """The PatchMatchOptions fields the filter reads (patch_match_options.h:135-141, 231)."""
    filter_geom_consistency_max_cost: float = 1.0
    max_normal_error: float = 15.0
    depth_diff_threshold: float = 0.02
    filter_min_num_consistent: int = 1

    def to_c(self) -> abi.pm_cross_options:
        return abi.pm_cross_options(float(self.filter_geom_consistency_max_cost), float(self.max_normal_error),
                                    float(self.depth_diff_threshold), int(self.filter_min_num_consistent))


@dataclass
class CrossImage:
    """One image: camera (fx, cx, fy, cy), world->camera R, T; its depth map
    (H, W) and its normal map, either (3, H, W) floats or (H, W) uint32
    NormalMap codes."""
    K: Sequence[float]
    R: np.ndarray
    T: np.ndarray
    depth: np.ndarray
    normal: np.ndarray


class CrossInputs:
    """Owns the numpy buffers and ctypes arrays behind one call."""

    def __init__(self, images: List[CrossImage], views: List[Tuple[int, Sequence[int]]]):
        from . import io as pmio
        self._keep = []
        n = len(images)
        self.images = (abi.pm_cross_image * n)()
        self.shapes = []
        for i, im in enumerate(images):
            d = np.ascontiguousarray(im.depth, np.float32)
            nm = np.asarray(im.normal)
            codes = np.ascontiguousarray(nm if nm.dtype == np.uint32 and nm.ndim == 2 else pmio.encode_normals(nm))
            self._keep += [d, codes]
            c = abi.pm_camera()
            for k in range(4):
                c.K[k] = float(im.K[k])
            R = np.asarray(im.R, np.float64).reshape(9)
            T = np.asarray(im.T, np.float64).reshape(3)
            for k in range(9):
                c.R[k] = float(R[k])
            for k in range(3):
                c.T[k] = float(T[k])
            self.images[i].cam = c
            self.images[i].height, self.images[i].width = d.shape
            self.images[i].depth = d.ctypes.data_as(C.POINTER(C.c_float))
            self.images[i].normal_codes = codes.ctypes.data_as(C.POINTER(C.c_uint32))
            self.shapes.append(d.shape)
        self.views = (abi.pm_cross_view * len(views))()
        self.outputs = []
        for k, (ref, srcs) in enumerate(views):
            s = (C.c_int32 * len(srcs))(*[int(x) for x in srcs])
            out = np.zeros(self.shapes[ref], np.float32)
            self._keep += [s]
            self.outputs.append(out)
            self.views[k].ref = int(ref)
            self.views[k].num_src = len(srcs)
            self.views[k].src = C.cast(s, C.POINTER(C.c_int32))
            self.views[k].refined = out.ctypes.data_as(C.POINTER(C.c_float))


class DedupInputs:
    """Owns the buffers behind one deduplication call: images
    (CrossImage, ref_flag) and views (ref, sources); outputs per image
    (visibility words, deduplicated depth) and per view (delt map)."""

    def __init__(self, images: List[Tuple[CrossImage, bool]], views: List[Tuple[int, Sequence[int]]]):
        self._keep = []
        n = len(images)
        self.images = (abi.pm_dedup_image * n)()
        self.visibility, self.dedup_depth, self.delt = [], [], []
        for i, (im, flag) in enumerate(images):
            d = np.ascontiguousarray(im.depth, np.float32)
            vis = np.zeros(d.shape, np.uint32)
            dd = np.zeros(d.shape, np.float32)
            self._keep += [d]
            self.visibility.append(vis)
            self.dedup_depth.append(dd)
            img = self.images[i]
            img.cam = _camera(im)
            img.height, img.width = d.shape
            img.depth = d.ctypes.data_as(C.POINTER(C.c_float))
            img.ref_flag = int(bool(flag))
            img.visibility = vis.ctypes.data_as(C.POINTER(C.c_uint32))
            img.dedup_depth = dd.ctypes.data_as(C.POINTER(C.c_float))
        self.views = (abi.pm_dedup_view * max(len(views), 1))()
        self.num_views = len(views)
        for k, (ref, srcs) in enumerate(views):
            s = (C.c_int32 * max(len(srcs), 1))(*[int(x) for x in srcs])
            out = np.zeros(np.asarray(images[ref][0].depth).shape, np.float32)
            self._keep.append(s)
            self.delt.append(out)
            self.views[k].ref = int(ref)
            self.views[k].num_src = len(srcs)
            self.views[k].src = C.cast(s, C.POINTER(C.c_int32))
            self.views[k].delt = out.ctypes.data_as(C.POINTER(C.c_float))


def _camera(im: CrossImage) -> abi.pm_camera:
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


def dedup(images: List[Tuple[CrossImage, bool]], views: List[Tuple[int, Sequence[int]]],
          options: Optional[CrossFilterOptions] = None, engine=None) -> dict:
    """CrossFilterCuda::DedupRun over views (ref index, source indices) in
    order; images are (CrossImage, ref_flag) with the depth map DedupRun
    reads (normals are not used).  Returns {"visibility": [per image (H, W)
    uint32], "dedup_depth": [per image (H, W)], "delt": [per view (H, W)]}."""
    from .patch_match import default_engine
    eng = engine or default_engine()
    opt = (options or CrossFilterOptions()).to_c()
    inp = DedupInputs(images, views)
    abi.check(eng.lib.pm_cross_dedup(eng.handle, C.byref(opt), len(images), inp.images, inp.num_views, inp.views))
    return {"visibility": inp.visibility, "dedup_depth": inp.dedup_depth, "delt": inp.delt}


def cross_filter(images: List[CrossImage], views: List[Tuple[int, Sequence[int]]],
                 options: Optional[CrossFilterOptions] = None, engine=None) -> List[np.ndarray]:
    """Refined depth map of every view (ref index, source indices), in one
    device pass."""
    from .patch_match import default_engine
    eng = engine or default_engine()
    opt = (options or CrossFilterOptions()).to_c()
    inp = CrossInputs(images, views)
    abi.check(eng.lib.pm_cross_filter(eng.handle, C.byref(opt), len(images), inp.images, len(views), inp.views))
    return inp.outputs


def cross_filter_problems(problems, views: List[Tuple[int, Sequence[int]]],
                          options: Optional[CrossFilterOptions] = None, write_back: bool = False,
                          engine=None) -> List[Optional[np.ndarray]]:
    """The filter over PatchMatchACMMCuda problems whose maps are on the device
    (pm_cross_filter_problems).  views index `problems`.  Returns each view's
    refined map, or, with write_back, replaces each view's problem depth map
    on the device and returns Nones."""
    from .patch_match import default_engine
    eng = engine or default_engine()
    opt = (options or CrossFilterOptions()).to_c()
    handles = (C.c_void_p * len(problems))(*[p.handle.value for p in problems])
    cv = (abi.pm_cross_view * len(views))()
    keep, outs = [], []
    for k, (ref, srcs) in enumerate(views):
        s = (C.c_int32 * len(srcs))(*[int(x) for x in srcs])
        keep.append(s)
        cv[k].ref = int(ref)
        cv[k].num_src = len(srcs)
        cv[k].src = C.cast(s, C.POINTER(C.c_int32))
        if write_back:
            cv[k].refined = None
            outs.append(None)
        else:
            out = np.zeros((problems[ref].H, problems[ref].W), np.float32)
            keep.append(out)
            outs.append(out)
            cv[k].refined = out.ctypes.data_as(C.POINTER(C.c_float))
    abi.check(eng.lib.pm_cross_filter_problems(eng.handle, C.byref(opt), len(problems), handles, len(views), cv))
    return outs
