"""The sparse model ahead of the patch-match path: source-view selection and
depth ranges on the device.

`Model` mirrors mvs::Model (src/mvs/model.h:22-101) as far as the
patch-match setup reads it: the registered images' world->camera poses and
rig flags, and the sparse points with their tracks.  Its two methods are the
ones PatchMatchController calls before building each Problem
(patch_match_controller.cc:3153, 3171, 3294):

  GetMaxOverlappingImages(max_num_images, min_triangulation_angle,
                          unique_image_idxs=None)   model.cc:474-699 / :701-934
  ComputeDepthRanges()                               model.cc:941-960

Both run through the C-ABI (pm_model_max_overlapping_images,
pm_model_depth_ranges).
"""
from __future__ import annotations

import ctypes as C
from typing import Iterable, List, Optional, Sequence, Tuple

import numpy as np

from . import abi


class Model:
    """Images (R (I,3,3), T (I,3), is_rig (I,)) and points (xyz (P,3)) with
    CSR tracks: point p is observed by track_images[track_offsets[p]:
    track_offsets[p+1]]."""

    def __init__(self, R, T, xyz, track_offsets, track_images, is_rig=None):
        self.R = np.ascontiguousarray(np.asarray(R, np.float32).reshape(-1, 9))
        self.T = np.ascontiguousarray(np.asarray(T, np.float32).reshape(-1, 3))
        self.xyz = np.ascontiguousarray(np.asarray(xyz, np.float32).reshape(-1, 3))
        self.track_offsets = np.ascontiguousarray(np.asarray(track_offsets, np.int64))
        self.track_images = np.ascontiguousarray(np.asarray(track_images, np.int32))
        I = self.R.shape[0]
        self.is_rig = None if is_rig is None else np.ascontiguousarray(np.asarray(is_rig, np.uint8).reshape(I))
        if self.T.shape[0] != I:
            raise ValueError("R and T disagree on the number of images")
        if self.track_offsets.shape != (self.xyz.shape[0] + 1,):
            raise ValueError("track_offsets must hold num_points + 1 entries")
        self.last_overlap_stats: Optional[abi.pm_overlap_stats] = None

    @classmethod
    def from_tracks(cls, R, T, xyz, tracks: Sequence[Sequence[int]], is_rig=None) -> "Model":
        """Build from one list of image indices per point (Model::Point::track)."""
        off = np.zeros(len(tracks) + 1, np.int64)
        off[1:] = np.cumsum([len(t) for t in tracks])
        flat = np.fromiter((int(i) for t in tracks for i in t), np.int32, count=int(off[-1]))
        return cls(R, T, xyz, off, flat, is_rig)

    @property
    def num_images(self) -> int:
        return self.R.shape[0]

    @property
    def num_points(self) -> int:
        return self.xyz.shape[0]

    def to_c(self) -> abi.pm_model:
        P = C.POINTER
        m = abi.pm_model()
        m.num_images = self.num_images
        m.R = self.R.ctypes.data_as(P(C.c_float))
        m.T = self.T.ctypes.data_as(P(C.c_float))
        m.is_rig = self.is_rig.ctypes.data_as(P(C.c_uint8)) if self.is_rig is not None else None
        m.num_points = self.num_points
        m.xyz = self.xyz.ctypes.data_as(P(C.c_float))
        m.track_offsets = self.track_offsets.ctypes.data_as(P(C.c_int64))
        m.track_images = self.track_images.ctypes.data_as(P(C.c_int32))
        return m

    def subset_mask(self, unique_image_idxs: Optional[Iterable[int]]) -> Optional[np.ndarray]:
        if unique_image_idxs is None:
            return None
        mask = np.zeros(self.num_images, np.uint8)
        idx = np.fromiter((int(i) for i in unique_image_idxs), np.int64)
        if idx.size and (idx.min() < 0 or idx.max() >= self.num_images):
            raise ValueError("image index out of range in unique_image_idxs")
        mask[idx] = 1
        return mask

    def GetMaxOverlappingImages(self, max_num_images: int, min_triangulation_angle: float,
                                unique_image_idxs: Optional[Iterable[int]] = None, engine=None) -> List[List[int]]:
        """Per image, its ranked source candidates (pm_model_max_overlapping_images).
        With unique_image_idxs, only those images are ranked and only they
        count as co-observers (model.cc:701-934); the others get []."""
        from .patch_match import default_engine
        eng = engine or default_engine()
        I = self.num_images
        mask = self.subset_mask(unique_image_idxs)
        off = np.zeros(I + 1, np.int32)
        max_num_images = min(int(max_num_images), I)  # a list holds at most I - 1 images
        out = np.zeros(max(I * max_num_images, 1), np.int32)
        st = abi.pm_overlap_stats()
        m = self.to_c()
        abi.check(eng.lib.pm_model_max_overlapping_images(
            eng.handle, C.byref(m), int(max_num_images), float(min_triangulation_angle),
            mask.ctypes.data_as(C.POINTER(C.c_uint8)) if mask is not None else None,
            off.ctypes.data_as(C.POINTER(C.c_int32)), out.ctypes.data_as(C.POINTER(C.c_int32)), C.byref(st)))
        self.last_overlap_stats = st
        return [out[off[i]:off[i + 1]].tolist() for i in range(I)]

    def ComputeDepthRanges(self, engine=None) -> List[Tuple[float, float]]:
        """(depth_min, depth_max) per image from the model's bounding box
        (pm_model_depth_ranges)."""
        from .patch_match import default_engine
        eng = engine or default_engine()
        out = np.zeros((self.num_images, 2), np.float32)
        m = self.to_c()
        abi.check(eng.lib.pm_model_depth_ranges(eng.handle, C.byref(m), out.ctypes.data_as(C.POINTER(C.c_float))))
        return [(float(a), float(b)) for a, b in out]


def problem_sources(overlapping: Sequence[int], unique_image_idxs: Optional[Iterable[int]],
                    max_num_src_images: int) -> List[int]:
    """The controller's cut of one image's ranked list into
    Problem::src_image_idxs (patch_match_controller.cc:3300-3319): members of
    the cluster only, at most max_num_src_images."""
    keep = None if unique_image_idxs is None else set(int(i) for i in unique_image_idxs)
    src = [int(i) for i in overlapping if keep is None or int(i) in keep]
    return src[:max(0, int(max_num_src_images))]
