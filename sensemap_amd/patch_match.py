"""Python mirror of SenseMap's PatchMatch engine API over the C-ABI.

Names follow the reference (src/controllers/patch_match_options.h,
src/mvs/patch_match_cuda.h): ``PatchMatchOptions``, ``Problem``, ``Image``
and ``PatchMatchACMMCuda`` with ``Run()``, ``GetDepthMap()``,
``GetNormalMap()``, ``GetConfMap()``.  The engine itself is the gfx950 HIP
library; this module only marshals numpy arrays into the C structs.
"""
from __future__ import annotations

import atexit
import ctypes as C
import math
import os
import weakref
from dataclasses import dataclass, field, fields
from typing import List, Optional, Sequence

import numpy as np

from . import abi


@dataclass
class PatchMatchOptions:
    """Engine fields of PatchMatchOptions (patch_match_options.h:42-304), same defaults."""
    window_radius: int = 5
    window_step: int = 2
    sigma_spatial: float = 5.0
    sigma_color: float = 0.2
    num_samples: int = 15
    num_iterations: int = 5
    geom_consistency: bool = True
    geom_consistency_regularizer: float = 0.3
    geom_consistency_max_cost: float = 3.0
    est_curvature: bool = False
    random_optimization: bool = True
    local_optimization: bool = True
    random_depth_ratio: float = 0.004
    random_angle1_range: float = math.radians(20.0)
    random_angle2_range: float = math.radians(12.0)
    init_ncc_matching_cost: float = 0.8
    max_ncc_matching_cost: float = 1.2
    alpha: float = 90.0
    beta: float = 0.3
    thk: int = 4
    num_good_hypothesis: int = 2
    num_bad_hypothesis: int = 3
    init_depth_random: Optional[bool] = None  # None = !geom_consistency, as the reference
    filter: bool = True
    conf_filter: bool = False
    median_filter: bool = False
    conf_threshold: float = 0.5
    depth_diff_threshold: float = 0.02
    filter_geom_consistency_max_cost: float = 1.0
    filter_min_num_consistent: int = 1
    propagate_depth: bool = False  # not supported: rejected by Check() / the engine
    has_prior_depth: bool = False  # not supported: rejected by Check() / the engine
    # extensions
    init_samples: int = 1
    compat_lo_geo_bug: bool = True
    image_storage: int = abi.PM_STORAGE_F16
    accum_mode: int = abi.PM_ACCUM_PLAIN
    rcp_mode: int = abi.PM_RCP_EXACT
    seed: int = 0

    def to_c(self) -> abi.pm_options:
        o = abi.pm_options()
        for f in fields(self):
            v = getattr(self, f.name)
            if f.name == "init_depth_random" and v is None:
                v = not self.geom_consistency
            setattr(o, f.name, int(v) if isinstance(v, bool) else v)
        return o

    def Check(self) -> bool:
        st = abi.load_engine().pm_options_check(C.byref(self.to_c()))
        return st == abi.PM_OK


@dataclass
class Image:
    """mvs::Image: u8 grey bitmap, K = (fx, cx, fy, cy), world->camera R, T."""
    bitmap: np.ndarray
    K: Sequence[float]
    R: np.ndarray
    T: np.ndarray

    def GetWidth(self) -> int:
        return int(self.bitmap.shape[1])

    def GetHeight(self) -> int:
        return int(self.bitmap.shape[0])


@dataclass
class Problem:
    """Problem (patch_match_options.h:306-353) with the controller-owned
    vectors given directly."""
    ref_image_idx: int
    src_image_idxs: List[int]
    images: List[Image]
    depth_min: float
    depth_max: float
    depth_maps: Optional[list] = None    # per image index, float32 (H, W) (geometric pass)
    normal_maps: Optional[list] = None   # per image index, float32 (3, H, W)
    mask_maps: Optional[list] = None     # per image index, uint8 (H, W) or None (Bitmap masks)


def _camera(img: Image) -> abi.pm_camera:
    c = abi.pm_camera()
    for i in range(4):
        c.K[i] = float(img.K[i])
    R = np.asarray(img.R, dtype=np.float64).reshape(9)
    T = np.asarray(img.T, dtype=np.float64).reshape(3)
    for i in range(9):
        c.R[i] = float(R[i])
    for i in range(3):
        c.T[i] = float(T[i])
    return c


class ViewDesc:
    """Owns the numpy buffers behind one pm_view_desc."""

    def __init__(self, problem: Problem, options: PatchMatchOptions):
        imgs = problem.images
        ref = imgs[problem.ref_image_idx]
        self._keep = []
        self.ref_bitmap = np.ascontiguousarray(ref.bitmap, dtype=np.uint8)
        self._keep.append(self.ref_bitmap)
        d = abi.pm_view_desc()
        d.view_id = int(problem.ref_image_idx)
        d.num_src = len(problem.src_image_idxs)
        d.ref = self._image(self.ref_bitmap)
        d.ref_cam = _camera(ref)
        n = len(problem.src_image_idxs)
        srcs = (abi.pm_image * max(n, 1))()
        cams = (abi.pm_camera * max(n, 1))()
        for k, idx in enumerate(problem.src_image_idxs):
            b = np.ascontiguousarray(imgs[idx].bitmap, dtype=np.uint8)
            self._keep.append(b)
            srcs[k] = self._image(b, cache_id=int(idx) + 1)
            cams[k] = _camera(imgs[idx])
        self._keep += [srcs, cams]
        d.src = C.cast(srcs, C.POINTER(abi.pm_image))
        d.src_cams = C.cast(cams, C.POINTER(abi.pm_camera))
        d.depth_min = float(problem.depth_min)
        d.depth_max = float(problem.depth_max)
        if options.geom_consistency:
            ptrs = (C.POINTER(C.c_float) * max(n, 1))()
            for k, idx in enumerate(problem.src_image_idxs):
                dm = np.ascontiguousarray(problem.depth_maps[idx], dtype=np.float32)
                self._keep.append(dm)
                ptrs[k] = dm.ctypes.data_as(C.POINTER(C.c_float))
            self._keep.append(ptrs)
            d.src_depth_maps = C.cast(ptrs, C.POINTER(C.POINTER(C.c_float)))
        init_random = options.init_depth_random
        if init_random is None:
            init_random = not options.geom_consistency
        if not init_random:
            dm = np.ascontiguousarray(problem.depth_maps[problem.ref_image_idx], dtype=np.float32)
            nm = np.ascontiguousarray(problem.normal_maps[problem.ref_image_idx], dtype=np.float32)
            self._keep += [dm, nm]
            d.init_depth = dm.ctypes.data_as(C.POINTER(C.c_float))
            d.init_normal = nm.ctypes.data_as(C.POINTER(C.c_float))
        mask = problem.mask_maps[problem.ref_image_idx] if problem.mask_maps is not None else None
        if mask is not None:
            mb = np.ascontiguousarray(mask, dtype=np.uint8)
            self._keep.append(mb)
            d.ref_mask = self._image(mb)
        self.desc = d
        self.width = ref.GetWidth()
        self.height = ref.GetHeight()
        self.num_src = n

    @staticmethod
    def _image(b: np.ndarray, cache_id: int = 0) -> abi.pm_image:
        """cache_id > 0: the image's key in a serving loop's source cache
        (here the Problem's image index + 1)."""
        im = abi.pm_image()
        im.cache_id = int(cache_id)
        im.data = b.ctypes.data_as(C.POINTER(C.c_uint8))
        im.height, im.width = int(b.shape[0]), int(b.shape[1])
        im.stride = int(b.strides[0])
        return im


class Engine:
    """A device and its HIP stream (pm_engine)."""

    def __init__(self, device: int = 0):
        self.lib = abi.load_engine()
        h = C.c_void_p()
        abi.check(self.lib.pm_engine_create(int(device), C.byref(h)))
        self.handle = h

    def set_counting(self, enable: bool) -> None:
        abi.check(self.lib.pm_engine_set_counting(self.handle, int(enable)))

    def set_tuning(self, phase: int, options: PatchMatchOptions, num_src: int, tile_lw: int,
                   blocks_per_cu: int = 0) -> None:
        """Launch shape of one kernel class (pm_engine_set_tuning)."""
        t = abi.pm_tuning(int(tile_lw), int(blocks_per_cu))
        o = options.to_c()
        abi.check(self.lib.pm_engine_set_tuning(self.handle, int(phase), C.byref(o), int(num_src), C.byref(t)))

    def get_tuning(self, phase: int, options: PatchMatchOptions, num_src: int):
        t = abi.pm_tuning()
        o = options.to_c()
        abi.check(self.lib.pm_engine_get_tuning(self.handle, int(phase), C.byref(o), int(num_src), C.byref(t)))
        return t.tile_lw, t.blocks_per_cu

    def save_tuning(self, path: str) -> None:
        abi.check(self.lib.pm_engine_save_tuning(self.handle, os.fsencode(path)))

    def load_tuning(self, path: str) -> None:
        abi.check(self.lib.pm_engine_load_tuning(self.handle, os.fsencode(path)))

    def close(self) -> None:
        if self.handle:
            self.lib.pm_engine_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_default_engine: Optional[Engine] = None


def default_engine() -> Engine:
    global _default_engine
    if _default_engine is None:
        _default_engine = Engine(0)
    return _default_engine


class PatchMatchACMMCuda:
    """Same surface as SenseMap's PatchMatchACMMCuda (patch_match_cuda.h:154-203):
    the constructor uploads the problem, Run() runs the iteration schedule,
    Get*Map() read the results back."""

    def __init__(self, options: PatchMatchOptions, problem: Problem, engine: Optional[Engine] = None):
        self.engine = engine or default_engine()
        self.lib = self.engine.lib
        self.options = options
        self.view = ViewDesc(problem, options)
        self._copt = options.to_c()
        h = C.c_void_p()
        abi.check(self.lib.pm_problem_create(self.engine.handle, C.byref(self._copt), C.byref(self.view.desc),
                                             C.byref(h)))
        self.handle = h
        self.W, self.H, self.N = self.view.width, self.view.height, self.view.num_src

    def Run(self) -> bool:
        abi.check(self.lib.pm_problem_run(self.handle))
        return True

    def _get(self, fn, shape, dtype):
        out = np.empty(shape, dtype=dtype)
        ctype = {np.float32: C.c_float, np.uint32: C.c_uint32, np.uint16: C.c_uint16}[dtype]
        abi.check(fn(self.handle, out.ctypes.data_as(C.POINTER(ctype))))
        return out

    def GetDepthMap(self) -> np.ndarray:
        return self._get(self.lib.pm_problem_get_depth_map, (self.H, self.W), np.float32)

    def GetNormalMap(self) -> np.ndarray:
        return self._get(self.lib.pm_problem_get_normal_map, (3, self.H, self.W), np.float32)

    def GetConfMap(self) -> np.ndarray:
        return self._get(self.lib.pm_problem_get_conf_map, (self.H, self.W), np.float32)

    def GetCostMap(self) -> np.ndarray:
        return self._get(self.lib.pm_problem_get_cost_map, (self.N, self.H, self.W), np.float32)

    def GetViewSelMap(self) -> np.ndarray:
        return self._get(self.lib.pm_problem_get_view_sel, (self.H, self.W), np.uint32)

    def GetCurvatureMap(self) -> np.ndarray:
        """2 x H x W uint16 {gradient, curvature} (patch_match_cuda.cu:444-446);
        raises unless the run had geom_consistency and est_curvature."""
        return self._get(self.lib.pm_problem_get_curvature_map, (2, self.H, self.W), np.uint16)

    def GetStats(self) -> abi.pm_run_stats:
        s = abi.pm_run_stats()
        abi.check(self.lib.pm_problem_get_stats(self.handle, C.byref(s)))
        return s

    def GetPoses(self):
        poses = np.zeros((self.N, abi.PM_POSE_STRIDE), dtype=np.float32)
        k8 = np.zeros(8, dtype=np.float32)
        abi.check(self.lib.pm_problem_get_poses(self.handle, poses.ctypes.data_as(C.POINTER(C.c_float)),
                                                k8.ctypes.data_as(C.POINTER(C.c_float))))
        return poses, k8

    def InitFrom(self, other: "PatchMatchACMMCuda") -> None:
        """Initial depth/normal from another problem's current maps (on device)."""
        abi.check(self.lib.pm_problem_init_from(self.handle, other.handle))

    def RunFilter(self, deduplication_flag: bool = False) -> bool:
        """RunFilter (patch_match_cuda.cu:1074-1108) on the view's given depth map."""
        abi.check(self.lib.pm_problem_run_filter(self.handle, int(bool(deduplication_flag))))
        return True

    def SetSrcDepthFrom(self, src_index: int, other: "PatchMatchACMMCuda") -> None:
        """Source `src_index`'s depth map from another problem's current depth map (on device)."""
        abi.check(self.lib.pm_problem_set_src_depth(self.handle, int(src_index), other.handle))

    def GetNormalCodes(self) -> np.ndarray:
        """The normal map in NormalMap's uint32 file encoding, encoded on the device."""
        return self._get(self.lib.pm_problem_get_normal_codes, (self.H, self.W), np.uint32)

    def WriteMaps(self, depth_path: Optional[str] = None, normal_path: Optional[str] = None) -> None:
        """DepthMap::Write / NormalMap::Write of the current maps."""
        abi.check(self.lib.pm_problem_write_maps(self.handle, os.fsencode(depth_path) if depth_path else None,
                                                 os.fsencode(normal_path) if normal_path else None))

    def Autotune(self) -> abi.pm_tune_report:
        """Time every launch shape on this view and keep the fastest in the
        engine (pm_problem_autotune); the maps are undefined until the next Run()."""
        r = abi.pm_tune_report()
        abi.check(self.lib.pm_problem_autotune(self.handle, C.byref(r)))
        return r

    def close(self) -> None:
        if getattr(self, "handle", None):
            self.lib.pm_problem_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class _ProblemView:
    """Read-only accessors on a problem handle owned by a batch.  The view
    keeps its batch alive: the handle is freed with the batch."""

    def __init__(self, batch, handle, W, H, N):
        self._batch = batch
        self.lib, self.handle, self.W, self.H, self.N = batch.lib, handle, W, H, N

    _get = PatchMatchACMMCuda._get
    GetDepthMap = PatchMatchACMMCuda.GetDepthMap
    GetNormalMap = PatchMatchACMMCuda.GetNormalMap
    GetConfMap = PatchMatchACMMCuda.GetConfMap
    GetCostMap = PatchMatchACMMCuda.GetCostMap
    GetViewSelMap = PatchMatchACMMCuda.GetViewSelMap
    GetCurvatureMap = PatchMatchACMMCuda.GetCurvatureMap


class PatchMatchBatch:
    """Batched form (pm_batch_*): many views of uneven size, source count and
    window shape advanced together, one launch per phase and kernel shape.
    items: list of (PatchMatchOptions, Problem)."""

    def __init__(self, items, engine: Optional[Engine] = None, graph: bool = False, autotune: bool = False):
        self.engine = engine or default_engine()
        self.lib = self.engine.lib
        self.descs = [ViewDesc(prob, opt) for opt, prob in items]
        n = len(items)
        self._copts = (abi.pm_options * n)(*[opt.to_c() for opt, _ in items])
        self._cdescs = (abi.pm_view_desc * n)(*[d.desc for d in self.descs])
        h = C.c_void_p()
        abi.check(self.lib.pm_batch_create(self.engine.handle, n, self._copts, self._cdescs, C.byref(h)))
        self.handle = h
        if graph:
            abi.check(self.lib.pm_batch_set_graph(self.handle, 1))
        if autotune:
            abi.check(self.lib.pm_batch_set_autotune(self.handle, 1))
        self.views = []
        for i, d in enumerate(self.descs):
            vh = C.c_void_p()
            abi.check(self.lib.pm_batch_get_view(self.handle, i, C.byref(vh)))
            self.views.append(_ProblemView(self, vh, d.width, d.height, d.num_src))

    def Run(self) -> bool:
        abi.check(self.lib.pm_batch_run(self.handle))
        return True

    def GetStats(self) -> abi.pm_run_stats:
        s = abi.pm_run_stats()
        abi.check(self.lib.pm_batch_get_stats(self.handle, C.byref(s)))
        return s

    def close(self) -> None:
        if getattr(self, "handle", None):
            self.lib.pm_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class ServerJob:
    """One depth-map job of a PatchMatchServer: owns the view's input buffers
    and the output maps the server writes."""

    MAPS = ("depth", "normal", "conf", "cost", "view_sel", "curvature")

    def __init__(self, server: "PatchMatchServer", options: PatchMatchOptions, problem: Problem, maps=None,
                 cache_ids=None):
        """maps: the names of the maps to return (default all); the others
        are neither allocated nor copied back.  cache_ids: per source, the
        key the server caches it under (> 0; <= 0 never cached); by default
        the Problem's image index + 1, so one server serves one image set."""
        self.server = server
        self.options = options
        self.desc = ViewDesc(problem, options)
        if cache_ids is not None:
            if len(cache_ids) != self.desc.num_src:
                raise ValueError("need one cache id per source")
            for k, cid in enumerate(cache_ids):
                self.desc.desc.src[k].cache_id = int(cid)
        W, H, N = self.desc.width, self.desc.height, self.desc.num_src
        want = set(self.MAPS if maps is None else maps)
        unknown = want - set(self.MAPS)
        if unknown:
            raise ValueError(f"unknown maps {sorted(unknown)}")
        if not (options.geom_consistency and options.est_curvature):
            want.discard("curvature")
        shapes = {"depth": ((H, W), np.float32), "normal": ((3, H, W), np.float32), "conf": ((H, W), np.float32),
                  "cost": ((N, H, W), np.float32), "view_sel": ((H, W), np.uint32),
                  "curvature": ((2, H, W), np.uint16)}
        ctypes_of = {np.float32: C.c_float, np.uint32: C.c_uint32, np.uint16: C.c_uint16}
        self._out = abi.pm_outputs()
        for name in self.MAPS:
            arr = None
            if name in want:
                shape, dt = shapes[name]
                arr = np.empty(shape, dt)
                setattr(self._out, name, arr.ctypes.data_as(C.POINTER(ctypes_of[dt])))
            setattr(self, name, arr)
        self._copt = options.to_c()
        t = C.c_uint64()
        abi.check(server.lib.pm_server_submit(server.handle, C.byref(self._copt), C.byref(self.desc.desc),
                                              C.byref(self._out), C.byref(t)))
        self.ticket = t.value
        # the server borrows this job's buffers until pm_server_wait returns:
        # it holds the job until then, whatever the caller keeps
        server._inflight[self.ticket] = self
        self.stats: Optional[abi.pm_job_stats] = None
        self._error: Optional[abi.PMError] = None

    def done(self) -> bool:
        if self.stats is not None:
            return True
        d = C.c_int32()
        abi.check(self.server.lib.pm_server_poll(self.server.handle, self.ticket, C.byref(d)))
        return bool(d.value)

    def _finish(self) -> None:
        """pm_server_wait once; keeps the stats and the job's error."""
        st = abi.pm_job_stats()
        rc = self.server.lib.pm_server_wait(self.server.handle, self.ticket, C.byref(st))
        self.stats = st
        if rc != abi.PM_OK:
            self._error = abi.PMError(rc, self.server.lib.pm_last_error().decode(errors="replace"))
        self.server._inflight.pop(self.ticket, None)

    def wait(self) -> abi.pm_job_stats:
        """Block until the maps are written; raises if the job failed (on
        every call, also after the server is closed)."""
        if self.stats is None:
            self._finish()
        if self._error is not None:
            raise self._error
        return self.stats

    def maps(self) -> dict:
        self.wait()
        return {name: getattr(self, name) for name in self.MAPS if getattr(self, name) is not None}


_live_servers: "weakref.WeakSet[PatchMatchServer]" = weakref.WeakSet()


@atexit.register
def _close_live_servers() -> None:
    """At interpreter exit, before module teardown: close every open server,
    so each finishes its jobs and joins its threads while the jobs' buffers
    and the HIP runtime are still there (CPython does not finalise every
    object alive at exit)."""
    for sv in list(_live_servers):
        try:
            sv.close()
        except Exception:
            pass


class PatchMatchServer:
    """Serving loop (pm_server_*): jobs admitted under an HBM budget, source
    images cached on the device, phases of all admitted jobs batched into
    shared launches.  Uses the engine's launch shapes."""

    def __init__(self, engine: Optional[Engine] = None, hbm_budget_bytes: int = 16 << 30,
                 max_active_views: int = 4):
        self.engine = engine or default_engine()
        self.lib = self.engine.lib
        self._inflight = {}  # ticket -> ServerJob not yet waited for
        cfg = abi.pm_server_config(int(hbm_budget_bytes), int(max_active_views), 0)
        h = C.c_void_p()
        abi.check(self.lib.pm_server_create(self.engine.handle, C.byref(cfg), C.byref(h)))
        self.handle = h
        _live_servers.add(self)

    def submit(self, options: PatchMatchOptions, problem: Problem, maps=None, cache_ids=None) -> ServerJob:
        """Queue a job; maps: the names of the maps to return (default all);
        cache_ids: the sources' cache keys (see ServerJob)."""
        return ServerJob(self, options, problem, maps, cache_ids)

    def stats(self) -> abi.pm_server_stats:
        s = abi.pm_server_stats()
        abi.check(self.lib.pm_server_get_stats(self.handle, C.byref(s)))
        return s

    def close(self) -> None:
        """Finish every submitted job (each keeps its maps and status for
        its wait()), then free the server."""
        if getattr(self, "handle", None):
            for job in list(self._inflight.values()):
                job._finish()
            self.lib.pm_server_destroy(self.handle)
            self.handle = None
        self._inflight = {}

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
