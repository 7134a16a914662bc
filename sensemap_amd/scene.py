"""Seeded synthetic multi-view scenes with exact ground-truth depth and normals.

Scenes are triangle soups rendered by libpm_scene.so (z-buffer, FP64 ray /
triangle intersection, 3D value-noise texture in world space).  Builders:

  planes_scene       3 seeded planes (a background plane and two bounded
                     quads), 1 reference + ring of sources
  heightfield_scene  64x64 heightfield mesh with Gaussian bumps seen by a
                     pitched camera, depth ~3..30, 1 reference + ring of sources
  orbit_cameras      cameras on a circle around a scene centre

Cameras are pinhole, world->camera R, T, with K = (fx, cx, fy, cy); pixel i
sits at coordinate i.
"""
from __future__ import annotations

import ctypes as C
import math
import os
from dataclasses import dataclass
from typing import List, Tuple

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_lib = None


class _ScnParams(C.Structure):
    _fields_ = [
        ("K", C.c_double * 4), ("R", C.c_double * 9), ("T", C.c_double * 3),
        ("width", C.c_int32), ("height", C.c_int32),
        ("tex_spacing", C.c_double), ("tex_amplitude", C.c_double), ("tex_offset", C.c_double),
        ("tex_seed", C.c_uint64), ("noise_sigma", C.c_double), ("noise_seed", C.c_uint64),
    ]


def _load():
    global _lib
    if _lib is None:
        path = os.path.join(HERE, "lib", "libpm_scene.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} missing: run __graft_entry__.build()")
        lib = C.CDLL(path)
        lib.scn_render.restype = C.c_int
        lib.scn_render.argtypes = [C.POINTER(C.c_double), C.c_int32, C.POINTER(_ScnParams),
                                   C.POINTER(C.c_uint8), C.POINTER(C.c_float), C.POINTER(C.c_float),
                                   C.POINTER(C.c_int32)]
        _lib = lib
    return _lib


@dataclass
class Camera:
    K: Tuple[float, float, float, float]  # fx, cx, fy, cy
    R: np.ndarray                         # 3x3 world -> camera
    T: np.ndarray                         # 3
    width: int
    height: int

    @property
    def center(self) -> np.ndarray:
        return -self.R.T @ self.T


@dataclass
class RenderedView:
    camera: Camera
    image: np.ndarray   # (H, W) uint8
    depth: np.ndarray   # (H, W) float32, 0 = no surface
    normal: np.ndarray  # (3, H, W) float32, camera frame


@dataclass
class Texture:
    spacing: float
    amplitude: float = 40.0
    offset: float = 128.0
    seed: int = 1
    noise_sigma: float = 0.0
    noise_seed: int = 0


def look_at(center, target, down=(0.0, 1.0, 0.0)) -> Tuple[np.ndarray, np.ndarray]:
    """World->camera (R, T) for a camera at `center` looking at `target`
    (x right, y down, z forward)."""
    c = np.asarray(center, dtype=np.float64)
    f = np.asarray(target, dtype=np.float64) - c
    f /= np.linalg.norm(f)
    d = np.asarray(down, dtype=np.float64)
    x = np.cross(d, f)
    x /= np.linalg.norm(x)
    y = np.cross(f, x)
    R = np.stack([x, y, f])
    return R, -R @ c


def render(tris: np.ndarray, cam: Camera, tex: Texture) -> RenderedView:
    lib = _load()
    tris = np.ascontiguousarray(tris, dtype=np.float64).reshape(-1, 9)
    p = _ScnParams()
    for i in range(4):
        p.K[i] = cam.K[i]
    for i in range(9):
        p.R[i] = float(cam.R.reshape(9)[i])
    for i in range(3):
        p.T[i] = float(cam.T[i])
    p.width, p.height = cam.width, cam.height
    p.tex_spacing, p.tex_amplitude, p.tex_offset = tex.spacing, tex.amplitude, tex.offset
    p.tex_seed, p.noise_sigma, p.noise_seed = tex.seed, tex.noise_sigma, tex.noise_seed
    img = np.zeros((cam.height, cam.width), dtype=np.uint8)
    depth = np.zeros((cam.height, cam.width), dtype=np.float32)
    normal = np.zeros((3, cam.height, cam.width), dtype=np.float32)
    lib.scn_render(tris.ctypes.data_as(C.POINTER(C.c_double)), int(tris.shape[0]), C.byref(p),
                   img.ctypes.data_as(C.POINTER(C.c_uint8)), depth.ctypes.data_as(C.POINTER(C.c_float)),
                   normal.ctypes.data_as(C.POINTER(C.c_float)), None)
    return RenderedView(cam, img, depth, normal)


def _quad(center, normal, half, up_hint=(0.0, 1.0, 0.0)) -> np.ndarray:
    n = np.asarray(normal, dtype=np.float64)
    n /= np.linalg.norm(n)
    a = np.cross(np.asarray(up_hint, dtype=np.float64), n)
    a /= np.linalg.norm(a)
    b = np.cross(n, a)
    c = np.asarray(center, dtype=np.float64)
    p = [c + half * (sa * a + sb * b) for sa, sb in ((-1, -1), (1, -1), (1, 1), (-1, 1))]
    return np.array([np.concatenate([p[0], p[1], p[2]]), np.concatenate([p[0], p[2], p[3]])])


def _tilted_normal(rng: np.random.Generator, max_deg: float) -> np.ndarray:
    """Unit normal facing the camera (-z) within max_deg of the optical axis."""
    t = math.radians(max_deg) * math.sqrt(rng.uniform())
    ph = rng.uniform(0, 2 * math.pi)
    return np.array([math.sin(t) * math.cos(ph), math.sin(t) * math.sin(ph), -math.cos(t)])


def ring_cameras(ref: Camera, n: int, radius: float, target, phase: float = 0.0, f=None,
                 size=None) -> List[Camera]:
    """n cameras on a circle of `radius` in the reference camera's x-y plane,
    all looking at `target` (world coordinates)."""
    cams = []
    C0 = ref.center
    ax = ref.R[0]
    ay = ref.R[1]
    down = ref.R[1]
    for k in range(n):
        a = phase + 2 * math.pi * k / n
        c = C0 + radius * (math.cos(a) * ax + math.sin(a) * ay)
        R, T = look_at(c, target, down)
        w, h = size or (ref.width, ref.height)
        K = ref.K if f is None else (f, (w - 1) / 2.0, f, (h - 1) / 2.0)
        cams.append(Camera(tuple(K), R, T, w, h))
    return cams


@dataclass
class Scene:
    ref: RenderedView
    srcs: List[RenderedView]
    depth_min: float
    depth_max: float
    tris: np.ndarray
    texture: Texture


def _depth_range(depth: np.ndarray) -> Tuple[float, float]:
    v = depth[depth > 0]
    return float(v.min()) * 0.9, float(v.max()) * 1.1


def planes_scene(seed: int = 1, width: int = 640, height: int = 480, f: float = 500.0, num_src: int = 4,
                 tex_amplitude: float = 40.0, noise_sigma: float = 0.0) -> Scene:
    """3 seeded planes at depths U[2, 10]: a background plane and two bounded
    quads, normals within 35 deg of the optical axis; sources on a ring of
    radius 0.1 x median depth looking at the median-depth point."""
    rng = np.random.default_rng(seed)
    depths = np.sort(rng.uniform(2.0, 10.0, size=3))[::-1]
    cx, cy = (width - 1) / 2.0, (height - 1) / 2.0
    tris = []
    # background plane through the optical axis at the largest depth
    tris.append(_quad([0, 0, depths[0]], _tilted_normal(rng, 35.0), 20.0 * depths[0]))
    for d in depths[1:]:
        u, v = rng.uniform(0.15, 0.85) * width, rng.uniform(0.15, 0.85) * height
        centre = np.array([(u - cx) / f * d, (v - cy) / f * d, d])
        half = d * rng.uniform(0.12, 0.22) * width / f
        tris.append(_quad(centre, _tilted_normal(rng, 35.0), half))
    tris = np.concatenate(tris)
    ref_cam = Camera((f, cx, f, cy), np.eye(3), np.zeros(3), width, height)
    dmed = float(np.median(depths))
    tex = Texture(spacing=8.0 * dmed / f, amplitude=tex_amplitude, seed=seed * 7919 + 1,
                  noise_sigma=noise_sigma, noise_seed=seed * 104729 + 3)
    ref = render(tris, ref_cam, tex)
    gt = ref.depth[ref.depth > 0]
    med = float(np.median(gt))
    srcs_c = ring_cameras(ref_cam, num_src, 0.1 * med, [0.0, 0.0, med], phase=rng.uniform(0, 2 * math.pi))
    srcs = [render(tris, c, Texture(tex.spacing, tex.amplitude, tex.offset, tex.seed, noise_sigma,
                                    tex.noise_seed + 1000 * (k + 1)))
            for k, c in enumerate(srcs_c)]
    dmin, dmax = _depth_range(ref.depth)
    return Scene(ref, srcs, dmin, dmax, tris, tex)


def heightfield_tris(rng: np.random.Generator, x_range, z_range, ground_y: float, grid: int = 64,
                     num_bumps: int = 40, sigma: float = 0.5, amp: float = 0.6) -> np.ndarray:
    """Heightfield mesh Y = ground_y - sum of Gaussian bumps over (X, Z)
    (Y points down), grid x grid vertices."""
    xs = np.linspace(x_range[0], x_range[1], grid)
    zs = np.linspace(z_range[0], z_range[1], grid)
    X, Z = np.meshgrid(xs, zs)
    bx = rng.uniform(x_range[0], x_range[1], num_bumps)
    bz = rng.uniform(z_range[0], z_range[1], num_bumps)
    ba = rng.uniform(-amp, amp, num_bumps)
    Hh = np.zeros_like(X)
    for a, x0, z0 in zip(ba, bx, bz):
        Hh += a * np.exp(-((X - x0) ** 2 + (Z - z0) ** 2) / (2 * sigma * sigma))
    Y = ground_y - Hh
    V = np.stack([X, Y, Z], axis=-1)
    tris = []
    for i in range(grid - 1):
        for j in range(grid - 1):
            a, b, c, d = V[i, j], V[i, j + 1], V[i + 1, j + 1], V[i + 1, j]
            tris.append(np.concatenate([a, b, c]))
            tris.append(np.concatenate([a, c, d]))
    return np.array(tris)


def heightfield_scene(seed: int = 2, width: int = 1920, height: int = 1080, f: float = 1500.0,
                      num_src: int = 10, tex_amplitude: float = 40.0, grid: int = 64) -> Scene:
    """Heightfield (grid x grid, Gaussian bumps sigma 0.5 m) seen by a camera
    pitched so that the depth spans about 3..30; sources on a ring of radius
    0.1 x median depth looking at the median-depth point."""
    rng = np.random.default_rng(seed)
    cy = (height - 1) / 2.0
    cx = (width - 1) / 2.0
    t = cy / f
    phi = math.atan(11.0 * t / 9.0)            # bottom row depth 3, top row depth 30
    h_cam = 3.0 * (math.sin(phi) + t * math.cos(phi))
    tris = heightfield_tris(rng, (-26.0, 26.0), (0.5, 40.0), h_cam, grid=grid)
    # camera at the origin, pitched down by phi about the x axis
    cp, sp = math.cos(phi), math.sin(phi)
    R = np.array([[1.0, 0.0, 0.0], [0.0, cp, -sp], [0.0, sp, cp]])
    ref_cam = Camera((f, cx, f, cy), R, np.zeros(3), width, height)
    tex = Texture(spacing=8.0 * 10.0 / f, amplitude=tex_amplitude, seed=seed * 7919 + 1)
    ref = render(tris, ref_cam, tex)
    med = float(np.median(ref.depth[ref.depth > 0]))
    target = ref_cam.R.T @ np.array([0.0, 0.0, med])
    srcs_c = ring_cameras(ref_cam, num_src, 0.1 * med, target, phase=rng.uniform(0, 2 * math.pi))
    srcs = [render(tris, c, tex) for c in srcs_c]
    dmin, dmax = _depth_range(ref.depth)
    return Scene(ref, srcs, dmin, dmax, tris, tex)


def to_problem(scene: Scene, with_gt_maps: bool = False):
    """Problem with image 0 = reference and images 1..N = sources."""
    from .patch_match import Image, Problem
    views = [scene.ref] + scene.srcs
    images = [Image(v.image, v.camera.K, v.camera.R, v.camera.T) for v in views]
    prob = Problem(ref_image_idx=0, src_image_idxs=list(range(1, len(views))), images=images,
                   depth_min=scene.depth_min, depth_max=scene.depth_max)
    if with_gt_maps:
        prob.depth_maps = [v.depth for v in views]
        prob.normal_maps = [v.normal for v in views]
    return prob


@dataclass
class Orbit:
    views: List[RenderedView]
    tris: np.ndarray
    texture: Texture


def orbit_scene(seed: int = 3, num_cams: int = 200, radius: float = 12.0, height: float = 12.0,
                min_px: float = 640 * 480, max_px: float = 3000 * 2000, grid: int = 128,
                extent: float = 45.0) -> Orbit:
    """`num_cams` cameras on a circle of `radius` m at `height` m above a
    heightfield (grid x grid mesh over [-extent, extent]^2, Gaussian bumps
    sigma 0.5 m), all looking at the centre.  Each camera has its own size,
    log-uniform in pixel count between min_px and max_px with aspect 4:3 or
    3:2, and focal length 0.8 x width."""
    rng = np.random.default_rng(seed)
    tris = heightfield_tris(rng, (-extent, extent), (-extent, extent), 0.0, grid=grid,
                            num_bumps=int(160 * (extent / 26.0) ** 2 / 4), sigma=0.5, amp=0.6)
    dist = math.hypot(radius, height)
    tex = Texture(spacing=8.0 * dist / 1000.0, amplitude=40.0, seed=seed * 7919 + 1)
    views = []
    for k in range(num_cams):
        a = 2 * math.pi * k / num_cams
        c = np.array([radius * math.cos(a), -height, radius * math.sin(a)])
        R, T = look_at(c, [0.0, 0.0, 0.0], down=(0.0, 1.0, 0.0))
        px = math.exp(rng.uniform(math.log(min_px), math.log(max_px)))
        aspect = 4.0 / 3.0 if rng.uniform() < 0.5 else 1.5
        w = int(round(math.sqrt(px * aspect)))
        h = int(round(w / aspect))
        f = 0.8 * w
        cam = Camera((f, (w - 1) / 2.0, f, (h - 1) / 2.0), R, T, w, h)
        views.append(render(tris, cam, tex))
    return Orbit(views, tris, tex)


def batch_problems(orbit: Orbit, num_views: int = 64, seed: int = 3, n_min: int = 5, n_max: int = 20,
                   iterations: int = 6, **opt_overrides):
    """Reference views drawn from the orbit, each with its N ~ U{n_min..n_max}
    nearest orbit neighbours as sources and a window of radius 5 (step 2) or
    7 (step 1) drawn per view.  Returns [(PatchMatchOptions, Problem)]."""
    from .patch_match import Image, PatchMatchOptions, Problem
    rng = np.random.default_rng(seed + 1000)
    M = len(orbit.views)
    images = [Image(v.image, v.camera.K, v.camera.R, v.camera.T) for v in orbit.views]
    refs = rng.choice(M, size=num_views, replace=False)
    items = []
    for r in refs:
        n = int(rng.integers(n_min, n_max + 1))
        order = sorted((d for d in range(1, M // 2 + 1)), key=lambda d: d)
        src = []
        for d in order:
            for s in ((r + d) % M, (r - d) % M):
                if len(src) < n and s != r and s not in src:
                    src.append(int(s))
            if len(src) >= n:
                break
        wide = rng.uniform() < 0.5
        dmin, dmax = _depth_range(orbit.views[r].depth)
        opts = PatchMatchOptions(window_radius=7 if wide else 5, window_step=1 if wide else 2, sigma_spatial=3.0,
                                 sigma_color=0.2, num_iterations=iterations, geom_consistency=False, thk=2,
                                 seed=seed, **opt_overrides)
        items.append((opts, Problem(int(r), src, images, dmin, dmax)))
    return items


@dataclass
class SparseScene:
    """Cameras and sparse points with tracks: the input of the source-view
    selection (sensemap_amd.model.Model)."""
    R: np.ndarray          # (I, 3, 3) world -> camera
    T: np.ndarray          # (I, 3)
    is_rig: np.ndarray     # (I,) uint8
    xyz: np.ndarray        # (P, 3) float32
    track_offsets: np.ndarray  # (P + 1,) int64
    track_images: np.ndarray   # int32
    blind: np.ndarray      # indices of the cameras that observe no point

    def model(self):
        from .model import Model
        return Model(self.R, self.T, self.xyz, self.track_offsets, self.track_images, self.is_rig)


def sparse_scene(seed: int = 5, num_cams: int = 500, num_points: int = 200_000, radius: float = 12.0,
                 height: float = 12.0, extent: float = 20.0, width: int = 1600, track_mean: float = 6.0,
                 rig_fraction: float = 0.3, num_blind: int = 3, chunk: int = 20_000) -> SparseScene:
    """`num_cams` cameras on a jittered orbit (radius, height, looking near
    the centre), image width x 0.75 width, focal 0.8 width; `num_points`
    points on a bumpy ground over [-extent, extent]^2.  A point's track is a
    run of 2 + Poisson(track_mean - 2) cameras, consecutive in orbit order
    among the cameras that see it (in front, inside the image), as feature
    tracks over neighbouring views are.  `num_blind` cameras observe
    nothing; a `rig_fraction` of the cameras carry the rig flag."""
    rng = np.random.default_rng(seed)
    I = int(num_cams)
    h_img = int(round(width * 0.75))
    f = 0.8 * width
    Rs = np.zeros((I, 3, 3))
    Ts = np.zeros((I, 3))
    for k in range(I):
        a = 2 * math.pi * k / I + rng.uniform(-0.2, 0.2) * 2 * math.pi / I
        r = radius * (1.0 + 0.05 * rng.uniform(-1, 1))
        hh = height * (1.0 + 0.1 * rng.uniform(-1, 1))
        c = np.array([r * math.cos(a), -hh, r * math.sin(a)])
        tgt = np.array([rng.uniform(-2, 2), 0.0, rng.uniform(-2, 2)])
        Rs[k], Ts[k] = look_at(c, tgt)
    Rs = Rs.astype(np.float32)
    Ts = Ts.astype(np.float32)
    is_rig = (rng.uniform(size=I) < rig_fraction).astype(np.uint8)
    blind = rng.choice(I, size=min(num_blind, I), replace=False) if num_blind > 0 else np.zeros(0, np.int64)
    RT = np.ascontiguousarray(Rs.transpose(2, 0, 1).reshape(3, 3 * I))  # X @ RT = [R_c X]_c
    xs, offs, imgs = [], [np.zeros(1, np.int64)], []
    total = 0
    left = int(num_points)
    while left > 0:
        n = min(chunk * 2, left * 2)
        X = np.empty((n, 3))
        X[:, 0] = rng.uniform(-extent, extent, n)
        X[:, 2] = rng.uniform(-extent, extent, n)
        X[:, 1] = -0.6 * np.sin(0.7 * X[:, 0]) * np.cos(0.5 * X[:, 2]) - 0.3 * np.sin(1.3 * X[:, 2] + 0.4)
        cam = (X.astype(np.float32) @ RT).reshape(-1, I, 3) + Ts[None]
        z = cam[..., 2]
        with np.errstate(divide="ignore", invalid="ignore"):
            u = f * cam[..., 0] / z + (width - 1) / 2.0
            v = f * cam[..., 1] / z + (h_img - 1) / 2.0
        vis = (z > 0.1) & (u >= 0) & (u <= width - 1) & (v >= 0) & (v <= h_img - 1)
        vis[:, blind] = False
        cv = vis.sum(1)
        keep = cv >= 2
        X, vis, cv = X[keep][:left], vis[keep][:left], cv[keep][:left]
        m = X.shape[0]
        if m == 0:
            continue
        L = np.minimum(cv, 2 + rng.poisson(max(track_mean - 2.0, 0.0), m))
        k0 = np.floor(rng.uniform(size=m) * cv).astype(np.int64)
        cs = np.cumsum(vis, axis=1) - 1
        rank = np.mod(cs - k0[:, None], cv[:, None])
        sel = vis & (rank < L[:, None])
        rows, cols = np.nonzero(sel)
        xs.append(X.astype(np.float32))
        offs.append(total + np.cumsum(np.bincount(rows, minlength=m)).astype(np.int64))
        imgs.append(cols.astype(np.int32))
        total += int(sel.sum())
        left -= m
    return SparseScene(Rs, Ts, is_rig, np.concatenate(xs), np.concatenate(offs), np.concatenate(imgs),
                       np.sort(blind).astype(np.int64))
