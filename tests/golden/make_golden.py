"""Writes the golden fixtures of this directory from the CPU transcription
(oracle/): run from the repository root as `python tests/golden/make_golden.py`.

- scene_<name>.npz: a tiny rendered scene (reference first, then sources:
  8-bit images, cameras K = (fx, cx, fy, cy), world-to-camera R and T, the
  depth range, ground-truth depth and normal maps) and the transcription's
  maps (depth, normal, conf, per-source cost, view_sel) after 1 and 6
  photometric iterations, and after 2 geometric iterations fed with the
  ground-truth depth maps scaled by a seeded 1% noise;
- ncc_kat.json: the NCC cost of the ground-truth hypothesis at 24 pixels of
  each scene against each source, in FP64 and in the canonical FP32;
- philox_kat.json: Philox4x32-10 blocks and the engine's keyed uniforms.

The reference ships no fixtures for this path, so these are the build's own;
tests/test_golden.py checks the transcription still reproduces them and
tests/test_gpu_golden.py checks the device against them."""
from __future__ import annotations

import ctypes as C
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import binding  # noqa: E402
from sensemap_amd import patch_match as pmm  # noqa: E402
from sensemap_amd import scene  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))

SCENES = {
    "planes_n2": dict(seed=11, width=64, height=48, f=52.0, num_src=2),
    "planes_n4": dict(seed=12, width=48, height=40, f=40.0, num_src=4),
}
MAPS = ("depth", "normal", "conf", "cost", "view_sel")


def options(iterations: int, geom: bool = False) -> pmm.PatchMatchOptions:
    return pmm.PatchMatchOptions(window_radius=5, window_step=2, sigma_spatial=3.0, sigma_color=0.2,
                                 num_iterations=iterations, geom_consistency=geom, init_depth_random=not geom,
                                 thk=2, seed=1)


def problem_from(f, noisy_depths: bool = False) -> pmm.Problem:
    """The Problem of a fixture file (image 0 = reference)."""
    n = f["images"].shape[0]
    images = [pmm.Image(f["images"][k], tuple(float(x) for x in f["K"][k]), f["R"][k], f["T"][k]) for k in range(n)]
    prob = pmm.Problem(0, list(range(1, n)), images, float(f["depth_min"]), float(f["depth_max"]))
    if noisy_depths:
        prob.depth_maps = [f["noisy_depth"][k] for k in range(n)]
        prob.normal_maps = [f["gt_normal"][k] for k in range(n)]
    return prob


def kat_pixels(H: int, W: int, count: int = 24):
    rng = np.random.default_rng(H * 1000 + W)
    return [(int(r), int(c)) for r, c in zip(rng.integers(6, H - 6, count), rng.integers(6, W - 6, count))]


def main() -> None:
    kat = {}
    for name, kw in SCENES.items():
        sc = scene.planes_scene(**kw)
        views = [sc.ref] + sc.srcs
        rng = np.random.default_rng(kw["seed"])
        data = {
            "images": np.stack([v.image for v in views]).astype(np.uint8),
            "K": np.array([v.camera.K for v in views], np.float64),
            "R": np.stack([np.asarray(v.camera.R, np.float64) for v in views]),
            "T": np.stack([np.asarray(v.camera.T, np.float64) for v in views]),
            "depth_min": np.float64(sc.depth_min),
            "depth_max": np.float64(sc.depth_max),
            "gt_depth": np.stack([v.depth for v in views]).astype(np.float32),
            "gt_normal": np.stack([v.normal for v in views]).astype(np.float32),
        }
        data["noisy_depth"] = (data["gt_depth"] * (1 + 0.01 * rng.standard_normal(data["gt_depth"].shape))).astype(
            np.float32)
        prob = problem_from(data)
        for it in (1, 6):
            o = options(it)
            out = binding.run_view(o, pmm.ViewDesc(prob, o))
            for k in MAPS:
                data[f"photo{it}_{k}"] = out[k]
        og = options(2, geom=True)
        gprob = problem_from(data, noisy_depths=True)
        out = binding.run_view(og, pmm.ViewDesc(gprob, og))
        for k in MAPS:
            data[f"geom2_{k}"] = out[k]
        np.savez_compressed(os.path.join(HERE, f"scene_{name}.npz"), **data)

        o = options(1)
        c32, c64 = binding.cost_maps(o, pmm.ViewDesc(prob, o), data["gt_depth"][0], data["gt_normal"][0])
        H, W = data["gt_depth"][0].shape
        rows = []
        for (r, c) in kat_pixels(H, W):
            for s in range(c32.shape[0]):
                rows.append({"row": r, "col": c, "source": s,
                             "depth": float(data["gt_depth"][0][r, c]),
                             "normal": [float(x) for x in data["gt_normal"][0][:, r, c]],
                             "cost_f64": float(c64[s, r, c]),
                             "cost_f32_bits": int(np.float32(c32[s, r, c]).view(np.uint32))})
        kat[name] = {"options": {"window_radius": 5, "window_step": 2, "sigma_spatial": 3.0, "sigma_color": 0.2},
                     "cases": rows}
    with open(os.path.join(HERE, "ncc_kat.json"), "w") as fh:
        json.dump(kat, fh, indent=1)

    lib = binding.load()
    blocks = []
    for ctr, key in (((0, 0, 0, 0), (0, 0)), ((0xffffffff,) * 4, (0xffffffff,) * 2),
                     ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0)),
                     ((1, 2, 3, 4), (5, 6))):
        c = (C.c_uint32 * 4)(*ctr)
        k = (C.c_uint32 * 2)(*key)
        o = (C.c_uint32 * 4)()
        lib.pmo_philox(c, k, o)
        blocks.append({"ctr": list(ctr), "key": list(key), "out": list(o)})
    uniforms = []
    for seed, view, pixel, pas, draw in ((0, 0, 0, 0, 0), (1, 2, 3, 4, 5), (2, 0, 1919 + 1079 * 1920, 7, 3),
                                         (12345, 17, 99, 1000, 63)):
        u = float(lib.pmo_uniform(seed, view, pixel, pas, draw))
        uniforms.append({"seed": seed, "view": view, "pixel": pixel, "pass": pas, "draw": draw,
                         "u_bits": int(np.float32(u).view(np.uint32))})
    with open(os.path.join(HERE, "philox_kat.json"), "w") as fh:
        json.dump({"philox4x32_10": blocks, "uniform": uniforms}, fh, indent=1)


if __name__ == "__main__":
    main()
