"""Precision study on a low-texture scene: the 640x480 three-plane scene with
texture amplitude 2 grey levels and Gaussian sensor noise sigma 0.5, so most
bilateral windows have a variance near the 1e-5 cut
(kernel_functions.cu:1331).

For each storage (u8, f16) and moment accumulation (plain FP32, compensated
FP32) the GPU engine runs the 6-iteration photometric pass; then, at every
pixel's final hypothesis and for every source, the FP64 restatement of the
NCC cost is compared with the FP32 cost the engine stored.  Reported per
run: cost error against FP64 (median / p99 / max over evaluations that
neither precision cuts), the variance-cut flip rate (one precision returns
the maximal cost 2, the other does not), and the depth error against the
rendered ground truth.  FP16 storage holds the 8-bit intensities exactly,
so its maps equal the u8 maps bit for bit.  Set PM_PRECISION_REPORT to a
path to write the numbers as JSON.

Tolerances (stated here, checked below):
  plain FP32 cost vs FP64        p99 <= 2e-2, max <= 0.1
  compensated FP32 cost vs FP64  p99 <= 2e-3, median <= plain median / 10
  flip rate                      <= 1e-2 (plain), <= plain (compensated)

Depth against the FP64 run: the transcription's whole run with every NCC
cost computed by its FP64 restatement and rounded to FP32 for the decisions
(geometry and random draws as in the FP32 run), on this scene and on the
textured 640x480 three-plane scene (amplitude 40, no noise).  Reported: the
share of pixels whose depth differs, the median relative depth difference,
the share within 1%, and each run's depth error against the ground truth.
Tolerances: on the textured scene >= 99% of the engine's depths (plain and
compensated) within 1% of the FP64 run's and median relative difference
<= 5e-4; on both scenes the FP64 run's median error against the ground
truth within 10% of the engine's (precision is not what limits it).
"""
import json
import os

import numpy as np
import pytest

from oracle import binding
from sensemap_amd import abi, scene
from sensemap_amd import patch_match as pmm
from tests import helpers

pytestmark = pytest.mark.gpu


def run(prob, opt):
    e = pmm.PatchMatchACMMCuda(opt, prob)
    e.Run()
    out = {"depth": e.GetDepthMap(), "normal": e.GetNormalMap(), "cost": e.GetCostMap(),
           "conf": e.GetConfMap(), "view_sel": e.GetViewSelMap()}
    e.close()
    return out


def cost_stats(gpu_cost, c32, c64):
    valid = np.isfinite(c64)
    both = valid & (gpu_cost < 2.0) & (c64 < 2.0)
    err = np.abs(gpu_cost[both].astype(np.float64) - c64[both])
    flips = valid & ((gpu_cost == 2.0) != (c64 == 2.0))
    return {
        "evaluations": int(valid.sum()),
        "uncut_evaluations": int(both.sum()),
        "cost_err_median": float(np.median(err)),
        "cost_err_p99": float(np.percentile(err, 99)),
        "cost_err_max": float(err.max()),
        "flip_rate": float(flips.sum() / valid.sum()),
        "cut_rate_fp32": float((gpu_cost[valid] == 2.0).mean()),
        "cut_rate_fp64": float((c64[valid] == 2.0).mean()),
        "gpu_equals_oracle_fp32": float(np.mean(gpu_cost[valid].view(np.uint32) == c32[valid].view(np.uint32))),
    }


def test_precision_low_texture():
    sc = helpers.small_planes(seed=1, width=640, height=480, f=500.0, amp=2.0, noise=0.5)
    prob = scene.to_problem(sc)
    gt = sc.ref.depth
    report = {"scene": "640x480 three planes, texture amplitude 2, noise sigma 0.5, 4 sources, "
                       "window 11x11 step 2, 6 iterations, best of 8 init samples"}
    maps = {}
    for storage, sname in ((abi.PM_STORAGE_U8, "u8"), (abi.PM_STORAGE_F16, "f16")):
        for accum, aname in ((abi.PM_ACCUM_PLAIN, "plain"), (abi.PM_ACCUM_COMPENSATED, "compensated")):
            name = f"{sname}_{aname}"
            print("[precision] run", name, flush=True)
            opt = helpers.c1_options(init_samples=8, image_storage=storage, accum_mode=accum)
            m = run(prob, opt)
            maps[name] = m
            view = pmm.ViewDesc(prob, opt)
            c32, c64 = binding.cost_maps(opt, view, m["depth"], m["normal"])
            st = cost_stats(m["cost"], c32, c64)
            err = helpers.depth_errors(m["depth"], gt)
            st.update({"depth_median_rel_err": float(np.median(err)), "depth_frac_within_1pct": float(np.mean(err < 0.01))})
            report[name] = st
            print("[precision]", name, json.dumps(st), flush=True)
    for aname in ("plain", "compensated"):
        diff = helpers.maps_equal(maps[f"u8_{aname}"], maps[f"f16_{aname}"])
        assert all(d == 0 for d in diff.values()), (aname, diff)
    valid = gt > 0
    dp, dc = maps["u8_plain"]["depth"], maps["u8_compensated"]["depth"]
    report["plain_vs_compensated_depth_median_rel_diff"] = float(
        np.median(np.abs(dp[valid] - dc[valid]) / gt[valid]))
    report["plain_vs_compensated_pixels_differing"] = float(np.mean(dp != dc))
    path = os.environ.get("PM_PRECISION_REPORT")
    if path:
        with open(path, "w") as f:
            json.dump(report, f, indent=1)
    p, c = report["u8_plain"], report["u8_compensated"]
    assert p["gpu_equals_oracle_fp32"] > 0.999 and c["gpu_equals_oracle_fp32"] > 0.999
    assert p["cost_err_p99"] <= 2e-2 and p["cost_err_max"] <= 0.1
    assert c["cost_err_p99"] <= 2e-3
    assert c["cost_err_median"] <= p["cost_err_median"] / 10
    assert p["flip_rate"] <= 1e-2 and c["flip_rate"] <= p["flip_rate"]


def test_depth_against_the_fp64_cost_run():
    report = {}
    cases = [("low_texture", dict(amp=2.0, noise=0.5)), ("textured", dict(amp=40.0, noise=0.0))]
    for name, kw in cases:
        sc = helpers.small_planes(seed=1, width=640, height=480, f=500.0, **kw)
        prob = scene.to_problem(sc)
        gt = sc.ref.depth
        valid = gt > 0
        opt = helpers.c1_options(init_samples=8, image_storage=abi.PM_STORAGE_U8, accum_mode=abi.PM_ACCUM_PLAIN)
        print("[precision] FP64 run,", name, flush=True)
        binding.set_cost_fp64(True)
        try:
            d64 = binding.run_view(opt, pmm.ViewDesc(prob, opt))["depth"]
        finally:
            binding.set_cost_fp64(False)
        rep = {"fp64_run_depth_median_rel_err_gt": float(np.median(helpers.depth_errors(d64, gt)))}
        for accum, aname in ((abi.PM_ACCUM_PLAIN, "plain"), (abi.PM_ACCUM_COMPENSATED, "compensated")):
            d = run(prob, helpers.c1_options(init_samples=8, image_storage=abi.PM_STORAGE_U8, accum_mode=accum))["depth"]
            rel = np.abs(d[valid] - d64[valid]) / np.maximum(np.abs(d64[valid]), 1e-12)
            rep[aname] = {"pixels_differing": float(np.mean(d[valid] != d64[valid])),
                          "median_rel_diff_vs_fp64_run": float(np.median(rel)),
                          "within_1pct_of_fp64_run": float(np.mean(rel < 0.01)),
                          "depth_median_rel_err_gt": float(np.median(helpers.depth_errors(d, gt)))}
        report[name] = rep
        print("[precision]", name, json.dumps(rep), flush=True)
    path = os.environ.get("PM_PRECISION_REPORT")
    if path:
        with open(path.replace(".json", "_fp64_run.json"), "w") as f:
            json.dump(report, f, indent=1)
    for name in ("low_texture", "textured"):
        rep = report[name]
        for aname in ("plain", "compensated"):
            e32, e64 = rep[aname]["depth_median_rel_err_gt"], rep["fp64_run_depth_median_rel_err_gt"]
            assert 0.9 * e32 <= e64 <= 1.1 * e32, (name, aname, e32, e64)
    for aname in ("plain", "compensated"):
        t = report["textured"][aname]
        assert t["within_1pct_of_fp64_run"] >= 0.99 and t["median_rel_diff_vs_fp64_run"] <= 5e-4, t
