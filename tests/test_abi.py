"""The C-ABI library loads, exports every function include/sensemap_pm.h
declares, its structs have the layout the header gives them, and the option
defaults / checks follow PatchMatchOptions (patch_match_options.h:42-304,
patch_match_options.cc:75-101).  No GPU needed."""
import ctypes as C
import math
import os
import re
import subprocess

import pytest

from sensemap_amd import abi

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "sensemap_pm.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pm_[a-z0-9_]+)\s*\(", text)))


def test_exports_every_declared_symbol():
    lib = abi.load_engine()
    names = declared_functions()
    assert len(names) >= 18
    out = subprocess.run(["nm", "-D", "--defined-only", abi.ENGINE_LIB], capture_output=True, text=True).stdout
    exported = set(re.findall(r"\bT (pm_[a-z0-9_]+)", out))
    missing = [n for n in names if n not in exported]
    assert not missing, missing
    for n in names:
        assert getattr(lib, n) is not None
    assert set(abi._SIGNATURES) == set(names)


def test_struct_layout_matches_header(tmp_path):
    src = tmp_path / "layout.c"
    structs = ["pm_options", "pm_camera", "pm_image", "pm_view_desc", "pm_outputs", "pm_run_stats", "pm_tuning",
               "pm_tune_report", "pm_server_config", "pm_job_stats", "pm_server_stats", "pm_cross_options",
               "pm_cross_image", "pm_cross_view", "pm_model", "pm_overlap_stats"]
    lines = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{HEADER}"', "int main(void) {"]
    for s in structs:
        lines.append(f'  printf("{s} %zu\\n", sizeof({s}));')
        for f, _ in getattr(abi, s)._fields_:
            lines.append(f'  printf("{s}.{f} %zu\\n", offsetof({s}, {f}));')
    lines.append("  return 0; }")
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", str(src), "-o", str(exe)], check=True)
    got = dict(l.split() for l in subprocess.run([str(exe)], capture_output=True, text=True).stdout.splitlines())
    for s in structs:
        cs = getattr(abi, s)
        assert int(got[s]) == C.sizeof(cs), s
        for f, _ in cs._fields_:
            assert int(got[f"{s}.{f}"]) == getattr(cs, f).offset, f"{s}.{f}"


def test_option_defaults_follow_patch_match_options():
    o = abi.default_options()
    assert (o.window_radius, o.window_step, o.num_samples, o.num_iterations) == (5, 2, 15, 5)
    assert o.sigma_spatial == 5.0 and abs(o.sigma_color - 0.2) < 1e-7
    assert o.geom_consistency == 1 and o.init_depth_random == 0
    assert abs(o.geom_consistency_regularizer - 0.3) < 1e-7 and o.geom_consistency_max_cost == 3.0
    assert abs(o.random_depth_ratio - 0.004) < 1e-9
    assert abs(o.random_angle1_range - math.radians(20)) < 1e-6
    assert abs(o.random_angle2_range - math.radians(12)) < 1e-6
    assert abs(o.init_ncc_matching_cost - 0.8) < 1e-7 and abs(o.max_ncc_matching_cost - 1.2) < 1e-7
    assert (o.alpha, o.thk, o.num_good_hypothesis, o.num_bad_hypothesis) == (90.0, 4, 2, 3)
    assert abs(o.beta - 0.3) < 1e-7
    assert o.init_samples == 1 and o.compat_lo_geo_bug == 1
    assert o.propagate_depth == 0 and o.has_prior_depth == 0


@pytest.mark.parametrize("field,value", [
    ("window_radius", 0), ("window_radius", 33), ("window_step", 3), ("window_step", 0),
    ("sigma_color", 0.0), ("num_samples", 0), ("thk", 0), ("thk", 5), ("geom_consistency_max_cost", -1.0),
    ("image_storage", 7), ("accum_mode", 3), ("init_samples", 0),
    ("propagate_depth", 1), ("has_prior_depth", 1),
])
def test_option_check_rejects(field, value):
    lib = abi.load_engine()
    o = abi.default_options(**{field: value})
    assert lib.pm_options_check(C.byref(o)) == 1  # PM_ERR_INVALID_ARG
    assert lib.pm_last_error()


def test_option_check_accepts_defaults_and_configs():
    lib = abi.load_engine()
    for kw in [{}, dict(window_radius=7, window_step=1), dict(window_radius=32, window_step=2)]:
        assert lib.pm_options_check(C.byref(abi.default_options(**kw))) == 0


def test_null_arguments_return_errors():
    lib = abi.load_engine()
    assert lib.pm_problem_run(None) == 1
    assert lib.pm_engine_set_counting(None, 1) == 1
    assert lib.pm_problem_create(None, None, None, None) == 1
    assert lib.pm_model_max_overlapping_images(None, None, 8, 1.0, None, None, None, None) == 1
    assert lib.pm_model_depth_ranges(None, None, None) == 1
    assert lib.pm_abi_version() == 6
