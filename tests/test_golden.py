"""The committed golden fixtures (tests/golden/, written by make_golden.py)
against the CPU transcription: it must still reproduce every map bit for
bit, the NCC known answers and the Philox known answers.  Independent pins:
the Philox block at the pi-digit counter is Random123's published answer,
the ground truth lies in the scenes' depth ranges, more iterations lower the
depth error, and the geometric pass from noisy ground truth lands within 1%."""
import ctypes as C
import json
import os

import numpy as np
import pytest

from oracle import binding
from sensemap_amd import patch_match as pmm
from tests import helpers
from tests.golden import make_golden as mg

GOLDEN = os.path.dirname(mg.__file__)


def fixture(name):
    return np.load(os.path.join(GOLDEN, f"scene_{name}.npz"))


@pytest.mark.parametrize("name", sorted(mg.SCENES))
def test_transcription_reproduces_the_scene_fixtures(name):
    f = fixture(name)
    prob = mg.problem_from(f)
    for it in (1, 6):
        o = mg.options(it)
        out = binding.run_view(o, pmm.ViewDesc(prob, o))
        diff = helpers.maps_equal(out, {k: f[f"photo{it}_{k}"] for k in mg.MAPS})
        assert all(v == 0 for v in diff.values()), (it, diff)
    og = mg.options(2, geom=True)
    out = binding.run_view(og, pmm.ViewDesc(mg.problem_from(f, noisy_depths=True), og))
    diff = helpers.maps_equal(out, {k: f[f"geom2_{k}"] for k in mg.MAPS})
    assert all(v == 0 for v in diff.values()), ("geom", diff)


@pytest.mark.parametrize("name", sorted(mg.SCENES))
def test_fixture_scenes_are_sane(name):
    f = fixture(name)
    gt = f["gt_depth"][0]
    m = gt > 0
    assert m.mean() > 0.95
    assert gt[m].min() >= f["depth_min"] and gt[m].max() <= f["depth_max"]
    med = {k: float(np.median(helpers.depth_errors(f[f"{k}_depth"], gt))) for k in ("photo1", "photo6", "geom2")}
    assert med["photo6"] < med["photo1"]
    assert med["geom2"] < 0.01


def test_ncc_known_answers():
    kat = json.load(open(os.path.join(GOLDEN, "ncc_kat.json")))
    for name, block in kat.items():
        f = fixture(name)
        prob = mg.problem_from(f)
        o = mg.options(1)
        c32, c64 = binding.cost_maps(o, pmm.ViewDesc(prob, o), f["gt_depth"][0], f["gt_normal"][0])
        assert len(block["cases"]) >= 24
        for case in block["cases"]:
            r, c, s = case["row"], case["col"], case["source"]
            assert abs(c64[s, r, c] - case["cost_f64"]) <= 1e-12
            assert int(np.float32(c32[s, r, c]).view(np.uint32)) == case["cost_f32_bits"]
            assert 0.0 <= case["cost_f64"] <= 2.0
        # the ground-truth hypothesis matches well where the window is textured
        assert np.median([cs["cost_f64"] for cs in block["cases"]]) < 0.5


def test_philox_known_answers():
    kat = json.load(open(os.path.join(GOLDEN, "philox_kat.json")))
    lib = binding.load()
    for b in kat["philox4x32_10"]:
        o = (C.c_uint32 * 4)()
        lib.pmo_philox((C.c_uint32 * 4)(*b["ctr"]), (C.c_uint32 * 2)(*b["key"]), o)
        assert list(o) == b["out"]
    # Random123's published known answer (kat_vectors, philox4x32_10, pi digits)
    pi = [b for b in kat["philox4x32_10"] if b["ctr"][0] == 0x243F6A88][0]
    assert pi["out"] == [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]
    for u in kat["uniform"]:
        v = lib.pmo_uniform(u["seed"], u["view"], u["pixel"], u["pass"], u["draw"])
        assert int(np.float32(v).view(np.uint32)) == u["u_bits"]
        assert 0.0 <= v < 1.0
