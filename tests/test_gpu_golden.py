"""The device against the committed golden fixtures (tests/golden/): every
map after 1 and 6 photometric iterations and after 2 geometric iterations,
bit for bit, without running the transcription on the box."""
import os

import numpy as np
import pytest

from sensemap_amd import patch_match as pmm
from tests import helpers
from tests.golden import make_golden as mg

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", sorted(mg.SCENES))
def test_device_reproduces_the_golden_maps(name):
    f = np.load(os.path.join(os.path.dirname(mg.__file__), f"scene_{name}.npz"))
    cases = [(f"photo{it}", mg.options(it), mg.problem_from(f)) for it in (1, 6)]
    cases.append(("geom2", mg.options(2, geom=True), mg.problem_from(f, noisy_depths=True)))
    for tag, o, prob in cases:
        print(f"[golden] {name} {tag}", flush=True)
        e = pmm.PatchMatchACMMCuda(o, prob)
        e.Run()
        got = {"depth": e.GetDepthMap(), "normal": e.GetNormalMap(), "conf": e.GetConfMap(), "cost": e.GetCostMap(),
               "view_sel": e.GetViewSelMap()}
        e.close()
        diff = helpers.maps_equal(got, {k: f[f"{tag}_{k}"] for k in mg.MAPS})
        assert all(v == 0 for v in diff.values()), (tag, diff)
