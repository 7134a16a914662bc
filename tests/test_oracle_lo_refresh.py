"""LocalOptimization's cost refresh (kernel_functions.cu:2736-2931) in the
CPU transcription: after a trial wins, the costs of the sources the pixel
did not select are recomputed at the new hypothesis.  In a photometric run
every stored per-source cost must therefore equal a fresh FP32 NCC at the
final maps, for every pixel and source; the local optimisation changes
pixels in these runs, and a transcription that skips the refresh leaves
1-2% of the costs stale (checked by planting that mistake)."""
import os

import numpy as np
import pytest

from oracle import binding
from sensemap_amd import patch_match as pmm
from tests.golden import make_golden as mg


@pytest.mark.parametrize("name", sorted(mg.SCENES))
def test_stored_costs_equal_a_fresh_evaluation_after_local_optimisation(name):
    f = np.load(os.path.join(os.path.dirname(mg.__file__), f"scene_{name}.npz"))
    prob = mg.problem_from(f)
    o = mg.options(6)
    out = binding.run_view(o, pmm.ViewDesc(prob, o))
    no_lo = mg.options(6)
    no_lo.local_optimization = False
    out2 = binding.run_view(no_lo, pmm.ViewDesc(prob, no_lo))
    assert (out["depth"] != out2["depth"]).sum() > 500  # the local optimisation acted
    c32, _ = binding.cost_maps(o, pmm.ViewDesc(prob, o), out["depth"], out["normal"])
    assert np.array_equal(out["cost"].view(np.uint32), c32.view(np.uint32))
