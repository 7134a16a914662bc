"""LocalOptimization's choice among its trials and the iteration schedule's
threshold (oracle/pm_oracle.c::pmo_lo_decide, pmo_th_mc;
kernel_functions.cu:2860-2900, patch_match_cuda.cu:778-779) pinned by hand:
the mean is over the selected sources only, a trial must beat the running
conf, the least mean wins with the first of equal means kept, and the
threshold starts at init_ncc_matching_cost and decays as exp(-iter^2/alpha)
(iteration 2: 0.8 e^(-4/90) = 0.76522, not e^(-2/90))."""
import math

import numpy as np
import pytest

from oracle import binding
from sensemap_amd.patch_match import PatchMatchOptions


def test_mean_over_the_selected_sources_and_sequential_acceptance():
    # sources 0 and 2 selected (sel 0b101); column 1 holds values that would
    # change every mean if it were counted
    M = np.array([[0.5, 9.0, 0.5],    # mean 0.5
                  [0.2, 9.0, 0.4],    # mean 0.3 < 0.4: taken
                  [0.3, 0.0, 0.3],    # mean 0.3, not < 0.3: the earlier one stays
                  [0.6, 0.0, 0.6],
                  [0.1, 9.0, 0.3]],   # mean 0.2: taken
                 np.float32)
    best, conf = binding.lo_decide(M, 0b101, 0.4)
    assert best == 4 and conf == pytest.approx(0.2)
    # without the last trial, the first of the two 0.3 means is kept
    M4 = M.copy()
    M4[4] = [0.9, 0.0, 0.9]
    best, conf = binding.lo_decide(M4, 0b101, 0.4)
    assert best == 1 and conf == pytest.approx(0.3)
    # nothing below the pixel's conf: no change
    best, conf = binding.lo_decide(M, 0b101, 0.1)
    assert best == -1 and conf == pytest.approx(0.1)
    # a trial not evaluated in full is skipped
    best, _ = binding.lo_decide(M, 0b101, 0.4, live=0b01111)
    assert best == 1


def test_threshold_schedule():
    o = PatchMatchOptions(init_ncc_matching_cost=0.8, alpha=90.0)
    assert binding.th_mc(o, 0) == np.float32(0.8)
    assert binding.th_mc(o, 2) == pytest.approx(0.8 * math.exp(-4 / 90), rel=1e-6)
    assert abs(binding.th_mc(o, 2) - 0.8 * math.exp(-2 / 90)) > 1e-2
    vals = [binding.th_mc(o, i) for i in range(12)]
    assert all(a > b for a, b in zip(vals, vals[1:]))
