"""Candidate gathering of the sweep (oracle/pm_oracle.c::pmo_gather_candidates,
kernel_functions.cu:2232-2327) pinned on a designed state: fronto-parallel
normals make PropagateDepth return a neighbour's own depth, and every pixel
has its own depth, so each candidate names the pixel it came from.  The conf
map puts each directional group's minimum at a chosen (row, col) offset of
the reference's dirSamples table (kernel_functions.cu:27-44, pairs are
(row, col) offsets as SweepFromCheckerBoardOpt reads them, :2279-2284), with
a tie that the first offset must win, and orders the 4-neighbours so the two
least-conf ones come first."""
import numpy as np

from oracle import binding

H = W = 60
R0 = C0 = 30
IK = (1 / 50.0, -29.5 / 50.0, 1 / 50.0, -29.5 / 50.0)

# (group, (row, col) offset holding the group's least conf), from the table
PICKS = [(0, (2, -3)), (1, (-3, 4)), (2, (-4, 3)), (3, (5, -4)), (4, (0, -11)), (5, (0, 21)), (6, (-3, 0)),
         (7, (7, 0))]


def state():
    depth = (1000.0 + 100.0 * np.arange(H)[:, None] + np.arange(W)[None, :]).astype(np.float32)
    normal = np.zeros((3, H, W), np.float32)
    normal[2] = -1.0
    conf = np.ones((H, W), np.float32)
    # 4-neighbours (row, col offsets): left 0.5, up 0.2, right 0.9, down 0.3
    for (dr, dc), c in [((0, -1), 0.5), ((-1, 0), 0.2), ((0, 1), 0.9), ((1, 0), 0.3)]:
        conf[R0 + dr, C0 + dc] = c
    for _, (dr, dc) in PICKS:
        conf[R0 + dr, C0 + dc] = 0.1
    conf[R0 + 15, C0] = 0.1  # same conf later in group 7: the first one wins
    return conf, depth, normal


def at(depth, dr, dc, r0=R0, c0=C0):
    return depth[r0 + dr, c0 + dc]


def test_candidates_name_the_reference_picks():
    conf, depth, normal = state()
    rows = binding.gather_candidates(conf, depth, normal, True, IK, R0, C0)
    assert rows.shape[0] == 11
    assert rows[0, 3] == depth[R0, C0]
    want = [at(depth, -1, 0), at(depth, 1, 0)] + [at(depth, dr, dc) for _, (dr, dc) in PICKS]
    assert np.array_equal(rows[1:, 3], np.array(want, np.float32))
    assert np.all(rows[:, 2] == -1.0)


def test_gradient_gate_drops_the_neighbours():
    conf, depth, normal = state()
    rows = binding.gather_candidates(conf, depth, normal, False, IK, R0, C0)
    assert rows.shape[0] == 9
    assert np.array_equal(rows[1:, 3], np.array([at(depth, dr, dc) for _, (dr, dc) in PICKS], np.float32))


def test_border_skips_offsets_outside_the_image():
    # at row 1, groups 2 and 6 (all row offsets <= -2) have no sample
    # inside; the up-neighbour (row 0) is still in
    conf, depth, normal = state()
    conf[0, C0] = 0.05
    rows = binding.gather_candidates(conf, depth, normal, True, IK, 1, C0)
    assert rows.shape[0] == 1 + 2 + 6
    assert rows[1, 3] == depth[0, C0]
