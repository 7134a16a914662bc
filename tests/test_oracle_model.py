"""The sparse-model transcription (oracle/pm_model_oracle.c:
Model::GetMaxOverlappingImages, model.cc:474-934, and ComputeDepthRanges,
model.cc:941-960) pinned by hand-built models whose answers follow from
the geometry: the 75th percentile angle threshold at its closed-form value,
the shared-point ranking with the 1.5x rig factor truncated, the image
subset, repeated observations counted per observation, the spatial fallback
on a line of cameras with a known mean spacing, and corner depths of the
bounding box; plus an independent vectorised restatement on a random
model.  The reference leaves equal scores in hash-map order; the
transcription ranks the lower image index first."""
import itertools
import math

import numpy as np
import pytest

from oracle import binding
from sensemap_amd import scene
from sensemap_amd.model import Model, problem_sources


def line_cams(xs, yaw_deg=None):
    """Cameras at (x, 0, 0) looking along +z, optionally turned about y."""
    I = len(xs)
    R = np.zeros((I, 3, 3))
    T = np.zeros((I, 3))
    for i, x in enumerate(xs):
        a = math.radians(0.0 if yaw_deg is None else yaw_deg[i])
        Ri = np.array([[math.cos(a), 0, -math.sin(a)], [0, 1, 0], [math.sin(a), 0, math.cos(a)]])
        R[i] = Ri
        T[i] = -Ri @ np.array([x, 0.0, 0.0])
    return R, T


def test_percentile_angle_threshold_closed_form():
    # cameras at x = -1, +1; points on the axis at depths 2, 4, 8, 16: the
    # pair's angles are 2 atan(1/d); the 75th percentile is element
    # round(0.75 * 3) = 2 of the sorted angles, d = 4: 28.0725 degrees.
    R, T = line_cams([-1.0, 1.0])
    d = [2.0, 4.0, 8.0, 16.0]
    m = Model.from_tracks(R, T, [[0, 0, z] for z in d], [[0, 1]] * 4)
    a = math.degrees(2 * math.atan(1 / 4))
    assert binding.max_overlapping_images(m, 8, a - 0.001) == [[1], [0]]
    assert binding.max_overlapping_images(m, 8, a + 0.001) == [[], []]
    # 5 points: element round(0.75 * 4) = 3: d = 4 of [32, 16, 8, 4, 2]
    m5 = Model.from_tracks(R, T, [[0, 0, z] for z in d + [32.0]], [[0, 1]] * 5)
    assert binding.max_overlapping_images(m5, 8, a - 0.001) == [[1], [0]]
    assert binding.max_overlapping_images(m5, 8, a + 0.001) == [[], []]


def ranking_model(rig):
    # ref 0 shares 5 points with 1, 4 with 2, 3 with 3, 4 with 4
    R, T = line_cams([0.0, 1.0, 2.0, 3.0, 4.0])
    tracks, xyz = [], []
    for j, n in [(1, 5), (2, 4), (3, 3), (4, 4)]:
        for k in range(n):
            tracks.append([0, j])
            xyz.append([0.3 * j + 0.05 * k, 0.1 * k, 6.0 + k])
    return Model.from_tracks(R, T, xyz, tracks, rig)


def test_shared_point_ranking_and_rig_factor():
    # no rig flags: 5, 4, 3, 4 -> [1, 2, 4, 3] (2 before 4: lower index)
    assert binding.max_overlapping_images(ranking_model(None), 8, 0.0)[0] == [1, 2, 4, 3]
    # rig: 0, 3 and 4 share the flag: 3 -> int(4.5) = 4, 4 -> 6
    rig = [1, 0, 0, 1, 1]
    assert binding.max_overlapping_images(ranking_model(rig), 8, 0.0)[0] == [4, 1, 2, 3]
    assert binding.max_overlapping_images(ranking_model(rig), 2, 0.0)[0] == [4, 1]
    assert binding.max_overlapping_images(ranking_model(rig), 0, 0.0) == [[]] * 5


def test_subset_limits_ranked_and_counted_images():
    out = binding.max_overlapping_images(ranking_model(None), 8, 0.0, unique_image_idxs=[0, 1, 2])
    assert out == [[1, 2], [0], [0], [], []]
    assert problem_sources([4, 1, 2, 3], [0, 1, 2], 8) == [1, 2]
    assert problem_sources([4, 1, 2, 3], None, 3) == [4, 1, 2]


def test_repeated_observations_count_each_time():
    R, T = line_cams([0.0, 1.0, 2.0])
    # image 0 twice in one track: 2 co-observations with image 1
    m = Model.from_tracks(R, T, [[0.5, 0, 6], [1, 0, 7], [1.2, 0.1, 8]], [[0, 0, 1], [0, 2], [0, 2]])
    assert binding.max_overlapping_images(m, 8, 0.0)[0] == [1, 2]
    m = Model.from_tracks(R, T, [[0.5, 0, 6], [1, 0, 7], [1.2, 0.1, 8], [1.3, 0.2, 9]],
                          [[0, 0, 1], [0, 2], [0, 2], [0, 2]])
    assert binding.max_overlapping_images(m, 8, 0.0)[0] == [2, 1]


def fallback_model():
    # image 0 observes nothing; 1..5 on x = 1..5 chained by shared points
    # (nearest co-observer 1 each: mean distance 1); 6 sits 0.05 from 1,
    # 7 is turned 70 degrees away, 8 turned 30 degrees at x = -1.5.
    xs = [0.0, 1.0, 2.0, 3.0, 4.0, 5.0, 1.05, -2.0, -1.5]
    yaw = [0, 0, 0, 0, 0, 0, 0, 70, 30]
    R, T = line_cams(xs, yaw)
    tracks = [[1, 2], [2, 3], [3, 4], [4, 5]]
    xyz = [[1.5, 0, 10], [2.5, 0, 10], [3.5, 0, 10], [4.5, 0, 10]]
    return Model.from_tracks(R, T, xyz, tracks)


def test_spatial_fallback_on_a_line_of_cameras():
    # scores (1.1 - cos) * max(d, 1): 1: 0.1, 6: 0.105 (within 0.1 of 1:
    # skipped), 2: 0.2, 3: 0.3, 8: 0.234 * 1.5 = 0.351, 4: 0.4, 5: 0.5
    out = binding.max_overlapping_images(fallback_model(), 4, 0.0)
    assert out[0] == [1, 2, 3, 8]
    out = binding.max_overlapping_images(fallback_model(), 10, 0.0)
    assert out[0] == [1, 2, 3, 8, 4, 5]
    # the subset does not limit the fallback's candidates
    assert binding.max_overlapping_images(fallback_model(), 4, 0.0, unique_image_idxs=[0])[0] == [1, 2, 3, 8]


def test_depth_ranges_are_box_corner_depths():
    rng = np.random.default_rng(4)
    R, T = line_cams([0.0, 3.0, -2.0], [0.0, 25.0, -40.0])
    xyz = rng.uniform([-4, -2, 5], [6, 3, 12], size=(50, 3))
    xyz[7] = [np.nan, 0, 8]  # NaN coordinates are skipped by std::min / std::max
    m = Model.from_tracks(R, T, xyz, [[0, 1]] * 50)
    out = binding.depth_ranges(m)
    pts = np.concatenate([xyz[~np.isnan(xyz).any(1)], [-(R[i].T @ T[i]) for i in range(3)]])
    lo, hi = pts.min(0), pts.max(0)
    for i in range(3):
        z = [(R[i] @ np.array(c) + T[i])[2] for c in itertools.product(*zip(lo, hi))]
        assert out[i, 1] == pytest.approx(max(z), rel=1e-5, abs=1e-5)
        assert out[i, 0] == pytest.approx(max(min(z), 1e-6), rel=1e-5, abs=1e-7)
    # the box holds every camera centre, so every near depth clamps
    assert np.all(out[:, 0] == np.float32(1e-6))


def brute_force(m, max_num, min_deg, subset=None):
    """Vectorised restatement over all (observation, co-observation) pairs;
    images that observe nothing are left out (None)."""
    I = m.num_images
    C = np.stack([-(m.R[i].reshape(3, 3).astype(np.float64).T @ m.T[i].astype(np.float64)) for i in range(I)])
    off, img = m.track_offsets, m.track_images
    L = np.diff(off)
    pt = np.repeat(np.arange(m.num_points), L)
    rows = []
    for p in range(m.num_points):
        t = img[off[p]:off[p + 1]]
        a, b = np.meshgrid(np.arange(len(t)), np.arange(len(t)), indexing="ij")
        k = a != b
        rows.append(np.stack([t[a[k]], t[b[k]], np.full(k.sum(), p)], 1))
    rec = np.concatenate(rows)
    rec = rec[rec[:, 0] != rec[:, 1]]
    if subset is not None:
        s = np.zeros(I, bool)
        s[list(subset)] = True
        rec = rec[s[rec[:, 0]] & s[rec[:, 1]]]
    X = m.xyz[rec[:, 2]].astype(np.float64)
    c1, c2 = C[rec[:, 0]], C[rec[:, 1]]
    b2 = ((c1 - c2) ** 2).sum(1)
    r1 = ((X - c1) ** 2).sum(1)
    r2 = ((X - c2) ** 2).sum(1)
    ang = np.abs(np.arccos((r1 + r2 - b2) / (2 * np.sqrt(r1 * r2))))
    ang = np.minimum(ang, np.pi - ang).astype(np.float32)
    observed = np.bincount(img, minlength=I) > 0
    rig = m.is_rig if m.is_rig is not None else np.zeros(I, np.uint8)
    out = []
    for i in range(I):
        if (subset is not None and i not in subset) or not observed[i]:
            out.append(None if observed[i] is False or (subset is None or i in subset) and not observed[i] else [])
            continue
        sel = rec[:, 0] == i
        j, a = rec[sel, 1], ang[sel]
        cand = []
        for jj in np.unique(j):
            aa = np.sort(a[j == jj])
            n = len(aa)
            pct = aa[int(math.floor(0.75 * (n - 1) + 0.5))]
            if pct >= np.float32(min_deg * math.pi / 180):
                cand.append((-(n * 3 // 2 if rig[jj] == rig[i] else n), int(jj)))
        out.append([jj for _, jj in sorted(cand)][:max_num])
    return out


@pytest.mark.parametrize("subset", [None, "half"])
def test_random_model_matches_independent_restatement(subset):
    s = scene.sparse_scene(seed=11, num_cams=40, num_points=3000, num_blind=2)
    m = s.model()
    sub = None if subset is None else set(range(0, 40, 2))
    got = binding.max_overlapping_images(m, 12, 2.0, unique_image_idxs=sub)
    ref = brute_force(m, 12, 2.0, sub)
    for i in range(m.num_images):
        if ref[i] is not None:
            assert got[i] == ref[i], i
    for b in s.blind:
        if sub is None or int(b) in sub:
            assert len(got[int(b)]) > 0  # spatial fallback


def test_point_order_does_not_matter():
    s = scene.sparse_scene(seed=12, num_cams=30, num_points=2000, num_blind=1)
    m = s.model()
    perm = np.random.default_rng(0).permutation(m.num_points)
    tracks = [m.track_images[m.track_offsets[p]:m.track_offsets[p + 1]][::-1] for p in perm]
    m2 = Model.from_tracks(m.R.reshape(-1, 3, 3), m.T, m.xyz[perm], tracks, m.is_rig)
    assert binding.max_overlapping_images(m, 10, 1.0) == binding.max_overlapping_images(m2, 10, 1.0)
    assert np.array_equal(binding.depth_ranges(m), binding.depth_ranges(m2))
