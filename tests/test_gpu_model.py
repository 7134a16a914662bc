"""Source-view selection and depth ranges on the device
(pm_model_max_overlapping_images / pm_model_depth_ranges) equal the
transcription (oracle/pm_model_oracle.c) exactly: random orbit models with
and without an image subset, several list lengths and angle thresholds,
blind cameras (spatial fallback), a model without points, the hand-built
ranking and fallback cases, and a 500-camera model; bad tracks are
rejected with an error before any device work."""
import numpy as np
import pytest

from oracle import binding
from sensemap_amd import abi, scene
from sensemap_amd.model import Model

from test_oracle_model import fallback_model, ranking_model

pytestmark = pytest.mark.gpu


def same(m, max_n, deg, subset=None):
    got = m.GetMaxOverlappingImages(max_n, deg, subset)
    want = binding.max_overlapping_images(m, max_n, deg, subset)
    bad = [i for i in range(m.num_images) if got[i] != want[i]]
    print("[model] images", m.num_images, "points", m.num_points, "max", max_n, "deg", deg,
          "subset", None if subset is None else len(subset), "mismatch", bad[:5], flush=True)
    assert not bad
    return got


@pytest.mark.parametrize("seed,cams,points", [(21, 40, 3000), (22, 120, 20000)])
def test_selection_equals_transcription(seed, cams, points):
    s = scene.sparse_scene(seed=seed, num_cams=cams, num_points=points, num_blind=3)
    m = s.model()
    for max_n, deg in [(8, 1.0), (40, 0.5), (1, 3.0), (0, 1.0)]:
        same(m, max_n, deg)
    sub = set(range(0, cams, 3)) | {int(b) for b in s.blind}
    got = same(m, 16, 1.0, sub)
    for i in range(cams):
        if i not in sub:
            assert got[i] == []
    st = m.last_overlap_stats
    L = np.diff(m.track_offsets)
    assert st.pair_records == int((L * (L - 1)).sum())
    assert st.fallback_images == len(s.blind)


def test_hand_built_cases():
    assert ranking_model([1, 0, 0, 1, 1]).GetMaxOverlappingImages(8, 0.0)[0] == [4, 1, 2, 3]
    assert ranking_model(None).GetMaxOverlappingImages(8, 0.0, [0, 1, 2]) == [[1, 2], [0], [0], [], []]
    assert fallback_model().GetMaxOverlappingImages(4, 0.0)[0] == [1, 2, 3, 8]
    same(fallback_model(), 10, 0.0)


def test_model_without_points_takes_the_fallback_everywhere():
    s = scene.sparse_scene(seed=23, num_cams=30, num_points=10, num_blind=0)
    m = Model(s.R, s.T, np.zeros((0, 3), np.float32), np.zeros(1, np.int64), np.zeros(0, np.int32), s.is_rig)
    got = same(m, 6, 1.0)
    assert m.last_overlap_stats.fallback_images == 30
    # no co-observers: the mean distance is 0/0, every distance test passes
    # and every score is NaN, so the cameras within 60 degrees come in index order
    assert all(len(g) > 0 and g == sorted(g) for g in got)


def test_depth_ranges_equal_transcription():
    for seed in (24, 25):
        m = scene.sparse_scene(seed=seed, num_cams=50, num_points=5000).model()
        got = np.asarray(m.ComputeDepthRanges(), np.float32)
        want = binding.depth_ranges(m)
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


def test_large_orbit_model():
    s = scene.sparse_scene(seed=26, num_cams=500, num_points=150_000, num_blind=4)
    m = s.model()
    same(m, 40, 1.0)
    st = m.last_overlap_stats
    print("[model] records", st.pair_records, "pairs", st.image_pairs, "ms", st.total_ms, flush=True)
    assert st.image_pairs > 500


def test_bad_tracks_are_rejected():
    R = np.stack([np.eye(3)] * 3)
    T = np.zeros((3, 3))
    m = Model.from_tracks(R, T, [[0, 0, 5]], [[0, 3]])
    with pytest.raises(abi.PMError):
        m.GetMaxOverlappingImages(4, 1.0)
    m = Model(R, T, np.zeros((2, 3)), np.array([0, 2, 1]), np.array([0, 1], np.int32))
    with pytest.raises(abi.PMError):
        m.GetMaxOverlappingImages(4, 1.0)
