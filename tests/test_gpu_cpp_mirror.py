"""The C++ mirror of the reference interface (sensemap_amd/host/patch_match_cuda.h)
drives the same engine: its depth and normal maps equal the Python path's
bitwise on the same view."""
import os
import subprocess

import numpy as np
import pytest

from sensemap_amd import scene
from sensemap_amd import patch_match as pmm
from tests import helpers

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "bin", "shim_run")


@pytest.mark.gpu
@pytest.mark.parametrize("with_mask,geom", [(False, False), (True, False), (True, True)])
def test_cpp_mirror_matches_python_path(tmp_path, with_mask, geom):
    assert os.path.exists(BIN), "tests/cpp/bin/shim_run is built by __graft_entry__.build()"
    sc = helpers.small_planes()
    opts = helpers.c1_options(num_iterations=2 if geom else 3)
    if geom:
        opts = helpers.c1_options(num_iterations=2, geom_consistency=True, init_depth_random=False,
                                  est_curvature=True, median_filter=True)
    views = [sc.ref] + sc.srcs
    lines = [f"{160} {120} {len(sc.srcs)} {sc.depth_min!r} {sc.depth_max!r} {opts.window_radius} {opts.window_step} "
             f"{opts.sigma_spatial!r} {opts.sigma_color!r} {opts.num_iterations} {opts.thk} {opts.seed}"]
    for k, v in enumerate(views):
        c = v.camera
        vals = list(c.K) + [float(x) for x in np.asarray(c.R, np.float32).reshape(9)] + \
            [float(x) for x in np.asarray(c.T, np.float32)]
        lines.append(" ".join(repr(float(np.float32(x))) for x in vals))
        v.image.tofile(str(tmp_path / f"img{k}.u8"))
    (tmp_path / "meta.txt").write_text("\n".join(lines) + "\n")
    prob = scene.to_problem(sc, with_gt_maps=geom)
    if geom:
        (tmp_path / "geom.txt").write_text("1\n")
        for k, v in enumerate(views):
            np.ascontiguousarray(prob.depth_maps[k], np.float32).tofile(str(tmp_path / f"depth{k}.f32"))
        np.ascontiguousarray(prob.normal_maps[0], np.float32).tofile(str(tmp_path / "normal0.f32"))
    if with_mask:
        mask = np.full((120, 160), 255, np.uint8)
        mask[20:70, 30:110] = 0
        mask.tofile(str(tmp_path / "mask.u8"))
        prob.mask_maps = [mask] + [None] * len(prob.src_image_idxs)
    subprocess.run([BIN, str(tmp_path)], check=True, timeout=120)
    d_cpp = np.fromfile(str(tmp_path / "depth.f32"), np.float32).reshape(120, 160)
    n_cpp = np.fromfile(str(tmp_path / "normal.f32"), np.float32).reshape(3, 120, 160)
    e = pmm.PatchMatchACMMCuda(opts, prob)
    e.Run()
    assert np.array_equal(d_cpp.view(np.uint32), e.GetDepthMap().view(np.uint32))
    assert np.array_equal(n_cpp.view(np.uint32), e.GetNormalMap().view(np.uint32))
    if with_mask:
        assert np.all(d_cpp[20:70, 30:110] == 0)
    if geom:
        c_cpp = np.fromfile(str(tmp_path / "curvature.u16"), np.uint16).reshape(2, 120, 160)
        assert np.array_equal(c_cpp, e.GetCurvatureMap())


def test_cpp_mirror_header_compiles(tmp_path):
    src = tmp_path / "inc.cpp"
    src.write_text('#include "sensemap_amd/host/patch_match_cuda.h"\nint main() { '
                   'sensemap_amd::mvs::PatchMatchOptions o; return o.window_radius == 5 ? 0 : 1; }\n')
    subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-I", ROOT, str(src)], check=True)


@pytest.mark.gpu
def test_cpp_cross_filter_matches_python_path(tmp_path):
    from sensemap_amd import io as pmio
    from sensemap_amd.cross_filter import CrossImage, cross_filter
    binary = os.path.join(ROOT, "tests", "cpp", "bin", "shim_cross")
    assert os.path.exists(binary), "tests/cpp/bin/shim_cross is built by __graft_entry__.build()"
    o = scene.orbit_scene(seed=9, num_cams=10, min_px=100 * 75, max_px=160 * 120, grid=32, extent=20.0)
    lines = [str(len(o.views))]
    ims = []
    for k, v in enumerate(o.views):
        c = v.camera
        vals = [c.width, c.height] + list(c.K) + [float(x) for x in np.asarray(c.R, np.float32).reshape(9)] + \
            [float(x) for x in np.asarray(c.T, np.float32)]
        lines.append(" ".join(repr(x) if isinstance(x, int) else repr(float(np.float32(x))) for x in vals))
        pmio.write_depth_map(str(tmp_path / f"depth{k}.bin"), v.depth)
        pmio.write_normal_map(str(tmp_path / f"normal{k}.bin"), v.normal)
        # the images as the C++ side sees them: float32 cameras; normals read
        # from the file (decoded) and encoded again, as InitNormalMaps does
        ims.append(CrossImage(tuple(np.float32(x) for x in c.K), np.asarray(c.R, np.float32),
                              np.asarray(c.T, np.float32), v.depth,
                              pmio.read_normal_map(str(tmp_path / f"normal{k}.bin"))))
    views = [(k, [(k + d) % 10 for d in (-2, -1, 1, 2)]) for k in range(0, 10, 2)]
    lines.append(str(len(views)))
    for r, srcs in views:
        lines.append(f"{r} {len(srcs)} " + " ".join(map(str, srcs)))
    lines.append("1.0 15.0 0.02 1")
    (tmp_path / "meta.txt").write_text("\n".join(lines) + "\n")
    subprocess.run([binary, str(tmp_path)], check=True, timeout=120)
    want = cross_filter(ims, views)
    for (r, _), w in zip(views, want):
        got = pmio.read_depth_map(str(tmp_path / f"refined{r}.bin"))
        assert (got > 0).sum() > 0
        assert np.array_equal(got.view(np.uint32), w.view(np.uint32)), r
    # DedupRun after Run: the references' refined maps, the others' own maps
    from sensemap_amd.cross_filter import dedup
    refined = {r: w for (r, _), w in zip(views, want)}
    dims = [(CrossImage(im.K, im.R, im.T, refined.get(k, im.depth), im.normal), k in refined)
            for k, im in enumerate(ims)]
    d = dedup(dims, views)
    for k in range(len(ims)):
        got_dd = pmio.read_depth_map(str(tmp_path / f"dedup{k}.bin"))
        got_vis = pmio.read_mat(str(tmp_path / f"vis{k}.bin"), np.uint32)
        got_delt = pmio.read_depth_map(str(tmp_path / f"delt{k}.bin"))
        assert np.array_equal(got_dd.view(np.uint32), d["dedup_depth"][k].view(np.uint32)), k
        assert np.array_equal(got_vis.reshape(d["visibility"][k].shape), d["visibility"][k]), k
        if k in refined:
            want_delt = d["delt"][[r for r, _ in views].index(k)]
            assert np.array_equal(got_delt.view(np.uint32), want_delt.view(np.uint32)), k
        else:
            assert got_delt.shape == (1, 1)


@pytest.mark.gpu
@pytest.mark.parametrize("subset", [False, True])
def test_cpp_model_selection_matches_transcription(tmp_path, subset):
    """mvs::Model::GetMaxOverlappingImages / ComputeDepthRanges through the
    C++ mirror (tests/cpp/shim_model.cpp) equal the transcription."""
    from oracle import binding as ob
    from sensemap_amd import scene as sc
    binary = os.path.join(ROOT, "tests", "cpp", "bin", "shim_model")
    assert os.path.exists(binary), "tests/cpp/bin/shim_model is built by __graft_entry__.build()"
    s = sc.sparse_scene(seed=31, num_cams=36, num_points=2500, num_blind=2)
    m = s.model()
    sub = sorted(set(range(0, 36, 2)) | {int(b) for b in s.blind}) if subset else None
    lines = [str(m.num_images)]
    for i in range(m.num_images):
        lines.append(" ".join(repr(float(x)) for x in m.R[i]) + " " + " ".join(repr(float(x)) for x in m.T[i]) +
                     f" {int(m.is_rig[i])}")
    lines.append(str(m.num_points))
    for p in range(m.num_points):
        t = m.track_images[m.track_offsets[p]:m.track_offsets[p + 1]]
        lines.append(" ".join(repr(float(x)) for x in m.xyz[p]) + f" {len(t)} " + " ".join(str(int(x)) for x in t))
    lines.append("12 1.5 " + ("-1" if sub is None else f"{len(sub)} " + " ".join(map(str, sub))))
    (tmp_path / "model.txt").write_text("\n".join(lines) + "\n")
    subprocess.run([binary, str(tmp_path)], check=True, timeout=120)
    got = [[int(x) for x in l.split()[1:]] for l in (tmp_path / "lists.txt").read_text().splitlines()]
    assert got == ob.max_overlapping_images(m, 12, 1.5, sub)
    rng = np.array([[int(x) for x in l.split()] for l in (tmp_path / "ranges.txt").read_text().splitlines()],
                   np.uint32)
    assert np.array_equal(rng, ob.depth_ranges(m).view(np.uint32))


@pytest.mark.gpu
def test_cpp_stereo_fusion_matches_transcription(tmp_path):
    """mvs::StereoFusion through the C++ mirror (tests/cpp/shim_fusion.cpp):
    points, scores and visibility equal the transcription's bit for bit."""
    from oracle import binding
    from sensemap_amd import io as pmio
    from sensemap_amd.fusion import FusionImage, StereoFusionOptions
    from tests.test_gpu_fusion import orbit_images
    binary = os.path.join(ROOT, "tests", "cpp", "bin", "shim_fusion")
    assert os.path.exists(binary), "tests/cpp/bin/shim_fusion is built by __graft_entry__.build()"
    src = orbit_images(cams=8)
    lines = [str(len(src))]
    ims = []
    for k, im in enumerate(src):
        H, W = im.depth.shape
        K = [float(np.float32(x)) for x in im.K]
        R = [float(x) for x in np.asarray(im.R, np.float32).reshape(9)]
        T = [float(x) for x in np.asarray(im.T, np.float32).reshape(3)]
        ov = list(im.overlapping)
        lines.append(" ".join([str(W), str(H)] + [repr(x) for x in K + R + T] +
                              ["0", "1", str(len(ov))] + [str(j) for j in ov]))
        pmio.write_depth_map(str(tmp_path / f"depth{k}.bin"), im.depth)
        pmio.write_normal_map(str(tmp_path / f"normal{k}.bin"), im.normal)
        (tmp_path / f"grey{k}.raw").write_bytes(np.ascontiguousarray(im.color, np.uint8).tobytes())
        # the inputs as the C++ side reads them: float32 cameras, normals
        # decoded from their file
        ims.append(FusionImage(tuple(K), np.asarray(R, np.float32).reshape(3, 3), np.asarray(T, np.float32),
                               pmio.read_depth_map(str(tmp_path / f"depth{k}.bin")),
                               pmio.read_normal_map(str(tmp_path / f"normal{k}.bin")), im.color, ov))
    o = StereoFusionOptions(min_num_pixels=3)
    lines.append(f"{o.step_size} {o.min_num_pixels} {o.max_num_pixels} {o.min_num_visible_images} "
                 f"{o.max_traversal_depth} {o.max_reproj_error!r} {o.max_depth_error!r} {int(o.with_normal)} "
                 f"{o.max_normal_error!r}")
    (tmp_path / "meta.txt").write_text("\n".join(lines) + "\n")
    subprocess.run([binary, str(tmp_path)], check=True, timeout=120)
    raw = (tmp_path / "points.bin").read_bytes()
    n = int(np.frombuffer(raw[:8], np.int64)[0])
    at = 8
    fl = np.frombuffer(raw[at:at + n * 24], np.float32).reshape(n, 6)
    at += n * 24
    rgb = np.frombuffer(raw[at:at + n * 3], np.uint8).reshape(n, 3)
    at += n * 3
    score = np.frombuffer(raw[at:at + n * 4], np.float32)
    at += n * 4
    off = np.frombuffer(raw[at:at + (n + 1) * 8], np.int64)
    at += (n + 1) * 8
    vis = np.frombuffer(raw[at:], np.uint32)
    want = binding.fusion(ims, o)
    print("[cpp mirror] fusion points", n, flush=True)
    assert n == len(want["score"]) and n > 100
    assert np.array_equal(fl[:, :3].view(np.uint32), want["xyz"].view(np.uint32))
    assert np.array_equal(fl[:, 3:].view(np.uint32), want["normal"].view(np.uint32))
    assert np.array_equal(rgb, want["rgb"])
    assert np.array_equal(score.view(np.uint32), want["score"].view(np.uint32))
    assert np.array_equal(off, want["vis_offsets"])
    assert np.array_equal(vis.astype(np.int32), want["vis_images"])
