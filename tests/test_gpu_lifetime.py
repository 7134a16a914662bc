"""Engine lifetime: problems, batches and servers keep their engine alive, so
the engine's handle may be dropped first, in any order, as a garbage
collector finalises (Python finalises the objects held in a module's
globals at exit in no fixed order)."""
import os
import subprocess
import sys
import textwrap

import numpy as np
import pytest

from sensemap_amd import patch_match as pmm
from tests import helpers
from tests.golden import make_golden as mg

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def fixture():
    return np.load(os.path.join(os.path.dirname(mg.__file__), "scene_planes_n2.npz"))


def maps(e):
    return {"depth": e.GetDepthMap(), "normal": e.GetNormalMap(), "conf": e.GetConfMap(), "cost": e.GetCostMap(),
            "view_sel": e.GetViewSelMap()}


def test_problem_runs_after_its_engine_is_closed():
    f = fixture()
    o = mg.options(1)
    eng = pmm.Engine(0)
    e = pmm.PatchMatchACMMCuda(o, mg.problem_from(f), eng)
    print("[lifetime] run, close the engine, run again", flush=True)
    e.Run()
    first = maps(e)
    eng.close()
    e.Run()
    diff = helpers.maps_equal(maps(e), first)
    e.close()
    assert all(v == 0 for v in diff.values()), diff
    diff = helpers.maps_equal(first, {k: f[f"photo1_{k}"] for k in mg.MAPS})
    assert all(v == 0 for v in diff.values()), diff


def test_module_globals_finalised_at_exit():
    # engine and problems held only in the script's globals: at exit the
    # engine's finaliser may run before the problems'
    script = textwrap.dedent("""
        import numpy as np
        from sensemap_amd import patch_match as pmm
        from tests.golden import make_golden as mg
        f = np.load("tests/golden/scene_planes_n2.npz")
        eng = pmm.Engine(0)
        probs = [pmm.PatchMatchACMMCuda(mg.options(1), mg.problem_from(f), eng) for _ in range(3)]
        def step():
            for p in probs:
                p.Run()
        step()
        print("ran", flush=True)
    """)
    print("[lifetime] child process with the engine in its globals", flush=True)
    r = subprocess.run([sys.executable, "-c", script], cwd=ROOT, capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr[-2000:], flush=True)
    assert r.returncode == 0
    assert "ran" in r.stdout


def _child(script):
    r = subprocess.run([sys.executable, "-c", textwrap.dedent(script)], cwd=ROOT, capture_output=True, text=True,
                       timeout=120)
    print(r.stdout, r.stderr[-2000:], flush=True)
    return r


def test_server_in_module_globals_at_exit():
    # a server with jobs still in flight, held only in the script's globals:
    # the module's exit hook closes it (jobs finish, threads joined) before
    # the interpreter tears its globals down
    script = """
        import numpy as np
        from sensemap_amd import patch_match as pmm
        from tests.golden import make_golden as mg
        f = np.load("tests/golden/scene_planes_n2.npz")
        eng = pmm.Engine(0)
        sv = pmm.PatchMatchServer(eng, hbm_budget_bytes=4 << 30, max_active_views=2)
        jobs = [sv.submit(mg.options(2), mg.problem_from(f)) for _ in range(6)]
        jobs[0].wait()
        print("ran", flush=True)
    """
    print("[lifetime] child process with a server and jobs in flight in its globals", flush=True)
    r = _child(script)
    assert r.returncode == 0
    assert "ran" in r.stdout
    assert "terminate" not in r.stderr


def test_server_never_destroyed_is_stopped_at_exit():
    # the server and its jobs are leaked (never finalised, not closed by the
    # Python hook), so only the library's own exit handler stops its threads
    # before the HIP runtime's teardown
    script = """
        import ctypes
        import numpy as np
        from sensemap_amd import patch_match as pmm
        from tests.golden import make_golden as mg
        f = np.load("tests/golden/scene_planes_n2.npz")
        eng = pmm.Engine(0)
        sv = pmm.PatchMatchServer(eng, hbm_budget_bytes=4 << 30, max_active_views=2)
        pmm._live_servers.discard(sv)
        jobs = [sv.submit(mg.options(2), mg.problem_from(f)) for _ in range(6)]
        jobs[0].wait()
        for obj in [sv, eng] + jobs:
            ctypes.pythonapi.Py_IncRef(ctypes.py_object(obj))
        print("ran", flush=True)
    """
    print("[lifetime] child process exiting with a leaked server", flush=True)
    r = _child(script)
    assert r.returncode == 0
    assert "ran" in r.stdout
    assert "terminate" not in r.stderr


def test_engine_with_staging_in_module_globals_at_exit():
    # an engine that copied host maps through its pinned staging buffers,
    # held only in the script's globals and never destroyed: its buffers are
    # freed before the HIP runtime's teardown
    script = """
        import numpy as np
        from sensemap_amd import patch_match as pmm
        from sensemap_amd.cross_filter import CrossImage, cross_filter
        eng = pmm.Engine(0)
        rng = np.random.default_rng(1)
        H, W = 480, 640
        K = (500.0, (W - 1) / 2, 500.0, (H - 1) / 2)
        ims = []
        for k in range(3):
            d = (5.0 + 0.1 * rng.random((H, W))).astype(np.float32)
            n = np.zeros((3, H, W), np.float32)
            n[2] = -1.0
            ims.append(CrossImage(K, np.eye(3), np.array([0.1 * k, 0.0, 0.0]), d, n))
        out = cross_filter(ims, [(0, [1, 2]), (1, [0, 2])], engine=eng)
        print("ran", out[0].shape, flush=True)
    """
    print("[lifetime] child process with a staging engine in its globals", flush=True)
    r = _child(script)
    assert r.returncode == 0
    assert "ran" in r.stdout
    assert "terminate" not in r.stderr
